/*
 * pyro_core.h -- C ABI of libpyro_core.so (mechanism-independent kernels).
 * Conventions as in pyro_hip.h: device pointers owned by the caller, float64,
 * asynchronous on `stream`, status 0 ok / -1 invalid argument / -2 launch
 * error / -3 unsupported size / -4 beyond a fixed-size workspace.
 */
#ifndef PYRO_CORE_H
#define PYRO_CORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* pyro_stream_t;

/* 1 if a compiled block-Thomas kernel exists for blocks of size nv. */
int pyro_block_thomas_supported(int nv);

/* Batched block-tridiagonal solve; replaces
 * pyrometheus/flamelets/linear_solver.py:6 block_thomas(lower, central,
 * upper, rhs) for many systems at once.
 *   offsets: int64[nsys+1]; system s owns block rows [offsets[s],
 *            offsets[s+1]) of the concatenated arrays (P = offsets[nsys]).
 *   C: (P, nv, nv) diagonal blocks; R: (P, nv) right-hand sides;
 *   L, U: diag == 0: (P, nv, nv) with L[g] = lower block coupling row g+1 to
 *         row g and U[g] = upper block coupling row g to row g+1 (the entry
 *         at the last row of each system is ignored);
 *         diag != 0: (P,) scalars, the blocks being L[g]*I and U[g]*I.
 *   X: (P, nv) solution.  work: (P, nv, nv+1) scratch.
 *   *n_singular (device int, may be NULL) counts systems that met a zero
 *   or NaN pivot (the result is then inf/NaN, as LAPACK would give). */
int pyro_block_thomas_batched(const double* L, const double* C,
                              const double* U, const double* R, double* X,
                              const int64_t* offsets, int64_t nsys, int nv,
                              int diag, double* work, int* n_singular,
                              pyro_stream_t stream);

/* Doubles in the fixed workspace of pyro_block_thomas_batched_noalloc
 * (8 Mi doubles = 64 MiB per device). */
int64_t pyro_block_thomas_scratch_doubles(void);

/* pyro_block_thomas_batched without an allocation or a workspace argument.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a fixed array inside
 * the library, one per device, held until the process ends.
 *   n_rows: P = offsets[nsys], the block rows of the batch; calls with
 *           n_rows * nv * (nv + 1) > pyro_block_thomas_scratch_doubles()
 *           return -4 and launch nothing.  A system whose rows reach past
 *           n_rows is not solved: its X is NaN and it counts in
 *           *n_singular.
 * Every call on one device shares that device's workspace (it holds the
 * last call's factors; nothing in it needs to start at zero), so calls
 * that overlap on one device race.  Otherwise as pyro_block_thomas_batched:
 * same kernels, same results bit for bit. */
int pyro_block_thomas_batched_noalloc(const double* L, const double* C,
                                      const double* U, const double* R,
                                      double* X, const int64_t* offsets,
                                      int64_t nsys, int64_t n_rows, int nv,
                                      int diag, int* n_singular,
                                      pyro_stream_t stream);

/* The per-flamelet control flow of a batched flamelet solve on the device:
 * the decision of FlameletSolver.solve's hybrid strategy (flamelets/
 * solver.py:103-147 Newton loop, :387-445 time march, :476-600 restarts)
 * after one batch update, from the update's per-flamelet sums of squares
 * norm2 (nsys, 2) [enthalpy row, species rows], with the update norm
 * sqrt(a + b) or, when settings[11] != 0, sqrt(a / h_scale^2 + b).
 * settings: 12 doubles in HOST memory (newton_maxiter, newton_tol,
 * bdf_newton_maxiter, bdf_newton_tol, bdf_time_step, bdf_maxsteps,
 * try_newton, max_attempts, equilibrate_steps, equilibrate_dt0,
 * equilibrate_growth, scaled_norm).  ctl: (9, nsys) int32 device array
 * (phase 0 Newton try / 1 time march / 2 Newton / 3 converged / 4 failed /
 * 5 equilibration, it, step, inner, attempt, n_newton, n_steps, eq_step,
 * fail), err_prev (nsys,) doubles; updated in place.  Writes the next
 * update's controls modes, active, dt (as pyro_flamelet_assemble and
 * pyro_flamelet_update read them) and per-flamelet action codes for
 * pyro_flamelet_state_copy. */
int pyro_flamelet_control(const double* settings, int* ctl, double* err_prev,
                          const double* norm2, const double* h_scale,
                          int* modes, int* active, double* dt, int* action,
                          int64_t nsys, pyro_stream_t stream);

/* The state copies the control step asked for, per point of each system:
 * action 1 phi = phi_guess, 2 phi = phi_save, 3 phi_guess = phi,
 * 4 phi_save = phi (all (nv, ld), n points, sys: system of each point). */
int pyro_flamelet_state_copy(double* phi, double* phi_guess, double* phi_save,
                             int64_t ld, int nv, const int* sys,
                             const int* action, int64_t n,
                             pyro_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PYRO_CORE_H */
