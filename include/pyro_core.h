/*
 * pyro_core.h -- C ABI of libpyro_core.so (mechanism-independent kernels).
 * Conventions as in pyro_hip.h: device pointers owned by the caller, float64,
 * asynchronous on `stream`, status 0 ok / -1 invalid argument / -2 launch
 * error / -3 unsupported size.
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

#ifdef __cplusplus
}
#endif

#endif /* PYRO_CORE_H */
