// Mechanism-independent library: batched block-tridiagonal solver.
#include <mutex>

#include "pyro_block_thomas.hpp"
#include "pyro_flamelet_control.hpp"

#define PYRO_API extern "C" __attribute__((visibility("default")))

namespace {

template <int NV>
int launch_bt(const double* L, const double* C, const double* U,
              const double* R, double* X, const int64_t* offsets,
              int64_t nsys, int diag, double* work, int64_t work_rows,
              int* n_singular, hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(nsys)), block(64);
  if (diag)
    pyro_bt::k_block_thomas<NV, true><<<grid, block, 0, stream>>>(
        L, C, U, R, X, offsets, work, work_rows, n_singular);
  else
    pyro_bt::k_block_thomas<NV, false><<<grid, block, 0, stream>>>(
        L, C, U, R, X, offsets, work, work_rows, n_singular);
  return pyro::launch_status();
}

int dispatch_bt(const double* L, const double* C, const double* U,
                const double* R, double* X, const int64_t* offsets,
                int64_t nsys, int nv, int diag, double* work,
                int64_t work_rows, int* n_singular, hipStream_t stream) {
#define PYRO_BT_CASE(N)                                                    \
  case N:                                                                  \
    return launch_bt<N>(L, C, U, R, X, offsets, nsys, diag, work,          \
                        work_rows, n_singular, stream);
  switch (nv) {
    PYRO_BT_CASE(1) PYRO_BT_CASE(2) PYRO_BT_CASE(3) PYRO_BT_CASE(4)
    PYRO_BT_CASE(5) PYRO_BT_CASE(6) PYRO_BT_CASE(7) PYRO_BT_CASE(8)
    PYRO_BT_CASE(9) PYRO_BT_CASE(10) PYRO_BT_CASE(11) PYRO_BT_CASE(12)
    PYRO_BT_CASE(13) PYRO_BT_CASE(14) PYRO_BT_CASE(15) PYRO_BT_CASE(16)
    PYRO_BT_CASE(20) PYRO_BT_CASE(24) PYRO_BT_CASE(32) PYRO_BT_CASE(33)
    PYRO_BT_CASE(54)
    default:
      return pyro::kUnsupported;
  }
#undef PYRO_BT_CASE
}

// Workspace of the allocation-free entry point: 64 MiB per device, held by
// the loaded code object until the process ends.  Its address on each
// device comes from hipGetSymbolAddress on the first call made with that
// device current; host code never uses the symbol itself as an address.
constexpr int64_t kScratchDoubles = int64_t(8) << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return pyro::kLaunchError;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return pyro::kLaunchError;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return pyro::kOk;
}

}  // namespace

// Block sizes with a compiled kernel.
PYRO_API int pyro_block_thomas_supported(int nv) {
  return (nv >= 1 && nv <= 16) || nv == 20 || nv == 24 || nv == 32 ||
         nv == 33 || nv == 54;
}

PYRO_API int pyro_block_thomas_batched(const double* L, const double* C,
                                       const double* U, const double* R,
                                       double* X, const int64_t* offsets,
                                       int64_t nsys, int nv, int diag,
                                       double* work, int* n_singular,
                                       hipStream_t stream) {
  if (nsys == 0) return pyro::kOk;
  if (nsys < 0 || nsys > 0x7fffffff || !C || !R || !X || !offsets || !work ||
      !L || !U)
    return pyro::kInvalidArgument;
  return dispatch_bt(L, C, U, R, X, offsets, nsys, nv, diag, work, INT64_MAX,
                     n_singular, stream);
}

PYRO_API int64_t pyro_block_thomas_scratch_doubles(void) {
  return kScratchDoubles;
}

PYRO_API int pyro_block_thomas_batched_noalloc(
    const double* L, const double* C, const double* U, const double* R,
    double* X, const int64_t* offsets, int64_t nsys, int64_t n_rows, int nv,
    int diag, int* n_singular, hipStream_t stream) {
  if (nsys == 0) return pyro::kOk;
  if (nsys < 0 || nsys > 0x7fffffff || n_rows < 0 || !C || !R || !X ||
      !offsets || !L || !U)
    return pyro::kInvalidArgument;
  if (!pyro_block_thomas_supported(nv)) return pyro::kUnsupported;
  const int64_t per_row = static_cast<int64_t>(nv) * (nv + 1);
  if (n_rows > kScratchDoubles / per_row) return pyro::kTooLarge;
  double* work = nullptr;
  const int st = scratch_on_current_device(&work);
  if (st != pyro::kOk) return st;
  return dispatch_bt(L, C, U, R, X, offsets, nsys, nv, diag, work, n_rows,
                     n_singular, stream);
}

// Flamelet control on the device (pyro_flamelet_control.hpp).  settings
// (host memory, 12 doubles): newton_maxiter, newton_tol,
// bdf_newton_maxiter, bdf_newton_tol, bdf_time_step, bdf_maxsteps,
// try_newton, max_attempts, equilibrate_steps, equilibrate_dt0,
// equilibrate_growth, scaled_norm.  ctl (device, 9 x nsys int32): phase,
// it, step, inner, attempt, n_newton, n_steps, eq_step, fail.
PYRO_API int pyro_flamelet_control(const double* settings, int* ctl,
                                   double* err_prev, const double* norm2,
                                   const double* h_scale, int* modes,
                                   int* active, double* dt, int* action,
                                   int64_t nsys, hipStream_t stream) {
  if (nsys == 0) return pyro::kOk;
  if (nsys < 0 || nsys > 0x7fffffff || !settings || !ctl || !err_prev ||
      !norm2 || !modes || !active || !dt || !action ||
      (settings[11] != 0.0 && !h_scale))
    return pyro::kInvalidArgument;
  pyro_ctl::Settings c;
  c.newton_maxiter = static_cast<int>(settings[0]);
  c.newton_tol = settings[1];
  c.bdf_newton_maxiter = static_cast<int>(settings[2]);
  c.bdf_newton_tol = settings[3];
  c.bdf_time_step = settings[4];
  c.bdf_maxsteps = static_cast<int>(settings[5]);
  c.try_newton = static_cast<int>(settings[6]);
  c.max_attempts = static_cast<int>(settings[7]);
  c.equilibrate_steps = static_cast<int>(settings[8]);
  c.equilibrate_dt0 = settings[9];
  c.equilibrate_growth = settings[10];
  c.scaled_norm = static_cast<int>(settings[11]);
  pyro_ctl::State st;
  st.phase = ctl;
  st.it = ctl + nsys;
  st.step = ctl + 2 * nsys;
  st.inner = ctl + 3 * nsys;
  st.attempt = ctl + 4 * nsys;
  st.n_newton = ctl + 5 * nsys;
  st.n_steps = ctl + 6 * nsys;
  st.eq_step = ctl + 7 * nsys;
  st.fail = ctl + 8 * nsys;
  st.err_prev = err_prev;
  pyro_ctl::k_flamelet_control<<<pyro::grid_for(nsys, 256), 256, 0,
                                 stream>>>(c, st, norm2, h_scale, modes,
                                           active, dt, action,
                                           static_cast<int>(nsys));
  return pyro::launch_status();
}

PYRO_API int pyro_flamelet_state_copy(double* phi, double* phi_guess,
                                      double* phi_save, int64_t ld, int nv,
                                      const int* sys, const int* action,
                                      int64_t n, hipStream_t stream) {
  if (n == 0) return pyro::kOk;
  if (n < 0 || ld < n || nv <= 0 || !phi || !phi_guess || !phi_save ||
      !sys || !action)
    return pyro::kInvalidArgument;
  pyro_ctl::k_flamelet_state_copy<<<pyro::grid_for(n * nv, 256), 256, 0,
                                    stream>>>(phi, phi_guess, phi_save, ld,
                                              nv, sys, action, n);
  return pyro::launch_status();
}
