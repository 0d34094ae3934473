// Mechanism-independent library: batched block-tridiagonal solver.
#include <mutex>

#include "pyro_block_thomas.hpp"

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
