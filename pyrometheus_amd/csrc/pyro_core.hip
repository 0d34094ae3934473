// Mechanism-independent library: batched block-tridiagonal solver.
#include "pyro_block_thomas.hpp"

#define PYRO_API extern "C" __attribute__((visibility("default")))

namespace {

template <int NV>
int launch_bt(const double* L, const double* C, const double* U,
              const double* R, double* X, const int64_t* offsets,
              int64_t nsys, int diag, double* work, int* n_singular,
              hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(nsys)), block(64);
  if (diag)
    pyro_bt::k_block_thomas<NV, true><<<grid, block, 0, stream>>>(
        L, C, U, R, X, offsets, work, n_singular);
  else
    pyro_bt::k_block_thomas<NV, false><<<grid, block, 0, stream>>>(
        L, C, U, R, X, offsets, work, n_singular);
  return pyro::launch_status();
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
#define PYRO_BT_CASE(N) \
  case N:               \
    return launch_bt<N>(L, C, U, R, X, offsets, nsys, diag, work, n_singular, stream);
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
