// Common device helpers shared by every generated mechanism translation unit
// and by the mechanism-independent flamelet library.
//
// Target: gfx950 (MI355X), wave64, FP64 arithmetic throughout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pyro {

// Status codes returned across the C ABI.  0 = ok, < 0 = error; positive
// values are reserved for "number of lanes that did not converge".
enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kLaunchError = -2,
  kUnsupported = -3,
  kTooLarge = -4,  // beyond a fixed-size workspace
};

constexpr double kGasConstant = 8314.46261815324;  // J/(kmol K)
constexpr double kOneAtm = 101325.0;               // Pa
constexpr double kLn10 = 2.302585092994045684;

// One cell per lane: flat 1-D grid over the cell index.  Launches stay far
// below 2^31 blocks for any cell count that fits in HBM.
__device__ __forceinline__ int64_t cell_index() {
  return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

inline int launch_status() {
  return hipGetLastError() == hipSuccess ? kOk : kLaunchError;
}

inline unsigned grid_for(int64_t n, int block) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// Integer power for small exponents, exact in the same sense as repeated
// multiplication (used for reaction orders and Arrhenius T^b with integer b).
template <int P>
__device__ __forceinline__ double ipow(double x) {
  if constexpr (P == 0) {
    return 1.0;
  } else if constexpr (P < 0) {
    return 1.0 / ipow<-P>(x);
  } else if constexpr (P == 1) {
    return x;
  } else {
    return x * ipow<P - 1>(x);
  }
}

// Pointer into the constant address space: uniform reads through it become
// scalar loads (one s_load_dwordx16 fetches eight doubles into SGPRs).
typedef const __attribute__((address_space(4))) double* cptr;
// Pointer into global memory (keeps loads through a pointer that passed an
// asm statement on the global path rather than the flat one).
typedef const __attribute__((address_space(1))) double* gptr;
// Writable pointer into global memory, for the same reason.
typedef __attribute__((address_space(1))) double* gmptr;
// Pointer into LDS (loads through it stay ds_read rather than flat).
typedef const __attribute__((address_space(3))) double* lptr;
// Writable pointer into LDS (stores through it stay ds_write).
typedef __attribute__((address_space(3))) double* lmptr;
// Store v at base + boff bytes: a uniform base and a 32-bit per-lane byte
// offset, the form of a global store with a scalar base address.
__device__ __forceinline__ void store_at(gmptr base, unsigned boff, double v) {
  *reinterpret_cast<gmptr>(
      reinterpret_cast<__attribute__((address_space(1))) char*>(base) + boff) = v;
}

// One cell's column of a species-major LDS table with S cells per row:
// entry k at p[k * S], one ds_read with an immediate offset.  It refers to
// the pointer variable, so an asm statement that passes that variable on
// ("+v") keeps later reads below it.
template <int S>
struct LdsColumn {
  const lptr& p;
  __device__ __forceinline__ double operator[](int k) const {
    return p[k * S];
  }
};

// Reciprocal from the hardware estimate (relative error below 5e-8 on
// gfx950, measured) and one Newton step: relative error below 2.3e-15.
__device__ __forceinline__ double recip1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// Reciprocal from the hardware estimate plus two Newton steps (within an
// ulp or so of 1/x for the positive normal operands it is used on).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

}  // namespace pyro
