// Shared device helpers of the dxtb_amd HIP library (gfx950, wave64, FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DXTB_MAXP 8  // max primitives per contracted shell (GFN1 H-2s has 7)

#define DXTB_CHECK(expr)                                   \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return (int)_e;                  \
  } while (0)

// Error codes returned by the C-ABI besides hipError_t values.
enum {
  DXTB_OK = 0,
  DXTB_EINVAL = 1001,   // bad argument (null pointer, size out of range)
  DXTB_EUNSUP = 1002,   // unsupported angular momentum / option
  DXTB_ENOCONV = 1003,  // iterative solve did not converge
  DXTB_ESCRATCH = 1004, // request exceeds the fixed scratch buffer of a no-allocation entry point
};

__device__ __forceinline__ double atomic_add_f64(double* p, double v) {
  return unsafeAtomicAdd(p, v);
}

// Wave64 reduction (sum over all 64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int L>
struct Cart {
  static constexpr int n = (L + 1) * (L + 2) / 2;
};

// Cartesian exponents in the reference's component order (impls/trafo.py NLM_CART):
// s: (0,0,0); p: y, z, x; d: xx, xy, xz, yy, yz, zz.
__device__ __host__ constexpr int cart_exp(int l, int c, int ax) {
  // clang-format off
  return l == 0 ? 0 :
         l == 1 ? ((c == 0) ? (ax == 1) : (c == 1) ? (ax == 2) : (ax == 0)) :
         /* l == 2 */
         (c == 0) ? (ax == 0) * 2 :
         (c == 1) ? (ax == 0 || ax == 1) :
         (c == 2) ? (ax == 0 || ax == 2) :
         (c == 3) ? (ax == 1) * 2 :
         (c == 4) ? (ax == 1 || ax == 2) :
                    (ax == 2) * 2;
  // clang-format on
}

// Cartesian -> real spherical transform, rows m = -l..l (impls/trafo.py TRAFO).
__device__ __host__ constexpr double trafo(int l, int m, int c) {
  // sqrt(3) and sqrt(3)/2
  return l == 0 ? 1.0 :
         l == 1 ? (m == c ? 1.0 : 0.0) :
         (m == 0) ? (c == 1 ? 1.7320508075688772 : 0.0) :
         (m == 1) ? (c == 4 ? 1.7320508075688772 : 0.0) :
         (m == 2) ? (c == 0 ? -0.5 : c == 3 ? -0.5 : c == 5 ? 1.0 : 0.0) :
         (m == 3) ? (c == 2 ? 1.7320508075688772 : 0.0) :
                    (c == 0 ? 0.8660254037844386 : c == 3 ? -0.8660254037844386 : 0.0);
}
