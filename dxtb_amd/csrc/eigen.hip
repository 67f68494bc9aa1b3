// Batched symmetric eigensolver for small matrices: two-sided cyclic Jacobi
// with the round-robin (circle) parallel ordering, one workgroup per matrix,
// the matrix resident in LDS (n <= 128, 132 KB with a padded stride), the
// rotation accumulator V in global memory (L2 resident).
//
// In the SCF the solver is warm-started: the host passes F'' = C^T F' C with
// C the previous iteration's eigenvectors, which is nearly diagonal, so a
// couple of sweeps reach machine precision (Jacobi converges quadratically).
//
// Rotation (Numerical Recipes convention): for the pair (p, q)
//   tau = (a_qq - a_pp) / (2 a_pq),  t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//   c = 1/sqrt(1 + t^2), s = t c;  A <- J^T A J, V <- V J with
//   J_pp = J_qq = c, J_pq = s, J_qp = -s.

#include "common.h"
#include "../../include/dxtb_hip.h"

namespace {

constexpr int JB = 256;          // threads per workgroup
constexpr int JMAX = 128;        // largest matrix held in LDS
constexpr int JSTRIDE = JMAX + 1;  // padded LDS row stride (doubles)

__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < JB / 64; ++k) r = fmax(r, sh[k]);
  return r;
}

// slot -> player of the circle method at step s for m players (m even)
__device__ __forceinline__ int player(int slot, int s, int m) {
  return slot == 0 ? 0 : ((slot - 1 + s) % (m - 1)) + 1;
}

__global__ __launch_bounds__(JB) void k_jacobi(int nmax, const int* __restrict__ nsize, const double* __restrict__ Ain,
                                               double* __restrict__ V, double* __restrict__ evals, int max_sweeps,
                                               double tol, int* __restrict__ sweeps_out) {
  extern __shared__ double lds[];
  double* A = lds;                          // n x JSTRIDE
  double* rc = lds + JMAX * JSTRIDE;        // cosines  (JMAX/2)
  double* rs = rc + JMAX / 2;               // sines    (JMAX/2)
  int* rp = (int*)(rs + JMAX / 2);          // p index  (JMAX/2)
  int* rq = rp + JMAX / 2;                  // q index  (JMAX/2)
  double* red = (double*)(rq + JMAX / 2);   // reduction scratch (JB/64)

  const int b = blockIdx.x;
  const int n = nsize[b];
  const size_t nn = (size_t)nmax * nmax;
  const double* Ab = Ain + (size_t)b * nn;
  double* Vb = V + (size_t)b * nn;
  double* eb = evals + (size_t)b * nmax;

  // load the leading n x n block; V = I on the full padded block
  for (int e = threadIdx.x; e < n * n; e += JB) {
    const int i = e / n, j = e - i * n;
    A[i * JSTRIDE + j] = Ab[(size_t)i * nmax + j];
  }
  for (int e = threadIdx.x; e < nmax * nmax; e += JB) {
    const int i = e / nmax, j = e - i * nmax;
    Vb[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

  const int m = n + (n & 1);  // even number of players, a dummy if n is odd
  const int npair = m / 2;
  int sweep = 0;
  // scale for the relative convergence test
  double amax = 0.0;
  for (int e = threadIdx.x; e < n; e += JB) amax = fmax(amax, fabs(A[e * JSTRIDE + e]));
  amax = block_max(amax, red);
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0;
    for (int e = threadIdx.x; e < n * n; e += JB) {
      const int i = e / n, j = e - i * n;
      if (i != j) off = fmax(off, fabs(A[i * JSTRIDE + j]));
    }
    off = block_max(off, red);
    if (off <= tol * fmax(amax, 1.0)) break;
    for (int s = 0; s < m - 1; ++s) {
      for (int k = threadIdx.x; k < npair; k += JB) {
        int p = player(k, s, m), q = player(m - 1 - k, s, m);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < n) {
          const double apq = A[p * JSTRIDE + q];
          const double app = A[p * JSTRIDE + p], aqq = A[q * JSTRIDE + q];
          if (fabs(apq) > 1e-300 && fabs(apq) > 1e-18 * (fabs(app) + fabs(aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = t * c;
          }
        }
        rc[k] = c;
        rs[k] = sn;
        rp[k] = p;
        rq[k] = (q < n) ? q : -1;
      }
      __syncthreads();
      // columns of A and V: A <- A J, V <- V J
      for (int e = threadIdx.x; e < npair * n; e += JB) {
        const int k = e / n, i = e - k * n;
        const int q = rq[k];
        if (q < 0) continue;
        const double sn = rs[k];
        if (sn == 0.0) continue;
        const int p = rp[k];
        const double c = rc[k];
        const double aip = A[i * JSTRIDE + p], aiq = A[i * JSTRIDE + q];
        A[i * JSTRIDE + p] = c * aip - sn * aiq;
        A[i * JSTRIDE + q] = sn * aip + c * aiq;
        double* vi = Vb + (size_t)i * nmax;
        const double vip = vi[p], viq = vi[q];
        vi[p] = c * vip - sn * viq;
        vi[q] = sn * vip + c * viq;
      }
      __syncthreads();
      // rows of A: A <- J^T A
      for (int e = threadIdx.x; e < npair * n; e += JB) {
        const int k = e / n, j = e - k * n;
        const int q = rq[k];
        if (q < 0) continue;
        const double sn = rs[k];
        if (sn == 0.0) continue;
        const int p = rp[k];
        const double c = rc[k];
        const double apj = A[p * JSTRIDE + j], aqj = A[q * JSTRIDE + j];
        A[p * JSTRIDE + j] = c * apj - sn * aqj;
        A[q * JSTRIDE + j] = sn * apj + c * aqj;
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nmax; i += JB) eb[i] = (i < n) ? A[i * JSTRIDE + i] : 0.0;
  if (threadIdx.x == 0 && sweeps_out != nullptr) sweeps_out[b] = sweep;
}

}  // namespace

extern "C" int dxtb_jacobi_max_n(void) { return JMAX; }

extern "C" int dxtb_jacobi_eigh(int nbatch, int nmax, const int* nsize, const double* A, double* V, double* evals,
                                int max_sweeps, double tol, int* sweeps, hipStream_t st) {
  if (nbatch <= 0 || nmax <= 0 || nmax > JMAX || !nsize || !A || !V || !evals || max_sweeps < 0) return DXTB_EINVAL;
  const size_t shmem = (size_t)JMAX * JSTRIDE * sizeof(double) + JMAX * sizeof(double) +
                       JMAX * sizeof(int) + (JB / 64) * sizeof(double);
  static bool attr_set = false;
  if (!attr_set) {
    DXTB_CHECK(hipFuncSetAttribute((const void*)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    attr_set = true;
  }
  hipLaunchKernelGGL(k_jacobi, dim3((unsigned)nbatch), dim3(JB), shmem, st, nmax, nsize, A, V, evals, max_sweeps, tol,
                     sweeps);
  return (int)hipGetLastError();
}
