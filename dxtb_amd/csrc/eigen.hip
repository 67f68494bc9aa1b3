// Batched symmetric eigensolvers for the SCF: two-sided cyclic Jacobi in two
// variants, the blocked one also in a split form (below).
//
// Scalar variant (n <= 128, A in LDS).  Every step of a sweep rotates n/2
// disjoint index pairs chosen by the round-robin (circle) ordering.  The whole
// similarity update A <- J^T A J of a step is one pass over the matrix: the
// work item (k, k') owns the 2x2 block (rows of pair k, columns of pair k')
// and applies the left rotation of pair k and the right rotation of pair k'
// to it.  The eigenvectors are accumulated transposed (VT = V^T, one
// eigenvector per row) in global memory: V <- V J rotates two rows of VT.
//
// Blocked variant (any n up to JMAX, A in global memory).  The matrix is cut
// into blocks of 16 rows; the circle ordering runs over blocks, so each outer
// step works on nb/2 disjoint block pairs (I, J).  For each pair the 32x32
// subproblem A[IJ, IJ] is staged in LDS and one cyclic Jacobi sweep is run on
// it by one wavefront (all pairs at the first step of an outer sweep, where
// every block has a partner; the pairs across I and J afterwards, so each
// index pair is rotated once per outer sweep), accumulating its 32x32
// rotation Q.  The rest
// of the matrix is then updated with dense products on the FP64 matrix cores
// (v_mfma_f64_16x16x4): every off-diagonal 32x32 tile becomes
// Q_k^T A[IJ_k, IJ_k'] Q_k' and every 32-column panel of VT rows IJ_k becomes
// Q_k^T VT[IJ_k, :].  An outer sweep thus reads and writes A and VT about
// nb times instead of n times (scalar Jacobi in global memory), and the
// flops run at matrix-core rate.  This is a cyclic Jacobi method in a blocked
// ordering; it converges like the scalar method.
//
// Warm start: in the SCF the host passes A = C^T F C with C the previous
// iteration's orbitals, which is nearly diagonal; Jacobi converges
// quadratically, so a few sweeps reach machine precision.
//
// Rotation (Numerical Recipes convention) for the pair (p, q):
//   tau = (a_qq - a_pp) / (2 a_pq),  t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//   c = 1/sqrt(1 + t^2), s = t c;  J_pp = J_qq = c, J_pq = s, J_qp = -s.

#include "common.h"
#include "../../include/dxtb_hip.h"

#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace {

constexpr int JB = 256;                // threads per workgroup of the scalar variant
constexpr int NWAVE = JB / 64;         // its wavefronts
constexpr int JBB = 512;               // threads per workgroup of the blocked variant
constexpr int NWB = JBB / 64;          // its wavefronts (one block pair or product item each)
constexpr int JMAX_LDS = 128;          // largest matrix of the scalar variant
constexpr int JMAX = 4096;             // largest matrix overall
constexpr int BS = 16;                 // block size of the blocked variant
constexpr int SB = 2 * BS;             // subproblem order
constexpr int SST = SB + 1;            // LDS stride of a subproblem

typedef double d4_t __attribute__((ext_vector_type(4)));

template <int NW = NWAVE>
__device__ __forceinline__ double block_max(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int k = 1; k < NW; ++k) r = fmax(r, sh[k]);
  return r;
}

// Orders LDS accesses of the lanes of one wavefront (lanes run in lockstep;
// the fences keep the compiler from moving LDS accesses across this point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// player at a slot of the circle method at step s for m players (m even)
__device__ __forceinline__ int player(int slot, int s, int m) {
  return slot == 0 ? 0 : ((slot - 1 + s) % (m - 1)) + 1;
}

__device__ __forceinline__ bool rotation(double app, double aqq, double apq, double thr, double& c, double& sn) {
  if (!(fabs(apq) > thr)) {
    c = 1.0;
    sn = 0.0;
    return false;
  }
  // The rotation t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 +
  // t^2), s = t c by its half angle: with r = sqrt(d^2 + 4 a_pq^2), cos 2theta
  // = |d| / r, so c^2 = h = (1 + |d| / r) / 2 and s = sign(d) a_pq / (r c).
  // Two reciprocal square roots and no division: about 16 dependent FP64
  // operations where the square root and division took about 45 (this chain
  // is the critical path of every round of a subproblem sweep).
  const double d = aqq - app;
  const double ir = rsqrt(fma(d, d, 4.0 * apq * apq));
  const double h = fma(0.5 * fabs(d), ir, 0.5);
  const double ih = rsqrt(h);
  c = h * ih;
  sn = (d >= 0.0 ? apq : -apq) * ir * ih;
  return true;
}

// Stopping rule of both variants: the largest off-diagonal element is below
// tol * max(1, max|a_ii|) (scale taken from the current diagonal, which tends
// to the eigenvalues), or it has stopped shrinking within a factor 1000 of
// that bar: Jacobi converges quadratically, so a sweep that removes less than
// half of it has reached the rounding floor of the matrix.
__device__ __forceinline__ bool stop_sweeps(double off, double off_prev, double thr, int sweep) {
  if (off <= thr) return true;
  return sweep > 0 && off > 0.5 * off_prev && off <= 1e3 * thr;
}

// ---------------------------------------------------------------- scalar ----

// LDS layout of the scalar variant for matrices up to order ncap =
// min(nmax, JMAX_LDS): A with an odd row stride (no bank conflicts down a
// column), then the rotations of one step (cosines, sines, p and q of up to
// np pairs) and the reduction slots.  Sized per call, so small matrices leave
// the CU's LDS to other kernels.
__host__ __device__ constexpr int scalar_stride(int ncap) { return ncap + 1 + (ncap & 1); }
__host__ __device__ constexpr int scalar_np(int ncap) { return (ncap + 2) / 2; }

__global__ __launch_bounds__(JB) void k_jacobi_scalar(int nmax, const int* __restrict__ nsize,
                                                      const double* __restrict__ Ag, double* __restrict__ VT,
                                                      double* __restrict__ evals, int max_sweeps, double tol,
                                                      const double* __restrict__ tol_b, int* __restrict__ sweeps_out) {
  extern __shared__ double lds[];
  const int ncap = min(nmax, JMAX_LDS);
  const int ls = scalar_stride(ncap), NP = scalar_np(ncap);
  const int b = blockIdx.x;
  const int n = min(max(nsize[b], 0), ncap);
  const size_t nn = (size_t)nmax * nmax;
  const double* Ab = Ag + (size_t)b * nn;
  double* Vb = VT + (size_t)b * nn;
  double* eb = evals + (size_t)b * nmax;

  double* A = lds;
  double* rc = lds + ncap * ls;  // cosines (NP)
  double* rs = rc + NP;                    // sines   (NP)
  int* rp = (int*)(rs + NP);               // p index (NP)
  int* rq = rp + NP;                       // q index (NP), -1 for the odd dummy
  double* red = (double*)(rq + NP);

  for (int e = threadIdx.x; e < n * n; e += JB) {
    const int i = e / n, j = e - i * n;
    A[i * ls + j] = Ab[(size_t)i * nmax + j];
  }
  for (int e = threadIdx.x; e < nmax * nmax; e += JB) {
    const int i = e / nmax, j = e - i * nmax;
    Vb[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

  const int m = n + (n & 1);
  const int npair = m / 2;
  int sweep = 0;
  double off_prev = 0.0;
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, amax = 0.0;
    for (int e = threadIdx.x; e < n * n; e += JB) {
      const int i = e / n, j = e - i * n;
      const double x = fabs(A[i * ls + j]);
      if (i != j) off = fmax(off, x);
      else amax = fmax(amax, x);
    }
    off = block_max(off, red);
    amax = block_max(amax, red);
    const double thr_conv = (tol_b ? tol_b[b] : tol) * fmax(amax, 1.0);
    if (stop_sweeps(off, off_prev, thr_conv, sweep)) break;
    off_prev = off;
    const double thr = fmax(1e-300, 1e-18 * fmax(amax, 1.0));
    for (int s = 0; s < m - 1; ++s) {
      for (int k = threadIdx.x; k < npair; k += JB) {
        int p = player(k, s, m), q = player(m - 1 - k, s, m);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < n) rotation(A[p * ls + p], A[q * ls + q], A[p * ls + q], thr, c, sn);
        rc[k] = c;
        rs[k] = sn;
        rp[k] = p;
        rq[k] = (q < n) ? q : -1;
      }
      __syncthreads();
      // A <- J^T A J, one 2x2 block per work item
      for (int e = threadIdx.x; e < npair * npair; e += JB) {
        const int k = e / npair, kk = e - k * npair;
        const double c1 = rc[k], s1 = rs[k], c2 = rc[kk], s2 = rs[kk];
        if (s1 == 0.0 && s2 == 0.0) continue;
        const int p = rp[k], q = rq[k], pp = rp[kk], qq = rq[kk];
        const double a = A[p * ls + pp];
        const double bq = (qq >= 0) ? A[p * ls + qq] : 0.0;
        const double cq = (q >= 0) ? A[q * ls + pp] : 0.0;
        const double d = (q >= 0 && qq >= 0) ? A[q * ls + qq] : 0.0;
        const double lpp = c1 * a - s1 * cq, lpq = c1 * bq - s1 * d;  // left (rows of pair k)
        const double lqp = s1 * a + c1 * cq, lqq = s1 * bq + c1 * d;
        A[p * ls + pp] = c2 * lpp - s2 * lpq;  // right (columns of pair kk)
        if (qq >= 0) A[p * ls + qq] = s2 * lpp + c2 * lpq;
        if (q >= 0) {
          A[q * ls + pp] = c2 * lqp - s2 * lqq;
          if (qq >= 0) A[q * ls + qq] = s2 * lqp + c2 * lqq;
        }
      }
      // VT <- J^T VT (rows p, q).  Four items per thread are loaded before any
      // is stored (the rows of one step are disjoint), so the global-memory
      // latency is paid once per four items.
      const int nitem = npair * n;
      for (int e0 = threadIdx.x; e0 < nitem; e0 += 4 * JB) {
        double va[4], vb[4];
        int op[4], oq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + u * JB;
          op[u] = -1;
          if (e < nitem) {
            const int k = e / n, i = e - k * n;
            if (rq[k] >= 0 && rs[k] != 0.0) {
              op[u] = rp[k] * nmax + i;
              oq[u] = rq[k] * nmax + i;
              va[u] = Vb[op[u]];
              vb[u] = Vb[oq[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (op[u] < 0) continue;
          const int k = (e0 + u * JB) / n;
          const double c = rc[k], sn = rs[k];
          Vb[op[u]] = c * va[u] - sn * vb[u];
          Vb[oq[u]] = sn * va[u] + c * vb[u];
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < nmax; i += JB) eb[i] = (i < n) ? A[i * ls + i] : 0.0;
  if (threadIdx.x == 0 && sweeps_out != nullptr) sweeps_out[b] = sweep;
}

size_t scalar_shmem(int nmax) {
  const int ncap = nmax < JMAX_LDS ? nmax : JMAX_LDS;
  const size_t np = scalar_np(ncap);
  return (size_t)ncap * scalar_stride(ncap) * sizeof(double) + np * (2 * sizeof(double) + 2 * sizeof(int)) +
         NWAVE * sizeof(double);
}

// --------------------------------------------------------------- blocked ----

struct WaveLds {
  double S[SB * SST];  // subproblem; reused as the staging tile of the products
  double Q[SB * SST];  // accumulated rotation of the subproblem
};

constexpr int KMAX = JMAX / SB;  // max block pairs
// Product kernel: 4-wavefront workgroups whose wavefronts hold only their
// 32x32 staging tile in LDS (the rotations are read from global memory):
// 32 KB per workgroup, so five fit on a CU (20 wavefronts, the register limit)
constexpr int PW = 4;               // wavefronts per product workgroup
constexpr int PROD_TILE = SB * SB;  // doubles per staging tile
constexpr size_t PROD_LDS = (size_t)PW * PROD_TILE * sizeof(double);

struct BlockedLds {
  WaveLds w[NWB];  // reused between sweeps as the VT panel of the deferred update
  double red[NWB];
};

// global row/column of local index r (0..31) of the block pair (bi, bj)
__device__ __forceinline__ int gidx(int r, int bi, int bj) { return r < BS ? bi * BS + r : bj * BS + (r - BS); }

// 1.0 or 0.0.  Out-of-range operands are loaded from a clamped (valid)
// address and multiplied by this mask; unlike a select, the product keeps the
// compiler from moving each load into a branch of its own, so all loads of a
// batch are in flight together.  The loaded values are finite matrix entries.
__device__ __forceinline__ double mask01(bool ok) { return ok ? 1.0 : 0.0; }

// circle method for the 32 indices of a subproblem (slot < 32, step s < 31)
__device__ __forceinline__ int player32(int slot, int s) {
  if (slot == 0) return 0;
  int t = slot - 1 + s;
  if (t >= SB - 1) t -= SB - 1;
  return t + 1;
}
// Pair j (< 16) of step s of a subproblem sweep.  Full sweep (31 steps): the
// circle method over all 32 indices.  Cross sweep (16 steps): only pairs
// between the two blocks, p = j, q = 16 + (j + s) mod 16.
__device__ __forceinline__ void pair32(int j, int s, bool cross, int& p, int& q) {
  if (cross) {
    p = j;
    q = BS + ((j + s) & (BS - 1));
    return;
  }
  const int a = player32(j, s), b = player32(SB - 1 - j, s);
  p = min(a, b);
  q = max(a, b);
}

__device__ __forceinline__ d4_t mfma(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row of element i of a lane in the accumulator of v_mfma_f64_16x16x4f64
// (lane l, element i holds row l/16 + 4 i, column l%16 of the 16x16 block).
__device__ __forceinline__ int acc_row(int lane, int i) { return (lane >> 4) + 4 * i; }

// One cyclic Jacobi sweep on the 32x32 subproblem in W.S, accumulating its
// rotations in W.Q (Q <- Q J): over all index pairs, or (cross) only over the
// pairs between its two blocks.  Executed by one wavefront; returns whether
// any rotation was applied (wave-uniform).  Each step is straight-line code:
// every lane computes the rotation of its column pair itself (4 lanes per
// pair, identical results), reads the other rotations it needs from LDS, then
// loads all of its S and Q elements before it writes any of them (the 2x2
// tiles and the Q row pairs of one step are disjoint).
// A double moved between the lanes of each row of 16 by the DPP control CTRL
// (two 32-bit moves).  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Value of lane L of this lane's row of 16 lanes (DPP row_share).
template <int L>
__device__ __forceinline__ double row_share(double v) { return dpp_f64<0x150 + L>(v); }

// Value of lane x + 1 (mod 16) of this lane's row of 16 lanes (DPP
// row_ror:15; checked on the MI355X: lane x of every row receives lane x + 1).
__device__ __forceinline__ double row_next(double v) { return dpp_f64<0x12F>(v); }

// The 2x2 tile update of a cross sweep, tile <- R_j^T tile R_x with
// R = (c, s; -s, c), and the update of a Q row pair, (a, b) <- (a, b) R_x.
// Written with explicit fma so that the one-wavefront and the four-wavefront
// forms round identically: the split and one-workgroup solvers then run the
// same arithmetic and agree bit for bit.
__device__ __forceinline__ void tile_rot(double c1, double s1, double c2, double s2, double& aii, double& bij,
                                         double& bji, double& bjj) {
  const double lpp = fma(c1, aii, -(s1 * bji)), lpq = fma(c1, bij, -(s1 * bjj));
  const double lqp = fma(s1, aii, c1 * bji), lqq = fma(s1, bij, c1 * bjj);
  aii = fma(c2, lpp, -(s2 * lpq));
  bij = fma(s2, lpp, c2 * lpq);
  bji = fma(c2, lqp, -(s2 * lqq));
  bjj = fma(s2, lqp, c2 * lqq);
}

__device__ __forceinline__ void qrow_rot(double c2, double s2, double& a, double& b) {
  const double a0 = a, b0 = b;
  a = fma(c2, a0, -(s2 * b0));
  b = fma(s2, a0, c2 * b0);
}

// One sweep of the 32x32 subproblem in W by one wavefront.  Lane j of row r
// (lane = 16 r + j) computes the rotation of pair kk = (j + r) mod 16 and owns
// the 2x2 tiles of S in column pair kk and row pairs r + 4u (u < 4); it also
// owns row (lane mod 32) of Q in the columns of pairs r + 4u and r + 1 + 4u.
// With this layout the rotation of pair r + 4u sits in lane 4u of the row and
// that of r + 1 + 4u in lane 4u + 1, for every row, so each lane fetches its
// coefficients with row_share moves instead of an LDS round trip.  Each S and
// Q element has one owner per round (disjoint pairs), so a round needs one
// wavefront barrier, before the next round reads the new diagonal.
template <bool CROSS>
__device__ bool inner_sweep(WaveLds& W, double thr, int lane) {
  constexpr bool cross = CROSS;
  bool any = false;
  const int r = lane >> 4;
  const int kk = ((lane & 15) + r) & 15;  // pair of this lane's rotation; its S column pair
  const int iq = lane & 31;               // Q row
  const int nsteps = cross ? BS : SB - 1;
  for (int s = 0; s < nsteps; ++s) {
    int pp, qq;
    pair32(kk, s, cross, pp, qq);
    double c2, s2;
    const bool rot = rotation(W.S[pp * SST + pp], W.S[qq * SST + qq], W.S[pp * SST + qq], thr, c2, s2);
    // a round of identity rotations changes nothing (common near convergence): skip its updates
    if (!__any(rot ? 1 : 0)) continue;
    any = true;
    double ca[4], sa[4], cb[4], sb[4];  // pairs r + 4u and r + 1 + 4u
    ca[0] = row_share<0>(c2);
    sa[0] = row_share<0>(s2);
    ca[1] = row_share<4>(c2);
    sa[1] = row_share<4>(s2);
    ca[2] = row_share<8>(c2);
    sa[2] = row_share<8>(s2);
    ca[3] = row_share<12>(c2);
    sa[3] = row_share<12>(s2);
    cb[0] = row_share<1>(c2);
    sb[0] = row_share<1>(s2);
    cb[1] = row_share<5>(c2);
    sb[1] = row_share<5>(s2);
    cb[2] = row_share<9>(c2);
    sb[2] = row_share<9>(s2);
    cb[3] = row_share<13>(c2);
    sb[3] = row_share<13>(s2);
    double x[4][4];
    int p1[4], q1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pair32(r + 4 * u, s, cross, p1[u], q1[u]);
      x[u][0] = W.S[p1[u] * SST + pp];
      x[u][1] = W.S[p1[u] * SST + qq];
      x[u][2] = W.S[q1[u] * SST + pp];
      x[u][3] = W.S[q1[u] * SST + qq];
    }
    double y[8][2];
    int pq[8], qv[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pair32(r + 4 * u, s, cross, pq[2 * u], qv[2 * u]);
      pair32((r + 1 + 4 * u) & 15, s, cross, pq[2 * u + 1], qv[2 * u + 1]);
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      y[v][0] = W.Q[iq * SST + pq[v]];
      y[v][1] = W.Q[iq * SST + qv[v]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // S <- J^T S J on the 2x2 tile (row pair r + 4u, column pair kk)
      tile_rot(ca[u], sa[u], c2, s2, x[u][0], x[u][1], x[u][2], x[u][3]);
      W.S[p1[u] * SST + pp] = x[u][0];
      W.S[p1[u] * SST + qq] = x[u][1];
      W.S[q1[u] * SST + pp] = x[u][2];
      W.S[q1[u] * SST + qq] = x[u][3];
    }
#pragma unroll
    for (int v = 0; v < 8; ++v) {  // Q <- Q J: columns p, q of row iq
      const double cq = (v & 1) ? cb[v >> 1] : ca[v >> 1], sq = (v & 1) ? sb[v >> 1] : sa[v >> 1];
      qrow_rot(cq, sq, y[v][0], y[v][1]);
      W.Q[iq * SST + pq[v]] = y[v][0];
      W.Q[iq * SST + qv[v]] = y[v][1];
    }
    wave_sync();
  }
  return any;
}

// out(32x32) = Q^T X for the 32x32 operand X given by a loader; Q is the
// 32x32 rotation at Qg (row-major).  acc[rb][cb] holds the 16x16 block
// (rb, cb) in the accumulator layout of acc_row.  All operand fragments are
// loaded before the first product so that their latencies overlap.
template <class Load>
__device__ __forceinline__ void qt_times(const double* __restrict__ Qg, Load load, int lane, d4_t acc[2][2]) {
  const int lr = lane & 15, lk = lane >> 4;
  double a0[8], a1[8], b0[8], b1[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int m = 4 * t + lk;
    // A operand (Q^T)[r][m] = Q[m][r]: lane l supplies row l%16, column l/16
    a0[t] = Qg[m * SB + lr];
    a1[t] = Qg[m * SB + BS + lr];
    // B operand X[m][c]: lane l supplies row l/16, column l%16
    b0[t] = load(m, lr);
    b1[t] = load(m, BS + lr);
  }
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    acc[0][0] = mfma(a0[t], b0[t], acc[0][0]);
    acc[0][1] = mfma(a0[t], b1[t], acc[0][1]);
    acc[1][0] = mfma(a1[t], b0[t], acc[1][0]);
    acc[1][1] = mfma(a1[t], b1[t], acc[1][1]);
  }
}

// block pair k of outer step s over nb blocks, bi < bj
__device__ __forceinline__ void block_pair(int k, int s, int nb, int& bi, int& bj) {
  const int a = player(k, s, nb), c = player(nb - 1 - k, s, nb);
  bi = min(a, c);
  bj = max(a, c);
}

// slot of block x at outer step s (player's inverse), its block pair, and
// the block it is paired with
__device__ __forceinline__ int block_slot(int x, int s, int nb) {
  if (x == 0) return 0;
  int t = (x - 1 - s) % (nb - 1);
  if (t < 0) t += nb - 1;
  return t + 1;
}
__device__ __forceinline__ int pair_of_block(int x, int s, int nb) {
  const int sl = block_slot(x, s, nb);
  return sl < nb / 2 ? sl : nb - 1 - sl;
}
__device__ __forceinline__ int partner_block(int x, int s, int nb) {
  return player(nb - 1 - block_slot(x, s, nb), s, nb);
}

// Whether the off-diagonal tile (k, k2) of step s - 1 holds the block A(I, J)
// of a block pair (I, J) of step s.  Every pair of step s straddles two pairs
// of step s - 1 (the circle ordering pairs no two blocks twice in a sweep),
// so it lies in exactly one such tile; above four blocks no tile holds two.
__device__ __forceinline__ bool tile_holds_next_pair(int nb, int s, int k, int k2) {
  int bi, bj;
  block_pair(k, s - 1, nb, bi, bj);
  return pair_of_block(partner_block(bi, s, nb), s - 1, nb) == k2 ||
         pair_of_block(partner_block(bj, s, nb), s - 1, nb) == k2;
}

// Deferred eigenvector update: up to this order the rotations of a whole
// outer sweep are kept (a log of (nb - 1) steps x npairs 32x32 blocks per
// matrix) and applied to VT once per sweep, a 32-column panel of all rows at
// a time in LDS (k_bj_apply, or the one-workgroup kernel between sweeps);
// above it every step updates VT panels in its product phase.
constexpr int DEFER_MAX = 512;
// ... and for batches of more than this many matrices (batch capacity): a
// lone matrix is bound by the latency of the sweep's launch chain, which the
// extra, MFMA-bound panel update lengthens, while full batches gain from
// reading and writing VT once per sweep instead of once per step
constexpr int DEFER_MIN_BATCH = 16;

__host__ __device__ constexpr bool defer_vt(int nmax) { return nmax <= DEFER_MAX; }
__host__ __device__ constexpr bool defer_call(int nmax, int nbcap) {
  return defer_vt(nmax) && nbcap > DEFER_MIN_BATCH;
}
// steps of rotations kept per matrix: a sweep's worth up to DEFER_MAX
// (else two, used by step parity: a launch that forms step s - 1's products
// and step s's rotations reads the one and writes the other)
__host__ __device__ constexpr int log_steps(int nmax) {
  return defer_vt(nmax) ? 2 * ((nmax + SB - 1) / SB) - 1 : 2;
}

// One matrix of the blocked variant.
struct BjMat {
  double* A;      // (nmax, nmax), in place
  double* V;      // eigenvector rows VT
  double* Qlog;   // rotations: log_steps x npad x 32 doubles (npairs 32x32 blocks per step)
  int* flags;     // per step and block pair: rotated anything (log_steps x KMAX)
  int n, nmax, nlast, npairs, nb, npad;
  bool defer;
};

__device__ __forceinline__ BjMat bj_mat(int b, int nmax, const int* nsize, double* Ag, double* VT, double* work,
                                       int* flags, int defer) {
  BjMat M;
  const int npad = (nmax + SB - 1) / SB * SB;
  const int nlog = log_steps(nmax);
  M.n = min(max(nsize[b], 0), nmax);
  M.nmax = nmax;
  M.nlast = max(M.n - 1, 0);
  M.npairs = (M.n + SB - 1) / SB;  // block pairs (n padded to 32)
  M.nb = 2 * M.npairs;             // blocks (even)
  M.npad = npad;
  M.defer = defer_vt(nmax) && defer != 0;
  M.A = Ag + (size_t)b * nmax * nmax;
  M.V = VT + (size_t)b * nmax * nmax;
  M.Qlog = work + (size_t)b * nlog * npad * SB;
  M.flags = flags + (size_t)b * nlog * KMAX;
  return M;
}

// rotations and flags of outer step s: a slot per step when deferred, else
// by step parity
__device__ __forceinline__ double* qstep(const BjMat& M, int s) {
  return M.Qlog + (size_t)(M.defer ? s : (s & 1)) * M.npad * SB;
}
__device__ __forceinline__ int* fstep(const BjMat& M, int s) { return M.flags + (M.defer ? s : (s & 1)) * KMAX; }

// Data pointers and batch size of one dxtb_jacobi_eigh call, written into the
// head of the workspace by k_bj_args before the call's other kernels run.
// The blocked kernels read them from there, so their launch arguments depend
// only on the workspace address and the call's sizes and sweep counts: the
// launch sequence of one such key can be replayed as a graph with other
// matrices.  Grids span a batch capacity (nbcap >= nbatch); workgroups of the
// rows beyond nbatch return at once.
struct JArgs {
  const int* nsize;
  double* A;
  double* VT;
  double* evals;
  const double* tol_b;
  int* sweeps;
  int nbatch;
  int defer;  // deferred eigenvector update (defer_call of the call's sizes)
};
constexpr size_t JARGS_BYTES = 256;  // head of the workspace (keeps the rest aligned)

// The blocks are read through the constant address space: the compiler may
// then load a field again where it would otherwise keep it in registers
// across the sweep loops (the block does not change while a kernel runs).
typedef const __attribute__((address_space(4))) JArgs* JArgsC;
__device__ __forceinline__ JArgsC jconst(const JArgs* ja) { return (JArgsC)ja; }

__global__ void k_bj_args(JArgs* __restrict__ ja, JArgs v) {
  if (threadIdx.x == 0) *ja = v;
}

// Cross sweep of block pair k at outer step s > 0 with the subproblem and its
// rotation in registers.  Round r (< 16) rotates the pairs (I_x, J_sigma(x)),
// sigma(x) = (x + r) mod 16, for every slot x < 16.  In slot space the four
// 16x16 blocks are
//   aii[j][x] = A(I_j, I_x), bij[j][x] = A(I_j, J_sigma(x)),
//   bji[j][x] = A(J_sigma(j), I_x), bjj[j][x] = A(J_sigma(j), J_sigma(x)),
// so the 2x2 tile touched by the rotations of pairs j (rows) and x (columns)
// sits at (j, x) of all four; likewise Q's rows i hold Q(i, I_x) and
// Q(i, J_sigma(x)) at slot x.  Lane l owns slot column x = l % 16 and the row
// slots j = 4 g + u (g = l / 16, u < 4) of the tiles, and the Q rows
// i = 8 g + v (v < 8).  Going from round r to r + 1 advances sigma by one:
// the J side of every slot moves one slot down (column x takes the value of
// x + 1, row j that of j + 1), by register moves and cross-lane shuffles.
// After 16 rounds sigma is the identity again.  The lane holding the diagonal
// tile (j = x) computes the rotation of pair x; the coefficients are exchanged
// through LDS (W.S as scratch).  Same rotations, in the same order, as the
// cyclic cross sweep of inner_sweep.
__device__ bool bj_inner_cross(double* cs, const BjMat& M, int k, int s, double thr, int lane) {
  int bi, bj;
  block_pair(k, s, M.nb, bi, bj);
  const int x = lane & 15, g = lane >> 4;
  const int n = M.n, nmax = M.nmax, last = M.nlast;
  const int gix = bi * BS + x, gjx = bj * BS + x;
  double aii[4], bij[4], bji[4], bjj[4], qi[8], qj[8];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = 4 * g + u;
    const int gij = bi * BS + j, gjj = bj * BS + j;
    const double* Ai = M.A + (size_t)min(gij, last) * nmax;
    const double* Aj = M.A + (size_t)min(gjj, last) * nmax;
    aii[u] = Ai[min(gix, last)] * mask01(gij < n && gix < n);
    bij[u] = Ai[min(gjx, last)] * mask01(gij < n && gjx < n);
    bji[u] = Aj[min(gix, last)] * mask01(gjj < n && gix < n);
    bjj[u] = Aj[min(gjx, last)] * mask01(gjj < n && gjx < n);
  }
  {  // a cross block at or below the bar rotates nothing (see bj_inner_cross4)
    bool need = false;
#pragma unroll
    for (int u = 0; u < 4; ++u) need |= fabs(bij[u]) > thr;
    if (!__any(need ? 1 : 0)) {
      wave_sync();
      return false;
    }
  }
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int i = 8 * g + v;
    qi[v] = (i == x) ? 1.0 : 0.0;
    qj[v] = (i == BS + x) ? 1.0 : 0.0;
  }
  // this lane holds the diagonal tile (x, x) when x is one of its row slots
  const int ud = x - 4 * g;
  const bool diag = ud >= 0 && ud < 4;
  const int down = (((g + 1) & 3) << 4) | x;        // row slot 4 g + 4, same column
  const int down_nxt = (((g + 1) & 3) << 4) | ((x + 1) & 15);
  // cs: (c, s) of the 16 pairs, 32 doubles of LDS scratch
  bool any = false;
  for (int r = 0; r < BS; ++r) {
    double app = aii[0], aqq = bjj[0], apq = bij[0];
#pragma unroll
    for (int u = 1; u < 4; ++u) {
      app = ud == u ? aii[u] : app;
      aqq = ud == u ? bjj[u] : aqq;
      apq = ud == u ? bij[u] : apq;
    }
    double c, sn;
    const bool rot = rotation(app, aqq, apq, thr, c, sn) && diag;
    if (__any(rot ? 1 : 0)) {
      any = true;
      if (diag) {
        cs[2 * x] = c;
        cs[2 * x + 1] = sn;
      }
      wave_sync();
      const double c2 = cs[2 * x], s2 = cs[2 * x + 1];
#pragma unroll
      for (int u = 0; u < 4; ++u) {  // tile (4 g + u, x) <- R_j^T tile R_x
        const double c1 = cs[2 * (4 * g + u)], s1 = cs[2 * (4 * g + u) + 1];
        tile_rot(c1, s1, c2, s2, aii[u], bij[u], bji[u], bjj[u]);
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) qrow_rot(c2, s2, qi[v], qj[v]);  // Q <- Q J: columns I_x and J_sigma(x)
      wave_sync();  // the coefficients are read before the next round writes them
    }
    // advance sigma: J-side columns take slot x + 1 (a rotation within the
    // row of 16 lanes, DPP), J-side rows slot j + 1 (the last one from the
    // next row group, the only moves across rows)
#pragma unroll
    for (int u = 0; u < 4; ++u) bij[u] = row_next(bij[u]);
#pragma unroll
    for (int v = 0; v < 8; ++v) qj[v] = row_next(qj[v]);
    const double tji = __shfl(bji[0], down, 64);
    const double tjj = __shfl(bjj[0], down_nxt, 64);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      bji[u] = bji[u + 1];
      bjj[u] = row_next(bjj[u + 1]);
    }
    bji[3] = tji;
    bjj[3] = tjj;
  }
  if (any) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = 4 * g + u;
      const int gij = bi * BS + j, gjj = bj * BS + j;
      if (gij < n && gix < n) M.A[(size_t)gij * nmax + gix] = aii[u];
      if (gij < n && gjx < n) M.A[(size_t)gij * nmax + gjx] = bij[u];
      if (gjj < n && gix < n) M.A[(size_t)gjj * nmax + gix] = bji[u];
      if (gjj < n && gjx < n) M.A[(size_t)gjj * nmax + gjx] = bjj[u];
    }
    double* Q = qstep(M, s) + (size_t)k * SB * SB;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int i = 8 * g + v;
      Q[i * SB + x] = qi[v];
      Q[i * SB + BS + x] = qj[v];
    }
  }
  wave_sync();
  return any;
}

// Coefficient of lane L of this lane's row of 16 for a wave-uniform L in
// {0, 4, 8, 12} (+1 with odd): row_share needs L at compile time.
__device__ __forceinline__ double row_share4(double v, int w, bool odd) {
  if (odd) {
    switch (w) {
      case 0: return row_share<1>(v);
      case 1: return row_share<5>(v);
      case 2: return row_share<9>(v);
      default: return row_share<13>(v);
    }
  }
  switch (w) {
    case 0: return row_share<0>(v);
    case 1: return row_share<4>(v);
    case 2: return row_share<8>(v);
    default: return row_share<12>(v);
  }
}

// The full sweep of inner_sweep (all 496 index pairs of the 32x32 subproblem
// in W, 31 circle-method steps) by four wavefronts: the same lane layout,
// rotations and update formulas, with wavefront w taking u = w of the
// one-wavefront form (the 2x2 tile of row pair r + 4w and column pair kk,
// and the Q columns of pairs r + 4w and r + 1 + 4w of row iq).  Every
// wavefront computes all 16 rotations of a step from the same LDS values,
// so they agree on them without an exchange; one workgroup barrier separates
// those reads from the step's writes (the diagonal tiles hold them), another
// orders the steps.  Returns whether this wavefront applied any rotation.
__device__ bool inner_sweep_full4(WaveLds& W, double thr, int tid) {
  {  // no off-diagonal element above the bar: no step would rotate (see bj_inner_cross4)
    bool need = false;
    for (int e = tid; e < SB * SB; e += 4 * 64)  // four wavefronts
      need |= (e >> 5) != (e & 31) && fabs(W.S[(e >> 5) * SST + (e & 31)]) > thr;
    if (!__syncthreads_or(need ? 1 : 0)) return false;
  }
  const int w = tid >> 6, lane = tid & 63;
  const int r = lane >> 4;
  const int kk = ((lane & 15) + r) & 15;
  const int iq = lane & 31;
  bool any = false;
  for (int s = 0; s < SB - 1; ++s) {
    int pp, qq;
    pair32(kk, s, false, pp, qq);
    double c2, s2;
    const bool rot = rotation(W.S[pp * SST + pp], W.S[qq * SST + qq], W.S[pp * SST + qq], thr, c2, s2);
    const bool step_any = __any(rot ? 1 : 0);  // the same in every wavefront (same values, same arithmetic)
    // every wavefront has read the diagonal entries of the step's pairs before
    // any wavefront overwrites them (the diagonal 2x2 tiles are updated below)
    __syncthreads();
    if (step_any) {
      any = true;
      const double ca = row_share4(c2, w, false), sa = row_share4(s2, w, false);
      const double cb = row_share4(c2, w, true), sb = row_share4(s2, w, true);
      int p1, q1, pa, qa, pb, qb;
      pair32(r + 4 * w, s, false, p1, q1);
      pair32(r + 4 * w, s, false, pa, qa);
      pair32((r + 1 + 4 * w) & 15, s, false, pb, qb);
      double x0 = W.S[p1 * SST + pp], x1 = W.S[p1 * SST + qq];
      double x2 = W.S[q1 * SST + pp], x3 = W.S[q1 * SST + qq];
      double ya0 = W.Q[iq * SST + pa], ya1 = W.Q[iq * SST + qa];
      double yb0 = W.Q[iq * SST + pb], yb1 = W.Q[iq * SST + qb];
      tile_rot(ca, sa, c2, s2, x0, x1, x2, x3);
      W.S[p1 * SST + pp] = x0;
      W.S[p1 * SST + qq] = x1;
      W.S[q1 * SST + pp] = x2;
      W.S[q1 * SST + qq] = x3;
      qrow_rot(ca, sa, ya0, ya1);
      qrow_rot(cb, sb, yb0, yb1);
      W.Q[iq * SST + pa] = ya0;
      W.Q[iq * SST + qa] = ya1;
      W.Q[iq * SST + pb] = yb0;
      W.Q[iq * SST + qb] = yb1;
    }
    __syncthreads();
  }
  return any;
}

// The cross sweep of bj_inner_cross spread over CW = 4 wavefronts (split
// form, outer steps s > 0).  Same slot space, rotations and update formulas;
// lane l of wavefront w owns the single tile at row slot j = 4 (l / 16) + w,
// column slot x = l % 16 of the four 16x16 blocks, and the Q rows
// i = 8 (l / 16) + 2 w + v (v < 2).  Per round: every lane computes the
// rotation of its column slot x from the round's inputs in LDS and takes
// that of its row slot j from lane j of its wavefront; the J-side column
// advance stays within each row of 16 lanes (DPP), the J-side row advance
// (slot j takes j + 1, which lives in the next wavefront) goes through LDS.
// The LDS exchanges are double buffered by round parity, so a round needs
// one barrier.
constexpr int CW = 4;

struct CrossLds {
  double xb[2][2][BS][BS + 1];   // bji and bjj rows for the row advance, by round parity
  double rin[2][BS][3];          // (a_pp, a_qq, a_pq) of the 16 pairs of a round, by round parity
};

__device__ bool bj_inner_cross4(CrossLds& X, const BjMat& M, int k, int s, double thr, int tid) {
  int bi, bj;
  block_pair(k, s, M.nb, bi, bj);
  const int w = tid >> 6, lane = tid & 63;
  const int x = lane & 15, g = lane >> 4;
  const int j = 4 * g + w;  // row slot
  const int n = M.n, nmax = M.nmax, last = M.nlast;
  const int gix = bi * BS + x, gjx = bj * BS + x;
  const int gij = bi * BS + j, gjj = bj * BS + j;
  double aii, bij, bji, bjj, qi[2], qj[2];
  {
    const double* Ai = M.A + (size_t)min(gij, last) * nmax;
    const double* Aj = M.A + (size_t)min(gjj, last) * nmax;
    aii = Ai[min(gix, last)] * mask01(gij < n && gix < n);
    bij = Ai[min(gjx, last)] * mask01(gij < n && gjx < n);
    bji = Aj[min(gix, last)] * mask01(gjj < n && gix < n);
    bjj = Aj[min(gjx, last)] * mask01(gjj < n && gjx < n);
  }
  // The 16 rotations of a round are computed from their inputs in LDS, which
  // the lanes holding them write there beside the row advance of the round
  // before (the diagonal tile's a_pp and a_pq, and a_qq of the next slot,
  // which the advance moves onto the diagonal).  Every wavefront computes all
  // 16 (each lane that of its column slot), so no second barrier hands them
  // out; the values, and so the results, are those of each diagonal lane
  // computing its own rotation.
  const bool diag = j == x;
  if (diag) {
    X.rin[0][x][0] = aii;
    X.rin[0][x][1] = bjj;
    X.rin[0][x][2] = bij;
  }
  // An element changes only through a rotation, and a rotation needs its
  // off-diagonal element above the bar: a pair whose cross block is at or
  // below the bar throughout rotates nothing in any round (the rounds would
  // leave everything as it is), so they are skipped.  (The barrier also
  // publishes the first round's inputs.)
  if (!__syncthreads_or(fabs(bij) > thr ? 1 : 0)) return false;
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const int i = 8 * g + 2 * w + v;
    qi[v] = (i == x) ? 1.0 : 0.0;
    qj[v] = (i == BS + x) ? 1.0 : 0.0;
  }
  const int jn = (j + 1) & (BS - 1), xn = (x + 1) & (BS - 1);
  bool any = false;
  for (int r = 0; r < BS; ++r) {
    const int par = r & 1;
    double c2, s2;
    const bool rot = rotation(X.rin[par][x][0], X.rin[par][x][1], X.rin[par][x][2], thr, c2, s2);
    // every wavefront holds all 16 column slots: the vote is the same in each
    if (__any(rot ? 1 : 0)) {  // uniform over the workgroup
      any = true;
      const double c1 = __shfl(c2, j), s1 = __shfl(s2, j);  // lane j holds column slot j
      tile_rot(c1, s1, c2, s2, aii, bij, bji, bjj);  // tile (j, x) <- R_j^T tile R_x
#pragma unroll
      for (int v = 0; v < 2; ++v) qrow_rot(c2, s2, qi[v], qj[v]);  // Q <- Q J: columns I_x and J_sigma(x)
    }
    // advance sigma: J-side columns take slot x + 1 (DPP within the row of 16
    // lanes), J-side rows slot j + 1 (through LDS)
    bij = row_next(bij);
#pragma unroll
    for (int v = 0; v < 2; ++v) qj[v] = row_next(qj[v]);
    X.xb[par][0][j][x] = bji;
    X.xb[par][1][j][x] = bjj;
    if (diag) {  // inputs of the next round: a_pp and a_pq of slot x, a_qq of slot x - 1
      X.rin[par ^ 1][x][0] = aii;
      X.rin[par ^ 1][x][2] = bij;
      X.rin[par ^ 1][(x + BS - 1) & (BS - 1)][1] = bjj;
    }
    __syncthreads();
    bji = X.xb[par][0][jn][x];
    bjj = X.xb[par][1][jn][xn];
  }
  if (any) {  // uniform over the workgroup
    if (gij < n && gix < n) M.A[(size_t)gij * nmax + gix] = aii;
    if (gij < n && gjx < n) M.A[(size_t)gij * nmax + gjx] = bij;
    if (gjj < n && gix < n) M.A[(size_t)gjj * nmax + gix] = bji;
    if (gjj < n && gjx < n) M.A[(size_t)gjj * nmax + gjx] = bjj;
    double* Q = qstep(M, s) + (size_t)k * SB * SB;
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int i = 8 * g + 2 * w + v;
      Q[i * SB + x] = qi[v];
      Q[i * SB + BS + x] = qj[v];
    }
  }
  return any;
}

// Inner phase of block pair k at outer step s, by one wavefront: stage the
// 32x32 subproblem, sweep it (all pairs at step 0, where every block has a
// partner; the pairs across the two blocks afterwards), write the diagonal
// tile and the rotation back.  Returns whether anything was rotated.
__device__ bool bj_inner_pair(WaveLds& W, const BjMat& M, int k, int s, double thr, int lane) {
  if (s > 0) return bj_inner_cross(W.S, M, k, s, thr, lane);
  int bi, bj;
  block_pair(k, s, M.nb, bi, bj);
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = lane + 64 * u;
    const int gr = gidx(e >> 5, bi, bj), gc = gidx(e & 31, bi, bj);
    const double x = M.A[(size_t)min(gr, M.nlast) * M.nmax + min(gc, M.nlast)];
    v[u] = x * mask01(gr < M.n && gc < M.n);
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = lane + 64 * u;
    const int r = e >> 5, c = e & 31;
    W.S[r * SST + c] = v[u];
    W.Q[r * SST + c] = (r == c) ? 1.0 : 0.0;
  }
  wave_sync();
  bool need = false;  // an off-diagonal element above the bar (else no step would rotate)
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = lane + 64 * u;
    need |= (e >> 5) != (e & 31) && fabs(v[u]) > thr;
  }
  // step 0: all pairs (later steps: bj_inner_cross)
  const bool any = __any(need ? 1 : 0) && inner_sweep<false>(W, thr, lane);
  if (any) {
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
      const int e = lane + 64 * u;
      const int r = e >> 5, c = e & 31;
      const int gr = gidx(r, bi, bj), gc = gidx(c, bi, bj);
      if (gr < M.n && gc < M.n) M.A[(size_t)gr * M.nmax + gc] = W.S[r * SST + c];
      qstep(M, s)[(size_t)k * SB * SB + e] = W.Q[r * SST + c];
    }
  }
  wave_sync();
  return any;
}

// Element (r, c) of a 32x32 staging tile stored without padding: the column
// is XORed with the row, so that a wavefront's accesses along a row and
// along a column (the transposed read) fall in distinct LDS banks.
__device__ __forceinline__ int stile(int r, int c) { return r * SB + (c ^ (r & (SB - 1))); }

// Staging tile index: swizzled 32x32 (SWZ) or padded 32x33 rows.
template <bool SWZ>
__device__ __forceinline__ int tix(int r, int c) { return SWZ ? stile(r, c) : r * SST + c; }

// tile index -> (k, k2), k < k2
__device__ __forceinline__ void tile_pairs(int item, int npairs, int& k, int& k2) {
  int rem = item;
  k = 0;
  while (rem >= npairs - 1 - k) {
    rem -= npairs - 1 - k;
    ++k;
  }
  k2 = k + 1 + rem;
}

// The off-diagonal tile (k < k2) of outer step s becomes Q_k^T A_kk2 Q_k2, by
// one wavefront (Sx: its staging tile in LDS, indexed by tix<SWZ>), written
// to both triangles.  rot: flags of the pairs.
template <bool SWZ>
__device__ void bj_product_tile(double* __restrict__ Sx, const BjMat& M, const int* rot, int k, int k2, int s,
                                int lane) {
  const int n = M.n, nmax = M.nmax, nlast = M.nlast;
  const bool rk = rot[k] != 0, rk2 = rot[k2] != 0;
  if (!rk && !rk2) return;
  int bi, bj, ci, cj;
  block_pair(k, s, M.nb, bi, bj);
  block_pair(k2, s, M.nb, ci, cj);
  double* A = M.A;
  auto loadA = [&](int m, int c) -> double {
    const int gr = gidx(m, bi, bj), gc = gidx(c, ci, cj);
    const double x = A[(size_t)min(gr, nlast) * nmax + min(gc, nlast)];
    return x * mask01(gr < n && gc < n);
  };
  d4_t acc[2][2];
  if (rk) {
    qt_times(qstep(M, s) + (size_t)k * SB * SB, loadA, lane, acc);
  } else {  // Q_k = 1: T is the tile itself
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[rb][cb][i] = loadA(rb * BS + acc_row(lane, i), cb * BS + (lane & 15));
  }
  // T = Q_k^T X -> LDS (swizzled rows)
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        Sx[tix<SWZ>(rb * BS + acc_row(lane, i), cb * BS + (lane & 15))] = acc[rb][cb][i];
  wave_sync();
  if (rk2) {
    // U^T = Q_k2^T T^T with U = T Q_k2: the same product, T read transposed
    auto loadT = [&](int m, int c) -> double { return Sx[tix<SWZ>(c, m)]; };
    qt_times(qstep(M, s) + (size_t)k2 * SB * SB, loadT, lane, acc);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int uc = rb * BS + acc_row(lane, i), ur = cb * BS + (lane & 15);
          const int gr = gidx(ur, bi, bj), gc = gidx(uc, ci, cj);
          if (gr < n && gc < n) {
            const double v = acc[rb][cb][i];
            A[(size_t)gr * nmax + gc] = v;
            A[(size_t)gc * nmax + gr] = v;
          }
        }
  } else {  // Q_k2 = 1: U = T
    for (int e = lane; e < SB * SB; e += 64) {
      const int r = e >> 5, c = e & 31;
      const int gr = gidx(r, bi, bj), gc = gidx(c, ci, cj);
      if (gr < n && gc < n) {
        const double v = Sx[tix<SWZ>(r, c)];
        A[(size_t)gr * nmax + gc] = v;
        A[(size_t)gc * nmax + gr] = v;
      }
    }
  }
  wave_sync();
}

// Product item `item` of outer step s, by one wavefront: an off-diagonal tile
// (bj_product_tile), or a 32-column panel of the VT rows of pair k, which
// becomes Q_k^T VT.
template <bool SWZ>
__device__ void bj_product_item(double* __restrict__ Sx, const BjMat& M, const int* rot, int item, int s, int lane) {
  const int npairs = M.npairs, n = M.n, nmax = M.nmax, nlast = M.nlast;
  const int ntile = npairs * (npairs - 1) / 2;
  const int npanel = (n + SB - 1) / SB;
  if (item < ntile) {
    int k, k2;
    tile_pairs(item, npairs, k, k2);
    bj_product_tile<SWZ>(Sx, M, rot, k, k2, s, lane);
  } else {
    const int w2 = item - ntile;
    const int k = w2 / npanel, pc = w2 - k * npanel;
    if (!rot[k]) return;
    int bi, bj;
    block_pair(k, s, M.nb, bi, bj);
    const int c0 = pc * SB;
    double* Vb = M.V;
    auto loadV = [&](int m, int c) -> double {
      const int gr = gidx(m, bi, bj), gc = c0 + c;
      const double x = Vb[(size_t)min(gr, nlast) * nmax + min(gc, nlast)];
      return x * mask01(gr < n && gc < n);
    };
    d4_t acc[2][2];
    qt_times(qstep(M, s) + (size_t)k * SB * SB, loadV, lane, acc);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int gr = gidx(rb * BS + acc_row(lane, i), bi, bj), gc = c0 + cb * BS + (lane & 15);
          if (gr < n && gc < n) Vb[(size_t)gr * nmax + gc] = acc[rb][cb][i];
        }
  }
}

// product items of a step: the off-diagonal tiles, and the VT panels unless
// the eigenvector update is deferred to the end of the sweep
__device__ __forceinline__ int bj_nwork(const BjMat& M) {
  return M.npairs * (M.npairs - 1) / 2 + (M.defer ? 0 : M.npairs * ((M.n + SB - 1) / SB));
}

// Deferred eigenvector update of the 32-column panel pc of VT: every step s of
// the sweep replaces the rows IJ_k of the panel by Q_k^T (rows IJ_k) for the
// block pairs that rotated, on the matrix cores, with the panel (all rows) in
// LDS at P (npad x PST doubles).  Executed by all nw wavefronts of the
// workgroup (barriers inside).
constexpr int PST = SB + 1;  // LDS row stride of a panel
static_assert((size_t)DEFER_MAX * PST * sizeof(double) <= sizeof(BlockedLds),
              "the deferred VT panel must fit the one-workgroup kernel's LDS");

__device__ void bj_apply_panel(const BjMat& M, int pc, double* P, int wv, int nw, int lane) {
  const int n = M.n, nmax = M.nmax, c0 = pc * SB;
  const int nthr = nw * 64, tid = wv * 64 + lane;
  for (int e = tid; e < M.npad * SB; e += nthr) {
    const int r = e >> 5, c = e & 31;
    const int rr = min(r, M.nlast), cc = min(c0 + c, M.nlast);
    P[r * PST + c] = M.V[(size_t)rr * nmax + cc] * mask01(r < n && c0 + c < n);
  }
  __syncthreads();
  for (int s = 0; s < M.nb - 1; ++s) {
    const int* fl = fstep(M, s);
    for (int k = wv; k < M.npairs; k += nw) {
      if (!fl[k]) continue;
      int bi, bj;
      block_pair(k, s, M.nb, bi, bj);
      auto loadP = [&](int m, int c) -> double { return P[gidx(m, bi, bj) * PST + c]; };
      d4_t acc[2][2];
      qt_times(qstep(M, s) + (size_t)k * SB * SB, loadP, lane, acc);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            P[gidx(rb * BS + acc_row(lane, i), bi, bj) * PST + cb * BS + (lane & 15)] = acc[rb][cb][i];
    }
    __syncthreads();
  }
  for (int e = tid; e < n * SB; e += nthr) {
    const int r = e >> 5, c = e & 31;
    if (c0 + c < n) M.V[(size_t)r * nmax + c0 + c] = P[r * PST + c];
  }
  __syncthreads();
}

// max |offdiag| and max |diag| of the matrix, by the whole workgroup
// (eight independent loads in flight per thread); ends with a barrier.
__device__ void bj_norms(const BjMat& M, double* red, double& off, double& amax) {
  off = 0.0;
  amax = 0.0;
  const int n = M.n, nn = n * n;
  for (int e0 = threadIdx.x; e0 < nn; e0 += 8 * JBB) {
    double v[8];
    bool dg[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = min(e0 + u * JBB, nn - 1);
      const int i = e / n, j = e - i * n;
      const double x = M.A[(size_t)i * M.nmax + j];
      v[u] = fabs(x) * mask01(e0 + u * JBB < nn);
      dg[u] = i == j;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      off = fmax(off, dg[u] ? 0.0 : v[u]);
      amax = fmax(amax, dg[u] ? v[u] : 0.0);
    }
  }
  off = block_max<NWB>(off, red);
  amax = block_max<NWB>(amax, red);
}

// One workgroup per matrix, all sweeps.  init_v = 0 continues from the VT
// the split kernels left; sweep_base is then their per-matrix state, whose
// fourth entry counts the sweeps already done.
__global__ __launch_bounds__(JBB) void k_jacobi_blocked(int nmax, const JArgs* __restrict__ ja,
                                                       double* __restrict__ work, int* __restrict__ flags,
                                                       int max_sweeps, double tol, int init_v,
                                                       const double* __restrict__ sweep_base) {
  extern __shared__ double lds_raw[];
  BlockedLds& L = *reinterpret_cast<BlockedLds*>(lds_raw);
  const int b = blockIdx.x;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch) return;
  const double* __restrict__ tol_b = jc->tol_b;
  int* __restrict__ sweeps_out = jc->sweeps;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const BjMat M = bj_mat(b, nmax, jc->nsize, jc->A, jc->VT, work, flags, jc->defer);
  double* eb = jc->evals + (size_t)b * nmax;
  const int base = sweep_base ? (int)sweep_base[4 * b + 3] : 0;

  if (init_v) {
    for (int e = threadIdx.x; e < nmax * nmax; e += JBB) {
      const int i = e / nmax, j = e - i * nmax;
      M.V[e] = (i == j) ? 1.0 : 0.0;
    }
  }
  if (M.n == 0) {
    for (int i = threadIdx.x; i < nmax; i += JBB) eb[i] = 0.0;
    if (threadIdx.x == 0 && sweeps_out != nullptr) sweeps_out[b] = base;
    return;
  }
  WaveLds& W = L.w[wv];
  int sweep = 0;
  double off_prev = sweep_base ? sweep_base[4 * b] : 0.0;
  if (sweep_base && sweep_base[4 * b + 2] != 0.0) max_sweeps = 0;  // converged in the split sweeps
  for (; sweep < max_sweeps; ++sweep) {
    double off, amax;
    bj_norms(M, L.red, off, amax);  // ends with a barrier: V init and last step visible
    const double thr = fmax(1e-300, (tol_b ? tol_b[b] : tol) * fmax(amax, 1.0));
    if (stop_sweeps(off, off_prev, thr, base + sweep)) break;
    off_prev = off;
    for (int s = 0; s < M.nb - 1; ++s) {
      for (int k = wv; k < M.npairs; k += NWB) {  // inner phase: one wavefront per block pair
        const bool any = bj_inner_pair(W, M, k, s, thr, lane);
        if (lane == 0) fstep(M, s)[k] = any ? 1 : 0;
      }
      __syncthreads();
      const int nwork = bj_nwork(M);  // product phase: tiles (and VT panels unless deferred)
      for (int item = wv; item < nwork; item += NWB) bj_product_item<false>(W.S, M, fstep(M, s), item, s, lane);
      __syncthreads();
    }
    if (M.defer) {  // the sweep's rotations onto VT, panel by panel (LDS reused)
      for (int pc = 0; pc < M.npairs; ++pc) bj_apply_panel(M, pc, lds_raw, wv, NWB, lane);
    }
  }
  for (int i = threadIdx.x; i < nmax; i += JBB) eb[i] = (i < M.n) ? M.A[(size_t)i * nmax + i] : 0.0;
  if (threadIdx.x == 0 && sweeps_out != nullptr) sweeps_out[b] = base + sweep;
}

// ---- split form: each outer step of the blocked variant as two grid-wide
// launches (one workgroup per (matrix, block pair), then workgroups over the
// matrix's product items).  For a few large matrices one workgroup each would
// leave the device idle and run a step's product items one after another; for
// a full batch it is about as fast as one workgroup per matrix alone, and its
// small workgroups share the device better with other streams' kernels.
// Per-matrix state in the workspace (4 doubles): off of the previous sweep,
// rotation bar, done flag, sweeps done.

constexpr int SPLIT_MIN_N = 128;  // auto: split form from this order, for batches below half of it
// step kernel: one-wavefront sweeps from this many block pairs per call
// (batch capacity x pairs).  Three sweeps, warm-start-like matrices, with
// one barrier per round of the four-wavefront sweep: up to 16 matrices the
// four-wavefront sweeps are 13-17 % faster, at 320 pairs (64 matrices of
// order 160) 5 %, at 512 (64 of order 240) 3 %; from 640 (64 of order 320)
// the one-wavefront sweeps win by 5-14 % (128 matrices of order 256: 2.27
// against 2.59 ms; 256 of order 384: 15.1 against 16.8 ms).
constexpr int WAVE1_MIN_PAIRS = 576;

// VT = 1 and the split state cleared: grid (matrix, chunk of INIT_ROWS rows),
// so that a lone matrix's identity is written by many workgroups at once
constexpr int INIT_ROWS = 8;
__host__ __device__ constexpr int init_chunks(int nmax) { return (nmax + INIT_ROWS - 1) / INIT_ROWS; }

__global__ __launch_bounds__(JBB) void k_bj_init(int nmax, const JArgs* __restrict__ ja, double* __restrict__ sst) {
  const int b = blockIdx.x, c = blockIdx.y;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch) return;
  double* Vb = jc->VT + (size_t)b * nmax * nmax;
  const int r0 = c * INIT_ROWS, r1 = min(nmax, r0 + INIT_ROWS);
  const int e0 = r0 * nmax, e1 = r1 * nmax;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += JBB) {
    const int i = e / nmax, j = e - i * nmax;
    Vb[e] = (i == j) ? 1.0 : 0.0;
  }
  if (c == 0 && threadIdx.x < 4) sst[4 * b + threadIdx.x] = 0.0;
}

// Convergence test of the split form, in two launches: k_bj_norms reduces
// chunks of whole rows of every matrix over many workgroups, k_bj_check
// combines a matrix's chunks and decides whether it stops.  (One workgroup
// reading a whole matrix took 2.2 ms at order 4002.)  A chunk holds about
// NORM_EL elements, at most 1024 chunks per matrix; every thread keeps NPT
// loads in flight at a time (a lone order-240 matrix: 25 us with one load at
// a time over chunks of 16384 elements).
constexpr int NRB = 256;  // threads per k_bj_norms workgroup
constexpr int NPT = 8;    // loads in flight per thread
constexpr int NORM_EL = NRB * NPT;
__host__ __device__ constexpr int norm_rows(int nmax) {
  return (NORM_EL / nmax > (nmax + 1023) / 1024) ? NORM_EL / nmax : (nmax + 1023) / 1024;
}
__host__ __device__ constexpr int norm_chunks(int nmax) { return (nmax + norm_rows(nmax) - 1) / norm_rows(nmax); }

__global__ __launch_bounds__(NRB) void k_bj_norms(int nmax, const JArgs* __restrict__ ja,
                                                 const double* __restrict__ sst, double* __restrict__ nrm) {
  __shared__ double red[NRB / 64];
  const int b = blockIdx.x, c = blockIdx.y;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const int n = min(max(jc->nsize[b], 0), nmax);
  const double* __restrict__ A = jc->A + (size_t)b * nmax * nmax;
  const int r0 = c * norm_rows(nmax), r1 = min(n, r0 + norm_rows(nmax));
  double off = 0.0, amax = 0.0;
  const int ncol = n, nel = max(r1 - r0, 0) * ncol;
  for (int base = threadIdx.x; base < nel; base += NRB * NPT) {
    // element indices past the chunk are clamped to its last element: read
    // again, they leave the maxima unchanged
    double x[NPT];
    bool dg[NPT];
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      const int e = min(base + u * NRB, nel - 1);
      const int q = e / ncol;
      const int i = r0 + q, j = e - q * ncol;
      x[u] = fabs(A[(size_t)i * nmax + j]);
      dg[u] = i == j;
    }
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
      off = fmax(off, dg[u] ? 0.0 : x[u]);
      amax = fmax(amax, dg[u] ? x[u] : 0.0);
    }
  }
  off = block_max<NRB / 64>(off, red);
  amax = block_max<NRB / 64>(amax, red);
  if (threadIdx.x == 0) {
    double* o = nrm + ((size_t)b * norm_chunks(nmax) + c) * 2;
    o[0] = off;
    o[1] = amax;
  }
}

__global__ __launch_bounds__(64) void k_bj_check(int nmax, const JArgs* __restrict__ ja, const double* __restrict__ nrm,
                                                double tol, double* __restrict__ sst, int sweep) {
  const int b = blockIdx.x, lane = threadIdx.x;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const double* __restrict__ tol_b = jc->tol_b;
  const int nch = norm_chunks(nmax);
  double off = 0.0, amax = 0.0;
  for (int base = lane; base < nch; base += 64 * 4) {
    double o[4], a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // clamped to the last chunk, as in k_bj_norms
      const int c = min(base + 64 * u, nch - 1);
      o[u] = nrm[((size_t)b * nch + c) * 2];
      a[u] = nrm[((size_t)b * nch + c) * 2 + 1];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      off = fmax(off, o[u]);
      amax = fmax(amax, a[u]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    off = fmax(off, __shfl_xor(off, o, 64));
    amax = fmax(amax, __shfl_xor(amax, o, 64));
  }
  const int n = min(max(jc->nsize[b], 0), nmax);
  const double thr = fmax(1e-300, (tol_b ? tol_b[b] : tol) * fmax(amax, 1.0));
  if (lane == 0) {
    if (n == 0 || stop_sweeps(off, sst[4 * b + 0], thr, sweep)) {
      sst[4 * b + 2] = 1.0;
    } else {
      sst[4 * b + 0] = off;
      sst[4 * b + 1] = thr;
      sst[4 * b + 3] += 1.0;
    }
  }
}

// Inner phase of the split form at outer step 0: one four-wavefront
// workgroup per (block pair, matrix) stages the 32x32 subproblem, runs the
// full sweep (inner_sweep_full4) and writes back the diagonal tile and the
// rotation, as bj_inner_pair does with one wavefront.
__global__ __launch_bounds__(CW * 64) void k_bj_inner_full(int nmax, const JArgs* __restrict__ ja,
                                                          double* __restrict__ work, const double* __restrict__ sst,
                                                          int* __restrict__ rot) {
  __shared__ WaveLds W;
  const int b = blockIdx.y, k = blockIdx.x, tid = threadIdx.x;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const BjMat M = bj_mat(b, nmax, jc->nsize, jc->A, jc->VT, work, rot, jc->defer);
  if (k >= M.npairs || M.nb < 2) return;
  int bi, bj;
  block_pair(k, 0, M.nb, bi, bj);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + CW * 64 * u;
    const int gr = gidx(e >> 5, bi, bj), gc = gidx(e & 31, bi, bj);
    const double x = M.A[(size_t)min(gr, M.nlast) * M.nmax + min(gc, M.nlast)];
    W.S[(e >> 5) * SST + (e & 31)] = x * mask01(gr < M.n && gc < M.n);
    W.Q[(e >> 5) * SST + (e & 31)] = ((e >> 5) == (e & 31)) ? 1.0 : 0.0;
  }
  __syncthreads();
  const bool any = __syncthreads_or(inner_sweep_full4(W, sst[4 * b + 1], tid) ? 1 : 0) != 0;
  if (any) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + CW * 64 * u;
      const int r = e >> 5, c = e & 31;
      const int gr = gidx(r, bi, bj), gc = gidx(c, bi, bj);
      if (gr < M.n && gc < M.n) M.A[(size_t)gr * M.nmax + gc] = W.S[r * SST + c];
      qstep(M, 0)[(size_t)k * SB * SB + e] = W.Q[r * SST + c];
    }
  }
  if (tid == 0) fstep(M, 0)[k] = any ? 1 : 0;
}

__global__ __launch_bounds__(PW * 64) __attribute__((amdgpu_waves_per_eu(5))) void k_bj_product(int nmax, const JArgs* __restrict__ ja, double* __restrict__ work,
                                                   const double* __restrict__ sst, int* __restrict__ rot,
                                                   int s) {
  extern __shared__ double lds_raw[];  // one staging tile per wavefront
  // grid (matrix, item group): consecutive workgroup ids go to the eight XCDs
  // in turn, so with a capacity that is a multiple of 8 all workgroups of one
  // matrix run on one XCD and share its rotations in that XCD's L2
  const int b = blockIdx.x;
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const BjMat M = bj_mat(b, nmax, jc->nsize, jc->A, jc->VT, work, rot, jc->defer);
  if (s >= M.nb - 1) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int item = blockIdx.y * PW + wv;
  if (item < bj_nwork(M)) bj_product_item<true>(lds_raw + (size_t)wv * PROD_TILE, M, fstep(M, s), item, s, lane);
}

// Outer step s > 0 of the split form in one launch: the product phase of
// step s - 1 and the cross sweeps of step s.  Every block pair (I, J) of step
// s is swept by the workgroup that forms the tile of step s - 1 holding
// A(I, J): with WAVE1 false, workgroup y < npairs(nmax) forms it (wavefront 0,
// bj_product_tile) and sweeps pair y with all four wavefronts
// (bj_inner_cross4); with WAVE1, wavefront w of workgroup y < ceil(npairs / 4)
// forms the tile of pair 4 y + w and sweeps the pair alone (bj_inner_cross,
// wavefront barriers only; a quarter of the workgroup slots per pair, for
// batches that fill the device).  A sweep reads the new tile and the diagonal
// blocks A(I, I), A(J, J), which the sweeps of step s - 1 wrote and no
// product changes.  The workgroups above take the other product items of
// step s - 1, PW per workgroup, as k_bj_product.  No workgroup writes what
// another reads: a pair's sweep writes its own blocks (in its own tile and in
// diagonal tiles of step s - 1, which no product reads) and the rotations and
// flags of step s (kept apart from those of step s - 1).  With four blocks
// the only tile holds both pairs: the workgroup (wavefront) of pair 0 sweeps
// them one after the other.  A matrix whose sweep ends before step s only
// forms its last products.  The two forms of the sweep, and the products,
// are those of the two-launch step, bit for bit.
__host__ __device__ constexpr int sweep_groups(int nmax, bool wave1) {
  return wave1 ? ((nmax + SB - 1) / SB + PW - 1) / PW : (nmax + SB - 1) / SB;
}

// (the one-wavefront sweep holds its 32x32 subproblem and rotation in 64
// registers per lane: three waves per SIMD, without spills; four spilled
// and ran 1-9 % slower)
template <bool WAVE1>
__global__ __launch_bounds__(PW * 64) __attribute__((amdgpu_waves_per_eu(WAVE1 ? 3 : 5))) void k_bj_step(
    int nmax, const JArgs* __restrict__ ja, double* __restrict__ work, const double* __restrict__ sst,
    int* __restrict__ rot, int s) {
  extern __shared__ double lds_raw[];  // staging tiles; a sweep's exchange area after the first
  const int b = blockIdx.x;            // the matrix: its workgroups on one XCD, as k_bj_product
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const BjMat M = bj_mat(b, nmax, jc->nsize, jc->A, jc->VT, work, rot, jc->defer);
  if (s < 1 || s > M.nb - 1) return;
  const bool sweep = s < M.nb - 1;  // step s exists for this matrix
  const int nsg = sweep_groups(nmax, WAVE1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int y = blockIdx.y;
  if (y < nsg) {
    const int p = WAVE1 ? PW * y + wv : y;  // the pair this wavefront (workgroup) sweeps
    // uniform over the wavefront (WAVE1) or the workgroup
    if (!sweep || p >= M.npairs || (M.nb == 4 && p == 1)) return;
    int bi, bj;
    block_pair(p, s, M.nb, bi, bj);
    const int ka = pair_of_block(bi, s - 1, M.nb), kb = pair_of_block(bj, s - 1, M.nb);
    const double thr = sst[4 * b + 1];
    const int klast = M.nb == 4 ? 1 : p;
    if (WAVE1) {
      double* Sx = lds_raw + (size_t)wv * PROD_TILE;
      bj_product_tile<true>(Sx, M, fstep(M, s - 1), min(ka, kb), max(ka, kb), s - 1, lane);
      wave_sync();  // the tile's writes to A seen by the whole wavefront; Sx free
      for (int k = p; k <= klast; ++k) {
        const bool any = bj_inner_cross(Sx, M, k, s, thr, lane);
        if (lane == 0) fstep(M, s)[k] = any ? 1 : 0;
      }
      return;
    }
    if (wv == 0) bj_product_tile<true>(lds_raw, M, fstep(M, s - 1), min(ka, kb), max(ka, kb), s - 1, lane);
    __syncthreads();  // the tile's writes to A seen by every wavefront
    CrossLds& X = *reinterpret_cast<CrossLds*>(lds_raw + PROD_TILE);
    for (int k = p; k <= klast; ++k) {
      const bool any = bj_inner_cross4(X, M, k, s, thr, threadIdx.x);
      if (threadIdx.x == 0) fstep(M, s)[k] = any ? 1 : 0;
      __syncthreads();  // X is reused by the next pair
    }
    return;
  }
  const int item = (y - nsg) * PW + wv;
  if (item >= bj_nwork(M)) return;
  if (sweep && item < M.npairs * (M.npairs - 1) / 2) {
    int k, k2;
    tile_pairs(item, M.npairs, k, k2);
    if (tile_holds_next_pair(M.nb, s, k, k2)) return;  // formed by the sweeper of that pair
  }
  bj_product_item<true>(lds_raw + (size_t)wv * PROD_TILE, M, fstep(M, s - 1), item, s - 1, lane);
}
static_assert((size_t)PROD_TILE * sizeof(double) + sizeof(CrossLds) <= PROD_LDS,
              "k_bj_step: a staging tile and the sweep's exchange area in the product LDS");

// Deferred eigenvector update of one sweep: one workgroup per (matrix, VT
// panel of 32 columns), the panel's rows in LDS.
// wavefronts per apply workgroup: the panel's LDS (all rows x 32 columns) lets
// one or two workgroups share a CU, so eight wavefronts take a step's block
// pairs one each (four took two each, one after the other)
constexpr int AW = 8;
// Above eight block pairs (orders 257-512) a step's pairs would take two
// passes of eight wavefronts, the second one partly idle: the workgroup has
// a wavefront per block pair instead (the panel's operations, and so its
// values, do not depend on which wavefront applies a pair).
constexpr int AW_MAX = 16;
__host__ __device__ constexpr int apply_waves(int npairs) {
  return npairs > AW ? (npairs < AW_MAX ? npairs : AW_MAX) : AW;
}

__global__ __launch_bounds__(AW_MAX * 64) void k_bj_apply(int nmax, int nbcap, const JArgs* __restrict__ ja,
                                                     double* __restrict__ work, const double* __restrict__ sst,
                                                     int* __restrict__ rot) {
  extern __shared__ double lds_raw[];
  // One workgroup per (matrix, panel) on a 1D grid.  Workgroup ids go to the
  // eight XCDs in turn (id = 8 slot + xcd); matrix b = 8 m + xcd takes slots
  // m npm .. m npm + npm - 1 of XCD xcd, so its panels are dispatched one
  // after another on one XCD and read the sweep's rotation log into that
  // XCD's L2 together (with the panels of a matrix nbcap ids apart, most of
  // the log was gone from the L2 when the next panel read it).
  const int npm = (nmax + SB - 1) / SB, id = blockIdx.x;
  int b, pc;
  if ((nbcap & 7) == 0) {
    const int slot = id >> 3, m = slot / npm;
    b = 8 * m + (id & 7);
    pc = slot - m * npm;
  } else {
    b = id / npm;
    pc = id - b * npm;
  }
  const JArgsC jc = jconst(ja);
  if (b >= jc->nbatch || sst[4 * b + 2] != 0.0) return;
  const BjMat M = bj_mat(b, nmax, jc->nsize, jc->A, jc->VT, work, rot, jc->defer);
  if (pc >= M.npairs) return;  // n <= 32 npairs: one panel per block pair
  bj_apply_panel(M, pc, lds_raw, threadIdx.x >> 6, (int)(blockDim.x >> 6), threadIdx.x & 63);
}

bool attr_done[64];

// Batch capacity of a call: the grids of the blocked kernels span it (rows
// beyond nbatch return at once), so that windows of nearby sizes share one
// recorded launch sequence.  Powers of two up to 8, then multiples of 8 up to
// 64, then multiples of 32.
int batch_capacity(int nbatch) {
  if (nbatch <= 8) {
    int c = 1;
    while (c < nbatch) c <<= 1;
    return c;
  }
  if (nbatch <= 64) return (nbatch + 7) / 8 * 8;
  return (nbatch + 31) / 32 * 32;
}

// workspace: argument block (JARGS_BYTES), rotations (nbcap x log_steps x
// npad x 32 doubles), split state (nbcap x 4 doubles), rotation flags (nbcap x
// log_steps x KMAX ints)
size_t rot_bytes(int nbcap, int nmax) {
  const size_t npad = (size_t)(nmax + SB - 1) / SB * SB;
  return (size_t)nbcap * log_steps(nmax) * npad * SB * sizeof(double);
}

size_t apply_lds(int nmax) { return (size_t)((nmax + SB - 1) / SB * SB) * PST * sizeof(double); }

struct Layout {
  JArgs* ja;
  double* rot_q;  // rotations of the current outer step
  double* sst;    // split state
  int* rot;       // rotation flags
  double* nrm;    // per matrix and row chunk: largest off-diagonal and diagonal magnitude
};

size_t flag_bytes(int nbcap, int nmax) {
  return ((size_t)nbcap * log_steps(nmax) * KMAX * sizeof(int) + 255) / 256 * 256;
}
size_t norm_bytes(int nbcap, int nmax) { return (size_t)nbcap * norm_chunks(nmax) * 2 * sizeof(double); }

Layout layout(double* work, int nbcap, int nmax) {
  Layout w;
  char* p = reinterpret_cast<char*>(work);
  w.ja = reinterpret_cast<JArgs*>(p);
  w.rot_q = reinterpret_cast<double*>(p + JARGS_BYTES);
  w.sst = w.rot_q + rot_bytes(nbcap, nmax) / sizeof(double);
  w.rot = reinterpret_cast<int*>(w.sst + 4 * (size_t)nbcap);
  w.nrm = reinterpret_cast<double*>(reinterpret_cast<char*>(w.rot) + flag_bytes(nbcap, nmax));
  return w;
}

// One-wavefront sweeps in the step kernel for batches whose step workgroups
// fill the device (environment DXTB_JACOBI_WAVE1: 0 never, 1 always).
int wave1_env() {
  static const int v = [] {
    const char* e = std::getenv("DXTB_JACOBI_WAVE1");
    return e ? std::atoi(e) : -1;
  }();
  return v;
}
bool step_wave1(int nbcap, int nmax) {
  const int e = wave1_env();
  if (e >= 0) return e != 0;
  return (long)nbcap * ((nmax + SB - 1) / SB) >= WAVE1_MIN_PAIRS;
}

// The blocked kernels of one call (after k_bj_args): nsw split sweeps, then
// the one-workgroup kernel for the remaining max_sweeps (and the eigenvalues).
// Every argument is fixed by (work, nbcap, nmax, nsw, max_sweeps, tol).
hipError_t enqueue_blocked(hipStream_t st, const Layout& w, int nbcap, int nmax, int nsw, int max_sweeps, double tol) {
  const int npairs = (nmax + SB - 1) / SB, nb = 2 * npairs;
  const bool defer = defer_call(nmax, nbcap);
  const int nwork = npairs * (npairs - 1) / 2 + (defer ? 0 : npairs * npairs);
  if (nsw > 0) {
    hipLaunchKernelGGL(k_bj_init, dim3((unsigned)nbcap, (unsigned)init_chunks(nmax)), dim3(JBB), 0, st, nmax, w.ja,
                       w.sst);
    for (int sw = 0; sw < nsw; ++sw) {
      hipLaunchKernelGGL(k_bj_norms, dim3((unsigned)nbcap, (unsigned)norm_chunks(nmax)), dim3(NRB), 0, st, nmax,
                         w.ja, w.sst, w.nrm);
      hipLaunchKernelGGL(k_bj_check, dim3((unsigned)nbcap), dim3(64), 0, st, nmax, w.ja, w.nrm, tol, w.sst, sw);
      // step 0: the full sweep of every subproblem (from LDS); steps 1 ..
      // nb - 2: the products of the step before and the cross sweeps in one
      // launch; then the products of the last step
      hipLaunchKernelGGL(k_bj_inner_full, dim3((unsigned)npairs, (unsigned)nbcap), dim3(CW * 64), 0, st, nmax, w.ja,
                         w.rot_q, w.sst, w.rot);
      const bool wave1 = step_wave1(nbcap, nmax);
      const unsigned ny = (unsigned)(sweep_groups(nmax, wave1) + (nwork + PW - 1) / PW);
      for (int s = 1; s < nb - 1; ++s) {
        if (wave1)
          hipLaunchKernelGGL(k_bj_step<true>, dim3((unsigned)nbcap, ny), dim3(PW * 64), PROD_LDS, st, nmax, w.ja,
                             w.rot_q, w.sst, w.rot, s);
        else
          hipLaunchKernelGGL(k_bj_step<false>, dim3((unsigned)nbcap, ny), dim3(PW * 64), PROD_LDS, st, nmax, w.ja,
                             w.rot_q, w.sst, w.rot, s);
      }
      if (nwork > 0)  // one block pair with deferred VT: nothing to multiply
        hipLaunchKernelGGL(k_bj_product, dim3((unsigned)nbcap, (unsigned)((nwork + PW - 1) / PW)), dim3(PW * 64),
                           PROD_LDS, st, nmax, w.ja, w.rot_q, w.sst, w.rot, nb - 2);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      if (defer)
        hipLaunchKernelGGL(k_bj_apply, dim3((unsigned)(nbcap * npairs)), dim3(apply_waves(npairs) * 64), apply_lds(nmax), st, nmax,
                           nbcap, w.ja, w.rot_q, w.sst, w.rot);
    }
  }
  hipLaunchKernelGGL(k_jacobi_blocked, dim3((unsigned)nbcap), dim3(JBB), sizeof(BlockedLds), st, nmax, w.ja, w.rot_q,
                     w.rot, max_sweeps - nsw, tol, nsw > 0 ? 0 : 1, nsw > 0 ? w.sst : nullptr);
  return hipGetLastError();
}

// Recorded launch sequences of the split form, one per (device, workspace,
// nbcap, nmax, nsw, max_sweeps, tol).  A split call is ~2 nmax/16 launches
// per sweep; replaying them as one graph costs the host one launch.  The
// cache only grows (up to GRAPH_CACHE_MAX entries; beyond that new keys are
// launched directly), so no graph is destroyed while a stream may run it.
struct GraphKey {
  int dev, nbcap, nmax, nsw, max_sweeps;
  double tol;
  const double* work;
  bool operator<(const GraphKey& o) const {
    return std::tie(dev, nbcap, nmax, nsw, max_sweeps, tol, work) <
           std::tie(o.dev, o.nbcap, o.nmax, o.nsw, o.max_sweeps, o.tol, o.work);
  }
};
constexpr size_t GRAPH_CACHE_MAX = 1024;
std::mutex graph_mu;
std::map<GraphKey, hipGraphExec_t> graph_cache;

bool graphs_enabled() {
  static const bool on = [] {
    const char* e = getenv("DXTB_JACOBI_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Replays (recording on first use) the split sequence of key k on st.
// Returns hipSuccess, or an error with nothing left capturing on st.
hipError_t launch_graph(hipStream_t st, const GraphKey& k, const Layout& w) {
  std::lock_guard<std::mutex> lock(graph_mu);
  auto it = graph_cache.find(k);
  if (it == graph_cache.end()) {
    if (graph_cache.size() >= GRAPH_CACHE_MAX)
      return enqueue_blocked(st, w, k.nbcap, k.nmax, k.nsw, k.max_sweeps, k.tol);
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t el = enqueue_blocked(st, w, k.nbcap, k.nmax, k.nsw, k.max_sweeps, k.tol);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (el != hipSuccess || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return el != hipSuccess ? el : e;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    it = graph_cache.emplace(k, x).first;
  }
  return hipGraphLaunch(it->second, st);
}

}  // namespace

// Workspace of dxtb_jacobi_eigh_noalloc: one copy of this array per GPU, held
// until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is a host-side stand-in): it uses the device address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[DXTB_JACOBI_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_addr[MAX_DEVICES];

// Device address of this GPU's scratch array, looked up once per device (the
// device must be current).
hipError_t scratch_for_device(int dev, double** out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_addr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    scratch_addr[dev] = static_cast<double*>(p);
  }
  *out = scratch_addr[dev];
  return hipSuccess;
}

int jacobi_impl(int nbatch, int nmax, const int* nsize, double* A, double* VT, double* evals, double* work,
                int variant, int max_sweeps, int split_sweeps, double tol, const double* tol_b, int* sweeps,
                hipStream_t st, bool allow_graphs);

}  // namespace

extern "C" int dxtb_jacobi_max_n(void) { return JMAX; }

extern "C" size_t dxtb_jacobi_work_size(int nbatch, int nmax) {
  if (nbatch <= 0 || nmax <= 0) return 0;
  const int nbcap = batch_capacity(nbatch);
  return JARGS_BYTES + rot_bytes(nbcap, nmax) + (size_t)nbcap * 4 * sizeof(double) + flag_bytes(nbcap, nmax) +
         norm_bytes(nbcap, nmax);
}

extern "C" int dxtb_jacobi_eigh(int nbatch, int nmax, const int* nsize, double* A, double* VT, double* evals,
                                double* work, int variant, int max_sweeps, int split_sweeps, double tol,
                                const double* tol_b, int* sweeps, hipStream_t st) {
  return jacobi_impl(nbatch, nmax, nsize, A, VT, evals, work, variant, max_sweeps, split_sweeps, tol, tol_b, sweeps,
                     st, true);
}

extern "C" int dxtb_jacobi_eigh_noalloc(int nbatch, int nmax, const int* nsize, double* A, double* VT,
                                        double* evals, int variant, int max_sweeps, int split_sweeps, double tol,
                                        const double* tol_b, int* sweeps, hipStream_t st) {
  if (nbatch <= 0 || nmax <= 0 || nmax > JMAX) return DXTB_EINVAL;
  if (dxtb_jacobi_work_size(nbatch, nmax) > DXTB_JACOBI_SCRATCH_BYTES) return DXTB_ESCRATCH;
  int dev = 0;
  DXTB_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return DXTB_EINVAL;
  double* work = nullptr;
  DXTB_CHECK(scratch_for_device(dev, &work));
  // direct launches only: recording a launch sequence as a graph is not for
  // callers that must not allocate
  return jacobi_impl(nbatch, nmax, nsize, A, VT, evals, work, variant, max_sweeps, split_sweeps, tol, tol_b, sweeps,
                     st, false);
}

namespace {

int jacobi_impl(int nbatch, int nmax, const int* nsize, double* A, double* VT, double* evals, double* work,
                int variant, int max_sweeps, int split_sweeps, double tol, const double* tol_b, int* sweeps,
                hipStream_t st, bool allow_graphs) {
  if (nbatch <= 0 || nmax <= 0 || nmax > JMAX || !nsize || !A || !VT || !evals || max_sweeps < 0 || variant < 0 ||
      variant > 3 || split_sweeps < 0)
    return DXTB_EINVAL;
  const bool scalar = variant == 1;
  if (scalar && nmax > JMAX_LDS) return DXTB_EINVAL;
  if (!scalar && !work) return DXTB_EINVAL;
  int dev = 0;
  DXTB_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return DXTB_EINVAL;
  if (!attr_done[dev]) {
    DXTB_CHECK(hipFuncSetAttribute((const void*)k_jacobi_scalar, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)scalar_shmem(JMAX_LDS)));
    DXTB_CHECK(hipFuncSetAttribute((const void*)k_jacobi_blocked, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)sizeof(BlockedLds)));
    DXTB_CHECK(hipFuncSetAttribute((const void*)k_bj_product, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)PROD_LDS));
    DXTB_CHECK(hipFuncSetAttribute((const void*)k_bj_apply, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)apply_lds(DEFER_MAX)));
    attr_done[dev] = true;
  }
  if (scalar) {
    hipLaunchKernelGGL(k_jacobi_scalar, dim3((unsigned)nbatch), dim3(JB), scalar_shmem(nmax), st, nmax, nsize, A, VT,
                       evals, max_sweeps, tol, tol_b, sweeps);
    return (int)hipGetLastError();
  }
  // split form for a few large matrices: one workgroup per matrix would
  // leave most compute units idle and run the product items of a step one
  // after another
  int nsw = min(max_sweeps, split_sweeps);
  const bool split = nsw > 0 && (variant == 3 || (variant == 0 && nmax >= SPLIT_MIN_N && 2 * nbatch < nmax));
  if (!split) nsw = 0;
  const int nbcap = batch_capacity(nbatch);
  const Layout w = layout(work, nbcap, nmax);
  const JArgs args{nsize, A, VT, evals, tol_b, sweeps, nbatch, defer_call(nmax, nbcap) ? 1 : 0};
  hipLaunchKernelGGL(k_bj_args, dim3(1), dim3(64), 0, st, w.ja, args);
  DXTB_CHECK(hipGetLastError());
  // the recorded sequence needs a stream that can be captured: not the null
  // stream, and not one that the caller is capturing already
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const bool graphs = allow_graphs && split && st != nullptr && graphs_enabled();
  if (graphs) DXTB_CHECK(hipStreamIsCapturing(st, &cs));
  if (graphs && cs == hipStreamCaptureStatusNone) {
    const GraphKey k{dev, nbcap, nmax, nsw, max_sweeps, tol, work};
    return (int)launch_graph(st, k, w);
  }
  return (int)enqueue_blocked(st, w, nbcap, nmax, nsw, max_sweeps, tol);
}

}  // namespace
