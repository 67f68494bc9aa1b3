// Batched FP64 matrix products of the SCF step on the matrix cores
// (v_mfma_f64_16x16x4f64): the Fock matrix in the previous orbital basis
// C^T F C, the orbital update C V, the accumulated rotation U V and its
// Newton-Schulz re-orthogonalisation, and the densities P = C diag(f) C^T and
// W = C diag(f e) C^T.
//
// Every matrix of a batch is a padded (n, n) row-major block whose leading
// (m_b, m_b) part, m_b = nsize[b], is the molecule; the padding rows and
// columns are decoupled (block-diagonal operands), so the sum over k runs to
// m_b only and the output tiles beyond m_b are not computed (they receive the
// padding value pad * delta_ij, or are left alone).  A pool whose molecules
// are smaller than its padded order therefore pays only for their own size.
//
// Work decomposition: one 256-thread workgroup per 64x64 output tile; its
// four wavefronts own 32x32 quadrants (2x2 blocks of 16x16 matrix-core
// accumulators).  The k dimension is staged through LDS in slabs of 32, the
// next slab loaded into registers while the current one is multiplied (with
// slabs of 16 a lone order-323 product spent most of its 34-45 us in the
// barriers of 21 slabs); the output tile is staged in the same LDS after the
// last slab.
// Symmetric products compute only the tiles on and above the diagonal and
// write each also transposed, staged through LDS so that both writes are
// coalesced.  Workgroups are numbered so that the tiles of one matrix run on
// one XCD (consecutive workgroup ids go to the eight XCDs in turn): the
// panels a tile re-reads stay in that XCD's L2.

#include "common.h"
#include "../../include/dxtb_hip.h"

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));

constexpr int GT = 64;        // output tile
constexpr int GKS = 5;        // log2 of the k slab
constexpr int GK = 1 << GKS;  // k slab
constexpr int GTH = 256;      // threads per workgroup
constexpr int GLD = GT + 4;   // LDS row stride of a staged slab
constexpr int GLU = GK * GT / GTH;  // slab elements of one operand per thread
constexpr int GSM = 2 * GK * GLD > GT * (GT + 1) ? 2 * GK * GLD : GT * (GT + 1);  // LDS doubles
constexpr int NXCD = 8;       // XCDs of the MI355X

struct GemmArgs {
  int nbatch, n, tiles;  // tiles per matrix row (ceil(n / GT))
  int sym, fill_pad;
  const int* nsize;
  const double* A;
  const double* B;
  const double* scale;  // (nbatch, n) factor of the k index, or null
  const double* Cin;    // beta term, or null
  double* Out;
  double alpha, beta, pad;
};

__device__ __forceinline__ d4_t mfma(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row of element i of a lane in the 16x16 accumulator (lane l, element i:
// row l/16 + 4 i, column l % 16).
__device__ __forceinline__ int acc_row(int lane, int i) { return (lane >> 4) + 4 * i; }

__device__ __forceinline__ double mask01(bool ok) { return ok ? 1.0 : 0.0; }

// tile t (< T (T + 1) / 2) of the upper triangle, row-major -> (ty, tx), tx >= ty
__device__ __forceinline__ void upper_tile(int t, int T, int& ty, int& tx) {
  int r = 0;
  while (t >= T - r) {
    t -= T - r;
    ++r;
  }
  ty = r;
  tx = r + t;
}

template <bool TA, bool TB>
__global__ __launch_bounds__(GTH) __attribute__((amdgpu_waves_per_eu(3))) void k_gemm(GemmArgs g) {
  __shared__ double smem[GSM];
  double* const As = smem;            // op(A) slab, [k][m]
  double* const Bs = smem + GK * GLD;  // op(B) slab, [k][c]
  double* const Ot = smem;            // the output tile, after the last slab

  const int T = g.tiles;
  const int per = g.sym ? T * (T + 1) / 2 : T * T;
  const int total = g.nbatch * per;
  // XCD-aware numbering: workgroup L runs on XCD L % 8; give each XCD a
  // contiguous range of (matrix, tile) items
  int L = blockIdx.x;
  if (total % NXCD == 0) L = (L % NXCD) * (total / NXCD) + L / NXCD;
  const int b = L / per;
  const int t = L - b * per;
  int ty, tx;
  if (g.sym) {
    upper_tile(t, T, ty, tx);
  } else {
    ty = t / T;
    tx = t - ty * T;
  }
  const int n = g.n;
  const size_t nn = (size_t)n * n;
  const int nb = g.nsize ? min(max(g.nsize[b], 0), n) : n;
  const int r0 = ty * GT, c0 = tx * GT;
  double* __restrict__ O = g.Out + (size_t)b * nn;
  const bool mirror = g.sym && tx != ty;
  const int tid = threadIdx.x;

  if (r0 >= nb || c0 >= nb) {  // padding only
    if (!g.fill_pad) return;
    for (int e = tid; e < GT * GT; e += GTH) {
      const int r = e >> 6, c = e & 63;
      const int row = r0 + r, col = c0 + c;
      if (row < n && col < n) O[(size_t)row * n + col] = row == col ? g.pad : 0.0;
      if (mirror) {
        const int mr = c0 + (e >> 6), mc = r0 + (e & 63);
        if (mr < n && mc < n) O[(size_t)mr * n + mc] = 0.0;
      }
    }
    return;
  }

  const double* __restrict__ A = g.A + (size_t)b * nn;
  const double* __restrict__ Bm = g.B + (size_t)b * nn;
  const double* __restrict__ sc = g.scale ? g.scale + (size_t)b * n : nullptr;
  const int lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
  const int last = nb - 1;

  double ra[GLU], rb[GLU];
  // op(A)[m][k] and op(B)[k][c] of slab k0 into registers: masked loads from
  // clamped (valid) addresses, all in flight together
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < GLU; ++u) {
      const int idx = tid + GTH * u;
      int m, k;
      if (TA) {  // A stored [k][m]: 64 contiguous m per k
        k = idx >> 6;
        m = idx & 63;
      } else {  // A stored [m][k]: GK contiguous k per m
        m = idx >> GKS;
        k = idx & (GK - 1);
      }
      const int row = r0 + m, kk = k0 + k;
      const int rr = min(row, last), kc = min(kk, last);
      const double x = TA ? A[(size_t)kc * n + rr] : A[(size_t)rr * n + kc];
      ra[u] = x * mask01(row < nb && kk < nb);
      int c, k2;
      if (TB) {  // B stored [c][k]
        c = idx >> GKS;
        k2 = idx & (GK - 1);
      } else {  // B stored [k][c]
        k2 = idx >> 6;
        c = idx & 63;
      }
      const int col = c0 + c, kb = k0 + k2;
      const int cc = min(col, last), kbc = min(kb, last);
      double y = TB ? Bm[(size_t)cc * n + kbc] : Bm[(size_t)kbc * n + cc];
      if (sc) y *= sc[kbc];
      rb[u] = y * mask01(col < nb && kb < nb);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < GLU; ++u) {
      const int idx = tid + GTH * u;
      const int m = TA ? (idx & 63) : (idx >> GKS), k = TA ? (idx >> 6) : (idx & (GK - 1));
      As[k * GLD + m] = ra[u];
      const int c = TB ? (idx >> GKS) : (idx & 63), k2 = TB ? (idx & (GK - 1)) : (idx >> 6);
      Bs[k2 * GLD + c] = rb[u];
    }
  };

  d4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

  load(0);
  for (int k0 = 0; k0 < nb; k0 += GK) {
    __syncthreads();  // the previous slab has been read
    store();
    __syncthreads();
    if (k0 + GK < nb) load(k0 + GK);  // next slab in flight during the products
#pragma unroll
    for (int kk = 0; kk < GK / 4; ++kk) {
      const int kr = (kk * 4 + (lane >> 4)) * GLD;
      const double a0 = As[kr + wr + (lane & 15)];
      const double a1 = As[kr + wr + 16 + (lane & 15)];
      const double b0 = Bs[kr + wc + (lane & 15)];
      const double b1 = Bs[kr + wc + 16 + (lane & 15)];
      acc[0][0] = mfma(a0, b0, acc[0][0]);
      acc[0][1] = mfma(a0, b1, acc[0][1]);
      acc[1][0] = mfma(a1, b0, acc[1][0]);
      acc[1][1] = mfma(a1, b1, acc[1][1]);
    }
  }

  // tile -> LDS (over the slabs: every wavefront has read the last one), then
  // coalesced writes (and the transposed copy)
  __syncthreads();
#pragma unroll
  for (int rb_ = 0; rb_ < 2; ++rb_)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        Ot[(wr + rb_ * 16 + acc_row(lane, i)) * (GT + 1) + wc + cb * 16 + (lane & 15)] = acc[rb_][cb][i];
  __syncthreads();
  const double* __restrict__ Ci = g.Cin ? g.Cin + (size_t)b * nn : nullptr;
  // a diagonal tile of a symmetric product takes its upper triangle for both
  // halves, so the output is exactly symmetric
  const bool diag_sym = g.sym && tx == ty;
  for (int e = tid; e < GT * GT; e += GTH) {
    const int r = e >> 6, c = e & 63;
    const int row = r0 + r, col = c0 + c;
    if (row >= n || col >= n) continue;
    const size_t o = (size_t)row * n + col;
    if (row < nb && col < nb) {
      double v = g.alpha * (diag_sym && r > c ? Ot[c * (GT + 1) + r] : Ot[r * (GT + 1) + c]);
      if (Ci) v += g.beta * Ci[o];
      O[o] = v;
    } else if (g.fill_pad) {
      O[o] = row == col ? g.pad : 0.0;
    }
  }
  if (!mirror) return;
  for (int e = tid; e < GT * GT; e += GTH) {
    const int r = e & 63, c = e >> 6;  // output row c0 + c, column r0 + r
    const int row = c0 + c, col = r0 + r;
    if (row >= n || col >= n) continue;
    const size_t o = (size_t)row * n + col;
    if (row < nb && col < nb) {
      double v = g.alpha * Ot[r * (GT + 1) + c];
      if (Ci) v += g.beta * Ci[o];
      O[o] = v;
    } else if (g.fill_pad) {
      O[o] = 0.0;  // off the diagonal: a mirrored tile never holds it
    }
  }
}

}  // namespace

extern "C" int dxtb_gemm_batched(int nbatch, int n, const int* nsize, int transa, int transb, double alpha,
                                 const double* A, const double* B, const double* scale, double beta,
                                 const double* Cin, double* Out, int sym, int fill_pad, double pad,
                                 hipStream_t stream) {
  if (nbatch < 0 || n < 0 || !A || !B || !Out || (beta != 0.0 && !Cin)) return DXTB_EINVAL;
  if (Out == A || Out == B) return DXTB_EINVAL;  // the operands are read by other workgroups
  if (nbatch == 0 || n == 0) return DXTB_OK;
  GemmArgs g;
  g.nbatch = nbatch;
  g.n = n;
  g.tiles = (n + GT - 1) / GT;
  g.sym = sym ? 1 : 0;
  g.fill_pad = fill_pad ? 1 : 0;
  g.nsize = nsize;
  g.A = A;
  g.B = B;
  g.scale = scale;
  g.Cin = beta != 0.0 ? Cin : nullptr;
  g.Out = Out;
  g.alpha = alpha;
  g.beta = beta;
  g.pad = pad;
  const long per = sym ? (long)g.tiles * (g.tiles + 1) / 2 : (long)g.tiles * g.tiles;
  const long total = per * nbatch;
  if (total > 0x7fffffffL) return DXTB_EINVAL;
  const dim3 grid((unsigned)total), block(GTH);
  const int key = (transa ? 2 : 0) | (transb ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((k_gemm<false, false>), grid, block, 0, stream, g); break;
    case 1: hipLaunchKernelGGL((k_gemm<false, true>), grid, block, 0, stream, g); break;
    case 2: hipLaunchKernelGGL((k_gemm<true, false>), grid, block, 0, stream, g); break;
    default: hipLaunchKernelGGL((k_gemm<true, true>), grid, block, 0, stream, g); break;
  }
  return (int)hipGetLastError();
}
