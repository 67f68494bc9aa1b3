// Broyden mixer of the SCF, one step per iteration on the device.
//
// The SCF looks for the root of y(x) = x - g(x) on each molecule's potential
// vector with the first ("good") Broyden method of xitorch's broyden1
// (exlibs/xitorch/_impls/optimize/root/rootsolver.py:42-261, _jacobian.py:
// 51-120, 250-316: no line search, inverse Jacobian G = -alpha I +
// sum_k c_k <d_k, .>, step x <- x + damp dx with dx = -G y), per molecule,
// with the inner product <u, v> = sum_i w_i u_i v_i of the orbital-resolved
// vector (a shell's potential weighted by its 2l+1 orbitals).
//
// One call advances every molecule of a batch by one iteration, after the
// caller has evaluated g at the point xe of the previous call:
//   ynew = xe - g(xe)
//   fresh molecule (first evaluation): x = xe, y = ynew
//   otherwise: the rank-one update with ddx = xe - x, ddy = ynew - y,
//     v = G^T ddx, c = ddx - G ddy, d = v / <ddy, v> (zero if that vanishes),
//     stored at the molecule's next history slot; x = xe, y = ynew;
//     converged when |dx| < x_tol (the step that led to xe) and |ynew| < f_tol
//   then the next point xe = x - damp G y with the updated history.
// Per molecule the state is ctl = (occupied, started, history length,
// iterations) and flag = 0 running, 1 converged, 2 iteration limit; a
// molecule that stops is frozen (occupied = 0) with x at its last point.
//
// Three kernels per iteration, so that a molecule's history is read by many
// wavefronts at once (a long history is the SCF's slow tail, where one
// workgroup per molecule was latency-bound):
//   k_bro_dots   one wavefront per history entry k: <c_k, w ddx>,
//                <d_k, w ddy>, <d_k, w ynew>;
//   k_bro_comb   per chunk of 64 elements and group of CS KPT consecutive
//                entries: the partial combinations sum_k <c_k, w ddx> d_k,
//                sum_k <d_k, w ddy> c_k, sum_k <d_k, w ynew> c_k (one slice of
//                entries per wavefront, added in LDS);
//   k_bro_update one workgroup per molecule: G^T ddx, G ddy, G ynew from the
//                partials, the rank-one update, the next point and the stop
//                test.
// Workspace per molecule: the coefficients (3 H) and the partial
// combinations (KGMAX groups x 3 L), KGMAX = ceil(H / (CS * KPT)).

#include "common.h"
#include "../../include/dxtb_hip.h"

namespace {

constexpr int MB = 512;  // threads per workgroup of k_bro_update
constexpr int MW = MB / 64;
constexpr int DW = 8;    // wavefronts per workgroup of k_bro_dots
constexpr int CE = 64;   // elements per workgroup of k_bro_comb
constexpr int CS = 8;    // slices of entries per workgroup of k_bro_comb (one wavefront each)
constexpr int KPT = 8;   // entries per thread of k_bro_comb a group aims at

__host__ __device__ inline int kg_max(int H) { return (H + CS * KPT - 1) / (CS * KPT); }
__host__ __device__ inline size_t work_stride(int L, int H) { return (size_t)3 * H + (size_t)kg_max(H) * 3 * L; }

__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < MW; ++k) r += red[k];
  return r;
}

struct MixArgs {
  int L, H, maxiter;
  int kg;  // entry groups of k_bro_comb (<= kg_max(H))
  double g0, damp, x_tol, f_tol;
  const double* w;   // (B, L)
  const double* gx;  // (B, L) g at xe
  double* xe;        // (B, L) point evaluated; becomes the next point
  double* x;         // (B, L)
  double* y;         // (B, L)
  double* hc;        // (B, H, L)
  double* hd;        // (B, H, L)
  double* work;      // (B, work_stride)
  int* ctl;          // (B, 4)
  int* flag;         // (B)
  double* dxn;       // (B) norm of the step that led to xe
};

// history length of a started molecule this iteration (a full history is
// forgotten: xitorch's restart reduction); -1: the molecule is not updated
__device__ __forceinline__ int history_len(const MixArgs& a, int b) {
  const int* ctl = a.ctl + 4 * b;
  if (!ctl[0] || !ctl[1]) return -1;
  const int K = ctl[2];
  return K >= a.H ? 0 : K;
}

__global__ __launch_bounds__(DW * 64) void k_bro_dots(MixArgs a) {
  const int b = blockIdx.x;
  const int K = history_len(a, b);
  if (K <= 0) return;
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.y * DW;
  const int L = a.L, H = a.H;
  const size_t base = (size_t)b * L;
  const double* __restrict__ w = a.w + base;
  const double* __restrict__ gx = a.gx + base;
  const double* __restrict__ xe = a.xe + base;
  const double* __restrict__ x = a.x + base;
  const double* __restrict__ y = a.y + base;
  double* __restrict__ coef = a.work + (size_t)b * work_stride(L, H);
  for (int k = blockIdx.y * DW + (threadIdx.x >> 6); k < K; k += nw) {
    const double* __restrict__ ck = a.hc + ((size_t)b * H + k) * L;
    const double* __restrict__ dk = a.hd + ((size_t)b * H + k) * L;
    double sa = 0.0, sb = 0.0, se = 0.0;
    for (int i = lane; i < L; i += 64) {
      const double wi = w[i], yv = xe[i] - gx[i];
      const double d = dk[i];
      sa += ck[i] * (wi * (xe[i] - x[i]));
      sb += d * (wi * (yv - y[i]));
      se += d * (wi * yv);
    }
    sa = wave_sum(sa);
    sb = wave_sum(sb);
    se = wave_sum(se);
    if (lane == 0) {
      coef[k] = sa;
      coef[H + k] = sb;
      coef[2 * H + k] = se;
    }
  }
}

__global__ __launch_bounds__(CE * CS) void k_bro_comb(MixArgs a) {
  __shared__ double red[3][CS - 1][CE];
  const int b = blockIdx.x, kg = blockIdx.z;
  const int K = history_len(a, b);
  if (K < 0) return;
  const int e = threadIdx.x & (CE - 1), sl = threadIdx.x / CE;
  const int L = a.L, H = a.H;
  const int i = blockIdx.y * CE + e;
  const double* __restrict__ coef = a.work + (size_t)b * work_stride(L, H);
  double v = 0.0, gdy = 0.0, gy = 0.0;
  if (i < L) {
    // group kg sums the entries [kg CS KPT, (kg + 1) CS KPT) (the last group
    // also any beyond), slice sl every CS-th of them: a fixed summation order
    // for every entry whatever the number of groups, so a result does not
    // depend on the caller's history bound (extra groups add zeros)
    const int k1 = (kg == a.kg - 1) ? K : min(K, (kg + 1) * CS * KPT);
    for (int k = kg * CS * KPT + sl; k < k1; k += CS) {
      const double c = a.hc[((size_t)b * H + k) * L + i];
      const double d = a.hd[((size_t)b * H + k) * L + i];
      v += coef[k] * d;
      gdy += coef[H + k] * c;
      gy += coef[2 * H + k] * c;
    }
  }
  if (sl > 0) {
    red[0][sl - 1][e] = v;
    red[1][sl - 1][e] = gdy;
    red[2][sl - 1][e] = gy;
  }
  __syncthreads();
  if (sl > 0 || i >= L) return;
  for (int q = 0; q < CS - 1; ++q) {
    v += red[0][q][e];
    gdy += red[1][q][e];
    gy += red[2][q][e];
  }
  double* __restrict__ part = a.work + (size_t)b * work_stride(L, H) + 3 * (size_t)H + (size_t)kg * 3 * L;
  part[i] = v;
  part[L + i] = gdy;
  part[2 * L + i] = gy;
}

__global__ __launch_bounds__(MB) void k_bro_update(MixArgs a) {
  __shared__ double red[MW];
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  int* ctl = a.ctl + 4 * b;
  const int occupied = ctl[0], started = ctl[1];
  const int iters = ctl[3];
  if (!occupied) return;  // uniform over the workgroup
  const int L = a.L, H = a.H;
  const size_t base = (size_t)b * L;
  const double* __restrict__ w = a.w + base;
  const double* __restrict__ gx = a.gx + base;
  double* __restrict__ xe = a.xe + base;
  double* __restrict__ x = a.x + base;
  double* __restrict__ y = a.y + base;
  const double g0 = a.g0;
  const double dxn_prev = a.dxn[b];

  if (!started) {  // first evaluation: no update, the step from the initial residual
    double y2 = 0.0;
    for (int i = tid; i < L; i += MB) {
      const double yv = xe[i] - gx[i];
      const double dx = -g0 * yv;
      x[i] = xe[i];
      y[i] = yv;
      xe[i] = xe[i] + a.damp * dx;
      y2 += w[i] * dx * dx;
    }
    const double n2 = block_sum(y2, red);
    if (tid == 0) {
      ctl[1] = 1;
      a.dxn[b] = sqrt(n2);
    }
    return;
  }
  const int K = history_len(a, b);
  double* __restrict__ part = a.work + (size_t)b * work_stride(L, H) + 3 * (size_t)H;
  // G^T ddx, G ddy, G ynew (into the first group's partials), the residual
  // norm and the sums for the update's denominator and step
  double yn2 = 0.0, den = 0.0, evy = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double ddx = xe[i] - x[i], ddy = yv - y[i];
    double v = g0 * ddx, gdy = g0 * ddy, gy = g0 * yv;
    if (K > 0) {
      for (int g = 0; g < a.kg; ++g) {
        const double* pg = part + (size_t)g * 3 * L;
        v += pg[i];
        gdy += pg[L + i];
        gy += pg[2 * L + i];
      }
    }
    const double wi = w[i];
    yn2 += wi * yv * yv;
    den += wi * ddy * v;
    evy += wi * v * yv;
    part[i] = v;
    part[L + i] = gdy;
    part[2 * L + i] = gy;
  }
  const double yn = sqrt(block_sum(yn2, red));
  den = block_sum(den, red);
  evy = block_sum(evy, red);
  const bool ok = den != 0.0;
  const double en = ok ? evy / den : 0.0;  // <d_new, ynew> of the new entry
  double* __restrict__ cK = a.hc + ((size_t)b * H + K) * L;
  double* __restrict__ dK = a.hd + ((size_t)b * H + K) * L;
  const int it = iters + 1;
  const bool conv = dxn_prev < a.x_tol && yn < a.f_tol;
  const bool stop = conv || it >= a.maxiter;
  double dx2 = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double cn = ok ? (xe[i] - x[i]) - part[L + i] : 0.0;
    cK[i] = cn;
    dK[i] = ok ? part[i] / den : 0.0;
    const double dx = -(part[2 * L + i] + en * cn);  // -G ynew with the new entry
    x[i] = xe[i];
    y[i] = yv;
    if (!stop) xe[i] = xe[i] + a.damp * dx;
    dx2 += w[i] * dx * dx;
  }
  dx2 = block_sum(dx2, red);
  if (tid == 0) {
    ctl[2] = K + 1;
    ctl[3] = it;
    a.dxn[b] = sqrt(dx2);
    if (stop) {
      ctl[0] = 0;
      a.flag[b] = conv ? 1 : 2;
    }
  }
}

}  // namespace

extern "C" size_t dxtb_broyden_work_size(int nbatch, int L, int H) {
  if (nbatch <= 0 || L <= 0 || H <= 0) return 0;
  return (size_t)nbatch * work_stride(L, H) * sizeof(double);
}

extern "C" int dxtb_broyden_step(int nbatch, int L, int H, const double* w, const double* gx, double* xe, double* x,
                                 double* y, double* hc, double* hd, int* ctl, int* flag, double* dxn, double* work,
                                 double alpha, double damp, double x_tol, double f_tol, int maxiter, int kbound,
                                 hipStream_t stream) {
  if (nbatch < 0 || nbatch > 65535 || L <= 0 || H <= 0 || H > 4096 || !w || !gx || !xe || !x || !y || !hc ||
      !hd || !ctl || !flag || !dxn || !work)
    return DXTB_EINVAL;
  if (nbatch == 0) return DXTB_OK;
  const int kb = (kbound <= 0 || kbound > H) ? H : kbound;  // sizes the grids only
  MixArgs a;
  a.L = L;
  a.H = H;
  a.maxiter = maxiter;
  a.kg = min(kg_max(H), (kb + CS * KPT - 1) / (CS * KPT));
  a.g0 = -alpha;
  a.damp = damp;
  a.x_tol = x_tol;
  a.f_tol = f_tol;
  a.w = w;
  a.gx = gx;
  a.xe = xe;
  a.x = x;
  a.y = y;
  a.hc = hc;
  a.hd = hd;
  a.work = work;
  a.ctl = ctl;
  a.flag = flag;
  a.dxn = dxn;
  const unsigned B = (unsigned)nbatch;
  hipLaunchKernelGGL(k_bro_dots, dim3(B, (unsigned)((kb + DW - 1) / DW)), dim3(DW * 64), 0, stream, a);
  DXTB_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_bro_comb, dim3(B, (unsigned)((L + CE - 1) / CE), (unsigned)a.kg), dim3(CE * CS), 0, stream,
                     a);
  DXTB_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_bro_update, dim3(B), dim3(MB), 0, stream, a);
  return (int)hipGetLastError();
}
