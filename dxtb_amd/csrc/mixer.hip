// Broyden mixer of the SCF, one fused step per iteration on the device.
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
// One workgroup per molecule.  The history (nbatch, H, L) is read twice per
// iteration: once for the dot products <c_k, ddx>, <d_k, ddy>, <d_k, ynew>
// (one wavefront per history entry), once for the three combinations G^T
// ddx, G ddy, G ynew (one thread per vector element, all entries).

#include "common.h"
#include "../../include/dxtb_hip.h"

namespace {

constexpr int MB = 512;  // threads per workgroup
constexpr int MW = MB / 64;

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
  double g0, damp, x_tol, f_tol;
  const double* w;   // (B, L)
  const double* gx;  // (B, L) g at xe
  double* xe;        // (B, L) point evaluated; becomes the next point
  double* x;         // (B, L)
  double* y;         // (B, L)
  double* hc;        // (B, H, L)
  double* hd;        // (B, H, L)
  double* scratch;   // (B, 3, L)
  int* ctl;          // (B, 4)
  int* flag;         // (B)
  double* dxn;       // (B) norm of the step that led to xe
};

__global__ __launch_bounds__(MB) void k_broyden(MixArgs a) {
  extern __shared__ double coef[];  // 3 H: <c_k, ddx>, <d_k, ddy>, <d_k, ynew>
  __shared__ double red[MW];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int* ctl = a.ctl + 4 * b;
  const int occupied = ctl[0], started = ctl[1];
  int K = ctl[2];
  const int iters = ctl[3];
  if (!occupied) return;  // uniform over the workgroup
  const int L = a.L, H = a.H;
  const size_t base = (size_t)b * L;
  const double* __restrict__ w = a.w + base;
  const double* __restrict__ gx = a.gx + base;
  double* __restrict__ xe = a.xe + base;
  double* __restrict__ x = a.x + base;
  double* __restrict__ y = a.y + base;
  double* __restrict__ s = a.scratch + (size_t)b * 3 * L;
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
  if (K >= H) K = 0;  // history full: forget it (xitorch's restart reduction)
  // weighted ddx, ddy, ynew for the dot products; the residual norm
  double yn2 = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double wi = w[i];
    yn2 += wi * yv * yv;
    s[i] = wi * (xe[i] - x[i]);
    s[L + i] = wi * (yv - y[i]);
    s[2 * L + i] = wi * yv;
  }
  const double yn = sqrt(block_sum(yn2, red));  // (its barrier also publishes s)
  for (int k = wv; k < K; k += MW) {
    const double* __restrict__ ck = a.hc + ((size_t)b * H + k) * L;
    const double* __restrict__ dk = a.hd + ((size_t)b * H + k) * L;
    double sa = 0.0, sb = 0.0, se = 0.0;
    for (int i = lane; i < L; i += 64) {
      const double c = ck[i], d = dk[i];
      sa += c * s[i];
      sb += d * s[L + i];
      se += d * s[2 * L + i];
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
  __syncthreads();
  // G^T ddx, G ddy, G ynew; each thread keeps to its own elements from here
  double den = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double ddx = xe[i] - x[i], ddy = yv - y[i];
    double v = g0 * ddx, gdy = g0 * ddy, gy = g0 * yv;
    for (int k = 0; k < K; ++k) {
      const double c = a.hc[((size_t)b * H + k) * L + i];
      const double d = a.hd[((size_t)b * H + k) * L + i];
      v += coef[k] * d;
      gdy += coef[H + k] * c;
      gy += coef[2 * H + k] * c;
    }
    den += w[i] * ddy * v;
    s[i] = v;
    s[L + i] = gdy;
    s[2 * L + i] = gy;
  }
  den = block_sum(den, red);
  const bool ok = den != 0.0;
  double* __restrict__ cK = a.hc + ((size_t)b * H + K) * L;
  double* __restrict__ dK = a.hd + ((size_t)b * H + K) * L;
  double en = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double cn = ok ? (xe[i] - x[i]) - s[L + i] : 0.0;
    const double dn = ok ? s[i] / den : 0.0;
    cK[i] = cn;
    dK[i] = dn;
    en += w[i] * dn * yv;
  }
  en = block_sum(en, red);
  const int it = iters + 1;
  const bool conv = dxn_prev < a.x_tol && yn < a.f_tol;
  const bool stop = conv || it >= a.maxiter;
  double dx2 = 0.0;
  for (int i = tid; i < L; i += MB) {
    const double yv = xe[i] - gx[i];
    const double cn = ok ? (xe[i] - x[i]) - s[L + i] : 0.0;
    const double dx = -(s[2 * L + i] + en * cn);  // -G ynew with the new entry
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

extern "C" size_t dxtb_broyden_work_size(int nbatch, int L) {
  if (nbatch <= 0 || L <= 0) return 0;
  return (size_t)nbatch * 3 * L * sizeof(double);
}

extern "C" int dxtb_broyden_step(int nbatch, int L, int H, const double* w, const double* gx, double* xe, double* x,
                                 double* y, double* hc, double* hd, int* ctl, int* flag, double* dxn, double* work,
                                 double alpha, double damp, double x_tol, double f_tol, int maxiter,
                                 hipStream_t stream) {
  if (nbatch < 0 || L <= 0 || H <= 0 || H > 4096 || !w || !gx || !xe || !x || !y || !hc || !hd || !ctl || !flag ||
      !dxn || !work)
    return DXTB_EINVAL;
  if (nbatch == 0) return DXTB_OK;
  MixArgs a;
  a.L = L;
  a.H = H;
  a.maxiter = maxiter;
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
  a.scratch = work;
  a.ctl = ctl;
  a.flag = flag;
  a.dxn = dxn;
  hipLaunchKernelGGL(k_broyden, dim3((unsigned)nbatch), dim3(MB), (size_t)3 * H * sizeof(double), stream, a);
  return (int)hipGetLastError();
}
