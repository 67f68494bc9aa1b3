// SCF step kernels: Fock build, Mulliken charges and atomic multipoles, Fermi
// smearing.  FP64, streaming over the padded (nbatch, nao_max, nao_max)
// matrices with one thread per element (Fock) or per AO column and chunk of
// rows (charges).

#include "common.h"
#include "pairfuncs.h"
#include "../../include/dxtb_hip.h"

namespace {

inline dim3 grid1(long n, int blk) { return dim3((unsigned)((n + blk - 1) / blk)); }

// F_ij = H0_ij - S_ij (v_i + v_j)/2 - (D^j_ij.vd_j + D^i_ij.vd_i)/2 - (Q^j_ij.vq_j + Q^i_ij.vq_i)/2
// (scf/pure/conversions.py:255-297).  The stored D/Q element (i, j) is
// centred on the ket atom j; the (j, i) element, centred on atom i, follows
// from the same element by the shift identities with delta = R_j - R_i:
//   D^i = D^j + delta S
//   Q^i = Q^j + 3/2 (delta_a D^j_b + delta_b D^j_a + delta_a delta_b S)
//             - (delta.D^j + |delta|^2 S / 2) [a == b]        (traceless form)
// so every thread reads only its own element of each matrix (coalesced).
template <bool MP>
__global__ __launch_bounds__(256) void k_fock(int B, int n, const int* __restrict__ nao,
                                              const int* __restrict__ ao_shell, const int* __restrict__ ao_atom,
                                              int nsh_max, int nat_max, const double* __restrict__ pos,
                                              const double* __restrict__ H0, const double* __restrict__ S,
                                              const double* __restrict__ D, const double* __restrict__ Q,
                                              const double* __restrict__ vsh, const double* __restrict__ vat,
                                              const double* __restrict__ vdp, const double* __restrict__ vqp,
                                              double pad_diag, double* __restrict__ F) {
  const size_t nn = (size_t)n * n;
  const int b = blockIdx.y;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int i = (int)(e / n), j = (int)(e - (size_t)i * n);
  const size_t off = (size_t)b * nn + e;
  const int nb = nao[b];
  if (i >= nb || j >= nb) {
    F[off] = (i == j) ? pad_diag : 0.0;
    return;
  }
  const int si = ao_shell[b * n + i], sj = ao_shell[b * n + j];
  const int ai = ao_atom[b * n + i], aj = ao_atom[b * n + j];
  const double vi = vsh[b * nsh_max + si] + vat[b * nat_max + ai];
  const double vj = vsh[b * nsh_max + sj] + vat[b * nat_max + aj];
  const double s = S[off];
  double f = H0[off] - 0.5 * s * (vi + vj);
  if constexpr (MP) {
    const double* pi = pos + ((size_t)b * nat_max + ai) * 3;
    const double* pj = pos + ((size_t)b * nat_max + aj) * 3;
    const double dl[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
    const size_t db = (size_t)b * 3 * nn + e;
    const double d[3] = {D[db], D[db + nn], D[db + 2 * nn]};
    const double* vdi = vdp + ((size_t)b * nat_max + ai) * 3;
    const double* vdj = vdp + ((size_t)b * nat_max + aj) * 3;
    double td = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) td += d[k] * vdj[k] + (d[k] + dl[k] * s) * vdi[k];
    const size_t qb = (size_t)b * 6 * nn + e;
    double q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = Q[qb + k * nn];
    const double dd = dl[0] * d[0] + dl[1] * d[1] + dl[2] * d[2];
    const double d2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
    const double tr = dd + 0.5 * d2 * s;
    // component order xx, yx, yy, zx, zy, zz -> (a, b) axes
    const int ca[6] = {0, 1, 1, 2, 2, 2}, cb[6] = {0, 0, 1, 0, 1, 2};
    const double* vqi = vqp + ((size_t)b * nat_max + ai) * 6;
    const double* vqj = vqp + ((size_t)b * nat_max + aj) * 6;
    double tq = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int x = ca[k], y = cb[k];
      double qi = q[k] + 1.5 * (dl[x] * d[y] + dl[y] * d[x] + dl[x] * dl[y] * s);
      if (x == y) qi -= tr;
      tq += q[k] * vqj[k] + qi * vqi[k];
    }
    f -= 0.5 * (td + tq);
  }
  F[off] = f;
}

// Per AO column j: population sum_k P_kj S_kj, band energy sum_k P_kj H0_kj
// and the AO contributions to the atomic multipoles -sum_k P_kj M_kj (the
// column index is the ket AO, i.e. the atom the moment is centred on).  A
// workgroup takes CS_COLS consecutive columns (consecutive lanes, coalesced
// rows) and splits the rows of the molecule into CS_CHUNKS chunks, one
// wavefront each; the chunks' sums are added in LDS.  A few molecules then
// still spread over the device, and a lone molecule's row loop is CS_CHUNKS
// times shorter.
constexpr int CS_COLS = 64;
constexpr int CS_CHUNKS = 8;

template <bool MP>
__global__ __launch_bounds__(CS_COLS * CS_CHUNKS) void k_colsums(int n, const int* __restrict__ nao,
                                                                 const double* __restrict__ P,
                                                                 const double* __restrict__ S,
                                                                 const double* __restrict__ H0,
                                                                 const double* __restrict__ D,
                                                                 const double* __restrict__ Q,
                                                                 double* __restrict__ out) {
  constexpr int NO = MP ? 11 : 2;
  __shared__ double part[CS_CHUNKS - 1][NO][CS_COLS];
  const int b = blockIdx.y;
  const int c = threadIdx.x & (CS_COLS - 1), r = threadIdx.x / CS_COLS;
  const int j = blockIdx.x * CS_COLS + c;
  const int nb = nao[b];
  double acc[NO];
#pragma unroll
  for (int t = 0; t < NO; ++t) acc[t] = 0.0;
  const size_t nn = (size_t)n * n;
  const int rows = (nb + CS_CHUNKS - 1) / CS_CHUNKS;
  const int k0 = r * rows, k1 = min(nb, k0 + rows);
  if (j < nb) {
    const double* Pb = P + (size_t)b * nn;
    const double* Sb = S + (size_t)b * nn;
    const double* Hb = H0 + (size_t)b * nn;
    for (int k = k0; k < k1; ++k) {
      const size_t e = (size_t)k * n + j;
      const double p = Pb[e];
      acc[0] += p * Sb[e];
      acc[1] += p * Hb[e];
      if constexpr (MP) {
        const double* Db = D + (size_t)b * 3 * nn + e;
        const double* Qb = Q + (size_t)b * 6 * nn + e;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) acc[2 + cc] -= p * Db[cc * nn];
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) acc[5 + cc] -= p * Qb[cc * nn];
      }
    }
  }
  if (r > 0) {
#pragma unroll
    for (int t = 0; t < NO; ++t) part[r - 1][t][c] = acc[t];
  }
  __syncthreads();
  if (r > 0 || j >= n) return;
  for (int q = 0; q < CS_CHUNKS - 1; ++q) {
#pragma unroll
    for (int t = 0; t < NO; ++t) acc[t] += part[q][t][c];
  }
  double* o = out + ((size_t)b * n + j) * NO;
#pragma unroll
  for (int t = 0; t < NO; ++t) o[t] = acc[t];
}

// Reduce AO column sums to shells and atoms.  One thread per atom.
template <bool MP>
__global__ void k_reduce_atoms(dxtb_basis_t bs, const int* __restrict__ nat, const int* __restrict__ at_sh0,
                               const int* __restrict__ at_nsh, const double* __restrict__ n0,
                               const double* __restrict__ col, double* __restrict__ qsh,
                               double* __restrict__ qat, double* __restrict__ qdp, double* __restrict__ qqp,
                               double* __restrict__ eh0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bs.nbatch * bs.nat_max) return;
  const int b = t / bs.nat_max, a = t - b * bs.nat_max;
  constexpr int NO = MP ? 11 : 2;
  double q = 0.0, eh = 0.0, dp[3] = {0, 0, 0}, qp[6] = {0, 0, 0, 0, 0, 0};
  if (a < nat[b]) {
    const int s0 = at_sh0[t], ns = at_nsh[t];
    const int n = bs.nao_max;
    for (int s = s0; s < s0 + ns; ++s) {
      const int ao0 = bs.sh_ao[b * bs.nsh_max + s];
      const int l = bs.ush_l[bs.sh_ush[b * bs.nsh_max + s]];
      double qs = 0.0;
      for (int m = 0; m < 2 * l + 1; ++m) {
        const int mu = ao0 + m;
        const double* c = col + ((size_t)b * n + mu) * NO;
        qs += n0[(size_t)b * n + mu] - c[0];
        eh += c[1];
        if constexpr (MP) {
          for (int k = 0; k < 3; ++k) dp[k] += c[2 + k];
          for (int k = 0; k < 6; ++k) qp[k] += c[5 + k];
        }
      }
      qsh[b * bs.nsh_max + s] = qs;
      q += qs;
    }
  }
  qat[t] = q;
  eh0[t] = eh;
  if constexpr (MP) {
    for (int k = 0; k < 3; ++k) qdp[(size_t)t * 3 + k] = dp[k];
    for (int k = 0; k < 6; ++k) qqp[(size_t)t * 6 + k] = qp[k];
  }
}

// ---- Fermi smearing ---------------------------------------------------------
// One wavefront per molecule: the sums over the orbitals are wave reductions
// (no workgroup barriers in the root search).
constexpr int FB = 64;

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double fermi_fn(double e, double mu, double kt) {
  const double x = (e - mu) / kt;
  return x > 0.0 ? exp(-x) / (1.0 + exp(-x)) : 1.0 / (1.0 + exp(x));
}

// For every spin channel the chemical potential is the root of
// N(mu) = sum_i f_i(mu) - n (monotonic): bracketed bisection until the
// bracket is within a few kT, then safeguarded Newton steps to machine
// precision (the converged result of filling.py:428-742).  Channels with the
// same electron count (closed shells) share one search.
__global__ __launch_bounds__(FB) void k_fermi(int n, const int* __restrict__ nao, const double* __restrict__ emo,
                                              const double* __restrict__ nab, double kt,
                                              double* __restrict__ occ, double* __restrict__ ts) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int nb = nao[b];
  const double* e = emo + (size_t)b * n;
  // bracket from the extreme levels (the eigensolver need not sort them)
  double emin = 1e300, emax = -1e300;
  for (int i = lane; i < nb; i += FB) {
    emin = fmin(emin, e[i]);
    emax = fmax(emax, e[i]);
  }
  emin = -wave_max(-emin);
  emax = wave_max(emax);
  const bool same = nab[b * 2] == nab[b * 2 + 1];
  double ent = 0.0;
  for (int spin = 0; spin < (same ? 1 : 2); ++spin) {
    const double target = nab[b * 2 + spin];
    double* o = occ + ((size_t)b * 2 + spin) * n;
    double* o2 = same ? occ + ((size_t)b * 2 + 1) * n : o;
    if (target <= 1e-12 || nb == 0) {
      for (int i = lane; i < n; i += FB) {
        o[i] = 0.0;
        o2[i] = 0.0;
      }
      continue;
    }
    double lo = emin - 1.0, hi = emax + 1.0;
    double mu = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
      double sm = 0.0, ds = 0.0;
      for (int i = lane; i < nb; i += FB) {
        const double f = fermi_fn(e[i], mu, kt);
        sm += f;
        ds += f * (1.0 - f);
      }
      sm = wave_sum(sm) - target;
      ds = wave_sum(ds) / kt;
      if (sm > 0.0) hi = mu; else lo = mu;
      if (fabs(sm) < 1e-13 * (1.0 + target)) break;
      double nm = mu - sm / ds;
      if (!(ds > 0.0) || !(nm > lo && nm < hi) || (hi - lo) > 50.0 * kt) nm = 0.5 * (lo + hi);
      if (nm == mu) break;
      mu = nm;
    }
    double es = 0.0;
    for (int i = lane; i < n; i += FB) {
      double f = 0.0;
      if (i < nb) {
        f = fermi_fn(e[i], mu, kt);
        // -TS: f ln f + (1 - f) ln(1 - f) with clamps at eps (scf/base.py:554-619)
        const double a = fmax(f, dxtb::EPS), c = fmax(1.0 - f, dxtb::EPS);
        es += a * log(a) + c * log(c);
      }
      o[i] = f;
      o2[i] = f;
    }
    ent += same ? 2.0 * es : es;
  }
  ent = wave_sum(ent);
  if (lane == 0) ts[b] = ent * kt;
}

}  // namespace

extern "C" int dxtb_fock(const dxtb_basis_t* bas, const int* nao, const int* ao_shell, const int* ao_atom,
                         const double* pos, const double* H0, const double* S, const double* D, const double* Q,
                         const double* vsh, const double* vat, const double* vdp, const double* vqp,
                         double pad_diag, double* F, hipStream_t st) {
  if (!bas || !nao || !ao_shell || !ao_atom || !pos || !H0 || !S || !vsh || !vat || !F) return DXTB_EINVAL;
  const bool mp = D != nullptr;
  if (mp && (!Q || !vdp || !vqp)) return DXTB_EINVAL;
  const int n = bas->nao_max;
  const dim3 grid((unsigned)(((size_t)n * n + 255) / 256), (unsigned)bas->nbatch);
  if (mp)
    hipLaunchKernelGGL(k_fock<true>, grid, dim3(256), 0, st, bas->nbatch, n, nao, ao_shell, ao_atom,
                       bas->nsh_max, bas->nat_max, pos, H0, S, D, Q, vsh, vat, vdp, vqp, pad_diag, F);
  else
    hipLaunchKernelGGL(k_fock<false>, grid, dim3(256), 0, st, bas->nbatch, n, nao, ao_shell, ao_atom,
                       bas->nsh_max, bas->nat_max, pos, H0, S, D, Q, vsh, vat, vdp, vqp, pad_diag, F);
  return (int)hipGetLastError();
}

extern "C" size_t dxtb_charges_work_size(int nbatch, int nao_max) {
  if (nbatch <= 0 || nao_max <= 0) return 0;
  return (size_t)nbatch * nao_max * 11 * sizeof(double);
}

extern "C" int dxtb_charges(const dxtb_basis_t* bas, const int* nao, const int* nat, const int* at_sh0,
                            const int* at_nsh, const double* n0, const double* P, const double* S,
                            const double* H0, const double* D, const double* Q, double* qsh, double* qat,
                            double* qdp, double* qqp, double* eh0, double* work, hipStream_t st) {
  if (!bas || !nao || !nat || !at_sh0 || !at_nsh || !n0 || !P || !S || !H0 || !qsh || !qat || !eh0 || !work)
    return DXTB_EINVAL;
  const bool mp = D != nullptr;
  if (mp && (!Q || !qdp || !qqp)) return DXTB_EINVAL;
  if (bas->nbatch <= 0) return DXTB_OK;
  if (bas->nbatch > 65535) return DXTB_EINVAL;  // grid y
  const int n = bas->nao_max;
  const dim3 g((unsigned)((n + CS_COLS - 1) / CS_COLS), (unsigned)bas->nbatch);
  const dim3 blk(CS_COLS * CS_CHUNKS);
  const dim3 ga = grid1((long)bas->nbatch * bas->nat_max, 128);
  if (mp) {
    hipLaunchKernelGGL(k_colsums<true>, g, blk, 0, st, n, nao, P, S, H0, D, Q, work);
    DXTB_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_atoms<true>, ga, dim3(128), 0, st, *bas, nat, at_sh0, at_nsh, n0, work, qsh, qat,
                       qdp, qqp, eh0);
  } else {
    hipLaunchKernelGGL(k_colsums<false>, g, blk, 0, st, n, nao, P, S, H0, D, Q, work);
    DXTB_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_atoms<false>, ga, dim3(128), 0, st, *bas, nat, at_sh0, at_nsh, n0, work, qsh,
                       qat, qdp, qqp, eh0);
  }
  return (int)hipGetLastError();
}

extern "C" int dxtb_fermi(int nbatch, int nao_max, const int* nao, const double* emo, const double* nab,
                          double kt, double* occ, double* ts, hipStream_t st) {
  if (nbatch <= 0 || nao_max <= 0 || !nao || !emo || !nab || !occ || !ts || !(kt > 0.0)) return DXTB_EINVAL;
  hipLaunchKernelGGL(k_fermi, dim3((unsigned)nbatch), dim3(FB), 0, st, nao_max, nao, emo, nab, kt, occ, ts);
  return (int)hipGetLastError();
}
