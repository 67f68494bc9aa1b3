// Nuclear gradient of the converged SCF energy, all terms except the
// integral derivatives (those are in multipole.hip):
//   * H0 distance polynomial and self-energy (CN) terms  sum P S df/dR, dE/dh
//   * ES2 gamma(R)                                        1/2 q dgamma/dR q
//   * AES2 interaction matrices and multipole radii       (GFN2)
//   * repulsion
//   * coordination-number chain dE/dCN * dCN/dR
// At self-consistency the free energy is stationary in the density, so these
// are explicit derivatives at fixed charges/density (the structure of the
// analytic GFN1 gradient, xtb/gfn1.py:185-408, and of the ES2 gradient,
// coulomb/secondorder.py:445-690).  A group of lanes per shell or atom
// splits the partner loop over the molecule's own shells/atoms.

#include "common.h"
#include "pairfuncs.h"
#include "../../include/dxtb_hip.h"

using namespace dxtb;

namespace {

inline dim3 grid1(long n, int blk) { return dim3((unsigned)((n + blk - 1) / blk)); }

// Each shell (atom) is served by a group of FG lanes that split its partner
// loop; the partial sums are combined by shuffles and lane 0 of the group
// writes (and adds the gradient atomically).
constexpr int FG = 16;
constexpr int FBLK = 256;

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = FG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// per shell A: ps_AB = sum_{mu in A, nu in B} P S ; gradient 2 sum_B ps dfsh unit, dE/dh_A = 2 sum_B ps hfsh
__global__ __launch_bounds__(FBLK) void k_h0_grad(dxtb_basis_t bs, const int* __restrict__ nsh,
                                                  const double* __restrict__ pos, const double* __restrict__ P,
                                                  const double* __restrict__ S, const double* __restrict__ dfsh,
                                                  const double* __restrict__ hfsh, double* __restrict__ dEdh,
                                                  double* __restrict__ grad) {
  const int ns = bs.nsh_max;
  const long tl = ((long)blockIdx.x * FBLK + threadIdx.x) / FG;
  const int lg = threadIdx.x & (FG - 1);
  const bool live = tl < (long)bs.nbatch * ns;
  const int t = live ? (int)tl : 0;
  const int b = t / ns, sa = t - b * ns;
  const int nb = live ? nsh[b] : 0;
  const bool act = sa < nb;  // uniform over the group
  double g[3] = {0, 0, 0}, dh = 0.0;
  int aa = 0;
  if (act) {
    const int n = bs.nao_max;
    const size_t nn = (size_t)n * n;
    const int base = b * ns;
    aa = bs.sh_atom[base + sa];
    const int la = bs.ush_l[bs.sh_ush[base + sa]];
    const int r0 = bs.sh_ao[base + sa];
    const double* pa = pos + ((size_t)b * bs.nat_max + aa) * 3;
    for (int sb = lg; sb < nb; sb += FG) {
      const int ab = bs.sh_atom[base + sb];
      const int lb = bs.ush_l[bs.sh_ush[base + sb]];
      const int c0 = bs.sh_ao[base + sb];
      double ps = 0.0;
      for (int i = 0; i < 2 * la + 1; ++i) {
        const size_t row = (size_t)b * nn + (size_t)(r0 + i) * n + c0;
        for (int j = 0; j < 2 * lb + 1; ++j) ps += P[row + j] * S[row + j];
      }
      const size_t pr = ((size_t)b * ns + sa) * ns + sb;
      dh += 2.0 * ps * hfsh[pr];
      if (ab != aa) {
        const double* pb = pos + ((size_t)b * bs.nat_max + ab) * 3;
        const double d[3] = {pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]};
        const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double w = 2.0 * ps * dfsh[pr] / R;
        for (int k = 0; k < 3; ++k) g[k] += w * d[k];
      }
    }
  }
  dh = group_sum(dh);
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = group_sum(g[k]);
  if (!live || lg != 0) return;
  dEdh[t] = dh;  // 0 for padding shells
  if (!act) return;
  double* ga = grad + ((size_t)b * bs.nat_max + aa) * 3;
  for (int k = 0; k < 3; ++k) atomic_add_f64(ga + k, g[k]);
}

// per atom: ES2, AES2, repulsion gradients and dE/dCN (H0 self-energies + AES2 radii)
template <bool G2>
__global__ __launch_bounds__(FBLK) void k_atom_grad(dxtb_basis_t bs, dxtb_params_t par,
                                                    const int* __restrict__ numbers, const double* __restrict__ pos,
                                                    const int* __restrict__ at_sh0, const int* __restrict__ at_nsh,
                                                    const int* __restrict__ nsh, const double* __restrict__ dEdh,
                                                    const double* __restrict__ cn, const double* __restrict__ mrad,
                                                    const double* __restrict__ qsh, const double* __restrict__ qat,
                                                    const double* __restrict__ qdp, const double* __restrict__ qqp,
                                                    double* __restrict__ dEdcn, double* __restrict__ grad) {
  const int na = bs.nat_max, ns = bs.nsh_max;
  const long tl = ((long)blockIdx.x * FBLK + threadIdx.x) / FG;
  const int lg = threadIdx.x & (FG - 1);
  const bool live = tl < (long)bs.nbatch * na;
  const int t = live ? (int)tl : 0;
  const int b = t / na, i = t - b * na;
  const int zi = live ? numbers[t] : 0;  // uniform over the group
  double g[3] = {0, 0, 0}, dEdm = 0.0;
  if (zi != 0) {
    const double* pi = pos + (size_t)t * 3;
    const int s0 = at_sh0[t], sn = at_nsh[t];
    // ES2: sum_{s in i} sum_{u not in i} q_s q_u dgamma/dR unit
    const int nb = nsh[b];
    for (int s = s0; s < s0 + sn; ++s) {
      const double qs = qsh[b * ns + s];
      const double hs = par.ush_hub[bs.sh_ush[b * ns + s]];
      for (int u = lg; u < nb; u += FG) {
        const int au = bs.sh_atom[b * ns + u];
        if (au == i) continue;
        const double* pu = pos + ((size_t)b * na + au) * 3;
        const double d[3] = {pi[0] - pu[0], pi[1] - pu[1], pi[2] - pu[2]};
        const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double gm, dg;
        es2_gamma_t<G2>(false, R, es2_avg(par.es2_average, hs, par.ush_hub[bs.sh_ush[b * ns + u]]), par.es2_gexp,
                        gm, dg);
        const double w = qs * qsh[b * ns + u] * dg / R;
        for (int k = 0; k < 3; ++k) g[k] += w * d[k];
      }
    }
    // repulsion and AES2 over partner atoms
    for (int j = lg; j < na; j += FG) {
      const int zj = numbers[b * na + j];
      if (zj == 0 || j == i) continue;
      const size_t tj = (size_t)b * na + j;
      const double* pj = pos + tj * 3;
      const double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
      const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (R <= REP_CUTOFF) {
        const double a = sqrt(par.z_arep[zi] * par.z_arep[zj] + TINY);
        const double k = (par.rep_light && zi <= 2 && zj <= 2) ? par.rep_klight : par.rep_kexp;
        double e, de;
        rep_pair(R, a, k, par.z_zeff[zi] * par.z_zeff[zj], e, de);
        for (int c = 0; c < 3; ++c) g[c] += de * d[c] / R;
      }
      if (par.gfn == 2) {
        double E, ga[3], dm;
        const double mui[3] = {qdp[(size_t)t * 3], qdp[(size_t)t * 3 + 1], qdp[(size_t)t * 3 + 2]};
        const double muj[3] = {qdp[tj * 3], qdp[tj * 3 + 1], qdp[tj * 3 + 2]};
        double thi[6], thj[6];
        for (int c = 0; c < 6; ++c) {
          thi[c] = qqp[(size_t)t * 6 + c];
          thj[c] = qqp[tj * 6 + c];
        }
        aes2_pair(d, mrad[t], mrad[tj], qat[t], qat[tj], mui, muj, thi, thj, par.aes_dmp3, par.aes_dmp5, E, ga,
                  dm);
        for (int c = 0; c < 3; ++c) g[c] += ga[c];
        dEdm += dm;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = group_sum(g[k]);
  dEdm = group_sum(dEdm);
  if (!live || lg != 0) return;
  if (zi == 0) {
    dEdcn[t] = 0.0;
    return;
  }
  // dE/dCN from the shell self-energies h = lev - kcn CN
  double dcn = 0.0;
  const int s0 = at_sh0[t], sn = at_nsh[t];
  for (int s = s0; s < s0 + sn; ++s) dcn -= par.ush_kcn[bs.sh_ush[b * ns + s]] * dEdh[b * ns + s];
  if (par.gfn == 2)
    dcn += dEdm * dmrad_dcn(cn[t], par.z_mprad[zi], par.z_mpvcn[zi], par.aes_kexp, par.aes_shift, par.aes_rmax);
  dEdcn[t] = dcn;
  double* gi = grad + (size_t)t * 3;
  for (int k = 0; k < 3; ++k) atomic_add_f64(gi + k, g[k]);
}

// CN chain: dE/dR_i = sum_j (dE/dCN_i + dE/dCN_j) dcount_ij/dR unit_ij
__global__ __launch_bounds__(FBLK) void k_cn_grad(int B, int na, int gfn, const int* __restrict__ numbers,
                                                  const double* __restrict__ pos, const double* __restrict__ rcov,
                                                  const double* __restrict__ dEdcn, double* __restrict__ grad) {
  const long tl = ((long)blockIdx.x * FBLK + threadIdx.x) / FG;
  const int lg = threadIdx.x & (FG - 1);
  const bool live = tl < (long)B * na;
  const int t = live ? (int)tl : 0;
  const int b = t / na, i = t - b * na;
  const int zi = live ? numbers[t] : 0;
  double g[3] = {0, 0, 0};
  if (zi != 0) {
    const double* pi = pos + (size_t)t * 3;
    for (int j = lg; j < na; j += FG) {
      const int zj = numbers[b * na + j];
      if (zj == 0 || j == i) continue;
      const double* pj = pos + ((size_t)b * na + j) * 3;
      const double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
      const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (R > CN_CUTOFF) continue;
      double c, dc;
      cn_count(gfn, R, rcov[zi] + rcov[zj], c, dc);
      const double w = (dEdcn[t] + dEdcn[b * na + j]) * dc / R;
      for (int k = 0; k < 3; ++k) g[k] += w * d[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = group_sum(g[k]);
  if (!live || lg != 0 || zi == 0) return;
  double* gi = grad + (size_t)t * 3;
  for (int k = 0; k < 3; ++k) atomic_add_f64(gi + k, g[k]);
}

}  // namespace

extern "C" int dxtb_forces_classical(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers,
                                     const double* pos, const int* nsh, const int* at_sh0, const int* at_nsh,
                                     const double* P, const double* S, const double* cn, const double* mrad,
                                     const double* dfsh, const double* hfsh, const double* qsh, const double* qat,
                                     const double* qdp, const double* qqp, double* dEdh, double* dEdcn,
                                     double* grad, hipStream_t st) {
  if (!bas || !par || !numbers || !pos || !nsh || !at_sh0 || !at_nsh || !P || !S || !cn || !dfsh || !hfsh ||
      !qsh || !qat || !dEdh || !dEdcn || !grad)
    return DXTB_EINVAL;
  if (par->gfn == 2 && (!mrad || !qdp || !qqp)) return DXTB_EINVAL;
  const dxtb_basis_t& b = *bas;
  hipLaunchKernelGGL(k_h0_grad, grid1((long)b.nbatch * b.nsh_max * FG, FBLK), dim3(FBLK), 0, st, b, nsh, pos, P,
                     S, dfsh, hfsh, dEdh, grad);
  DXTB_CHECK(hipGetLastError());
  if (par->es2_gexp == 2.0)
    hipLaunchKernelGGL(k_atom_grad<true>, grid1((long)b.nbatch * b.nat_max * FG, FBLK), dim3(FBLK), 0, st, b, *par,
                       numbers, pos, at_sh0, at_nsh, nsh, dEdh, cn, mrad, qsh, qat, qdp, qqp, dEdcn, grad);
  else
    hipLaunchKernelGGL(k_atom_grad<false>, grid1((long)b.nbatch * b.nat_max * FG, FBLK), dim3(FBLK), 0, st, b, *par,
                       numbers, pos, at_sh0, at_nsh, nsh, dEdh, cn, mrad, qsh, qat, qdp, qqp, dEdcn, grad);
  DXTB_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_cn_grad, grid1((long)b.nbatch * b.nat_max * FG, FBLK), dim3(FBLK), 0, st, b.nbatch,
                     b.nat_max, par->gfn, numbers, pos, par->z_rcov, dEdcn, grad);
  return (int)hipGetLastError();
}
