// Per-geometry set-up (coordination numbers, H0 shell-pair factors, repulsion,
// AES2 radii) and the per-iteration electrostatic potentials (ES2, ES3, AES2).
// All O(nat^2)/O(nsh^2) streaming work in FP64; the set-up kernels take one
// thread per atom or shell pair, the potentials a group of lanes per shell or
// atom.  Partner loops run over the molecule's own atoms/shells.

#include "common.h"
#include "pairfuncs.h"
#include "../../include/dxtb_hip.h"

using namespace dxtb;

namespace {

__device__ __forceinline__ void load3(const double* p, double (&r)[3]) {
  r[0] = p[0];
  r[1] = p[1];
  r[2] = p[2];
}

// ---- coordination numbers (tad-mctc cn_d3 with exp / gfn2 counting) -------
__global__ void k_cn(int B, int nat_max, int gfn, const int* __restrict__ numbers,
                     const double* __restrict__ pos, const double* __restrict__ rcov,
                     double* __restrict__ cn) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * nat_max) return;
  const int b = t / nat_max;
  const int za = numbers[t];
  double acc = 0.0;
  if (za != 0) {
    double ra[3];
    load3(pos + (size_t)t * 3, ra);
    const int a = t - b * nat_max;
    for (int j = 0; j < nat_max; ++j) {
      const int zb = numbers[b * nat_max + j];
      if (zb == 0 || j == a) continue;
      const double* pb = pos + ((size_t)b * nat_max + j) * 3;
      const double dx = ra[0] - pb[0], dy = ra[1] - pb[1], dz = ra[2] - pb[2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > CN_CUTOFF) continue;
      double c, dc;
      cn_count(gfn, r, rcov[za] + rcov[zb], c, dc);
      acc += c;
    }
  }
  cn[t] = acc;
}

// ---- H0 shell-pair factor: f, df/dR, df/dh (xtb/base.py:255-372) ----------
__global__ void k_h0pair(dxtb_basis_t bs, dxtb_params_t par, const int* __restrict__ numbers,
                         const double* __restrict__ pos, const double* __restrict__ cn,
                         double* __restrict__ fsh, double* __restrict__ dfsh, double* __restrict__ hfsh,
                         const int* __restrict__ nsh) {
  const int ns = bs.nsh_max;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)bs.nbatch * ns * ns) return;
  const int b = (int)(t / ((long)ns * ns));
  const int rem = (int)(t - (long)b * ns * ns);
  const int sa = rem / ns, sb = rem - (rem / ns) * ns;
  double f = 0.0, df = 0.0, hf = 0.0;
  if (sa < nsh[b] && sb < nsh[b]) {
    const int base = b * ns;
    const int ua = bs.sh_ush[base + sa], ub = bs.sh_ush[base + sb];
    const int aa = bs.sh_atom[base + sa], ab = bs.sh_atom[base + sb];
    const int za = numbers[b * bs.nat_max + aa], zb = numbers[b * bs.nat_max + ab];
    const double ha = par.ush_lev[ua] - par.ush_kcn[ua] * cn[b * bs.nat_max + aa];
    const double hb = par.ush_lev[ub] - par.ush_kcn[ub] * cn[b * bs.nat_max + ab];
    const double hm = 0.5 * (ha + hb);
    if (aa == ab) {
      f = hm;
      hf = 0.5;
    } else {
      const double* pa = pos + ((size_t)b * bs.nat_max + aa) * 3;
      const double* pb = pos + ((size_t)b * bs.nat_max + ab) * 3;
      const double dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      double p, dp;
      h0_poly(R, par.z_rad[za] + par.z_rad[zb], par.ush_shpoly[ua], par.ush_shpoly[ub], p, dp);
      const double hs = par.hscale[(size_t)ua * bs.nush + ub];
      double k = hs;
      if (par.ush_valence[ua] && par.ush_valence[ub]) {
        const double den = par.z_en[za] - par.z_en[zb];
        k = hs * par.kpair[za * 87 + zb] * (1.0 + par.enscale * den * den);
      }
      f = p * k * hm;
      df = dp * k * hm;
      hf = 0.5 * p * k;
    }
  }
  fsh[t] = f;
  dfsh[t] = df;
  hfsh[t] = hf;
}

// ---- repulsion energy per atom (repulsion/base.py:293-358) ---------------
__device__ __forceinline__ bool rep_params(const dxtb_params_t& par, int za, int zb, double& a, double& k,
                                           double& z) {
  a = sqrt(par.z_arep[za] * par.z_arep[zb] + TINY);
  k = (par.rep_light && za <= 2 && zb <= 2) ? par.rep_klight : par.rep_kexp;
  z = par.z_zeff[za] * par.z_zeff[zb];
  return true;
}

__global__ void k_rep(int B, int nat_max, dxtb_params_t par, const int* __restrict__ numbers,
                      const double* __restrict__ pos, double* __restrict__ erep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * nat_max) return;
  const int b = t / nat_max, i = t - b * nat_max;
  const int za = numbers[t];
  double acc = 0.0;
  if (za != 0) {
    double ra[3];
    load3(pos + (size_t)t * 3, ra);
    for (int j = 0; j < nat_max; ++j) {
      const int zb = numbers[b * nat_max + j];
      if (zb == 0 || j == i) continue;
      const double* pb = pos + ((size_t)b * nat_max + j) * 3;
      const double dx = ra[0] - pb[0], dy = ra[1] - pb[1], dz = ra[2] - pb[2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      if (R > REP_CUTOFF) continue;
      double a, k, z, e, de;
      rep_params(par, za, zb, a, k, z);
      rep_pair(R, a, k, z, e, de);
      acc += 0.5 * e;
    }
  }
  erep[t] = acc;
}

// ---- AES2 multipole radii from the GFN2 CN (coulomb/multipole.py:320-333) -
__global__ void k_mrad(int B, int nat_max, dxtb_params_t par, const int* __restrict__ numbers,
                       const double* __restrict__ cn, double* __restrict__ mrad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * nat_max) return;
  const int z = numbers[t];
  if (z == 0) {
    mrad[t] = 0.0;
    return;
  }
  const double rad = par.z_mprad[z];
  const double t1 = exp(-par.aes_kexp * (cn[t] - par.z_mpvcn[z] - par.aes_shift));
  mrad[t] = rad + (par.aes_rmax - rad) / (1.0 + t1);
}

// ---- per-iteration potentials: PG lanes per shell or atom -----------------
// Each shell (atom) of a window is served by a group of PG lanes that split
// its partner loop; the group's partial sums are combined by shuffles.  A
// window of a few large molecules then still spreads over the device, and a
// lone molecule's partner loops run PG times shorter.
constexpr int PG = 16;        // lanes per shell / atom (a power of two within a wavefront)
constexpr int POT_BLK = 256;  // threads per workgroup: POT_BLK / PG targets

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = PG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---- shell potential: ES2 gamma*q (+ GFN2 shell ES3) ----------------------
template <bool G2>
__global__ __launch_bounds__(POT_BLK) void k_pot_shell(dxtb_basis_t bs, dxtb_params_t par,
                                                       const double* __restrict__ pos, const int* __restrict__ nsh,
                                                       const double* __restrict__ qsh, double* __restrict__ vsh,
                                                       double* __restrict__ esh) {
  const int ns = bs.nsh_max;
  const long t = ((long)blockIdx.x * POT_BLK + threadIdx.x) / PG;  // target shell of this lane's group
  const int g = threadIdx.x & (PG - 1);
  const bool live = t < (long)bs.nbatch * ns;
  const int b = live ? (int)(t / ns) : 0, s = live ? (int)(t - (long)b * ns) : 0;
  const int n = live ? nsh[b] : 0;
  const bool act = live && s < n;
  const int base = b * ns;
  double v = 0.0;
  double hs = 0.0;
  int us = 0;
  if (act) {  // the whole group takes this branch together
    us = bs.sh_ush[base + s];
    const int as = bs.sh_atom[base + s];
    hs = par.ush_hub[us];
    double ra[3];
    load3(pos + ((size_t)b * bs.nat_max + as) * 3, ra);
    for (int u = g; u < n; u += PG) {
      const int au = bs.sh_atom[base + u];
      const double* pb = pos + ((size_t)b * bs.nat_max + au) * 3;
      const double dx = ra[0] - pb[0], dy = ra[1] - pb[1], dz = ra[2] - pb[2];
      const double R = sqrt(dx * dx + dy * dy + dz * dz);
      double gm, dg;
      es2_gamma_t<G2>(au == as, R, es2_avg(par.es2_average, hs, par.ush_hub[bs.sh_ush[base + u]]), par.es2_gexp, gm,
                      dg);
      v += gm * qsh[base + u];
    }
  }
  v = group_sum(v);
  if (!live || g != 0) return;
  if (!act) {
    vsh[t] = 0.0;
    esh[t] = 0.0;
    return;
  }
  const double q = qsh[t];
  double e = 0.5 * q * v;
  if (par.es3_shell) {
    const double hd = par.ush_hd3[us];
    v += hd * q * q;
    e += hd * q * q * q / 3.0;
  }
  vsh[t] = v;
  esh[t] = e;
}

// ---- atom potentials: GFN1 atom ES3, GFN2 AES2 ------------------------------
__global__ __launch_bounds__(POT_BLK) void k_pot_atom(int B, int nat_max, dxtb_params_t par,
                                                      const int* __restrict__ numbers,
                                                      const double* __restrict__ pos, const double* __restrict__ mrad,
                                                      const double* __restrict__ qat, const double* __restrict__ qdp,
                                                      const double* __restrict__ qqp, double* __restrict__ vat,
                                                      double* __restrict__ vdp, double* __restrict__ vqp,
                                                      double* __restrict__ eat) {
  const long tl = ((long)blockIdx.x * POT_BLK + threadIdx.x) / PG;  // target atom of this lane's group
  const int g = threadIdx.x & (PG - 1);
  const bool live = tl < (long)B * nat_max;
  const int t = live ? (int)tl : 0;
  const int b = t / nat_max, i = t - b * nat_max;
  const int z = live ? numbers[t] : 0;
  // partner sums of the AES2 potentials (GFN2): monopole, dipole (charge and
  // dipole parts) and quadrupole potential at atom i
  double va = 0.0, vsd[3] = {0, 0, 0}, vdd[3] = {0, 0, 0}, vsq[6] = {0, 0, 0, 0, 0, 0};
  if (z != 0 && par.gfn == 2) {  // uniform over the group
    double ri[3];
    load3(pos + (size_t)t * 3, ri);
    const double mi = mrad[t];
    for (int j = g; j < nat_max; j += PG) {
      const int zj = numbers[b * nat_max + j];
      if (zj == 0 || j == i) continue;
      const size_t tj = (size_t)b * nat_max + j;
      const double* pj = pos + tj * 3;
      // r = R_i - R_j (the reference's rij for the (i, j) element)
      const double r[3] = {ri[0] - pj[0], ri[1] - pj[1], ri[2] - pj[2]};
      const double R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double R = sqrt(R2);
      const double g1 = 1.0 / R, g3 = g1 * g1 * g1, g5 = g3 * g1 * g1;
      const double rr = 0.5 * (mi + mrad[tj]) * g1;
      const double f3 = 1.0 / (1.0 + 6.0 * pow(rr, par.aes_dmp3));
      const double f5 = 1.0 / (1.0 + 6.0 * pow(rr, par.aes_dmp5));
      const double qj = qat[tj];
      const double muj[3] = {qdp[tj * 3], qdp[tj * 3 + 1], qdp[tj * 3 + 2]};
      double thj[6];
      for (int k = 0; k < 6; ++k) thj[k] = qqp[tj * 6 + k];
      // sd[i,j] = r g3 f3 ; sd[j,i] = -sd[i,j]
      const double s3 = g3 * f3;
      // sq[i,j] = sq[j,i] = (x^2, 2xy, y^2, 2xz, 2yz, z^2) g5 f5
      const double s5 = g5 * f5;
      const double sq[6] = {r[0] * r[0] * s5, 2 * r[0] * r[1] * s5, r[1] * r[1] * s5,
                            2 * r[0] * r[2] * s5, 2 * r[1] * r[2] * s5, r[2] * r[2] * s5};
      // monopole potential at i: sum_j sd[i,j].mu_j + sq[i,j].th_j
      va += s3 * (r[0] * muj[0] + r[1] * muj[1] + r[2] * muj[2]);
      for (int k = 0; k < 6; ++k) va += sq[k] * thj[k];
      // dipole potential at i: sum_j sd[j,i] q_j + dd[j,i] mu_j
      for (int k = 0; k < 3; ++k) vsd[k] += -r[k] * s3 * qj;
      const double rmu = r[0] * muj[0] + r[1] * muj[1] + r[2] * muj[2];
      for (int k = 0; k < 3; ++k) vdd[k] += g3 * f5 * muj[k] - 3.0 * r[k] * rmu * s5;
      // quadrupole potential at i: sum_j sq[j,i] q_j
      for (int k = 0; k < 6; ++k) vsq[k] += sq[k] * qj;
    }
  }
  va = group_sum(va);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    vsd[k] = group_sum(vsd[k]);
    vdd[k] = group_sum(vdd[k]);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) vsq[k] = group_sum(vsq[k]);
  if (!live || g != 0) return;
  double e = 0.0;
  double vd[3] = {0, 0, 0}, vq[6] = {0, 0, 0, 0, 0, 0};
  if (z != 0) {
    const double q = qat[t];
    if (!par.es3_shell) {
      const double hd = par.z_gam3[z];
      va += hd * q * q;
      e += hd * q * q * q / 3.0;
    }
    if (par.gfn == 2) {
      double mu[3], th[6];
      for (int k = 0; k < 3; ++k) mu[k] = qdp[(size_t)t * 3 + k];
      for (int k = 0; k < 6; ++k) th[k] = qqp[(size_t)t * 6 + k];
      const double dk = par.z_dkernel[z], qk = par.z_qkernel[z];
      const double scale[6] = {1, 2, 1, 2, 2, 1};
      for (int k = 0; k < 3; ++k) {
        vd[k] = vsd[k] + vdd[k] + 2.0 * dk * mu[k];
        e += dk * mu[k] * mu[k] + (vsd[k] + 0.5 * vdd[k]) * mu[k];
      }
      for (int k = 0; k < 6; ++k) {
        vq[k] = vsq[k] + 2.0 * qk * scale[k] * th[k];
        e += qk * scale[k] * th[k] * th[k] + vsq[k] * th[k];
      }
    }
  } else {
    va = 0.0;
  }
  vat[t] = va;
  eat[t] = e;
  if (vdp != nullptr)
    for (int k = 0; k < 3; ++k) vdp[(size_t)t * 3 + k] = vd[k];
  if (vqp != nullptr)
    for (int k = 0; k < 6; ++k) vqp[(size_t)t * 6 + k] = vq[k];
}

inline dim3 grid1(long n, int blk) { return dim3((unsigned)((n + blk - 1) / blk)); }

}  // namespace

extern "C" int dxtb_setup(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers,
                          const double* pos, const int* nsh, double* cn, double* fsh, double* dfsh,
                          double* hfsh, double* erep, double* mrad, hipStream_t st) {
  if (!bas || !par || !numbers || !pos || !nsh || !cn || !fsh || !dfsh || !hfsh || !erep) return DXTB_EINVAL;
  if (par->gfn != 1 && par->gfn != 2) return DXTB_EUNSUP;
  if (par->gfn == 2 && mrad == nullptr) return DXTB_EINVAL;
  const dxtb_basis_t& b = *bas;
  const int na = b.nbatch * b.nat_max;
  hipLaunchKernelGGL(k_cn, grid1(na, 256), dim3(256), 0, st, b.nbatch, b.nat_max, par->gfn, numbers, pos,
                     par->z_rcov, cn);
  DXTB_CHECK(hipGetLastError());
  const long np = (long)b.nbatch * b.nsh_max * b.nsh_max;
  hipLaunchKernelGGL(k_h0pair, grid1(np, 256), dim3(256), 0, st, b, *par, numbers, pos, cn, fsh, dfsh, hfsh,
                     nsh);
  DXTB_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_rep, grid1(na, 256), dim3(256), 0, st, b.nbatch, b.nat_max, *par, numbers, pos, erep);
  DXTB_CHECK(hipGetLastError());
  if (par->gfn == 2) {
    hipLaunchKernelGGL(k_mrad, grid1(na, 256), dim3(256), 0, st, b.nbatch, b.nat_max, *par, numbers, cn, mrad);
    DXTB_CHECK(hipGetLastError());
  }
  return 0;
}

extern "C" int dxtb_potential(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers,
                              const double* pos, const int* nsh, const double* mrad, const double* qsh,
                              const double* qat, const double* qdp, const double* qqp, double* vsh, double* vat,
                              double* vdp, double* vqp, double* esh, double* eat, hipStream_t st) {
  if (!bas || !par || !numbers || !pos || !nsh || !qsh || !qat || !vsh || !vat || !esh || !eat)
    return DXTB_EINVAL;
  if (par->gfn == 2 && (!mrad || !qdp || !qqp || !vdp || !vqp)) return DXTB_EINVAL;
  if (!(par->es2_gexp > 0.0)) return DXTB_EINVAL;
  const dxtb_basis_t& b = *bas;
  if (par->es2_gexp == 2.0)
    hipLaunchKernelGGL(k_pot_shell<true>, grid1((long)b.nbatch * b.nsh_max * PG, POT_BLK), dim3(POT_BLK), 0, st, b,
                       *par, pos, nsh, qsh, vsh, esh);
  else
    hipLaunchKernelGGL(k_pot_shell<false>, grid1((long)b.nbatch * b.nsh_max * PG, POT_BLK), dim3(POT_BLK), 0, st, b,
                       *par, pos, nsh, qsh, vsh, esh);
  DXTB_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_pot_atom, grid1((long)b.nbatch * b.nat_max * PG, POT_BLK), dim3(POT_BLK), 0, st, b.nbatch,
                     b.nat_max, *par, numbers, pos, mrad, qat, qdp, qqp, vat, vdp, vqp, eat);
  DXTB_CHECK(hipGetLastError());
  return 0;
}
