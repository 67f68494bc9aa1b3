// Per-pair physics functions shared by the kernels and the CPU derivative
// checks (tests/cpu_pairfuncs.cpp compiles this header on the host).
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define DXTB_HD __host__ __device__ __forceinline__
#else
#define DXTB_HD inline
#endif

namespace dxtb {

constexpr double EPS = 2.220446049250313e-16;    // FP64 machine epsilon
constexpr double TINY = 2.2250738585072014e-308;  // smallest normal FP64
constexpr double CN_CUTOFF = 25.0;                // Bohr
constexpr double REP_CUTOFF = 25.0;               // Bohr

// Exponential counting function 1/(1+exp(-k(r0/r-1))) and its r-derivative.
DXTB_HD void exp_count(double r, double r0, double k, double& c, double& dc) {
  const double t = exp(-k * (r0 / r - 1.0));
  c = 1.0 / (1.0 + t);
  dc = -c * c * t * k * r0 / (r * r);
}

// D3 coordination-number count: GFN1 exp (k = 16), GFN2 product of two
// exponential counts (k = 10 at r0, k = 20 at r0 + 2 Bohr).
DXTB_HD void cn_count(int gfn, double r, double r0, double& c, double& dc) {
  if (gfn == 1) {
    exp_count(r, r0, 16.0, c, dc);
  } else {
    double c1, d1, c2, d2;
    exp_count(r, r0, 10.0, c1, d1);
    exp_count(r, r0 + 2.0, 20.0, c2, d2);
    c = c1 * c2;
    dc = d1 * c2 + c1 * d2;
  }
}

// H0 distance polynomial Pi(R) = (1 + ka sqrt(R/Rab))(1 + kb sqrt(R/Rab)) and dPi/dR.
DXTB_HD void h0_poly(double R, double rab, double ka, double kb, double& p, double& dp) {
  const double sr = sqrt(R / rab);
  p = (1.0 + ka * sr) * (1.0 + kb * sr);
  dp = (ka * (1.0 + kb * sr) + kb * (1.0 + ka * sr)) * 0.5 / sqrt(R * rab);
}

// ES2 shell pair: gamma = ((R+eps)^2 + avg^-2)^(-1/2) for gexp = 2 (GFN1 and
// GFN2) and its R derivative; same atom: (eps + avg^-2)^(-1/2).
DXTB_HD double es2_avg(int mode, double ha, double hb) {
  ha += EPS;
  hb += EPS;
  if (mode == 0) return 2.0 / (1.0 / ha + 1.0 / hb);
  if (mode == 1) return 0.5 * (ha + hb);
  return sqrt(ha * hb);
}

DXTB_HD void es2_gamma(bool same_atom, double R, double avg, double& g, double& dg) {
  const double ia2 = 1.0 / (avg * avg);
  if (same_atom) {
    g = 1.0 / sqrt(EPS + ia2);
    dg = 0.0;
    return;
  }
  const double re = R + EPS;
  const double t = re * re + ia2;
  g = 1.0 / sqrt(t);
  dg = -re * g / t;
}

// The same for any exponent g > 0 (secondorder.py:799-870):
// gamma = ((R+eps)^g + avg^-g)^(-1/g); same atom: (eps + avg^-g)^(-1/g).
DXTB_HD void es2_gamma_gexp(bool same_atom, double R, double avg, double gexp, double& g, double& dg) {
  const double ia = pow(avg, -gexp);
  if (same_atom) {
    g = pow(EPS + ia, -1.0 / gexp);
    dg = 0.0;
    return;
  }
  const double re = R + EPS;
  const double rg = pow(re, gexp);
  const double t = rg + ia;
  g = pow(t, -1.0 / gexp);
  dg = -(rg / re) * g / t;  // -R^(g-1) gamma / t
}

// Kernels are instantiated for g = 2 (es2_gamma, the code of the GFN1/GFN2
// exponent on its own) and for any other exponent.
template <bool G2>
DXTB_HD void es2_gamma_t(bool same_atom, double R, double avg, double gexp, double& g, double& dg) {
  if (G2)
    es2_gamma(same_atom, R, avg, g, dg);
  else
    es2_gamma_gexp(same_atom, R, avg, gexp, g, dg);
}

// Repulsion pair energy z exp(-a R^k)/R and dE/dR (cutoff applied by caller).
DXTB_HD void rep_pair(double R, double a, double k, double z, double& e, double& de) {
  const double rk = pow(R, k);
  e = z * exp(-a * rk) / R;
  de = -e * (a * k * rk / R + 1.0 / R);
}

// AES2 damping f_n = 1/(1 + 6 rr^n), rr = 0.5 (ra + rb) / R; n integer (3 or 4).
DXTB_HD double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

}  // namespace dxtb

namespace dxtb {

// AES2 interaction of one atom pair i != j (coulomb/multipole.py:351-531),
// r = R_i - R_j, multipole radii m_i, m_j, charges q, dipoles mu, traceless
// quadrupoles th (xx, yx, yy, zx, zy, zz):
//   E = f3 R^-3 r.(q_i mu_j - q_j mu_i)
//     + f5 [R^-3 mu_i.mu_j - 3 R^-5 (r.mu_i)(r.mu_j)]
//     + f5 R^-5 r^T (q_i T(th_j) + q_j T(th_i)) r
// with f_n = 1/(1 + 6 rr^dmp_n), rr = (m_i + m_j)/(2R).  Returns E, dE/dr and
// dE/dm_i (= dE/dm_j).
DXTB_HD void aes2_pair(const double r[3], double mi, double mj, double qi, double qj, const double mui[3],
                       const double muj[3], const double thi[6], const double thj[6], double dmp3, double dmp5,
                       double& E, double g[3], double& dEdm) {
  const double R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double R = sqrt(R2), iR = 1.0 / R, iR2 = iR * iR;
  const double g3 = iR * iR2, g5 = g3 * iR2;
  const double m = 0.5 * (mi + mj);
  const double rr = m * iR;
  const double p3 = pow(rr, dmp3), p5 = pow(rr, dmp5);
  const double f3 = 1.0 / (1.0 + 6.0 * p3), f5 = 1.0 / (1.0 + 6.0 * p5);
  const double df3 = -6.0 * dmp3 * (p3 / rr) * f3 * f3;  // d f3 / d rr
  const double df5 = -6.0 * dmp5 * (p5 / rr) * f5 * f5;
  // d rr / d r = -m R^-3 r ; d rr / d m = 1/R
  const double drr = -m * g3;
  // charge-dipole
  double w[3], rw = 0.0;
  for (int k = 0; k < 3; ++k) {
    w[k] = qi * muj[k] - qj * mui[k];
    rw += r[k] * w[k];
  }
  const double A = f3 * g3 * rw;
  // dipole-dipole
  const double a = mui[0] * muj[0] + mui[1] * muj[1] + mui[2] * muj[2];
  const double ri = r[0] * mui[0] + r[1] * mui[1] + r[2] * mui[2];
  const double rj = r[0] * muj[0] + r[1] * muj[1] + r[2] * muj[2];
  const double Bs = g3 * a - 3.0 * g5 * ri * rj;
  // charge-quadrupole: M = q_i T(th_j) + q_j T(th_i)
  double M[3][3];
  {
    const double t[6] = {qi * thj[0] + qj * thi[0], qi * thj[1] + qj * thi[1], qi * thj[2] + qj * thi[2],
                         qi * thj[3] + qj * thi[3], qi * thj[4] + qj * thi[4], qi * thj[5] + qj * thi[5]};
    M[0][0] = t[0];
    M[0][1] = M[1][0] = t[1];
    M[1][1] = t[2];
    M[0][2] = M[2][0] = t[3];
    M[1][2] = M[2][1] = t[4];
    M[2][2] = t[5];
  }
  double Mr[3], phi = 0.0;
  for (int k = 0; k < 3; ++k) {
    Mr[k] = M[k][0] * r[0] + M[k][1] * r[1] + M[k][2] * r[2];
    phi += r[k] * Mr[k];
  }
  E = A + f5 * Bs + f5 * g5 * phi;
  for (int k = 0; k < 3; ++k) {
    const double gA = f3 * (g3 * w[k] - 3.0 * g5 * rw * r[k]) + g3 * rw * df3 * drr * r[k];
    const double gBs = -3.0 * g5 * a * r[k] - 3.0 * (-5.0 * g5 * iR2 * ri * rj * r[k] + g5 * (mui[k] * rj + muj[k] * ri));
    const double gB = f5 * gBs + Bs * df5 * drr * r[k];
    const double gC = f5 * (-5.0 * g5 * iR2 * phi * r[k] + 2.0 * g5 * Mr[k]) + g5 * phi * df5 * drr * r[k];
    g[k] = gA + gB + gC;
  }
  dEdm = 0.5 * iR * (g3 * rw * df3 + (Bs + g5 * phi) * df5);
}

// d mrad / d CN of the AES2 multipole radius (coulomb/multipole.py:331-333)
DXTB_HD double dmrad_dcn(double cn, double rad, double vcn, double kexp, double shift, double rmax) {
  const double t1 = exp(-kexp * (cn - vcn - shift));
  return (rmax - rad) * kexp * t1 / ((1.0 + t1) * (1.0 + t1));
}

// ---- generalized Born solvation (components/interactions/solvation) ----

// Descreening contribution of atom j (scaled radius rho = 0.8 R_j) to the
// integral psi_i of atom i (van der Waals radius ri) at distance r
// (born.py:163-239), and its derivative d/dr.  Overlapping branch when
// r < ri + rho and r + rho > ri, else the closed form of the
// non-overlapping spheres.
DXTB_HD void gb_psi_term(double r, double rho, double ri, double& t, double& dt) {
  const double dpl = r + rho, dmi = r - rho;
  if (r < ri + rho && dpl > ri) {
    const double ir = 1.0 / r, ipl = 1.0 / dpl, iri2 = 1.0 / (ri * ri);
    const double lg = log(dpl / ri);
    const double a = 0.5 * dmi * (ipl - dpl * iri2) - lg;
    const double da = 0.5 * (ipl - dpl * iri2) + 0.5 * dmi * (-ipl * ipl - iri2) - ipl;
    t = 1.0 / ri - ipl + 0.5 * ir * a;
    dt = ipl * ipl + 0.5 * ir * (da - a * ir);
    return;
  }
  const double pr = dpl * dmi;  // r^2 - rho^2
  t = rho / pr;
  dt = -2.0 * r * rho / (pr * pr);
  const double qq = dmi / dpl;
  if (qq > 0.0) {
    const double lq = log(qq);
    t += 0.5 * lq / r;
    dt += 0.5 * (2.0 * rho / (pr * r) - lq / (r * r));
  }
}

// Born radius from psi (Onufriev-Bashford-Case, born.py:140-160) and dB/dpsi.
DXTB_HD void gb_born_radius(double psi, double rvdw, double scale, double offset, const double obc[3], double& b,
                            double& db) {
  const double srvdw = rvdw - offset;
  const double r1 = 1.0 / rvdw, s1 = 1.0 / srvdw;
  const double p = 0.5 * psi * srvdw;
  const double arg = p * (obc[0] + p * (obc[2] * p - obc[1]));
  const double darg = 0.5 * srvdw * (obc[0] - 2.0 * obc[1] * p + 3.0 * obc[2] * p * p);
  const double th = tanh(arg);
  const double den = s1 - r1 * th;
  b = scale / den;
  db = scale * r1 * (1.0 - th * th) * darg / (den * den);
}

// Interaction kernels of the generalized Born model (alpb.py:94-141) with
// ab = B_i B_j: value and derivatives by r and by ab.  kind 1: P16,
// 1 / (r + s f^16), s = sqrt(ab), f = s / (s + 1.028 r / 16); kind 2: Still,
// 1 / sqrt(r^2 + ab exp(-r^2 / (4 ab))).
DXTB_HD void gb_kernel(int kind, double r, double ab, double& k, double& dkdr, double& dkdab) {
  if (kind == 1) {
    const double c = 1.028 / 16.0;
    const double s = sqrt(ab);
    const double f = s / (s + c * r);
    double f16 = f * f;
    f16 *= f16;
    f16 *= f16;
    f16 *= f16;  // f^16
    k = 1.0 / (r + s * f16);
    const double f17 = f16 * f;
    dkdr = -k * k * (1.0 - 16.0 * c * f17);
    dkdab = -k * k * (f16 + 16.0 * c * r * f17 / s) / (2.0 * s);
  } else {
    const double r2 = r * r;
    const double e = exp(-0.25 * r2 / ab);
    k = 1.0 / sqrt(r2 + ab * e);
    const double k3 = k * k * k;
    dkdr = -k3 * r * (1.0 - 0.25 * e);
    dkdab = -0.5 * k3 * e * (1.0 + 0.25 * r2 / ab);
  }
}

}  // namespace dxtb
