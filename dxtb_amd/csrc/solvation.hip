// Generalized Born implicit solvation (GBSA/ALPB electrostatics of the
// reference's GeneralizedBorn interaction, components/interactions/solvation/
// alpb.py and born.py): Born radii, the atom-resolved potential and energy,
// and the hand-written force contraction.
//
// One workgroup per molecule; the atoms' positions and radii are staged in
// LDS and every thread owns atoms i (strided) and loops over all partners j.
// Molecules with keps = 0 (no solvation) are skipped.
//
//   psi_i = sum_{j != i, r <= cutoff} I(r_ij; rho_j = descreening R_j, R_i)
//   B_i   = born(psi_i)                         (Onufriev-Bashford-Case)
//   V_i   = keps sum_j K(r_ij, B_i B_j) q_j     (j = i included: K = 1/B_i)
//   E     = 1/2 sum_i q_i V_i
//   dE/dR_a = sum_{j != a} [keps q_a q_j dK/dr + g_a I'(r; rho_j, R_a)
//                           + g_j I'(r; rho_a, R_j)] (R_a - R_j) / r,
//   g_i   = dE/dB_i dB_i/dpsi_i,  dE/dB_i = keps q_i sum_j q_j dK/d(ab) B_j.

#include "common.h"
#include "pairfuncs.h"
#include "../../include/dxtb_hip.h"

namespace {

constexpr int GT = 256;
constexpr int GB_MAX_NAT = 1024;  // LDS staging limit (7 doubles per atom)

struct GbPar {
  double descreening, cutoff, scale, offset, obc[3];
};

__global__ __launch_bounds__(GT) void k_gb_born(int na, const int* __restrict__ nat, const int* __restrict__ numbers,
                                                const double* __restrict__ pos, const double* __restrict__ keps,
                                                const double* __restrict__ rvdw_z, GbPar par,
                                                double* __restrict__ born, double* __restrict__ dborn) {
  extern __shared__ double sh[];
  const int b = blockIdx.x;
  if (keps[b] == 0.0) return;
  const int n = min(nat[b], na);
  double* X = sh;             // 3 na
  double* R = sh + 3 * na;    // na
  for (int i = threadIdx.x; i < n; i += GT) {
    for (int k = 0; k < 3; ++k) X[3 * i + k] = pos[((size_t)b * na + i) * 3 + k];
    R[i] = rvdw_z[numbers[(size_t)b * na + i]];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += GT) {
    const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2], ri = R[i];
    double psi = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = xi - X[3 * j], dy = yi - X[3 * j + 1], dz = zi - X[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > par.cutoff) continue;
      double t, dt;
      dxtb::gb_psi_term(r, par.descreening * R[j], ri, t, dt);
      psi += t;
    }
    double bi, dbi;
    dxtb::gb_born_radius(psi, ri, par.scale, par.offset, par.obc, bi, dbi);
    born[(size_t)b * na + i] = bi;
    dborn[(size_t)b * na + i] = dbi;
  }
}

__global__ __launch_bounds__(GT) void k_gb_potential(int na, const int* __restrict__ nat,
                                                     const double* __restrict__ pos, const double* __restrict__ keps,
                                                     const int* __restrict__ kind, const double* __restrict__ born,
                                                     const double* __restrict__ qat, double* __restrict__ vat,
                                                     double* __restrict__ eat) {
  extern __shared__ double sh[];
  const int b = blockIdx.x;
  const double ke = keps[b];
  if (ke == 0.0) return;
  const int n = min(nat[b], na), kd = kind[b];
  double* X = sh;
  double* Bq = sh + 3 * na;  // born
  double* Q = Bq + na;
  for (int i = threadIdx.x; i < n; i += GT) {
    for (int k = 0; k < 3; ++k) X[3 * i + k] = pos[((size_t)b * na + i) * 3 + k];
    Bq[i] = born[(size_t)b * na + i];
    Q[i] = qat[(size_t)b * na + i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += GT) {
    const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2], bi = Bq[i];
    double v = 0.0;
    for (int j = 0; j < n; ++j) {
      const double dx = xi - X[3 * j], dy = yi - X[3 * j + 1], dz = zi - X[3 * j + 2];
      const double r = (j == i) ? 0.0 : sqrt(dx * dx + dy * dy + dz * dz);
      double k, dkr, dkab;
      dxtb::gb_kernel(kd, r, bi * Bq[j], k, dkr, dkab);
      v += k * Q[j];
    }
    v *= ke;
    vat[(size_t)b * na + i] += v;
    eat[(size_t)b * na + i] += 0.5 * Q[i] * v;
  }
}

__global__ __launch_bounds__(GT) void k_gb_grad(int na, const int* __restrict__ nat, const int* __restrict__ numbers,
                                                const double* __restrict__ pos, const double* __restrict__ keps,
                                                const int* __restrict__ kind, const double* __restrict__ rvdw_z,
                                                GbPar par, const double* __restrict__ born,
                                                const double* __restrict__ dborn, const double* __restrict__ qat,
                                                double* __restrict__ grad) {
  extern __shared__ double sh[];
  const int b = blockIdx.x;
  const double ke = keps[b];
  if (ke == 0.0) return;
  const int n = min(nat[b], na), kd = kind[b];
  double* X = sh;
  double* Bq = sh + 3 * na;
  double* Q = Bq + na;
  double* R = Q + na;
  double* G = R + na;  // g_i = dE/dB_i dB_i/dpsi_i
  for (int i = threadIdx.x; i < n; i += GT) {
    for (int k = 0; k < 3; ++k) X[3 * i + k] = pos[((size_t)b * na + i) * 3 + k];
    Bq[i] = born[(size_t)b * na + i];
    Q[i] = qat[(size_t)b * na + i];
    R[i] = rvdw_z[numbers[(size_t)b * na + i]];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += GT) {
    const double xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2], bi = Bq[i];
    double dedb = 0.0;
    for (int j = 0; j < n; ++j) {
      const double dx = xi - X[3 * j], dy = yi - X[3 * j + 1], dz = zi - X[3 * j + 2];
      const double r = (j == i) ? 0.0 : sqrt(dx * dx + dy * dy + dz * dz);
      double k, dkr, dkab;
      dxtb::gb_kernel(kd, r, bi * Bq[j], k, dkr, dkab);
      dedb += Q[j] * dkab * Bq[j];
    }
    G[i] = ke * Q[i] * dedb * dborn[(size_t)b * na + i];
  }
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += GT) {
    const double xa = X[3 * a], ya = X[3 * a + 1], za = X[3 * a + 2], ba = Bq[a], ra = R[a], ga = G[a];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == a) continue;
      const double dx = xa - X[3 * j], dy = ya - X[3 * j + 1], dz = za - X[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      double k, dkr, dkab;
      dxtb::gb_kernel(kd, r, ba * Bq[j], k, dkr, dkab);
      double f = ke * Q[a] * Q[j] * dkr;
      if (r <= par.cutoff) {
        double t, dt;
        dxtb::gb_psi_term(r, par.descreening * R[j], ra, t, dt);
        f += ga * dt;
        dxtb::gb_psi_term(r, par.descreening * ra, R[j], t, dt);
        f += G[j] * dt;
      }
      f /= r;
      g0 += f * dx;
      g1 += f * dy;
      g2 += f * dz;
    }
    grad[((size_t)b * na + a) * 3 + 0] += g0;
    grad[((size_t)b * na + a) * 3 + 1] += g1;
    grad[((size_t)b * na + a) * 3 + 2] += g2;
  }
}

GbPar to_par(const dxtb_gb_params_t* p) {
  GbPar g;
  g.descreening = p->descreening;
  g.cutoff = p->cutoff;
  g.scale = p->born_scale;
  g.offset = p->born_offset;
  for (int k = 0; k < 3; ++k) g.obc[k] = p->obc[k];
  return g;
}

}  // namespace

extern "C" int dxtb_gb_born(int nbatch, int nat_max, const int* nat, const int* numbers, const double* pos,
                            const double* keps, const double* rvdw_z, const dxtb_gb_params_t* par, double* born,
                            double* dborn, hipStream_t st) {
  if (nbatch <= 0 || nat_max <= 0 || nat_max > GB_MAX_NAT || !nat || !numbers || !pos || !keps || !rvdw_z || !par ||
      !born || !dborn)
    return DXTB_EINVAL;
  const size_t shmem = (size_t)nat_max * 4 * sizeof(double);
  hipLaunchKernelGGL(k_gb_born, dim3((unsigned)nbatch), dim3(GT), shmem, st, nat_max, nat, numbers, pos, keps, rvdw_z,
                     to_par(par), born, dborn);
  return (int)hipGetLastError();
}

extern "C" int dxtb_gb_potential(int nbatch, int nat_max, const int* nat, const double* pos, const double* keps,
                                 const int* kind, const double* born, const double* qat, double* vat, double* eat,
                                 hipStream_t st) {
  if (nbatch <= 0 || nat_max <= 0 || nat_max > GB_MAX_NAT || !nat || !pos || !keps || !kind || !born || !qat || !vat ||
      !eat)
    return DXTB_EINVAL;
  const size_t shmem = (size_t)nat_max * 5 * sizeof(double);
  hipLaunchKernelGGL(k_gb_potential, dim3((unsigned)nbatch), dim3(GT), shmem, st, nat_max, nat, pos, keps, kind, born,
                     qat, vat, eat);
  return (int)hipGetLastError();
}

extern "C" int dxtb_gb_forces(int nbatch, int nat_max, const int* nat, const int* numbers, const double* pos,
                              const double* keps, const int* kind, const double* rvdw_z, const dxtb_gb_params_t* par,
                              const double* born, const double* dborn, const double* qat, double* grad,
                              hipStream_t st) {
  if (nbatch <= 0 || nat_max <= 0 || nat_max > GB_MAX_NAT || !nat || !numbers || !pos || !keps || !kind || !rvdw_z ||
      !par || !born || !dborn || !qat || !grad)
    return DXTB_EINVAL;
  const size_t shmem = (size_t)nat_max * 7 * sizeof(double);
  hipLaunchKernelGGL(k_gb_grad, dim3((unsigned)nbatch), dim3(GT), shmem, st, nat_max, nat, numbers, pos, keps, kind,
                     rvdw_z, to_par(par), born, dborn, qat, grad);
  return (int)hipGetLastError();
}
