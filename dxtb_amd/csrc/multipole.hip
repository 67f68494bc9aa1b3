// Overlap / dipole / quadrupole integrals over contracted Gaussians and their
// nuclear derivatives (gfx950, FP64).
//
// Work decomposition of the forward pass: one thread per (ordered shell pair
// (A bra, B ket), spherical bra component m_a).  Each thread owns one row
// segment of the (2l_B+1)-wide block of every output matrix, so the whole
// matrix (both triangles) is written exactly once and no scatter or mirror
// pass is needed.  The nuclear-derivative pass takes one thread per ordered
// shell pair and sums over the bra components in registers.
// The moments are taken about the ket centre B directly, which is the
// reference's r0 integral followed by its r0->rj shift
// (integral/types/dipole.py:40-86, quadrupole.py:308-389) computed without the
// cancellation of the large origin-centred moments: (x - B_x)^e phi_b raises
// the ket index of the 1D overlap table, <a|(x-B)^e|b>_x = s_x[i][j+e].
//
// 1D tables: two-index Obara-Saika overlap recursion per primitive pair
// (the e = 0 slice of the reference's three-index recursion,
// driver/pytorch/impls/kernels/os.py:61-163).
//
// Shell pairs arrive grouped by class (l_A, l_B) so every kernel instance is
// specialised on both angular momenta; all table indices are compile-time
// constants except the bra Cartesian exponents, which are selected from the
// table rows with compare chains (no private-memory indexing).

#include "common.h"
#include "../../include/dxtb_hip.h"

namespace {

constexpr double PI32 = 5.568327996831707845;  // pi^(3/2)
constexpr double S3 = 1.7320508075688772;

// bra spherical component -> up to 3 (cartesian, coefficient) pairs (d only)
__device__ __forceinline__ void d_combo(int m, int k, int& c, double& w) {
  // rows of the d transform (impls/trafo.py TRAFO[2]); selects instead of a
  // table so that the per-lane index never touches private memory
  // m=0: s3*xy   m=1: s3*yz   m=2: -xx/2 - yy/2 + zz   m=3: s3*xz   m=4: s3/2 (xx - yy)
  const int k0 = (m == 0) ? 1 : (m == 1) ? 4 : (m == 3) ? 2 : 0;
  const double w0 = (m == 2) ? -0.5 : (m == 4) ? 0.5 * S3 : S3;
  const int k1 = 3;
  const double w1 = (m == 2) ? -0.5 : (m == 4) ? -0.5 * S3 : 0.0;
  const int k2 = 5;
  const double w2 = (m == 2) ? 1.0 : 0.0;
  c = (k == 0) ? k0 : (k == 1) ? k1 : k2;
  w = (k == 0) ? w0 : (k == 1) ? w1 : w2;
}

template <int NI, int NJ>
__device__ __forceinline__ void os_table(double (&s)[NI][NJ], double pa, double pb, double x2p) {
  s[0][0] = 1.0;
#pragma unroll
  for (int j = 1; j < NJ; ++j) s[0][j] = pb * s[0][j - 1] + (j > 1 ? (j - 1) * x2p * s[0][j - 2] : 0.0);
#pragma unroll
  for (int i = 1; i < NI; ++i) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      double v = pa * s[i - 1][j];
      double low = 0.0;
      if (i > 1) low += (i - 1) * s[i - 2][j];
      if (j > 0) low += j * s[i - 1][j - 1];
      s[i][j] = v + x2p * low;
    }
  }
}

// row s[e][*] with a runtime e in [0, NI)
template <int NI, int NJ>
__device__ __forceinline__ void select_row(const double (&s)[NI][NJ], int e, double (&r)[NJ]) {
  if constexpr (NI <= 2) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) r[j] = (e == 1 && NI == 2) ? s[NI - 1][j] : s[0][j];
  } else {
    // arithmetic blend with exact 0/1 weights: a compare chain over three or
    // more rows is turned back into an indexed private-memory load by the
    // compiler, the blend keeps everything in registers
    double m[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) m[i] = (e == i) ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      double v = m[0] * s[0][j];
#pragma unroll
      for (int i = 1; i < NI; ++i) v = fma(m[i], s[i][j], v);
      r[j] = v;
    }
  }
}

struct PairCtx {
  int b, ia, ib, atA, atB, uA, uB, row, col0;
  double AB[3];
};

__device__ __forceinline__ bool load_pair(const dxtb_basis_t& bs, const double* __restrict__ pos,
                                          int item, PairCtx& c, int ma) {
  c.b = bs.item_mol[item];
  const int pr = bs.item_pair[item];
  c.ia = pr >> 16;
  c.ib = pr & 0xffff;
  const int sb = c.b * bs.nsh_max;
  c.atA = bs.sh_atom[sb + c.ia];
  c.atB = bs.sh_atom[sb + c.ib];
  c.uA = bs.sh_ush[sb + c.ia];
  c.uB = bs.sh_ush[sb + c.ib];
  c.row = bs.sh_ao[sb + c.ia] + ma;
  c.col0 = bs.sh_ao[sb + c.ib];
  const double* pA = pos + ((size_t)c.b * bs.nat_max + c.atA) * 3;
  const double* pB = pos + ((size_t)c.b * bs.nat_max + c.atB) * 3;
#pragma unroll
  for (int x = 0; x < 3; ++x) c.AB[x] = pB[x] - pA[x];
  return true;
}

template <int LA>
__device__ __forceinline__ void bra_combo(int ma, int k, int (&e)[3], double& w) {
  int c;
  if constexpr (LA == 2) {
    d_combo(ma, k, c, w);
  } else {
    c = ma;
    w = 1.0;
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) e[x] = cart_exp(LA, c, x);
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <int LA, int LB, bool MP>
__global__ __launch_bounds__(256) void k_multipole_fwd(dxtb_basis_t bs, int off, int nitems,
                                                       const double* __restrict__ pos,
                                                       const double* __restrict__ fsh,
                                                       double* __restrict__ S, double* __restrict__ D,
                                                       double* __restrict__ Q, double* __restrict__ H0,
                                                       int traceless) {
  constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1, NCB = Cart<LB>::n;
  constexpr int NI = LA + 1, NJ = LB + (MP ? 3 : 1);
  constexpr int NCOMB = (LA == 2) ? 3 : 1;
  constexpr int NC = MP ? 10 : 1;
  // component-major: consecutive lanes take consecutive pairs of the class
  // (consecutive ket shells of one bra row) with the same bra component, so
  // with the AO layout grouped by l a wavefront stores adjacent row segments
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nitems * NA) return;
  const int ma = tid / nitems;
  const int local = tid - ma * nitems;
  PairCtx c;
  load_pair(bs, pos, off + local, c, ma);
  const double r2 = c.AB[0] * c.AB[0] + c.AB[1] * c.AB[1] + c.AB[2] * c.AB[2];

  double acc[NCB][NC];
#pragma unroll
  for (int q = 0; q < NCB; ++q)
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[q][t] = 0.0;

  const int npA = bs.ush_np[c.uA], npB = bs.ush_np[c.uB];
  const double* alA = bs.ush_alpha + c.uA * DXTB_MAXP;
  const double* alB = bs.ush_alpha + c.uB * DXTB_MAXP;
  const double* coA = bs.ush_coeff + c.uA * DXTB_MAXP;
  const double* coB = bs.ush_coeff + c.uB * DXTB_MAXP;

  for (int i = 0; i < npA; ++i) {
    const double a = alA[i], ca = coA[i];
    for (int j = 0; j < npB; ++j) {
      const double bb = alB[j];
      const double p = a + bb, oop = 1.0 / p, x2p = 0.5 * oop;
      const double kab = ca * coB[j] * exp(-a * bb * oop * r2) * oop * sqrt(oop) * PI32;
      double s[3][NI][NJ];
#pragma unroll
      for (int x = 0; x < 3; ++x) os_table<NI, NJ>(s[x], bb * oop * c.AB[x], -a * oop * c.AB[x], x2p);
#pragma unroll
      for (int k = 0; k < NCOMB; ++k) {
        int ex[3];
        double ew;
        bra_combo<LA>(ma, k, ex, ew);
        double rx[NJ], ry[NJ], rz[NJ];
        select_row<NI, NJ>(s[0], ex[0], rx);
        select_row<NI, NJ>(s[1], ex[1], ry);
        select_row<NI, NJ>(s[2], ex[2], rz);
        const double w = kab * ew;
#pragma unroll
        for (int q = 0; q < NCB; ++q) {
          const int bx = cart_exp(LB, q, 0), by = cart_exp(LB, q, 1), bz = cart_exp(LB, q, 2);
          const double x0 = rx[bx], y0 = ry[by], z0 = rz[bz];
          const double wyz = w * y0 * z0;
          acc[q][0] += wyz * x0;
          if constexpr (MP) {
            const double x1 = rx[bx + 1], y1 = ry[by + 1], z1 = rz[bz + 1];
            const double x2 = rx[bx + 2], y2 = ry[by + 2], z2 = rz[bz + 2];
            const double wx0 = w * x0;
            acc[q][1] += wyz * x1;
            acc[q][2] += wx0 * y1 * z0;
            acc[q][3] += wx0 * y0 * z1;
            acc[q][4] += wyz * x2;
            acc[q][5] += w * x1 * y1 * z0;
            acc[q][6] += w * x1 * y0 * z1;
            acc[q][7] += wx0 * y2 * z0;
            acc[q][8] += wx0 * y1 * z1;
            acc[q][9] += wx0 * y0 * z2;
          }
        }
      }
    }
  }

  const int n = bs.nao_max;
  const size_t nn = (size_t)n * n;
  const size_t base = (size_t)c.b * nn + (size_t)c.row * n + c.col0;
  double f = 0.0;
  if (H0 != nullptr) f = fsh[((size_t)c.b * bs.nsh_max + c.ia) * bs.nsh_max + c.ib];
#pragma unroll
  for (int mb = 0; mb < NB; ++mb) {
    double v[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      double acc_t = 0.0;
#pragma unroll
      for (int q = 0; q < NCB; ++q) {
        constexpr int LBc = LB;
        const double tq = trafo(LBc, mb, q);
        if (tq != 0.0) acc_t += tq * acc[q][t];
      }
      v[t] = acc_t;
    }
    S[base + mb] = v[0];
    if (H0 != nullptr) H0[base + mb] = v[0] * f;
    if constexpr (MP) {
      if (D != nullptr) {
        const size_t db = (size_t)c.b * 3 * nn + (size_t)c.row * n + c.col0 + mb;
        D[db] = v[1];
        D[db + nn] = v[2];
        D[db + 2 * nn] = v[3];
      }
      if (Q != nullptr) {
        const size_t qb = (size_t)c.b * 6 * nn + (size_t)c.row * n + c.col0 + mb;
        if (traceless) {
          const double tr = 0.5 * (v[4] + v[7] + v[9]);
          Q[qb] = 1.5 * v[4] - tr;
          Q[qb + nn] = 1.5 * v[5];
          Q[qb + 2 * nn] = 1.5 * v[7] - tr;
          Q[qb + 3 * nn] = 1.5 * v[6];
          Q[qb + 4 * nn] = 1.5 * v[8];
          Q[qb + 5 * nn] = 1.5 * v[9] - tr;
        } else {
          Q[qb] = v[4];
          Q[qb + nn] = v[5];
          Q[qb + 2 * nn] = v[7];
          Q[qb + 3 * nn] = v[6];
          Q[qb + 4 * nn] = v[8];
          Q[qb + 5 * nn] = v[9];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// vector-Jacobian product: d/dR sum(adj * M)
// ---------------------------------------------------------------------------
// Derivative of a bra primitive w.r.t. its centre:
//   d/dA_x (x-A_x)^i e^{-a(x-A_x)^2} = 2a (x-A_x)^{i+1} e - i (x-A_x)^{i-1} e.
// The ket-centred moments depend on A - B only, so d/dB = -d/dA.
// Adjoint sources: either explicit adjoint matrices (adjS/adjD/adjQ) or, in
// SCF mode, the variational adjoints of the converged free energy built on
// the fly from the density P and the energy-weighted density W:
//   adjS_mn = P_mn (f_AB - v_n) - W_mn,  adjD_mn = -P_mn vd(atom n),
//   adjQ_mn = -P_mn vq(atom n)   (f_AB: H0 shell-pair factor, v: potentials)
struct AdjSrc {
  const double* adjS;
  const double* adjD;
  const double* adjQ;
  const double* P;
  const double* W;
  const double* fsh;
  const double* vsh;
  const double* vdp;
  const double* vqp;
};

// Sum of v over the run of consecutive lanes of the wavefront that share
// `key`, returned on the first lane of each run (other lanes: partial sums).
// Runs are numbered from a ballot of their first lanes, so a key that comes
// back later in the wavefront starts a new run.  Every lane of the wavefront
// must call it.
__device__ __forceinline__ double run_sum_to_head(double v, int key, int lane, bool& head) {
  const int kp = __shfl_up(key, 1, 64);
  head = lane == 0 || kp != key;
  const unsigned long long heads = __ballot(head);
  const int run = __popcll(heads & ((2ull << lane) - 1));  // runs starting at lanes 0..lane
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double vd = __shfl_down(v, d, 64);
    const int rd = __shfl_down(run, d, 64);
    if (lane + d < 64 && rd == run) v += vd;
  }
  return v;
}

// One thread per ordered shell pair (A, B) of the class; the thread loops over
// the bra components, so each pair's gradient g = d/dR_A (the ket atom gets
// -g) is summed in registers.  The bra-atom contributions of a wavefront are
// then summed over its runs of equal (molecule, bra atom) -- the pairs of a
// class are ordered by molecule, bra shell, ket shell -- and each run issues
// one atomic per direction; the ket atom takes one atomic per pair.  The
// (d, d) class, which has few pairs and the most registers, keeps one thread
// per (pair, bra component), component-major.
template <int LA, int LB, bool MP, bool SCF>
__global__ __launch_bounds__(256) void k_multipole_vjp(dxtb_basis_t bs, int off, int nitems,
                                                       const double* __restrict__ pos, AdjSrc src,
                                                       double* __restrict__ grad) {
  constexpr int NA = 2 * LA + 1, NB = 2 * LB + 1, NCB = Cart<LB>::n;
  constexpr int NI = LA + 2, NJ = LB + (MP ? 3 : 1);
  constexpr int NCOMB = (LA == 2) ? 3 : 1;
  constexpr int NC = MP ? 10 : 1;
  constexpr bool SPLIT = LA == 2 && LB == 2;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool active = tid < (SPLIT ? nitems * NA : nitems);
  const int ma0 = SPLIT ? tid / nitems : 0;
  const int local = SPLIT ? tid - ma0 * nitems : tid;
  const int ma1 = SPLIT ? ma0 + 1 : NA;
  double g[3] = {0.0, 0.0, 0.0};
  // run key (molecule, bra atom): sorted over the active lanes, which precede
  // the inactive ones (the grid's tail), whose keys are unique
  int key = -1 - lane;
  bool two = false;  // a pair on one atom adds g and -g to that atom: g = 0
  PairCtx c;
  if (active) {
    load_pair(bs, pos, off + local, c, 0);
    key = c.b * bs.nat_max + c.atA;
    two = c.atA != c.atB;
  }
  if (two) {
  const double r2 = c.AB[0] * c.AB[0] + c.AB[1] * c.AB[1] + c.AB[2] * c.AB[2];
  const int row0 = c.row;
  const int n = bs.nao_max;
  const size_t nn = (size_t)n * n;
  for (int ma = ma0; ma < ma1; ++ma) {
  c.row = row0 + ma;
  // adjoint of this row segment, transformed to ket Cartesians and raw moments
  double adj[NCB][NC];
#pragma unroll
  for (int q = 0; q < NCB; ++q)
#pragma unroll
    for (int t = 0; t < NC; ++t) adj[q][t] = 0.0;
  double fab = 0.0, vb = 0.0;
  const double* vdb = nullptr;
  const double* vqb = nullptr;
  if constexpr (SCF) {
    fab = src.fsh[((size_t)c.b * bs.nsh_max + c.ia) * bs.nsh_max + c.ib];
    vb = src.vsh[(size_t)c.b * bs.nsh_max + c.ib];
    if constexpr (MP) {
      vdb = src.vdp + ((size_t)c.b * bs.nat_max + c.atB) * 3;
      vqb = src.vqp + ((size_t)c.b * bs.nat_max + c.atB) * 6;
    }
  }
#pragma unroll
  for (int mb = 0; mb < NB; ++mb) {
    const size_t e0 = (size_t)c.b * nn + (size_t)c.row * n + c.col0 + mb;
    double a[NC];
    double pe = 0.0;
    if constexpr (SCF) {
      pe = src.P[e0];
      a[0] = pe * (fab - vb) - src.W[e0];
    } else {
      a[0] = src.adjS[e0];
    }
    if constexpr (MP) {
      const size_t db = (size_t)c.b * 3 * nn + (size_t)c.row * n + c.col0 + mb;
      const size_t qb = (size_t)c.b * 6 * nn + (size_t)c.row * n + c.col0 + mb;
      double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
      a[1] = a[2] = a[3] = 0.0;
      if constexpr (SCF) {
        a[1] = -pe * vdb[0];
        a[2] = -pe * vdb[1];
        a[3] = -pe * vdb[2];
        t0 = -pe * vqb[0];
        t1 = -pe * vqb[1];
        t2 = -pe * vqb[2];
        t3 = -pe * vqb[3];
        t4 = -pe * vqb[4];
        t5 = -pe * vqb[5];
      } else {
        if (src.adjD != nullptr) {
          a[1] = src.adjD[db];
          a[2] = src.adjD[db + nn];
          a[3] = src.adjD[db + 2 * nn];
        }
        if (src.adjQ != nullptr) {
          t0 = src.adjQ[qb];
          t1 = src.adjQ[qb + nn];
          t2 = src.adjQ[qb + 2 * nn];
          t3 = src.adjQ[qb + 3 * nn];
          t4 = src.adjQ[qb + 4 * nn];
          t5 = src.adjQ[qb + 5 * nn];
        }
      }
      // traceless transform is linear: adjoint onto raw xx, xy, xz, yy, yz, zz
      a[4] = t0 - 0.5 * t2 - 0.5 * t5;
      a[5] = 1.5 * t1;
      a[6] = 1.5 * t3;
      a[7] = -0.5 * t0 + t2 - 0.5 * t5;
      a[8] = 1.5 * t4;
      a[9] = -0.5 * t0 - 0.5 * t2 + t5;
    }
#pragma unroll
    for (int q = 0; q < NCB; ++q) {
      const double tq = trafo(LB, mb, q);
      if (tq != 0.0) {
#pragma unroll
        for (int t = 0; t < NC; ++t) adj[q][t] += tq * a[t];
      }
    }
  }

  const int npA = bs.ush_np[c.uA], npB = bs.ush_np[c.uB];
  const double* alA = bs.ush_alpha + c.uA * DXTB_MAXP;
  const double* alB = bs.ush_alpha + c.uB * DXTB_MAXP;
  const double* coA = bs.ush_coeff + c.uA * DXTB_MAXP;
  const double* coB = bs.ush_coeff + c.uB * DXTB_MAXP;
  for (int i = 0; i < npA; ++i) {
    const double a = alA[i], ca = coA[i];
    for (int j = 0; j < npB; ++j) {
      const double bb = alB[j];
      const double p = a + bb, oop = 1.0 / p, x2p = 0.5 * oop;
      const double kab = ca * coB[j] * exp(-a * bb * oop * r2) * oop * sqrt(oop) * PI32;
      double s[3][NI][NJ];
#pragma unroll
      for (int x = 0; x < 3; ++x) os_table<NI, NJ>(s[x], bb * oop * c.AB[x], -a * oop * c.AB[x], x2p);
#pragma unroll
      for (int k = 0; k < NCOMB; ++k) {
        int ex[3];
        double ew;
        bra_combo<LA>(ma, k, ex, ew);
        double r[3][NJ], d[3][NJ];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          double up[NJ], dn[NJ];
          select_row<NI, NJ>(s[x], ex[x], r[x]);
          select_row<NI, NJ>(s[x], ex[x] + 1, up);
          select_row<NI, NJ>(s[x], ex[x] > 0 ? ex[x] - 1 : 0, dn);
          const double lo = (double)ex[x];
#pragma unroll
          for (int jj = 0; jj < NJ; ++jj) d[x][jj] = 2.0 * a * up[jj] - lo * dn[jj];
        }
        double tx = 0.0, ty = 0.0, tz = 0.0;
#pragma unroll
        for (int q = 0; q < NCB; ++q) {
          const int bx = cart_exp(LB, q, 0), by = cart_exp(LB, q, 1), bz = cart_exp(LB, q, 2);
          // (ex, ey, ez) of the ten moments
          constexpr int E[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                    {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};
#pragma unroll
          for (int t = 0; t < NC; ++t) {
            const double av = adj[q][t];
            const int jx = bx + E[t][0], jy = by + E[t][1], jz = bz + E[t][2];
            tx += av * d[0][jx] * r[1][jy] * r[2][jz];
            ty += av * r[0][jx] * d[1][jy] * r[2][jz];
            tz += av * r[0][jx] * r[1][jy] * d[2][jz];
          }
        }
        const double w = kab * ew;
        g[0] += w * tx;
        g[1] += w * ty;
        g[2] += w * tz;
      }
    }
  }
  }  // bra components
  }  // pair on two atoms
  bool head = false;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double s = run_sum_to_head(g[x], key, lane, head);
    if (head && key >= 0) atomic_add_f64(grad + (size_t)key * 3 + x, s);
  }
  if (two) {
    double* gB = grad + ((size_t)c.b * bs.nat_max + c.atB) * 3;
#pragma unroll
    for (int x = 0; x < 3; ++x) atomic_add_f64(gB + x, -g[x]);
  }
}

template <int LA, int LB>
int launch_fwd(const dxtb_basis_t& bs, const double* pos, const double* fsh, double* S, double* D,
               double* Q, double* H0, int tl, hipStream_t st) {
  const int c = 3 * LA + LB;
  const int off = bs.class_off[c], cnt = bs.class_off[c + 1] - off;
  if (cnt <= 0) return 0;
  const long threads = (long)cnt * (2 * LA + 1);
  const int blk = 256;
  const dim3 grid((unsigned)((threads + blk - 1) / blk));
  if (D != nullptr || Q != nullptr)
    hipLaunchKernelGGL((k_multipole_fwd<LA, LB, true>), grid, dim3(blk), 0, st, bs, off, cnt, pos, fsh, S,
                       D, Q, H0, tl);
  else
    hipLaunchKernelGGL((k_multipole_fwd<LA, LB, false>), grid, dim3(blk), 0, st, bs, off, cnt, pos, fsh,
                       S, D, Q, H0, tl);
  return (int)hipGetLastError();
}

template <int LA, int LB>
int launch_vjp(const dxtb_basis_t& bs, const double* pos, const AdjSrc& src, bool mp, bool scf, double* grad,
               hipStream_t st) {
  const int c = 3 * LA + LB;
  const int off = bs.class_off[c], cnt = bs.class_off[c + 1] - off;
  if (cnt <= 0) return 0;
  const long threads = (LA == 2 && LB == 2) ? (long)cnt * 5 : cnt;  // per pair; (d, d): per (pair, component)
  const int blk = 256;
  const dim3 grid((unsigned)((threads + blk - 1) / blk));
  if (scf) {
    if (mp)
      hipLaunchKernelGGL((k_multipole_vjp<LA, LB, true, true>), grid, dim3(blk), 0, st, bs, off, cnt, pos, src,
                         grad);
    else
      hipLaunchKernelGGL((k_multipole_vjp<LA, LB, false, true>), grid, dim3(blk), 0, st, bs, off, cnt, pos, src,
                         grad);
  } else {
    if (mp)
      hipLaunchKernelGGL((k_multipole_vjp<LA, LB, true, false>), grid, dim3(blk), 0, st, bs, off, cnt, pos, src,
                         grad);
    else
      hipLaunchKernelGGL((k_multipole_vjp<LA, LB, false, false>), grid, dim3(blk), 0, st, bs, off, cnt, pos,
                         src, grad);
  }
  return (int)hipGetLastError();
}

int vjp_all(const dxtb_basis_t& b, const double* pos, const AdjSrc& src, bool mp, bool scf, double* grad,
            hipStream_t st) {
  int e = 0;
  if ((e = launch_vjp<0, 0>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<0, 1>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<0, 2>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<1, 0>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<1, 1>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<1, 2>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<2, 0>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<2, 1>(b, pos, src, mp, scf, grad, st))) return e;
  if ((e = launch_vjp<2, 2>(b, pos, src, mp, scf, grad, st))) return e;
  return 0;
}

// ---------------------------------------------------------------------------
// shell-pair screening bounds
// ---------------------------------------------------------------------------
// Upper bound of the largest |element| of the spherical block of a shell pair
// for the multipole orders e = 0, 1, 2 about the origin C (the reference's
// _bounds, impls/pairs.py:411-494): per primitive pair
//   |c_i c_j| exp(-mu R^2) sum_{i<=la, j<=lb, k<=e} C(la,i) C(lb,j) C(e,k)
//     |PA|^i |PB|^j |PC|^k 2 pi Gamma((n+3)/2) p^(-(n+3)/2),
// n = (la-i) + (lb-j) + (e-k), |PA| = b R / p, |PB| = a R / p and |PC| <=
// max(|AC|, |BC|), summed over the primitive pairs and scaled by the largest
// absolute row sums of the two spherical transforms (1, 1, 2 for s, p, d)
// times 1 + 1e-10.  Both orders of a pair are evaluated as the reference's
// lower triangle orders them (bra = the larger shell index), so (A, B) and
// (B, A) get bit-identical bounds.  One thread per work-list item; out[4 i +
// 3] is 1 when both shells sit at the same position.
__device__ __forceinline__ double binom2(int n, int k) { return (n == 2 && k == 1) ? 2.0 : 1.0; }
__device__ __forceinline__ double sph_row_norm(int l) { return l == 2 ? 2.0 : 1.0; }

__global__ __launch_bounds__(256) void k_pair_bounds(dxtb_basis_t bs, int nitems, const double* __restrict__ pos,
                                                     double cx, double cy, double cz, double* __restrict__ out) {
  const int item = blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= nitems) return;
  const int b = bs.item_mol[item];
  const int pr = bs.item_pair[item];
  const int i0 = pr >> 16, j0 = pr & 0xffff;
  const int sa = max(i0, j0), sb = min(i0, j0);
  const int base = b * bs.nsh_max;
  const int atA = bs.sh_atom[base + sa], atB = bs.sh_atom[base + sb];
  const int uA = bs.sh_ush[base + sa], uB = bs.sh_ush[base + sb];
  const int la = bs.ush_l[uA], lb = bs.ush_l[uB];
  const double* pA = pos + ((size_t)b * bs.nat_max + atA) * 3;
  const double* pB = pos + ((size_t)b * bs.nat_max + atB) * 3;
  const double c[3] = {cx, cy, cz};
  double r2 = 0.0, ac2 = 0.0, bc2 = 0.0;
  bool same = true;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double d = pB[x] - pA[x], da = pA[x] - c[x], db = pB[x] - c[x];
    r2 += d * d;
    ac2 += da * da;
    bc2 += db * db;
    same = same && pA[x] == pB[x];
  }
  const double r = sqrt(r2), xc = fmax(sqrt(ac2), sqrt(bc2));
  const double xcp[3] = {1.0, xc, xc * xc};
  // 2 pi Gamma((n + 3) / 2), n = 0..6
  constexpr double SQPI = 1.7724538509055160273;
  constexpr double TPI = 6.283185307179586477;
  const double g[7] = {TPI * 0.5 * SQPI, TPI * 1.0, TPI * 0.75 * SQPI, TPI * 2.0,
                       TPI * 1.875 * SQPI, TPI * 6.0, TPI * 6.5625 * SQPI};
  const int npA = bs.ush_np[uA], npB = bs.ush_np[uB];
  const double* alA = bs.ush_alpha + uA * DXTB_MAXP;
  const double* alB = bs.ush_alpha + uB * DXTB_MAXP;
  const double* coA = bs.ush_coeff + uA * DXTB_MAXP;
  const double* coB = bs.ush_coeff + uB * DXTB_MAXP;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < npA; ++i) {
    const double a = alA[i];
    for (int j = 0; j < npB; ++j) {
      const double bb = alB[j];
      const double p = a + bb, oop = 1.0 / p, sq = sqrt(oop);
      const double pref = fabs(coA[i] * coB[j]) * exp(-a * bb * oop * r * r);
      const double xa = bb * oop * r, xb = a * oop * r;
      double mom[7];
      double pw = oop * sq;  // p^(-3/2)
#pragma unroll
      for (int n = 0; n < 7; ++n) {
        mom[n] = g[n] * pw;
        pw *= sq;
      }
      const double xap[3] = {1.0, xa, xa * xa}, xbp[3] = {1.0, xb, xb * xb};
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        double poly = 0.0;
        for (int ii = 0; ii <= la; ++ii)
          for (int jj = 0; jj <= lb; ++jj)
#pragma unroll
            for (int k = 0; k <= e; ++k)
              poly += binom2(la, ii) * binom2(lb, jj) * binom2(e, k) * xap[ii] * xbp[jj] * xcp[k] *
                      mom[(la - ii) + (lb - jj) + (e - k)];
        acc[e] += pref * poly;
      }
    }
  }
  const double sph = sph_row_norm(la) * sph_row_norm(lb) * (1.0 + 1e-10);
#pragma unroll
  for (int e = 0; e < 3; ++e) out[(size_t)item * 4 + e] = acc[e] * sph;
  out[(size_t)item * 4 + 3] = same ? 1.0 : 0.0;
}

bool basis_ok(const dxtb_basis_t* b) {
  if (b == nullptr) return false;
  if (b->nbatch <= 0 || b->nao_max <= 0 || b->nsh_max <= 0 || b->nat_max <= 0) return false;
  if (b->nsh_max > 0xffff) return false;
  if (!b->sh_atom || !b->sh_ush || !b->sh_ao || !b->ush_np || !b->ush_alpha || !b->ush_coeff) return false;
  if (b->class_off[9] > 0 && (!b->item_mol || !b->item_pair)) return false;
  for (int c = 0; c < 9; ++c)
    if (b->class_off[c + 1] < b->class_off[c]) return false;
  return true;
}

// Padding of nmat stacked (n, n) matrices per molecule: rows and columns at
// or beyond nao[b].  One workgroup per matrix (with a workgroup per 16 rows
// the dispatch of the many small workgroups took longer than the writes),
// its four wavefronts taking the rows in turn, lanes along the row
// (coalesced writes); a row inside the molecule has only its columns from
// nao[b] on written.
__global__ __launch_bounds__(256) void k_zero_padding(int n, int nmat, const int* __restrict__ nao,
                                                      double* __restrict__ X) {
  const long mat = blockIdx.x;  // b * nmat + component
  const int m = min(max(nao[mat / nmat], 0), n);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = w; r < n; r += 4) {
    double* row = X + ((size_t)mat * n + r) * n;
    for (int c = (r < m ? m : 0) + lane; c < n; c += 64) row[c] = 0.0;
  }
}

}  // namespace

extern "C" int dxtb_zero_padding(int nbatch, int n, int nmat, const int* nao, double* X, hipStream_t stream) {
  if (nbatch < 0 || n < 0 || nmat <= 0 || (nbatch > 0 && n > 0 && (nao == nullptr || X == nullptr)))
    return DXTB_EINVAL;
  if (nbatch == 0 || n == 0) return 0;
  const long nblocks = (long)nbatch * nmat;
  if (nblocks > 0x7fffffffL) return DXTB_EINVAL;
  hipLaunchKernelGGL(k_zero_padding, dim3((unsigned)nblocks), dim3(256), 0, stream, n, nmat, nao, X);
  return (int)hipGetLastError();
}

extern "C" int dxtb_multipole_fwd(const dxtb_basis_t* bas, const double* pos, const double* fsh, double* S,
                                  double* D, double* Q, double* H0, int traceless, hipStream_t stream) {
  if (!basis_ok(bas) || pos == nullptr || S == nullptr) return DXTB_EINVAL;
  if (H0 != nullptr && fsh == nullptr) return DXTB_EINVAL;
  const dxtb_basis_t& b = *bas;
  int e = 0;
  if ((e = launch_fwd<0, 0>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<0, 1>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<0, 2>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<1, 0>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<1, 1>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<1, 2>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<2, 0>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<2, 1>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  if ((e = launch_fwd<2, 2>(b, pos, fsh, S, D, Q, H0, traceless, stream))) return e;
  return 0;
}

extern "C" int dxtb_multipole_vjp(const dxtb_basis_t* bas, const double* pos, const double* adjS,
                                  const double* adjD, const double* adjQ, double* grad, hipStream_t stream) {
  if (!basis_ok(bas) || pos == nullptr || adjS == nullptr || grad == nullptr) return DXTB_EINVAL;
  AdjSrc src{adjS, adjD, adjQ, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  return vjp_all(*bas, pos, src, adjD != nullptr || adjQ != nullptr, false, grad, stream);
}

extern "C" int dxtb_forces_multipole(const dxtb_basis_t* bas, const double* pos, const double* P, const double* W,
                                     const double* fsh, const double* vsh, const double* vdp, const double* vqp,
                                     double* grad, hipStream_t stream) {
  if (!basis_ok(bas) || !pos || !P || !W || !fsh || !vsh || !grad) return DXTB_EINVAL;
  if ((vdp == nullptr) != (vqp == nullptr)) return DXTB_EINVAL;
  AdjSrc src{nullptr, nullptr, nullptr, P, W, fsh, vsh, vdp, vqp};
  return vjp_all(*bas, pos, src, vdp != nullptr, true, grad, stream);
}

extern "C" int dxtb_pair_bounds(const dxtb_basis_t* bas, const double* pos, const double* origin, double* out,
                                hipStream_t stream) {
  if (!basis_ok(bas) || pos == nullptr || out == nullptr || bas->ush_l == nullptr) return DXTB_EINVAL;
  const int n = bas->class_off[9];
  if (n <= 0) return 0;
  const double cx = origin ? origin[0] : 0.0, cy = origin ? origin[1] : 0.0, cz = origin ? origin[2] : 0.0;
  hipLaunchKernelGGL(k_pair_bounds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *bas, n, pos, cx, cy,
                     cz, out);
  return (int)hipGetLastError();
}
