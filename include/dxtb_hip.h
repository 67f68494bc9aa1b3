/*
 * dxtb_hip -- C ABI of the MI355X (gfx950) GFN1/GFN2-xTB single-point hot path.
 *
 * Every entry point takes device pointers (FP64 data, int32 index tables) and
 * a hipStream_t, enqueues its kernels on that stream and returns 0 on success,
 * a hipError_t value on a HIP failure, or one of the DXTB_E* codes below.  No
 * entry point allocates device memory except where stated; workspaces are
 * passed in by the caller.  Batched layouts are zero padded:
 *   positions   (nbatch, nat_max, 3)
 *   matrices    (nbatch, [ncomp,] nao_max, nao_max), row-major
 *   shell/atom vectors (nbatch, nsh_max) / (nbatch, nat_max)
 * The AO order is the reference's: atoms in input order, shells in parameter
 * order, m = -l..l per shell (p as y, z, x).
 *
 * The Python host binds this header with ctypes (dxtb_amd/_lib.py).
 */
#ifndef DXTB_HIP_H
#define DXTB_HIP_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DXTB_OK 0
#define DXTB_EINVAL 1001
#define DXTB_EUNSUP 1002
#define DXTB_ENOCONV 1003
#define DXTB_ESCRATCH 1004 /* request exceeds the fixed scratch buffer of a no-allocation entry point */

/* Device-resident description of a padded batch of molecules and its
 * shell-pair work lists (built once per set of atomic numbers). */
typedef struct dxtb_basis {
  int nbatch, nat_max, nsh_max, nao_max;
  const int* sh_atom; /* (nbatch, nsh_max) atom of each shell */
  const int* sh_ush;  /* (nbatch, nsh_max) unique-shell id (row of the tables below) */
  const int* sh_ao;   /* (nbatch, nsh_max) first AO of each shell */
  int nush;           /* number of unique shells */
  const int* ush_l;   /* (nush) angular momentum */
  const int* ush_np;  /* (nush) primitives (<= 8) */
  const double* ush_alpha; /* (nush, 8) exponents */
  const double* ush_coeff; /* (nush, 8) normalized contraction coefficients */
  /* ordered shell pairs (bra A, ket B) grouped by class c = 3*l(A) + l(B);
   * items class_off[c] .. class_off[c+1]-1 */
  const int* item_mol;  /* molecule of each pair */
  const int* item_pair; /* (A << 16) | B */
  int class_off[10];
} dxtb_basis_t;

/* Overlap S, dipole D (3) centred on the ket atom, traceless quadrupole Q
 * (6: xx, yx, yy, zx, zy, zz) centred on the ket atom and optionally
 * H0 = S * fsh[shell pair].  Replaces the reference's pair builder
 * (src/dxtb/_src/integral/driver/pytorch/impls/pairs.py:228-399,
 * pipeline.py:151-235, kernels/os.py:61-163) followed by the container
 * post-processing (integral/container.py:182-535: normalisation, r0->rj
 * shift, traceless) and the H0 build (xtb/base.py:255-372).  D, Q, H0 and fsh
 * may be NULL.  traceless = 0 writes the raw ket-centred second moments in the
 * same component order instead of the traceless form.  Outputs must be zeroed
 * by the caller (padding is not written). */
int dxtb_multipole_fwd(const dxtb_basis_t* bas, const double* pos, const double* fsh,
                       double* S, double* D, double* Q, double* H0, int traceless, hipStream_t stream);

/* Zeroes the padding of X (nbatch, nmat, n, n): in each matrix of molecule b
 * the rows and columns at or beyond nao[b] (nao: (nbatch) device ints).  With
 * the complete shell-pair work lists of a batch, dxtb_multipole_fwd writes
 * every element of each molecule's leading (nao[b], nao[b]) block, so this
 * replaces zeroing the whole output beforehand (the reference allocates its
 * matrices as zeros, impls/pairs.py:218); screened work lists leave blocks
 * unwritten and need the zeroed output. */
int dxtb_zero_padding(int nbatch, int n, int nmat, const int* nao, double* X, hipStream_t stream);

/* Nuclear gradient of sum(adjS*S) + sum(adjD*D) + sum(adjQ*Q) (the VJP of
 * dxtb_multipole_fwd).  adjD/adjQ may be NULL.  grad (nbatch, nat_max, 3) is
 * accumulated into (caller zeroes it).  Replaces autograd through the
 * reference pair builder (src/dxtb/_src/integral/driver/pytorch/impls/pairs.py). */
int dxtb_multipole_vjp(const dxtb_basis_t* bas, const double* pos, const double* adjS,
                       const double* adjD, const double* adjQ, double* grad, hipStream_t stream);

/* Shell-pair screening bounds of the work lists (screening is off unless a
 * caller asks for it, as in the reference).  For every item of bas (class
 * order, class_off[9] items) out (nitems, 4) receives the rigorous upper
 * bounds of the largest |element| of the pair's spherical block for the
 * multipole orders 0, 1 and 2 about origin (a host pointer to 3 doubles, NULL:
 * the Cartesian origin), then 1.0 if both shells sit at the same position
 * (else 0.0).  Replaces the reference's _bounds and the same-position test of
 * _screen (src/dxtb/_src/integral/driver/pytorch/impls/pairs.py:411-541);
 * the caller keeps the items with any bound above its threshold or at one
 * position (select_pairs, pairs.py:544-614). */
int dxtb_pair_bounds(const dxtb_basis_t* bas, const double* pos, const double* origin, double* out,
                     hipStream_t stream);

/* Method parameters as device tables.  Per-unique-shell arrays use the row
 * order of dxtb_basis_t.ush_*; per-element arrays are indexed by the atomic
 * number (0..86, row 0 unused). */
typedef struct dxtb_params {
  int gfn;                 /* 1 or 2 */
  int es3_shell;           /* 1: shell-resolved third order (GFN2), 0: atom (GFN1) */
  int es2_average;         /* 0 harmonic, 1 arithmetic, 2 geometric */
  int rep_light;           /* 1: light-atom repulsion exponent rep_klight for H/He pairs */
  double enscale, es2_gexp, rep_kexp, rep_klight; /* es2_gexp > 0 (2 for GFN1/GFN2) */
  double aes_dmp3, aes_dmp5, aes_kexp, aes_shift, aes_rmax;
  const double* ush_lev;    /* (nush) shell level * EV2AU */
  const double* ush_kcn;    /* (nush) kcn * EV2AU */
  const double* ush_shpoly; /* (nush) */
  const double* ush_hub;    /* (nush) gam(Z) * lgam(shell) */
  const double* ush_hd3;    /* (nush) GFN2: gam3(Z) * shell scale */
  const int* ush_valence;   /* (nush) 1 valence / 0 not */
  const double* hscale;     /* (nush, nush) */
  const double* z_rad;      /* atomic radii of the H0 polynomial (Bohr) */
  const double* z_en;       /* electronegativity */
  const double* z_rcov;     /* D3 covalent radii (Bohr) */
  const double* kpair;      /* (87, 87) */
  const double* z_gam3;     /* GFN1 atom third-order */
  const double* z_arep;
  const double* z_zeff;
  const double* z_mprad;
  const double* z_mpvcn;
  const double* z_dkernel;
  const double* z_qkernel;
} dxtb_params_t;

/* Per-geometry set-up: coordination numbers cn (nbatch, nat_max); H0
 * shell-pair factor fsh with its radial derivative dfsh and d f / d h
 * (hfsh), all (nbatch, nsh_max, nsh_max); per-atom repulsion energy erep
 * (nbatch, nat_max); AES2 multipole radii mrad (nbatch, nat_max, GFN2 only,
 * may be NULL otherwise).
 * Replaces tad-mctc cn_d3 as called by xtb/gfn1.py:58-64 and xtb/gfn2.py:58-64,
 * the H0 factors of xtb/base.py:255-372, repulsion/base.py:293-358 and the
 * mrad of coulomb/multipole.py:320-333.  nsh: (nbatch) shells per molecule. */
int dxtb_setup(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers, const double* pos,
               const int* nsh, double* cn, double* fsh, double* dfsh, double* hfsh, double* erep,
               double* mrad, hipStream_t stream);

/* Fock matrix F (nbatch, nao_max, nao_max) from H0, S, ket-centred D and Q
 * and the shell/atom potentials vsh (nbatch, nsh_max), vat (nbatch, nat_max),
 * vdp (nbatch, nat_max, 3), vqp (nbatch, nat_max, 6).  D/Q/vdp/vqp may be
 * NULL (GFN1).  Padding rows/columns get F_ii = pad_diag.
 * Replaces scf/pure/conversions.py:255-297 (potential_to_hamiltonian).
 * ao_shell/ao_atom: (nbatch, nao_max) maps, nao: (nbatch). */
int dxtb_fock(const dxtb_basis_t* bas, const int* nao, const int* ao_shell, const int* ao_atom,
              const double* pos, const double* H0, const double* S, const double* D, const double* Q,
              const double* vsh, const double* vat, const double* vdp, const double* vqp, double pad_diag,
              double* F, hipStream_t stream);

/* Mulliken populations and atomic multipoles of a density P:
 * qsh (nbatch, nsh_max) = n0-sum shell charges, qat (nbatch, nat_max),
 * qdp (nbatch, nat_max, 3), qqp (nbatch, nat_max, 6) and the per-atom
 * band energy eh0 (nbatch, nat_max) = sum diag(P H0).  n0: (nbatch, nao_max)
 * reference occupations.  at_sh0/at_nsh: first shell / shell count per atom.
 * work: dxtb_charges_work_size(nbatch, nao_max) bytes of scratch (the
 * column sums).
 * Replaces scf/pure/conversions.py:203-252 (density_to_charges). */
size_t dxtb_charges_work_size(int nbatch, int nao_max);
int dxtb_charges(const dxtb_basis_t* bas, const int* nao, const int* nat, const int* at_sh0, const int* at_nsh,
                 const double* n0, const double* P, const double* S, const double* H0, const double* D,
                 const double* Q, double* qsh, double* qat, double* qdp, double* qqp, double* eh0,
                 double* work, hipStream_t stream);

/* Potentials and interaction energies of ES2 + ES3 (+ AES2 for GFN2) from
 * the charges: vsh (nbatch, nsh_max), vat (nbatch, nat_max), vdp, vqp;
 * per-shell esh and per-atom eat energies.
 * Replaces components/interactions/base.py:135-181 with coulomb/secondorder.py,
 * thirdorder.py and multipole.py. */
int dxtb_potential(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers, const double* pos,
                   const int* nsh, const double* mrad, const double* qsh, const double* qat, const double* qdp,
                   const double* qqp, double* vsh, double* vat, double* vdp, double* vqp, double* esh,
                   double* eat, hipStream_t stream);

/* Fermi-Dirac occupations occ (nbatch, 2, nao_max) of ascending orbital
 * energies emo (nbatch, nao_max) for nab (nbatch, 2) electrons at kT, and the
 * electronic free-energy term ts (nbatch) = kT sum f ln f + (1-f) ln(1-f).
 * Replaces wavefunction/filling.py:428-742 and scf/base.py:554-619. */
int dxtb_fermi(int nbatch, int nao_max, const int* nao, const double* emo, const double* nab, double kt,
               double* occ, double* ts, hipStream_t stream);

/* Integral part of the nuclear gradient of the converged SCF free energy:
 * sum over all AO pairs of adjS dS/dR + adjD dD/dR + adjQ dQ/dR with the
 * variational adjoints built on the fly from the density P and the
 * energy-weighted density W (nbatch, nao_max, nao_max):
 *   adjS = P (fsh - v_ket) - W, adjD = -P vdp(ket atom), adjQ = -P vqp(ket atom).
 * vsh: total shell potential (nbatch, nsh_max); vdp/vqp NULL for GFN1.
 * grad (nbatch, nat_max, 3) is accumulated into.  Replaces the autograd
 * backward of the reference through its integrals (calculators/types/autograd.py:82-193);
 * the formula is the one of the analytic GFN1 gradient (xtb/gfn1.py:311-313). */
int dxtb_forces_multipole(const dxtb_basis_t* bas, const double* pos, const double* P, const double* W,
                          const double* fsh, const double* vsh, const double* vdp, const double* vqp,
                          double* grad, hipStream_t stream);

/* Remaining gradient terms at fixed charges: H0 distance polynomial and
 * self-energy (CN) terms, ES2 gamma(R), AES2 matrices and radii (GFN2),
 * repulsion and the coordination-number chain.  dEdh (nbatch, nsh_max) and
 * dEdcn (nbatch, nat_max) are workspaces (outputs for inspection); grad is
 * accumulated into.  Replaces the autograd backward through
 * xtb/base.py:255-372, coulomb/secondorder.py:799-870, coulomb/multipole.py:270-449,
 * repulsion/base.py:293-430 and tad-mctc cn_d3. */
int dxtb_forces_classical(const dxtb_basis_t* bas, const dxtb_params_t* par, const int* numbers,
                          const double* pos, const int* nsh, const int* at_sh0, const int* at_nsh,
                          const double* P, const double* S, const double* cn, const double* mrad,
                          const double* dfsh, const double* hfsh, const double* qsh, const double* qat,
                          const double* qdp, const double* qqp, double* dEdh, double* dEdcn, double* grad,
                          hipStream_t stream);

/* ---- generalized Born implicit solvation ------------------------------
 * Replaces the reference's GeneralizedBorn interaction
 * (components/interactions/solvation/alpb.py:187-350: get_cache,
 * get_monopole_atom_potential/energy) and get_born_radii (solvation/born.py:
 * 52-239); the forces replace autograd through both.  Per molecule b:
 * keps[b] = (1/eps - 1) / (1 + alpbet) (0: no solvation, molecule skipped),
 * kind[b] = 1 (P16 kernel) or 2 (Still kernel).  rvdw_z: van der Waals radius
 * by atomic number (Bohr, device).  nat_max <= 1024. */
typedef struct {
  double descreening; /* 0.8 */
  double cutoff;      /* 66 Bohr */
  double born_scale;  /* 1 */
  double born_offset; /* 0 */
  double obc[3];      /* 1.0, 0.8, 4.85 */
} dxtb_gb_params_t;

/* Born radii born (B, nat_max) and dB/dpsi dborn (B, nat_max). */
int dxtb_gb_born(int nbatch, int nat_max, const int* nat, const int* numbers, const double* pos, const double* keps,
                 const double* rvdw_z, const dxtb_gb_params_t* par, double* born, double* dborn, hipStream_t stream);

/* Adds V_i = keps sum_j K(r_ij, B_i B_j) q_j to vat and q_i V_i / 2 to eat. */
int dxtb_gb_potential(int nbatch, int nat_max, const int* nat, const double* pos, const double* keps, const int* kind,
                      const double* born, const double* qat, double* vat, double* eat, hipStream_t stream);

/* Adds dE_GB/dR at fixed charges (direct pair terms and the Born-radius
 * chain) to grad (B, nat_max, 3). */
int dxtb_gb_forces(int nbatch, int nat_max, const int* nat, const int* numbers, const double* pos, const double* keps,
                   const int* kind, const double* rvdw_z, const dxtb_gb_params_t* par, const double* born,
                   const double* dborn, const double* qat, double* grad, hipStream_t stream);

/* Fixed device workspace of dxtb_jacobi_eigh_noalloc, per GPU (64 MiB). */
#define DXTB_JACOBI_SCRATCH_BYTES ((size_t)64 << 20)

/* dxtb_jacobi_eigh without a workspace argument and without any device
 * allocation: the workspace is a fixed array of DXTB_JACOBI_SCRATCH_BYTES in
 * device memory, one per GPU, kept until the process ends; the split form's
 * launches are always issued directly (no recorded graphs).  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Arguments and results as dxtb_jacobi_eigh.  A call whose
 * workspace (dxtb_jacobi_work_size(nbatch, nmax)) exceeds the array returns
 * DXTB_ESCRATCH and launches nothing, e.g. beyond 64 matrices of order 256
 * or 160 of order 1536.  All calls on one GPU share its array, so
 * calls that overlap on one GPU (different streams, or threads) race: order
 * them on one stream. */
int dxtb_jacobi_eigh_noalloc(int nbatch, int nmax, const int* nsize, double* A, double* VT, double* evals,
                             int variant, int max_sweeps, int split_sweeps, double tol, const double* tol_b,
                             int* sweeps, hipStream_t stream);

/* Batched FP64 product on the matrix cores, for every matrix b of a batch of
 * padded (n, n) row-major matrices:
 *   Out_b = alpha * op(A_b) diag(scale_b) op(B_b) + beta * Cin_b
 * op(X) = X or X^T (transa/transb != 0).  Only the leading (m, m) block,
 * m = nsize[b] (n if nsize is NULL), is computed, with the sum over k running
 * to m: the operands must be block diagonal with respect to the padding (the
 * SCF's matrices are).  Outside that block Out gets pad on the diagonal and
 * 0 elsewhere when fill_pad != 0, else it is left as it is.  sym != 0: the
 * product is symmetric; only the tiles on and above the diagonal are computed
 * and written to both triangles (Cin must then be symmetric too).  scale
 * (nbatch, n) may be NULL; Cin may be NULL when beta == 0 and may alias Out;
 * Out must not alias A or B.
 * Replaces the dense products of the SCF step: the reference's
 * L^-1 F L^-T and L^-T C of xitorch exacteig (exlibs/xitorch/_impls/linalg/
 * symeig.py:38-75 via scf/pure/ovlp_diag.py:36-62) in the previous orbital
 * basis, and the densities of scf/utils.py:29-52 (get_density). */
int dxtb_gemm_batched(int nbatch, int n, const int* nsize, int transa, int transb, double alpha, const double* A,
                      const double* B, const double* scale, double beta, const double* Cin, double* Out, int sym,
                      int fill_pad, double pad, hipStream_t stream);

/* Bytes of the workspace of dxtb_broyden_step for histories of H entries
 * (per molecule 3 H coefficients and the partial combinations,
 * ceil(H / 64) x 3 x L doubles). */
size_t dxtb_broyden_work_size(int nbatch, int L, int H);

/* One iteration of the per-molecule Broyden mixer of the SCF, after the
 * caller has evaluated the SCF map at xe (gx = g(xe)), for nbatch molecules
 * with potential vectors of length L (row-major (nbatch, L) arrays):
 * residual ynew = xe - gx; a fresh molecule (ctl started == 0) keeps x = xe,
 * y = ynew; otherwise the first-Broyden rank-one update of its inverse
 * Jacobian from ddx = xe - x, ddy = ynew - y (history (nbatch, H, L) in hc
 * and hd, its length in ctl), x = xe, y = ynew, and the convergence test
 * |dx| < x_tol and |ynew| < f_tol (dx the step that led to xe).  The next
 * point xe = x - damp G ynew is written for every molecule that goes on.
 * Inner products and norms are weighted by w (nbatch, L).
 * ctl (nbatch, 4) int32: occupied, started, history length, iterations;
 * a molecule with occupied == 0 is skipped; one that converges (flag 1) or
 * reaches maxiter iterations (flag 2) gets occupied = 0.  dxn (nbatch): norm
 * of the last step.  work: dxtb_broyden_work_size(nbatch, L, H) bytes.
 * H <= 4096, nbatch <= 65535.  kbound (<= 0: H) is an upper bound the caller
 * knows for the history lengths; it sizes the launch grids only (every
 * history length gives the same result).
 * Replaces the broyden1 root finder of the reference's implicit SCF
 * (exlibs/xitorch/_impls/optimize/root/rootsolver.py:42-261 with
 * _jacobian.py:51-120, 250-316; alpha = -0.5, damp = 0.5 as set in
 * scf/base.py:268-287), applied per molecule. */
int dxtb_broyden_step(int nbatch, int L, int H, const double* w, const double* gx, double* xe, double* x, double* y,
                      double* hc, double* hd, int* ctl, int* flag, double* dxn, double* work, double alpha,
                      double damp, double x_tol, double f_tol, int maxiter, int kbound, hipStream_t stream);

/* Largest matrix order handled by dxtb_jacobi_eigh. */
int dxtb_jacobi_max_n(void);

/* Bytes of the workspace dxtb_jacobi_eigh needs for the blocked variants
 * (sized for a batch capacity: nbatch rounded up to a power of two up to 8,
 * a multiple of 8 up to 64, else a multiple of 32). */
size_t dxtb_jacobi_work_size(int nbatch, int nmax);

/* Batched symmetric eigensolver, cyclic two-sided Jacobi.  variant 1: scalar
 * Jacobi with A in LDS, one workgroup per matrix (nmax <= 128); variant 2:
 * blocked Jacobi (16-row blocks, FP64 matrix-core updates, A in global
 * memory), one workgroup per matrix; variant 3: the blocked Jacobi whose
 * first split_sweeps sweeps run as grid-wide launches over many workgroups
 * per matrix (for a few large matrices; converged matrices skip the rest),
 * finished by variant 2; variant 0 picks 3 for nmax >= 128 and
 * nbatch < nmax / 2 (where it is faster on a lone batch), else 2 (the
 * scalar variant, slower than the blocked ones at every order measured, only
 * when asked for).  Variants 2 and 3 give bit-identical results for any
 * split_sweeps.  split_sweeps (>= 0) is best set to
 * the sweeps the caller expects (e.g. of the previous SCF iteration).
 * A (nbatch, nmax, nmax): only the leading nsize[b] block of every matrix is
 * used; the blocked variant overwrites it.  VT (nbatch, nmax, nmax) receives
 * the eigenvectors as ROWS (identity on the padding), evals (nbatch, nmax)
 * the unsorted eigenvalues (0 on the padding).  work: device buffer of
 * dxtb_jacobi_work_size bytes (may be NULL for variant 1).  Stops after
 * max_sweeps, when max|offdiag| <= tol * max(1, max|diag|), or when a sweep
 * no longer halves max|offdiag| within 1000x of that bar (rounding floor);
 * tol_b (nbatch, device, may be NULL) replaces tol per matrix; sweeps
 * (nbatch) may be NULL.  On a stream other than the null stream that is not
 * being captured, the split launches are recorded once per (workspace, batch
 * capacity, nmax, split sweeps, max_sweeps, tol) and replayed as one graph
 * (environment DXTB_JACOBI_GRAPHS=0: direct launches; DXTB_JACOBI_WAVE1=0 or
 * 1 fixes the split step's sweep form, four or one wavefront per block pair,
 * which the library otherwise picks by batch capacity x block pairs, same
 * results either way); calls that share a
 * workspace must be ordered (one stream).  Replaces the per-iteration eigensolve of
 * scf/pure/ovlp_diag.py:36-62 (xitorch exacteig / torch.linalg.eigh). */
int dxtb_jacobi_eigh(int nbatch, int nmax, const int* nsize, double* A, double* VT, double* evals, double* work,
                     int variant, int max_sweeps, int split_sweeps, double tol, const double* tol_b, int* sweeps,
                     hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
