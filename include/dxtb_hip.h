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
 * The Python host (dxtb_amd) binds this header with ctypes; INTEGRATION.md
 * shows the binding.
 */
#ifndef DXTB_HIP_H
#define DXTB_HIP_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DXTB_OK 0
#define DXTB_EINVAL 1001
#define DXTB_EUNSUP 1002
#define DXTB_ENOCONV 1003

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
 * may be NULL.  Outputs must be zeroed by the caller (padding is not written). */
int dxtb_multipole_fwd(const dxtb_basis_t* bas, const double* pos, const double* fsh,
                       double* S, double* D, double* Q, double* H0, hipStream_t stream);

/* Nuclear gradient of sum(adjS*S) + sum(adjD*D) + sum(adjQ*Q) (the VJP of
 * dxtb_multipole_fwd).  adjD/adjQ may be NULL.  grad (nbatch, nat_max, 3) is
 * accumulated into (caller zeroes it).  Replaces autograd through the
 * reference pair builder (src/dxtb/_src/integral/driver/pytorch/impls/pairs.py). */
int dxtb_multipole_vjp(const dxtb_basis_t* bas, const double* pos, const double* adjS,
                       const double* adjD, const double* adjQ, double* grad, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
