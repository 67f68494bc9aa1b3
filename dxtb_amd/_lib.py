"""
ctypes binding of the HIP library (C ABI of ``include/dxtb_hip.h``).

The library is loaded from the package directory.  If it is missing the
import of any compute function fails loudly: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_size_t, c_void_p
from pathlib import Path

import torch

LIBPATH = Path(__file__).resolve().parent / "libdxtb_hip.so"

DXTB_EINVAL = 1001
DXTB_EUNSUP = 1002
DXTB_ENOCONV = 1003
DXTB_ESCRATCH = 1004

_ERRORS = {
    DXTB_EINVAL: "invalid argument",
    DXTB_EUNSUP: "unsupported option",
    DXTB_ENOCONV: "iterative solve did not converge",
    DXTB_ESCRATCH: "request exceeds the fixed scratch buffer of the no-allocation entry point",
}


class DxtbBasis(ctypes.Structure):
    _fields_ = [
        ("nbatch", c_int),
        ("nat_max", c_int),
        ("nsh_max", c_int),
        ("nao_max", c_int),
        ("sh_atom", c_void_p),
        ("sh_ush", c_void_p),
        ("sh_ao", c_void_p),
        ("nush", c_int),
        ("ush_l", c_void_p),
        ("ush_np", c_void_p),
        ("ush_alpha", c_void_p),
        ("ush_coeff", c_void_p),
        ("item_mol", c_void_p),
        ("item_pair", c_void_p),
        ("class_off", c_int * 10),
    ]


class DxtbParams(ctypes.Structure):
    _fields_ = [
        ("gfn", c_int),
        ("es3_shell", c_int),
        ("es2_average", c_int),
        ("rep_light", c_int),
        ("enscale", c_double),
        ("es2_gexp", c_double),
        ("rep_kexp", c_double),
        ("rep_klight", c_double),
        ("aes_dmp3", c_double),
        ("aes_dmp5", c_double),
        ("aes_kexp", c_double),
        ("aes_shift", c_double),
        ("aes_rmax", c_double),
    ] + [
        (name, c_void_p)
        for name in (
            "ush_lev", "ush_kcn", "ush_shpoly", "ush_hub", "ush_hd3", "ush_valence", "hscale", "z_rad",
            "z_en", "z_rcov", "kpair", "z_gam3", "z_arep", "z_zeff", "z_mprad", "z_mpvcn", "z_dkernel",
            "z_qkernel",
        )
    ]


# name -> argtypes; every entry point returns int
class DxtbGbParams(ctypes.Structure):
    """dxtb_gb_params_t"""

    _fields_ = [("descreening", c_double), ("cutoff", c_double), ("born_scale", c_double), ("born_offset", c_double),
                ("obc", c_double * 3)]


_SIGNATURES: dict[str, list] = {
    "dxtb_multipole_fwd": [POINTER(DxtbBasis)] + [c_void_p] * 6 + [c_int, c_void_p],
    "dxtb_zero_padding": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p],
    "dxtb_multipole_vjp": [POINTER(DxtbBasis), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "dxtb_pair_bounds": [POINTER(DxtbBasis), c_void_p, c_void_p, c_void_p, c_void_p],
    "dxtb_setup": [POINTER(DxtbBasis), POINTER(DxtbParams)] + [c_void_p] * 9 + [c_void_p],
    "dxtb_fock": [POINTER(DxtbBasis)] + [c_void_p] * 12 + [c_double, c_void_p, c_void_p],
    "dxtb_charges_work_size": [c_int, c_int],
    "dxtb_charges": [POINTER(DxtbBasis)] + [c_void_p] * 17,
    "dxtb_potential": [POINTER(DxtbBasis), POINTER(DxtbParams)] + [c_void_p] * 14 + [c_void_p],
    "dxtb_fermi": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p],
    "dxtb_forces_multipole": [POINTER(DxtbBasis)] + [c_void_p] * 9,
    "dxtb_jacobi_max_n": [],
    "dxtb_gb_born": [c_int, c_int] + [c_void_p] * 5 + [POINTER(DxtbGbParams)] + [c_void_p] * 3,
    "dxtb_gb_potential": [c_int, c_int] + [c_void_p] * 9,
    "dxtb_gb_forces": [c_int, c_int] + [c_void_p] * 6 + [POINTER(DxtbGbParams)] + [c_void_p] * 5,
    "dxtb_jacobi_work_size": [c_int, c_int],
    "dxtb_jacobi_eigh": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                         c_double, c_void_p, c_void_p, c_void_p],
    "dxtb_jacobi_eigh_noalloc": [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                 c_double, c_void_p, c_void_p, c_void_p],
    "dxtb_forces_classical": [POINTER(DxtbBasis), POINTER(DxtbParams)] + [c_void_p] * 19,
    "dxtb_broyden_work_size": [c_int, c_int, c_int],
    "dxtb_broyden_step": [c_int, c_int, c_int] + [c_void_p] * 11 + [c_double] * 4 + [c_int, c_int, c_void_p],
    "dxtb_gemm_batched": [c_int, c_int, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_double,
                          c_void_p, c_void_p, c_int, c_int, c_double, c_void_p],
}

_lib = None


def exported_names() -> list[str]:
    return list(_SIGNATURES)


def load() -> ctypes.CDLL:
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(
                f"HIP library {LIBPATH} not found: build it with `python -m dxtb_amd.build` "
                "(there is no CPU fallback)."
            )
        lib = ctypes.CDLL(str(LIBPATH))
        for name, args in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, c_int)
        _lib = lib
    return _lib


_RESTYPES = {"dxtb_jacobi_work_size": c_size_t, "dxtb_broyden_work_size": c_size_t,
             "dxtb_charges_work_size": c_size_t}


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = _ERRORS.get(rc, f"HIP error {rc}")
        raise RuntimeError(f"{what} failed: {msg} (code {rc})")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor passed to the HIP library must be contiguous")
    return t.data_ptr()


def stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def basis_struct(bb) -> DxtbBasis:
    d = bb.dev
    s = DxtbBasis()
    s.nbatch = bb.nbatch
    s.nat_max = bb.nat_max
    s.nsh_max = bb.nsh_max
    s.nao_max = bb.nao_max
    s.sh_atom = ptr(d["sh_atom"])
    s.sh_ush = ptr(d["sh_ush"])
    s.sh_ao = ptr(d["sh_ao"])
    s.nush = int(d["ush_l"].numel())
    s.ush_l = ptr(d["ush_l"])
    s.ush_np = ptr(d["ush_np"])
    s.ush_alpha = ptr(d["ush_alpha"])
    s.ush_coeff = ptr(d["ush_coeff"])
    s.item_mol = ptr(d["item_mol"])
    s.item_pair = ptr(d["item_pair"])
    for i in range(10):
        s.class_off[i] = int(bb.class_off[i])
    return s


def gemm(A: torch.Tensor, B: torch.Tensor, *, transa: bool = False, transb: bool = False, alpha: float = 1.0,
         scale: torch.Tensor | None = None, beta: float = 0.0, cin: torch.Tensor | None = None,
         out: torch.Tensor | None = None, nsize: torch.Tensor | None = None, sym: bool = False,
         pad: float | None = None) -> torch.Tensor:
    """Batched FP64 product alpha op(A) diag(scale) op(B) + beta cin on the
    matrix cores (dxtb_gemm_batched) for (nbatch, n, n) tensors; ``nsize``
    (int32, device) limits each product to its molecule's leading block;
    ``pad`` (None: leave) fills the rest with pad * identity."""
    lib = load()
    if A.dim() != 3 or A.shape != B.shape or A.shape[1] != A.shape[2] or A.dtype != torch.float64 \
            or B.dtype != torch.float64:
        raise ValueError("gemm takes two FP64 (nbatch, n, n) tensors of one shape")
    if out is None:
        out = torch.empty_like(A)
    if scale is not None and (scale.shape != A.shape[:2] or scale.dtype != torch.float64):
        raise ValueError("scale must be an FP64 (nbatch, n) tensor")
    if nsize is not None and (nsize.dtype != torch.int32 or nsize.numel() < A.shape[0]):
        raise ValueError("nsize must be an int32 tensor with one entry per matrix")
    rc = lib.dxtb_gemm_batched(A.shape[0], A.shape[1], ptr(nsize), int(transa), int(transb), float(alpha), ptr(A),
                               ptr(B), ptr(scale), float(beta), ptr(cin), ptr(out), int(sym), int(pad is not None),
                               float(pad or 0.0), stream_ptr(A.device))
    check(rc, "dxtb_gemm_batched")
    return out
