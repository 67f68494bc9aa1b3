"""
Multipole integrals on the GPU: overlap S, ket-centred dipole D, traceless
ket-centred quadrupole Q (and H0 = S * f) for a padded batch, plus their
nuclear-coordinate VJP.

Output layouts follow the reference integral container
(``src/dxtb/_src/integral/container.py:182-535``): S (B, nao, nao),
D (B, 3, nao, nao), Q (B, 6, nao, nao) with components xx, yx, yy, zx, zy, zz.
"""

from __future__ import annotations

import torch

from . import _lib
from .basis import BatchBasis


def _check_positions(bb: BatchBasis, positions: torch.Tensor) -> torch.Tensor:
    if positions.dtype != torch.float64:
        raise TypeError("positions must be float64")
    if positions.device.type != "cuda":
        raise RuntimeError("the HIP integral path needs positions on a GPU device")
    pos = positions.reshape(bb.nbatch, -1, 3)
    if pos.shape[1] != bb.nat_max:
        raise ValueError(f"positions have {pos.shape[1]} atoms per molecule, the basis {bb.nat_max}")
    return pos.detach().contiguous()


def work_basis(bb: BatchBasis, positions: torch.Tensor, screening_threshold: float | None = None,
               pairs: torch.Tensor | None = None) -> BatchBasis:
    """The basis whose work lists the kernels run: ``bb`` itself, or with
    shell-pair screening (``screening.py``) only the pairs of the selection
    ``pairs`` or of ``screening_threshold`` on these positions (detached:
    dropped blocks stay zero and get no gradient), as the reference's
    assemble_matrix takes them (impls/pairs.py:228-340)."""
    if pairs is not None and screening_threshold is not None:
        raise ValueError("Pass either `screening_threshold` or `pairs`, not both.")
    if pairs is None and screening_threshold is None:
        return bb
    from .screening import screened_basis, select_pairs

    if pairs is None:
        pairs = select_pairs(bb, _check_positions(bb, positions), screening_threshold)
    return screened_basis(bb, pairs)


def multipole_forward(bb: BatchBasis, positions: torch.Tensor, multipoles: bool = True,
                      fsh: torch.Tensor | None = None, traceless: bool = True,
                      screening_threshold: float | None = None, pairs: torch.Tensor | None = None):
    """S, D, Q (None if ``multipoles`` is False) and H0 (if ``fsh`` is given).
    ``traceless=False`` returns the raw ket-centred second moments.
    ``screening_threshold`` or ``pairs`` skip the screened shell pairs
    (``work_basis``)."""
    lib = _lib.load()
    bb = work_basis(bb, positions, screening_threshold, pairs)
    pos = _check_positions(bb, positions)
    dev = pos.device
    n = bb.nao_max
    B = bb.nbatch
    # complete work lists write every element of each molecule's block: only
    # the padding is zeroed (screened lists leave blocks unwritten)
    alloc = torch.empty if bb.complete_pairs else torch.zeros
    S = alloc((B, n, n), dtype=torch.float64, device=dev)
    D = alloc((B, 3, n, n), dtype=torch.float64, device=dev) if multipoles else None
    Q = alloc((B, 6, n, n), dtype=torch.float64, device=dev) if multipoles else None
    H0 = alloc((B, n, n), dtype=torch.float64, device=dev) if fsh is not None else None
    stream = _lib.stream_ptr(dev)
    if bb.complete_pairs:
        for X, nmat in ((S, 1), (D, 3), (Q, 6), (H0, 1)):
            if X is not None:
                _lib.check(lib.dxtb_zero_padding(B, n, nmat, _lib.ptr(bb.dev["nao"]), _lib.ptr(X), stream),
                           "dxtb_zero_padding")
    st = _lib.basis_struct(bb)
    rc = lib.dxtb_multipole_fwd(
        st, _lib.ptr(pos), _lib.ptr(fsh), _lib.ptr(S), _lib.ptr(D), _lib.ptr(Q), _lib.ptr(H0),
        1 if traceless else 0, stream,
    )
    _lib.check(rc, "dxtb_multipole_fwd")
    return S, D, Q, H0


def multipole_vjp(bb: BatchBasis, positions: torch.Tensor, adjS, adjD=None, adjQ=None,
                  screening_threshold: float | None = None, pairs: torch.Tensor | None = None) -> torch.Tensor:
    """d/dR [sum(adjS*S) + sum(adjD*D) + sum(adjQ*Q)], shape (B, nat_max, 3)
    (of the screened matrices with ``screening_threshold`` or ``pairs``)."""
    lib = _lib.load()
    bb = work_basis(bb, positions, screening_threshold, pairs)
    pos = _check_positions(bb, positions)
    grad = torch.zeros_like(pos)
    st = _lib.basis_struct(bb)
    c = lambda t: None if t is None else t.contiguous()  # noqa: E731
    aS, aD, aQ = c(adjS), c(adjD), c(adjQ)
    rc = lib.dxtb_multipole_vjp(
        st, _lib.ptr(pos), _lib.ptr(aS), _lib.ptr(aD), _lib.ptr(aQ), _lib.ptr(grad), _lib.stream_ptr(pos.device)
    )
    _lib.check(rc, "dxtb_multipole_vjp")
    return grad


class MultipoleFunction(torch.autograd.Function):
    """Differentiable (S, D, Q) whose backward is the hand-written VJP kernel."""

    @staticmethod
    def forward(ctx, positions, bb: BatchBasis, multipoles: bool):
        S, D, Q, _ = multipole_forward(bb, positions, multipoles)
        ctx.bb = bb
        ctx.shape = positions.shape
        ctx.save_for_backward(positions)
        if multipoles:
            return S, D, Q
        return S, torch.zeros(0, dtype=S.dtype, device=S.device), torch.zeros(0, dtype=S.dtype, device=S.device)

    @staticmethod
    def backward(ctx, gS, gD, gQ):
        (positions,) = ctx.saved_tensors
        if gS is None:
            gS = torch.zeros((ctx.bb.nbatch, ctx.bb.nao_max, ctx.bb.nao_max), dtype=torch.float64,
                             device=positions.device)
        if gD is not None and gD.numel() == 0:
            gD = None
        if gQ is not None and gQ.numel() == 0:
            gQ = None
        g = multipole_vjp(ctx.bb, positions, gS, gD, gQ)
        return g.reshape(ctx.shape), None, None


def multipoles(bb: BatchBasis, positions: torch.Tensor, multipoles: bool = True,
               screening_threshold: float | None = None, pairs: torch.Tensor | None = None):
    """Differentiable (S, D, Q); with ``screening_threshold`` or ``pairs`` the
    screened matrices (the selection is made once, from the detached
    positions)."""
    return MultipoleFunction.apply(positions, work_basis(bb, positions, screening_threshold, pairs), multipoles)
