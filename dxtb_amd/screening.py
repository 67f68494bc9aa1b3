"""
Shell-pair screening of the integral work lists (off unless asked for, as in
the reference).

Mirrors ``src/dxtb/_src/integral/driver/pytorch/impls/pairs.py``:

* ``pair_bounds``: the rigorous upper bound of the largest |element| of every
  shell pair's spherical block for the multipole orders 0, 1 and 2 about a
  multipole origin (``_bounds``, pairs.py:411-494), on the device
  (``dxtb_pair_bounds``);
* ``select_pairs``: the pairs to keep -- both shells at one position, or the
  bound of any order above the threshold (``_screen``, pairs.py:497-541) --
  as the union over one or more sets of geometries (``select_pairs``,
  pairs.py:544-614), so that one selection serves the overlap, dipole and
  quadrupole alike and every dropped element is below the threshold;
* ``screened_basis``: the batch with only the kept pairs in its work lists.
  The integral kernels then never evaluate the dropped pairs: their blocks
  stay zero and get no gradient (pairs.py:261-283).

Both orders (A, B) and (B, A) of a pair get the same bound, so the screened
matrices stay symmetric.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .basis import BatchBasis


def pair_bounds(bb: BatchBasis, positions: torch.Tensor, origin=None) -> torch.Tensor:
    """(nitems, 4) for the work-list items of ``bb`` in class order: the
    bounds for the orders 0, 1, 2 about ``origin`` (3 numbers; default the
    Cartesian origin), then 1.0 where both shells sit at one position."""
    pos = positions.detach()
    if pos.dtype != torch.float64 or pos.device.type != "cuda":
        raise ValueError("positions must be a float64 CUDA tensor")
    pos = pos.reshape(bb.nbatch, bb.nat_max, 3).contiguous()
    n = int(bb.class_off[9])
    out = torch.empty((max(n, 1), 4), dtype=torch.float64, device=pos.device)
    org = None
    if origin is not None:
        o = [float(v) for v in torch.as_tensor(origin, dtype=torch.float64).reshape(-1)[:3].tolist()]
        org = ctypes.cast((ctypes.c_double * 3)(*o), ctypes.c_void_p)
    rc = _lib.load().dxtb_pair_bounds(_lib.basis_struct(bb), _lib.ptr(pos), org, _lib.ptr(out),
                                      _lib.stream_ptr(pos.device))
    _lib.check(rc, "dxtb_pair_bounds")
    return out[:n]


def select_pairs(bb: BatchBasis, positions: torch.Tensor, threshold: float, origin=None) -> torch.Tensor:
    """Boolean mask over the work-list items of ``bb``: the pairs to compute.
    ``positions`` is one set of geometries (nbatch, nat_max, 3) or several
    (ngeom, nbatch, nat_max, 3); a pair kept for any of them is kept (the
    union, so the error bound holds for every geometry)."""
    pos = positions.detach()
    sets = pos.reshape(-1, bb.nbatch, bb.nat_max, 3)
    keep = None
    for g in sets:
        bd = pair_bounds(bb, g, origin)
        k = (bd[:, 3] != 0.0) | (bd[:, :3] > threshold).any(-1)
        keep = k if keep is None else keep | k
    return keep


def screened_basis(bb: BatchBasis, keep: torch.Tensor) -> BatchBasis:
    """Copy of ``bb`` whose work lists hold only the items where ``keep`` is
    true (the maps and tables are shared)."""
    n = int(bb.class_off[9])
    if keep.numel() != n:
        raise ValueError(f"the selection has {keep.numel()} entries, the work lists {n}")
    kh = keep.detach().cpu().numpy().astype(bool)
    counts = np.array([kh[bb.class_off[c] : bb.class_off[c + 1]].sum() for c in range(9)], dtype=np.int32)
    class_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    kd = keep.to(device=bb.dev["item_mol"].device, dtype=torch.bool)
    dev = dict(bb.dev)
    dev["item_mol"] = bb.dev["item_mol"][kd].contiguous()
    dev["item_pair"] = bb.dev["item_pair"][kd].contiguous()
    return BatchBasis(bb.gfn, bb.numbers, bb.nat, bb.nsh, bb.nao, bb.nat_max, bb.nsh_max, bb.nao_max,
                      bb.sh_atom, bb.sh_ush, bb.sh_ao, bb.sh_l, bb.ao_atom, bb.ao_shell, bb.item_mol[kh],
                      bb.item_pair[kh], class_off, dev, bb.ao_order)
