"""
Basis sets and the device-resident description of a batch of molecules.

* Contracted Gaussians: STO-nG expansion of every parametrized Slater shell
  (``src/dxtb/_src/basis/slater.py:69-145``), the GFN1 H-2s Gram-Schmidt step
  (``basis/ortho.py:74-110``) and the exact unit-self-overlap rescaling of
  the pair builder (``integral/driver/pytorch/impls/pairs.py:181-204``).
  Built once per parametrization for all elements, in numpy (FP64).
* AO order (``basis/indexhelper.py``): atoms in input order, shells in
  parameter order, 2l+1 spherical functions per shell, p as (y, z, x).
* ``BatchBasis``: padded per-molecule shell maps plus the ordered shell-pair
  work lists of the integral kernels, grouped by (l_bra, l_ket), on device.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import param

MAXP = 8  # primitives per shell in the device tables
_DFACT = (1.0, 1.0, 3.0, 15.0, 105.0)
_ITYPE = (0, 4, 7, 9, 10)


def _self_overlap(alpha: np.ndarray, coeff: np.ndarray, l: int) -> float:
    p = alpha[:, None] + alpha[None, :]
    pref = (math.pi / p) ** 1.5 * _DFACT[l] / (2.0 * p) ** l
    return float((coeff[:, None] * coeff[None, :] * pref).sum())


def slater_to_gauss(ng: int, n: int, l: int, zeta: float) -> tuple[np.ndarray, np.ndarray]:
    """Normalized STO-nG expansion of one Slater shell (basis/slater.py:69-145)."""
    if not 1 <= ng <= 6 or n <= l or zeta <= 0.0:
        raise ValueError(f"invalid STO-nG request ng={ng} n={n} l={l} zeta={zeta}")
    itype = 15 + l if (n == 6 and ng == 6) else n + _ITYPE[l] - 1
    tab = param.stong_table(ng)
    coeff = tab[0, itype].copy()
    alpha = tab[1, itype] * zeta * zeta
    coeff *= (2.0 / math.pi * alpha) ** 0.75 * np.sqrt(4.0 * alpha) ** l / math.sqrt(_DFACT[l])
    return alpha, coeff / math.sqrt(_self_overlap(alpha, coeff, l))


def _s_overlap(ai, aj, ci, cj) -> float:
    p = ai[:, None] + aj[None, :]
    return float((np.sqrt(math.pi / p) ** 3 * ci[:, None] * cj[None, :]).sum())


def orthogonalize(a1, c1, a2, c2) -> tuple[np.ndarray, np.ndarray]:
    """Second s function orthonormalized against the first (basis/ortho.py:74-110)."""
    ov = _s_overlap(a1, a2, c1, c2)
    an = np.concatenate([a2, a1])
    cn = np.concatenate([c2, -ov * c1])
    return an, cn / math.sqrt(_s_overlap(an, an, cn, cn))


class ShellTable:
    """All contracted shells of one parametrization (unique shells)."""

    def __init__(self, gfn: str) -> None:
        par = param.get(gfn)
        self.gfn = gfn
        # the second s shell of H is orthogonalised to the first in GFN1
        # (basis/bas.py:147-163: meta.name of the GFN0/GFN1 family)
        gfn1 = par.kind == "gfn1"
        self.first = np.full(param.MAXZ + 1, -1, dtype=np.int64)  # first unique shell of Z
        self.count = np.zeros(param.MAXZ + 1, dtype=np.int64)
        ls, nps, al, co, zs = [], [], [], [], []
        eps = np.finfo(np.float64).eps
        for z in range(1, param.MAXZ + 1):
            if not par.has(z):
                continue
            e = par.element(z)
            val = par.valence(z)
            self.first[z] = len(ls)
            self.count[z] = len(e["shells"])
            prev = None
            for i, (n, l) in enumerate(par.shells(z)):
                a, c = slater_to_gauss(int(e["ngauss"][i]), n, l, float(e["slater"][i]))
                if gfn1 and not val[i]:
                    a, c = orthogonalize(prev[0], prev[1], a, c)
                prev = (a, c)
                # exact unit self overlap (impls/pairs.py:181-204)
                s = _self_overlap(a, c, l)
                cnorm = c / (math.sqrt(max(s, eps)) + eps)
                if len(a) > MAXP:
                    raise ValueError("too many primitives")
                ls.append(l)
                nps.append(len(a))
                al.append(np.pad(a, (0, MAXP - len(a))))
                co.append(np.pad(cnorm, (0, MAXP - len(a))))
                zs.append(z)
        self.l = np.array(ls, dtype=np.int32)
        self.np = np.array(nps, dtype=np.int32)
        self.alpha = np.array(al, dtype=np.float64)
        self.coeff = np.array(co, dtype=np.float64)
        self.z = np.array(zs, dtype=np.int32)
        self._dev: dict = {}

    def device(self, dev: torch.device) -> dict:
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = {
                "l": torch.as_tensor(self.l, device=dev),
                "np": torch.as_tensor(self.np, device=dev),
                "alpha": torch.as_tensor(self.alpha, device=dev).contiguous(),
                "coeff": torch.as_tensor(self.coeff, device=dev).contiguous(),
            }
        return self._dev[key]


_TABLES: dict = {}


def shell_table(gfn: str) -> ShellTable:
    if gfn not in _TABLES:
        _TABLES[gfn] = ShellTable(gfn)
    return _TABLES[gfn]


@dataclass
class BatchBasis:
    """Padded shell/AO maps of a batch and its shell-pair work lists."""

    gfn: str
    numbers: np.ndarray  # (B, nat_max) int, 0 = padding
    nat: np.ndarray  # (B,)
    nsh: np.ndarray
    nao: np.ndarray
    nat_max: int
    nsh_max: int
    nao_max: int
    sh_atom: np.ndarray  # (B, nsh_max) int32
    sh_ush: np.ndarray
    sh_ao: np.ndarray
    sh_l: np.ndarray
    ao_atom: np.ndarray  # (B, nao_max) int32 (0 for padding)
    ao_shell: np.ndarray
    item_mol: np.ndarray
    item_pair: np.ndarray
    class_off: np.ndarray  # (10,)
    dev: dict
    # AO layout of the matrices: "reference" (atoms in order, each atom's
    # shells in order) or "l" (the shells grouped by angular momentum: every
    # s shell, then every p, then every d, each group in reference order)
    ao_order: str = "reference"
    # the work lists hold every shell pair of every molecule (build_batch),
    # so the integral kernels write each molecule's whole leading block
    complete_pairs: bool = False

    @property
    def nbatch(self) -> int:
        return int(self.numbers.shape[0])


def to_device(a, dev, dtype: torch.dtype | None = None) -> torch.Tensor:
    """Host array to the device through pinned memory without waiting (a copy
    from pageable memory can wait for every stream of the device)."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    if torch.device(dev).type != "cuda":
        return t.to(dev)
    return t.pin_memory().to(dev, non_blocking=True)


def build_batch(numbers, gfn: str, device: torch.device | str, dims: tuple[int, int, int] | None = None,
                ao_order: str = "reference") -> BatchBasis:
    """Shell maps and ordered shell-pair lists for a (padded) batch.

    ``dims = (nat_max, nsh_max, nao_max)`` pads the batch to fixed sizes (slot
    pools of the serving engine hold molecules of different sizes in arrays
    of one shape); by default the sizes are the batch maxima.

    ``ao_order`` places the shells' basis functions in the matrices: in the
    reference's order (atoms in order, each atom's shells in order), or with
    ``"l"`` grouped by angular momentum (every s shell, then every p, then
    every d, each group in reference order; a shell's 2l+1 functions stay
    together in the reference's component order).  Shell indices, and every
    per-shell and per-atom quantity, are the same in both; AO-indexed
    matrices differ by a symmetric permutation, so energies, charges and
    forces do not depend on the choice.  With ``"l"`` the pairs of one class
    (l_A, l_B) fill a contiguous column range of each of their rows, which
    the integral kernels write and read in whole cache lines."""
    if ao_order not in ("reference", "l"):
        raise ValueError(f"ao_order must be 'reference' or 'l', not {ao_order!r}")
    nums = np.atleast_2d(np.asarray(numbers.cpu() if torch.is_tensor(numbers) else numbers, dtype=np.int64))
    if dims is not None:
        if nums.shape[1] > dims[0]:
            if (nums[:, dims[0]:] != 0).any():
                raise ValueError(f"a molecule has more than {dims[0]} atoms")
            nums = nums[:, : dims[0]]
        nums = np.pad(nums, ((0, 0), (0, dims[0] - nums.shape[1])))
    table = shell_table(gfn)
    B, nat_max = nums.shape
    nat = (nums != 0).sum(1)
    cnt = table.count[nums] * (nums != 0)
    if ((nums != 0) & (table.first[nums] < 0)).any():
        bad = sorted(set(nums[(nums != 0) & (table.first[nums] < 0)].tolist()))
        raise ValueError(f"Elements {bad} are not parametrized in {gfn}.")
    nsh = cnt.sum(1)
    nsh_max = int(max(nsh.max(), 1))
    if dims is not None:
        if nsh_max > dims[1]:
            raise ValueError(f"a molecule has more than {dims[1]} shells")
        nsh_max = int(dims[1])
    if nsh_max > 0xFFFF:
        raise ValueError("too many shells per molecule")
    sh_atom = np.zeros((B, nsh_max), dtype=np.int32)
    sh_ush = np.zeros((B, nsh_max), dtype=np.int32)
    sh_l = np.zeros((B, nsh_max), dtype=np.int32)
    # every atom of the batch repeated by its shell count (padding atoms have
    # none), vectorised over the batch
    reps = cnt.ravel()
    tot = int(reps.sum())
    if tot:
        atom = np.repeat(np.tile(np.arange(nat_max), B), reps)
        mol = np.repeat(np.repeat(np.arange(B), nat_max), reps)
        starts = np.cumsum(reps) - reps
        ush = np.repeat(table.first[nums.ravel()], reps) + (np.arange(tot) - np.repeat(starts, reps))
        pos = np.arange(tot) - np.repeat(np.cumsum(nsh) - nsh, nsh)
        sh_atom[mol, pos] = atom
        sh_ush[mol, pos] = ush
        sh_l[mol, pos] = table.l[ush]
    valid = np.arange(nsh_max)[None, :] < nsh[:, None]
    nsph = np.where(valid, 2 * sh_l + 1, 0)
    if ao_order == "l":
        # shells in (l, index) order, padding shells last; offsets in that order
        perm = np.argsort(np.where(valid, sh_l, 3) * nsh_max + np.arange(nsh_max)[None, :], axis=1, kind="stable")
        w = np.take_along_axis(nsph, perm, 1)
        sh_ao = np.zeros_like(w)
        np.put_along_axis(sh_ao, perm, np.cumsum(w, 1) - w, 1)
        sh_ao = sh_ao.astype(np.int32)
    else:
        sh_ao = (np.cumsum(nsph, 1) - nsph).astype(np.int32)
    nao = nsph.sum(1)
    nao_max = int(max(nao.max(), 1))
    if dims is not None:
        if nao_max > dims[2]:
            raise ValueError(f"a molecule has more than {dims[2]} basis functions")
        nao_max = int(dims[2])
    ao_atom = np.zeros((B, nao_max), dtype=np.int32)
    ao_shell = np.zeros((B, nao_max), dtype=np.int32)
    reps = nsph.ravel()  # 0 on padding shells
    tot = int(reps.sum())
    if tot:
        shell = np.repeat(np.tile(np.arange(nsh_max), B), reps)
        mol = np.repeat(np.repeat(np.arange(B), nsh_max), reps)
        comp = np.arange(tot) - np.repeat(np.cumsum(reps) - reps, reps)  # component within its shell
        pos = sh_ao[mol, shell] + comp
        ao_shell[mol, pos] = shell
        ao_atom[mol, pos] = sh_atom[mol, shell]

    # ordered shell pairs (A, B) of every molecule, grouped by class 3*lA + lB
    mols, pairs, cls = [], [], []
    for b in range(B):
        n = int(nsh[b])
        r = np.arange(n, dtype=np.int32)
        lb = sh_l[b, :n].astype(np.uint8)
        mols.append(np.full(n * n, b, dtype=np.int32))
        pairs.append(((r[:, None] << 16) | r[None, :]).ravel())
        cls.append((3 * lb[:, None] + lb[None, :]).ravel())
    mols = np.concatenate(mols)
    pairs = np.concatenate(pairs)
    cls = np.concatenate(cls)
    order = np.argsort(cls, kind="stable")  # 8-bit keys: numpy's stable sort is a radix sort
    item_mol = mols[order]
    item_pair = pairs[order]
    counts = np.bincount(cls, minlength=9)
    class_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)

    # first shell and shell count of every atom
    at_nsh = (table.count[nums] * (nums != 0)).astype(np.int32)
    at_sh0 = (np.cumsum(at_nsh, 1) - at_nsh).astype(np.int32)

    dev = torch.device(device)
    td = table.device(dev)
    t = lambda x: to_device(x, dev)  # noqa: E731
    d = {
        "sh_atom": t(sh_atom),
        "sh_ush": t(sh_ush),
        "sh_ao": t(sh_ao),
        "item_mol": t(item_mol),
        "item_pair": t(item_pair),
        "ush_l": td["l"],
        "ush_np": td["np"],
        "ush_alpha": td["alpha"],
        "ush_coeff": td["coeff"],
        "ao_atom": t(ao_atom).long(),
        "ao_shell": t(ao_shell).long(),
        "sh_atom_l": t(sh_atom).long(),
        "ao_atom_i": t(ao_atom),
        "ao_shell_i": t(ao_shell),
        "at_sh0": t(at_sh0),
        "at_nsh": t(at_nsh),
        "numbers_i": t(nums.astype(np.int32)),
        "nat": t(nat.astype(np.int32)),
        "nsh": t(nsh.astype(np.int32)),
        "nao": t(nao.astype(np.int32)),
    }
    return BatchBasis(
        gfn, nums, nat, nsh, nao, int(nat_max), nsh_max, nao_max, sh_atom, sh_ush, sh_ao, sh_l,
        ao_atom, ao_shell, item_mol, item_pair, class_off, d, ao_order, complete_pairs=True,
    )
