"""
Oracle: basis set and AO bookkeeping (test infrastructure only).

Follows ``src/dxtb/_src/basis/slater.py:69-145`` (STO-nG expansion),
``src/dxtb/_src/basis/ortho.py:74-110`` (GFN1 H-2s orthogonalisation),
``src/dxtb/_src/basis/bas.py:124-168`` (per-unique-shell CGTOs) and the AO
ordering of ``src/dxtb/_src/basis/indexhelper.py`` (atoms in input order,
shells in parameter order, 2l+1 spherical components per shell).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

from . import params

DFACT = [1.0, 1.0, 3.0, 15.0, 105.0, 945.0]  # (2l-1)!!, slater.py:40-43
_ITYPE_OFF = [0, 4, 7, 9, 10]


def slater_to_gauss(ng: int, n: int, l: int, zeta: float) -> tuple[torch.Tensor, torch.Tensor]:
    """STO-nG expansion of a Slater function (slater.py:69-145)."""
    table = params.load("stong")[str(ng)]
    itype = n + _ITYPE_OFF[l] - 1
    if n == 6 and ng == 6:
        itype = 15 + l
    coeff = torch.tensor(table[0][itype], dtype=torch.float64)
    alpha = torch.tensor(table[1][itype], dtype=torch.float64) * zeta**2
    coeff = coeff * ((2.0 / math.pi * alpha) ** 0.75 * torch.sqrt(4 * alpha) ** l / math.sqrt(DFACT[l]))
    p = alpha.unsqueeze(-1) + alpha.unsqueeze(-2)
    pref = (math.pi / p) ** 1.5 * DFACT[l] / (2 * p) ** l
    self_ovlp = (coeff.unsqueeze(-1) * coeff.unsqueeze(-2) * pref).sum()
    return alpha, coeff / torch.sqrt(self_ovlp)


def _ss(ai, aj, ci, cj):
    oij = 1.0 / (ai.unsqueeze(-1) + aj.unsqueeze(-2))
    return (torch.sqrt(math.pi * oij) ** 3 * ci.unsqueeze(-1) * cj.unsqueeze(-2)).sum()


def orthogonalize(alpha, coeff):
    """Orthonormalise the second s function against the first (ortho.py:74-110)."""
    (ai, aj), (ci, cj) = alpha, coeff
    ov = _ss(ai, aj, ci, cj)
    an = torch.cat((aj, ai))
    cn = torch.cat((cj, -ov * ci))
    return an, cn / torch.sqrt(_ss(an, an, cn, cn))


@dataclass
class Shell:
    atom: int
    z: int
    l: int
    alpha: torch.Tensor
    coeff: torch.Tensor
    ao: int  # first AO index
    ush: tuple  # (z, index of shell within the element)


@dataclass
class Basis:
    numbers: list
    gfn: str
    shells: list = field(default_factory=list)
    nao: int = 0

    @property
    def nsh(self) -> int:
        return len(self.shells)

    def ao_atom(self) -> torch.Tensor:
        out = []
        for sh in self.shells:
            out += [sh.atom] * (2 * sh.l + 1)
        return torch.tensor(out, dtype=torch.long)

    def ao_shell(self) -> torch.Tensor:
        out = []
        for i, sh in enumerate(self.shells):
            out += [i] * (2 * sh.l + 1)
        return torch.tensor(out, dtype=torch.long)

    def shell_atom(self) -> torch.Tensor:
        return torch.tensor([sh.atom for sh in self.shells], dtype=torch.long)


_CGTO_CACHE: dict = {}


def element_cgtos(gfn: str, z: int):
    """CGTOs of all shells of one element (bas.py:124-168)."""
    key = (gfn, z)
    if key in _CGTO_CACHE:
        return _CGTO_CACHE[key]
    e = params.elem(gfn, z)
    val = params.valence(e["shells"])
    out = []
    for i, sh in enumerate(e["shells"]):
        n, l = params.shell_lab(sh)
        a, c = slater_to_gauss(e["ngauss"][i], n, l, e["slater"][i])
        if params.kind(gfn) == "gfn1" and not val[i]:
            a, c = orthogonalize((out[i - 1][1], a), (out[i - 1][2], c))
        out.append((l, a, c))
    _CGTO_CACHE[key] = out
    return out


def build(numbers, gfn: str) -> Basis:
    nums = [int(z) for z in numbers if int(z) != 0]
    b = Basis(nums, gfn)
    ao = 0
    for iat, z in enumerate(nums):
        for ish, (l, a, c) in enumerate(element_cgtos(gfn, z)):
            b.shells.append(Shell(iat, z, l, a, c, ao, (z, ish)))
            ao += 2 * l + 1
    b.nao = ao
    return b
