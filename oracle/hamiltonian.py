"""
Oracle: coordination numbers and the core Hamiltonian H0 (test infrastructure only).

* CN: D3-type coordination number ``CN_A = sum_B count(R_AB, rcov_A + rcov_B)``
  with the 4/3-scaled Pyykko covalent radii and a 25 Bohr cutoff, as called
  from ``src/dxtb/_src/xtb/gfn1.py:58-64`` (``exp_count``, k=16) and
  ``src/dxtb/_src/xtb/gfn2.py:58-64`` (``gfn2_count``: the product of two
  exponential counts, k=10 at r0 and k=20 at r0+2 Bohr).  The counting
  functions live in tad-mctc 0.7.0 (absent); they are restated from the
  published D3 (Grimme 2010) and GFN2-xTB (Bannwarth 2019) definitions.
* H0: ``src/dxtb/_src/xtb/base.py:107-186`` (parameters) and ``:255-372``
  (build); shell-pair scaling ``xtb/gfn1.py:66-165`` and ``xtb/gfn2.py:67-156``.
"""

from __future__ import annotations

import math

import torch

from . import params
from .basis import Basis

CN_CUTOFF = 25.0


def exp_count(r, r0, k=16.0):
    return 1.0 / (1.0 + torch.exp(-k * (r0 / r - 1.0)))


def gfn2_count(r, r0):
    return exp_count(r, r0, 10.0) * exp_count(r, r0 + 2.0, 20.0)


def coordination_number(numbers, positions, gfn: str):
    nat = len(numbers)
    rcov = torch.tensor([params.COV_D3[z] for z in numbers], dtype=positions.dtype)
    rc = rcov.unsqueeze(-1) + rcov.unsqueeze(-2)
    diff = positions.unsqueeze(-2) - positions.unsqueeze(-3)
    eye = torch.eye(nat, dtype=torch.bool)
    r = torch.sqrt(torch.where(eye, torch.ones_like(rc), (diff * diff).sum(-1)))
    count = exp_count(r, rc) if params.kind(gfn) == "gfn1" else gfn2_count(r, rc)
    mask = (~eye) & (r <= CN_CUTOFF)
    return torch.where(mask, count, torch.zeros_like(count)).sum(-1)


def _shell_label(l):
    return "spdfg"[l]


def hscale(gfn: str, shells) -> torch.Tensor:
    """Shell-pair scaling K^{ll'} for a list of Shell objects."""
    p = params.par(gfn)["hamiltonian"]["xtb"]
    keys = p["shell"]
    n = len(shells)
    k = torch.ones((n, n), dtype=torch.float64)
    if params.kind(gfn) == "gfn1":
        kpol = p["kpol"]
        vals, kii = [], []
        for sh in shells:
            e = params.elem(gfn, sh.z)
            v = params.valence(e["shells"])[sh.ush[1]]
            vals.append(v)
            lab = _shell_label(sh.l) * 2
            kii.append(kpol if not v else keys.get(lab, 1.0))
        for i in range(n):
            for j in range(n):
                li, lj = _shell_label(shells[i].l), _shell_label(shells[j].l)
                if vals[i] and vals[j]:
                    if li + lj in keys:
                        v = keys[li + lj]
                    elif lj + li in keys:
                        v = keys[lj + li]
                    else:
                        v = 0.5 * (kii[i] + kii[j])
                else:
                    v = 0.5 * (kii[i] + kii[j])
                k[i, j] = v
        return k
    wexp = p["wexp"]
    for i in range(n):
        for j in range(n):
            li, lj = _shell_label(shells[i].l), _shell_label(shells[j].l)
            if li + lj in keys:
                kij = keys[li + lj]
            elif lj + li in keys:
                kij = keys[lj + li]
            else:
                kij = 0.5 * (keys[li + li] + keys[lj + lj])
            zi = params.elem(gfn, shells[i].z)["slater"][shells[i].ush[1]]
            zj = params.elem(gfn, shells[j].z)["slater"][shells[j].ush[1]]
            zmat = (2.0 * math.sqrt(zi * zj) / (zi + zj)) ** wexp
            k[i, j] = kij * zmat
    return k


def kpair(gfn: str, za: int, zb: int) -> float:
    kp = params.par(gfn)["hamiltonian"]["xtb"].get("kpair", {})
    a, b = params.SYMBOLS[za], params.SYMBOLS[zb]
    return kp.get(f"{a}-{b}", kp.get(f"{b}-{a}", 1.0))


def shell_selfenergy(basis: Basis, cn: torch.Tensor) -> torch.Tensor:
    gfn = basis.gfn
    lev, kcn = [], []
    for sh in basis.shells:
        e = params.elem(gfn, sh.z)
        lev.append(e["levels"][sh.ush[1]] * params.EV2AU)
        kcn.append(e["kcn"][sh.ush[1]] * params.EV2AU)
    lev = torch.tensor(lev, dtype=torch.float64)
    kcn = torch.tensor(kcn, dtype=torch.float64)
    return lev - kcn * cn[basis.shell_atom()]


def h0_shell_factor(basis: Basis, positions: torch.Tensor, cn: torch.Tensor) -> torch.Tensor:
    """Shell-pair factor f_ij such that H0 = S * spread(f) (xtb/base.py:255-372)."""
    gfn = basis.gfn
    p = params.par(gfn)["hamiltonian"]["xtb"]
    shells = basis.shells
    nsh = len(shells)
    sat = basis.shell_atom()
    nums = basis.numbers
    nat = len(nums)
    hsh = shell_selfenergy(basis, cn)
    rad = torch.tensor([params.ATOMIC_RADII[z] for z in nums], dtype=torch.float64)
    en = torch.tensor([params.elem(gfn, z)["en"] for z in nums], dtype=torch.float64)
    eye = torch.eye(nat, dtype=torch.bool)
    diff = positions.unsqueeze(-2) - positions.unsqueeze(-3)
    dist = torch.sqrt(torch.where(eye, torch.ones(nat, nat, dtype=positions.dtype), (diff * diff).sum(-1)))
    rr = dist / (rad.unsqueeze(-1) + rad.unsqueeze(-2))
    srr = torch.where(eye, torch.zeros_like(rr), torch.sqrt(rr))
    srr_sh = srr[sat][:, sat]
    shpoly = torch.tensor(
        [params.elem(gfn, sh.z)["shpoly"][sh.ush[1]] for sh in shells], dtype=torch.float64
    )
    var_pi = (1.0 + shpoly.unsqueeze(-1) * srr_sh) * (1.0 + shpoly.unsqueeze(-2) * srr_sh)
    diffatom = (~eye)[sat][:, sat]
    enscale = p.get("enscale", 0.0)
    en_sh = en[sat]
    var_x = torch.where(diffatom, 1.0 + enscale * (en_sh.unsqueeze(-1) - en_sh.unsqueeze(-2)) ** 2, 0.0)
    kp = torch.tensor([[kpair(gfn, a.z, b.z) for b in shells] for a in shells], dtype=torch.float64)
    hs = hscale(gfn, shells)
    if params.kind(gfn) == "gfn1":
        val = torch.tensor(
            [params.valence(params.elem(gfn, sh.z)["shells"])[sh.ush[1]] for sh in shells]
        )
    else:
        val = torch.ones(nsh, dtype=torch.bool)
    vv = val.unsqueeze(-1) & val.unsqueeze(-2)
    var_k = torch.where(vv, hs * kp * var_x, hs)
    var_h = 0.5 * (hsh.unsqueeze(-1) + hsh.unsqueeze(-2))
    return torch.where(diffatom, var_pi * var_k * var_h, var_h)


def h0_matrix(basis: Basis, positions: torch.Tensor, s: torch.Tensor, cn: torch.Tensor | None = None):
    if cn is None:
        cn = coordination_number(basis.numbers, positions, basis.gfn)
    f = h0_shell_factor(basis, positions, cn)
    aosh = basis.ao_shell()
    h = f[aosh][:, aosh]
    if s is not None:
        h = h * s
    return 0.5 * (h + h.T)
