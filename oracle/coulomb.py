"""
Oracle: electrostatics and repulsion (test infrastructure only).

* ES2 shell-resolved Coulomb matrix -- ``components/interactions/coulomb/secondorder.py:799-870``
  with the averaging functions of ``coulomb/average.py:41-98``; energy/potential ``:365-443``.
* ES3 on-site third order -- ``coulomb/thirdorder.py:183-358`` (GFN1 atom-resolved,
  GFN2 shell-resolved with the s/p/d scaling of the parametrization).
* AES2 anisotropic electrostatics (GFN2) -- ``coulomb/multipole.py:270-641``
  (multipole radii from the GFN2 CN, damped interaction matrices, on-site kernels).
* Repulsion -- ``components/classicals/repulsion/base.py:192-359`` and ``rep.py:65-141``.

(Paths relative to ``src/dxtb/_src``.)
"""

from __future__ import annotations

import torch

from . import params
from .basis import Basis
from .hamiltonian import coordination_number

EPS = torch.finfo(torch.float64).eps
TINY = torch.finfo(torch.float64).tiny
REP_CUTOFF = 25.0


def _dist(positions):
    nat = positions.shape[0]
    eye = torch.eye(nat, dtype=torch.bool)
    diff = positions.unsqueeze(-2) - positions.unsqueeze(-3)
    r = torch.sqrt(torch.where(eye, torch.ones(nat, nat, dtype=positions.dtype), (diff * diff).sum(-1)))
    return torch.where(eye, torch.zeros_like(r), r), eye, diff


def es2_matrix(basis: Basis, positions: torch.Tensor) -> torch.Tensor:
    """Shell-resolved gamma (secondorder.py:799-870)."""
    gfn = basis.gfn
    cp = params.par(gfn)["charge"]["effective"]
    g = cp["gexp"]
    h = torch.tensor(
        [params.elem(gfn, sh.z)["gam"] * params.elem(gfn, sh.z)["lgam"][sh.ush[1]] for sh in basis.shells],
        dtype=torch.float64,
    )
    sat = basis.shell_atom()
    dist, eye, _ = _dist(positions)
    maskat = ~eye
    dist_g = torch.where(maskat, torch.pow(dist + EPS, g), torch.full_like(dist, EPS))[sat][:, sat]
    he = h + EPS
    if cp["average"] == "harmonic":
        inv = 1.0 / he
        avg = 2.0 / (inv.unsqueeze(-1) + inv.unsqueeze(-2))
    elif cp["average"] == "arithmetic":
        avg = 0.5 * (he.unsqueeze(-1) + he.unsqueeze(-2))
    else:
        avg = torch.sqrt(he.unsqueeze(-1) * he.unsqueeze(-2))
    tmp = dist_g + torch.pow(avg, -g)
    return 1.0 / torch.pow(tmp, 1.0 / g)


def es3_hubbard_derivs(basis: Basis):
    """Atom (GFN1) or shell (GFN2) resolved third-order parameters (thirdorder.py:183-245)."""
    gfn = basis.gfn
    to = params.par(gfn)["thirdorder"]
    if to.get("shell") is False:
        return torch.tensor([params.elem(gfn, z)["gam3"] for z in basis.numbers], dtype=torch.float64), False
    scale = to["shell"]
    lab = "spdfg"
    return (
        torch.tensor(
            [params.elem(gfn, sh.z)["gam3"] * scale[lab[sh.l]] for sh in basis.shells], dtype=torch.float64
        ),
        True,
    )


def aes2_cache(basis: Basis, positions: torch.Tensor):
    """mrad and the damped interaction matrices (multipole.py:270-449)."""
    gfn = basis.gfn
    mp = params.par(gfn)["multipole"]["damped"]
    nums = basis.numbers
    cn = coordination_number(nums, positions, "gfn2")
    rad = torch.tensor([params.elem(gfn, z)["mprad"] for z in nums], dtype=torch.float64)
    vcn = torch.tensor([params.elem(gfn, z)["mpvcn"] for z in nums], dtype=torch.float64)
    t1 = torch.exp(-mp["kexp"] * (cn - vcn - mp["shift"]))
    mrad = rad + (mp["rmax"] - rad) / (1.0 + t1)
    dist, eye, diff = _dist(positions)
    g1 = torch.where(eye, torch.zeros_like(dist), 1.0 / torch.where(eye, torch.ones_like(dist), dist))
    g3 = g1 * g1 * g1
    g5 = g3 * g1 * g1
    rr = 0.5 * (mrad.unsqueeze(-1) + mrad.unsqueeze(-2)) * g1
    fd3 = 1.0 / (1.0 + 6.0 * rr ** mp["dmp3"])
    fd5 = 1.0 / (1.0 + 6.0 * rr ** mp["dmp5"])
    rij = torch.where(eye.unsqueeze(-1), torch.full_like(diff, EPS), diff)  # R_i - R_j
    sd = rij * (g3 * fd3).unsqueeze(-1)
    g5f = (g5 * fd5).unsqueeze(-1)
    unity = torch.eye(3, dtype=torch.float64)
    dd = unity * (g3 * fd5).unsqueeze(-1).unsqueeze(-1) - rij.unsqueeze(-1) * rij.unsqueeze(-2) * 3 * g5f.unsqueeze(-1)
    x, y, z = rij[..., 0], rij[..., 1], rij[..., 2]
    sq = torch.stack([x * x, 2 * x * y, y * y, 2 * x * z, 2 * y * z, z * z], -1) * g5f
    dk = torch.tensor([params.elem(gfn, zz)["dkernel"] for zz in nums], dtype=torch.float64)
    qk = torch.tensor([params.elem(gfn, zz)["qkernel"] for zz in nums], dtype=torch.float64)
    return {"mrad": mrad, "sd": sd, "dd": dd, "sq": sq, "dk": dk, "qk": qk, "cn": cn}


QSCALE = torch.tensor([1.0, 2.0, 1.0, 2.0, 2.0, 1.0], dtype=torch.float64)


def aes2_energy(c, qat, qdp, qqp):
    """Per-atom dipole + quadrupole energies (multipole.py:450-531)."""
    vdp_sd = torch.einsum("ijx,i->jx", c["sd"], qat)
    vdp_dd = 0.5 * torch.einsum("ijxy,ix->jy", c["dd"], qdp)
    edp = (c["dk"].unsqueeze(-1) * qdp * qdp).sum(-1) + ((vdp_sd + vdp_dd) * qdp).sum(-1)
    vqp = torch.einsum("ijx,i->jx", c["sq"], qat)
    eqp = (c["qk"].unsqueeze(-1) * qqp * QSCALE * qqp).sum(-1) + (vqp * qqp).sum(-1)
    return edp + eqp


def aes2_potential(c, qat, qdp, qqp):
    """Monopole, dipole and quadrupole atom potentials (multipole.py:533-640)."""
    vat = torch.einsum("ijx,jx->i", c["sd"], qdp) + torch.einsum("ijx,jx->i", c["sq"], qqp)
    vdp = torch.einsum("ijx,i->jx", c["sd"], qat) + torch.einsum("ijxy,ix->jy", c["dd"], qdp) + 2 * c["dk"].unsqueeze(-1) * qdp
    vqp = torch.einsum("ijx,i->jx", c["sq"], qat) + 2 * c["qk"].unsqueeze(-1) * qqp * QSCALE
    return vat, vdp, vqp


def repulsion_energy(numbers, positions, gfn: str) -> torch.Tensor:
    """Per-atom repulsion energy (repulsion/base.py:192-359, rep.py:96)."""
    rp = params.par(gfn)["repulsion"]["effective"]
    nat = len(numbers)
    arep = torch.tensor([params.elem(gfn, z)["arep"] for z in numbers], dtype=torch.float64)
    zeff = torch.tensor([params.elem(gfn, z)["zeff"] for z in numbers], dtype=torch.float64)
    eye = torch.eye(nat, dtype=torch.bool)
    mask = ~eye
    a = torch.where(mask, torch.sqrt(arep.unsqueeze(-1) * arep.unsqueeze(-2) + TINY), 0.0)
    if "enscale" in rp:
        en = torch.tensor([params.elem(gfn, z)["en"] for z in numbers], dtype=torch.float64)
        den2 = (en.unsqueeze(-1) - en.unsqueeze(-2)) ** 2
        a = a * (1.0 + (0.01 * den2 + 0.01 * den2**2) * rp["enscale"])
    zz = zeff.unsqueeze(-1) * zeff.unsqueeze(-2) * mask
    k = torch.full((nat, nat), rp["kexp"], dtype=torch.float64) * mask
    if "klight" in rp:
        light = torch.tensor([z <= 2 for z in numbers])
        both = light.unsqueeze(-1) & light.unsqueeze(-2)
        k = torch.where(~both, k, torch.full_like(k, rp["klight"])) * mask
    dist, _, _ = _dist(positions)
    dist = torch.where(mask, dist, torch.full_like(dist, EPS))
    e = zz * torch.exp(-a * torch.pow(dist, k)) / dist
    e = torch.where(mask & (dist <= REP_CUTOFF), e, torch.zeros_like(e))
    return 0.5 * e.sum(-1)
