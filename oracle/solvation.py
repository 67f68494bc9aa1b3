"""
Oracle (test infrastructure only): generalized Born implicit solvation.

CPU restatement of the reference's GeneralizedBorn interaction
(src/dxtb/_src/components/interactions/solvation/alpb.py:187-405) and its
Born radii (.../solvation/born.py:52-239, Onufriev-Bashford-Case integrator):

  psi_i = sum_{j != i, r_ij <= 66} I(r_ij; rho_j = 0.8 R_j, R_i)      (born.py:163-239)
  B_i   = 1 / (1/R_i - tanh(p (1.0 - 0.8 p + 4.85 p^2)) / R_i),  p = psi_i R_i / 2
  K_ij  = P16: 1 / (r + sqrt(ab) (sqrt(ab) / (sqrt(ab) + 1.028 r / 16))^16),  ab = B_i B_j
          Still: 1 / sqrt(r^2 + ab exp(-r^2 / (4 ab)))                    (alpb.py:94-141)
  mat   = keps K (+ keps alpbet adet with ALPB),  keps = (1/eps - 1) / (1 + alpbet)
  V_i   = sum_k mat_ik q_k,  E = 1/2 q . mat . q                          (alpb.py:300-350)

R_i are the D3 van der Waals radii (tad-mctc VDW_D3, which the reference
imports; restated below for the first 18 elements).  The diagonal of the
matrix is included (r = 0: K_ii = 1/B_i, the Born self energy).

Everything is torch so that forces follow by autograd.
"""

from __future__ import annotations

import torch

from .params import AA2AU as ANGSTROM_TO_BOHR

# D3 van der Waals radii in Angstrom, Z = 1..18 (tad-mctc VDW_D3 / the D3
# reference data); converted to Bohr on use.
VDW_D3_ANGSTROM = [
    1.09155, 0.86735,  # H He
    1.74780, 1.54910, 1.60800, 1.45515, 1.31125, 1.24085, 1.14980, 1.06870,  # Li-Ne
    1.85410, 1.74195, 2.00530, 1.89585, 1.75085, 1.65535, 1.55230, 1.45740,  # Na-Ar
]

ALPB_ALPHA = 0.571412  # alpb.py:83
OBC = (1.0, 0.8, 4.85)  # born.py:58
DESCREENING = 0.8
CUTOFF = 66.0


def vdw_radii(numbers) -> torch.Tensor:
    z = [int(x) for x in numbers]
    if max(z) > len(VDW_D3_ANGSTROM):
        raise ValueError(f"no van der Waals radius for Z = {max(z)}")
    return torch.tensor([VDW_D3_ANGSTROM[x - 1] for x in z], dtype=torch.float64) * ANGSTROM_TO_BOHR


def compute_psi(positions: torch.Tensor, rvdw: torch.Tensor, cutoff: float = CUTOFF,
                descreening: float = DESCREENING) -> torch.Tensor:
    """Descreening integral psi_i (born.py:163-239): pair (i, j) with the
    descreener j (scaled radius rho_j) and the atom i of radius R_i."""
    n = positions.shape[0]
    rho = rvdw * descreening
    psi = []
    for i in range(n):
        acc = positions.new_zeros(())
        for j in range(n):
            if i == j:
                continue
            r = torch.linalg.norm(positions[i] - positions[j])
            if r > cutoff:
                continue
            acc = acc + psi_term(r, rho[j], rvdw[i])
        psi.append(acc)
    return torch.stack(psi)


def psi_term(r, rho, ri):
    """One pair contribution: overlapping branch when r < R_i + rho and
    r + rho > R_i, otherwise the non-overlapping closed form."""
    d_pl = r + rho
    d_mi = r - rho
    if bool(r < ri + rho) and bool(d_pl > ri):
        return (1.0 / ri - 1.0 / d_pl
                + 0.5 / r * (0.5 * d_mi * (1.0 / d_pl - d_pl / ri**2) - torch.log(d_pl / ri)))
    out = rho / (d_pl * d_mi)
    q = d_mi / d_pl
    if bool(q > 0):
        out = out + 0.5 * torch.log(q) / r
    return out


def born_radii(numbers, positions: torch.Tensor, born_scale: float = 1.0, born_offset: float = 0.0) -> torch.Tensor:
    rvdw = vdw_radii(numbers)
    psi = compute_psi(positions, rvdw)
    srvdw = rvdw - born_offset
    p = 0.5 * psi * srvdw
    arg = p * (OBC[0] + p * (OBC[2] * p - OBC[1]))
    return born_scale / (1.0 / srvdw - torch.tanh(arg) / rvdw)


def p16_kernel(r, ab):
    s = torch.sqrt(ab)
    return 1.0 / (r + s * (s / (s + r * 1.028 / 16.0)) ** 16)


def still_kernel(r, ab):
    r2 = r * r
    return 1.0 / torch.sqrt(r2 + ab * torch.exp(-0.25 * r2 / ab))


def adet(positions: torch.Tensor, rad: torch.Tensor) -> torch.Tensor:
    """Shape descriptor of the ALPB correction (alpb.py:144-184)."""
    vol = rad**3
    center = (positions * vol.unsqueeze(-1)).sum(0) / vol.sum()
    d = positions - center
    diag = (d * d).sum(-1) + 0.4 * rad**2
    inertia = (vol.unsqueeze(-1).unsqueeze(-1) * (-d.unsqueeze(-1) * d.unsqueeze(-2)
                                                  + diag.unsqueeze(-1).unsqueeze(-1) * torch.eye(3, dtype=d.dtype))).sum(0)
    det = torch.linalg.det(inertia)
    return torch.sqrt(5.0 * det ** (1.0 / 3.0) / (2.0 * vol.sum()))


def gb_matrix(numbers, positions: torch.Tensor, eps: float, alpb: bool = False, kernel: str = "p16") -> torch.Tensor:
    """keps * K (+ ALPB shape term), the atom-resolved interaction matrix."""
    born = born_radii(numbers, positions)
    alpbet = ALPB_ALPHA / eps if alpb else 0.0
    keps = (1.0 / eps - 1.0) / (1.0 + alpbet)
    n = positions.shape[0]
    diff = positions.unsqueeze(1) - positions.unsqueeze(0)
    eye = torch.eye(n, dtype=torch.bool)
    r = torch.sqrt(torch.where(eye.unsqueeze(-1), 1.0, diff * diff).sum(-1))
    r = torch.where(eye, 0.0, r)
    ab = born.unsqueeze(-1) * born.unsqueeze(-2)
    k = {"p16": p16_kernel, "still": still_kernel}[kernel]
    mat = keps * k(r, ab)
    if alpb:
        mat = mat + keps * alpbet * adet(positions, vdw_radii(numbers))
    return mat
