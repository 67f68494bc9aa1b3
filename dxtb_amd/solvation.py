"""
Generalized Born implicit solvation (GBSA and ALPB electrostatics) on the HIP
path: settings object mirroring the reference's GeneralizedBorn interaction
(components/interactions/solvation/alpb.py:187-250) and the per-molecule
arrays the kernels of ``csrc/solvation.hip`` read.

    gb = GeneralizedBorn(numbers, dielectric_constant=78.4, alpb=False)   # GBSA
    calc = GFN2Calculator(numbers, interaction=[gb], device="cuda:0")

Per molecule: keps = (1/eps - 1) / (1 + alpbet) with alpbet = 0.571412/eps
for ALPB (0 for GBSA), the kernel (P16 by default, or Still), Born radii from
the Onufriev-Bashford-Case integrator with descreening 0.8 and a 66 Bohr
cutoff.  ALPB adds the shape term keps alpbet adet Q_tot to every atom's
potential (alpb.py:316-318, get_adet 144-184).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, param

ALPB_ALPHA = 0.571412
KERNELS = {"p16": 1, "still": 2}


@dataclass
class GeneralizedBorn:
    """Settings of the generalized Born model (reference defaults: ALPB with
    the P16 kernel, born_scale 1, born_offset 0)."""

    numbers: torch.Tensor | np.ndarray | None = None
    dielectric_constant: float = 80.3
    alpb: bool = True
    kernel: str = "p16"
    born_scale: float = 1.0
    born_offset: float = 0.0
    device: object = None
    dtype: object = None

    def __post_init__(self):
        if torch.is_tensor(self.dielectric_constant):
            self.dielectric_constant = float(self.dielectric_constant)
        if self.kernel not in KERNELS:
            raise ValueError(f"unknown generalized Born kernel '{self.kernel}'")
        if self.dielectric_constant <= 0:
            raise ValueError("dielectric constant must be positive")

    @property
    def alpbet(self) -> float:
        return ALPB_ALPHA / self.dielectric_constant if self.alpb else 0.0

    @property
    def keps(self) -> float:
        return (1.0 / self.dielectric_constant - 1.0) / (1.0 + self.alpbet)


def per_molecule(solvation, nbatch: int) -> list:
    """Normalize None / one GeneralizedBorn / a list (None or GeneralizedBorn
    per molecule) to a list of length nbatch."""
    if solvation is None:
        return [None] * nbatch
    if isinstance(solvation, GeneralizedBorn):
        return [solvation] * nbatch
    out = list(solvation)
    if len(out) != nbatch:
        raise ValueError(f"{len(out)} solvation settings for {nbatch} molecules")
    for s in out:
        if s is not None and not isinstance(s, GeneralizedBorn):
            raise TypeError("solvation entries must be GeneralizedBorn or None")
    return out


def gb_params(settings: list) -> _lib.DxtbGbParams:
    scales = {(s.born_scale, s.born_offset) for s in settings if s is not None}
    if len(scales) > 1:
        raise ValueError("one born_scale/born_offset per batch")
    scale, offset = scales.pop() if scales else (1.0, 0.0)
    p = _lib.DxtbGbParams()
    p.descreening, p.cutoff, p.born_scale, p.born_offset = 0.8, 66.0, scale, offset
    p.obc[0], p.obc[1], p.obc[2] = 1.0, 0.8, 4.85
    return p


_VDW_DEV: dict = {}


def vdw_device(device) -> torch.Tensor:
    dev = torch.device(device)
    if dev not in _VDW_DEV:
        _VDW_DEV[dev] = torch.as_tensor(np.nan_to_num(param.VDW_D3, nan=0.0), dtype=torch.float64, device=dev)
    return _VDW_DEV[dev]


def check_elements(numbers: np.ndarray, settings: list) -> None:
    for b, s in enumerate(settings):
        if s is None:
            continue
        z = numbers[b][numbers[b] != 0]
        bad = sorted({int(x) for x in z if not np.isfinite(param.VDW_D3[int(x)])})
        if bad:
            raise ValueError(f"no van der Waals radius for elements {bad} (generalized Born supports Z <= 18)")


def adet(pos: torch.Tensor, rad: torch.Tensor, mask: torch.Tensor) -> torch.Tensor:
    """ALPB shape descriptor of every molecule (alpb.py:144-184), batched;
    padding atoms carry zero volume."""
    vol = torch.where(mask, rad**3, 0.0)
    center = (pos * vol.unsqueeze(-1)).sum(-2) / vol.sum(-1, keepdim=True)
    d = torch.where(mask.unsqueeze(-1), pos - center.unsqueeze(-2), 0.0)
    diag = (d * d).sum(-1) + 0.4 * rad**2
    eye = torch.eye(3, dtype=pos.dtype, device=pos.device)
    inertia = (vol[..., None, None] * (diag[..., None, None] * eye - d.unsqueeze(-1) * d.unsqueeze(-2))).sum(-3)
    return torch.sqrt(5.0 * torch.linalg.det(inertia) ** (1.0 / 3.0) / (2.0 * vol.sum(-1)))
