"""
Parametrizations (GFN1-xTB, GFN2-xTB), element data and units.

The parameters are the reference's TOML files converted to JSON
(``data/make_params.py``); the element radii and unit constants, which the
reference takes from tad-mctc 0.7.0, are in ``data/elements.json``.
Mirrors the parameter access of ``src/dxtb/_src/param/module/utils.py``
(element parameters, pair parameters, valence shells).
"""

from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "data"


@lru_cache(maxsize=None)
def _json(name: str) -> dict:
    with open(DATA / f"{name}.json") as fh:
        return json.load(fh)


ELEMENTS = _json("elements")
SYMBOLS: list[str] = ["X"] + ELEMENTS["symbols"]
ATOMIC_NUMBER = {s: i for i, s in enumerate(SYMBOLS)}
MAXZ = len(SYMBOLS) - 1

BOHR_ANGSTROM = ELEMENTS["bohr_angstrom"]
AA2AU = 1.0 / BOHR_ANGSTROM
EV2AU = 1.0 / ELEMENTS["hartree_ev"]
KELVIN2AU = ELEMENTS["kb_hartree_per_kelvin"]

#: D3 covalent radii (4/3 x Pyykko 2009) in Bohr, index = Z
COV_D3 = np.array([0.0] + [4.0 / 3.0 * r * AA2AU for r in ELEMENTS["covalent_2009"]])
#: atomic radii of the H0 distance polynomial in Bohr, index = Z
ATOMIC_RADII = np.array([0.0] + [r * AA2AU for r in ELEMENTS["atomic"]])
#: D3 van der Waals radii (Angstrom, Z = 1..18) of the generalized Born
#: solvation model (the VDW_D3 table the reference takes from tad-mctc);
#: VDW_D3[Z] in Bohr, NaN where no radius is available
_VDW_D3_AA = [1.09155, 0.86735, 1.74780, 1.54910, 1.60800, 1.45515, 1.31125, 1.24085, 1.14980, 1.06870,
              1.85410, 1.74195, 2.00530, 1.89585, 1.75085, 1.65535, 1.55230, 1.45740]
VDW_D3 = np.full(MAXZ + 1, np.nan)
VDW_D3[1 : len(_VDW_D3_AA) + 1] = np.array(_VDW_D3_AA) * AA2AU

_LVAL = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}


class Parametrization:
    """Read-only view of one xTB parametrization ("gfn1" or "gfn2")."""

    def __init__(self, name: str) -> None:
        if name not in ("gfn1", "gfn2"):
            raise ValueError(f"Unknown parametrization '{name}' (expected 'gfn1' or 'gfn2').")
        self.name = name
        self.raw = _json(name)
        self.h0 = self.raw["hamiltonian"]["xtb"]
        self.elements = self.raw["element"]

    def has(self, z: int) -> bool:
        return 0 < z <= MAXZ and SYMBOLS[z] in self.elements

    def element(self, z: int) -> dict:
        try:
            return self.elements[SYMBOLS[z]]
        except (IndexError, KeyError):
            raise ValueError(f"Element Z={z} is not parametrized in {self.name}.") from None

    def shells(self, z: int) -> list[tuple[int, int]]:
        """(principal n, angular l) of every shell of an element."""
        return [(int(s[:-1]), _LVAL[s[-1]]) for s in self.element(z)["shells"]]

    def valence(self, z: int) -> list[bool]:
        """Only the first shell of each angular momentum is a valence shell."""
        seen: set[int] = set()
        out = []
        for _, l in self.shells(z):
            out.append(l not in seen)
            seen.add(l)
        return out

    def kpair(self, za: int, zb: int) -> float:
        kp = self.h0.get("kpair", {})
        a, b = SYMBOLS[za], SYMBOLS[zb]
        return float(kp.get(f"{a}-{b}", kp.get(f"{b}-{a}", 1.0)))


@lru_cache(maxsize=None)
def get(name: str) -> Parametrization:
    return Parametrization(name)


def stong_table(ng: int) -> np.ndarray:
    """STO-nG table (2, ntype, ng): coefficients and exponents."""
    return np.asarray(_json("stong")[str(ng)], dtype=np.float64)
