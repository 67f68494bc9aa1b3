"""Structure and charge readers: the inputs of a single point.

Turbomole ``coord`` files (Bohr), xyz files (Angstrom) and ``.CHRG`` files,
as the reference's tests and command line read them
(test/test_singlepoint/test_energy.py:70-72 reads ``mols/<name>/coord`` and
``.CHRG``; cli/driver.py:93-116 takes the charge from a ``.CHRG`` file beside
the input, else 0).  The reference's readers live in tad-mctc 0.7.0
(``tad_mctc.io.read``, ``read_chrg``), which is not available here; these
are written from the file formats:

* ``coord``: the ``$coord`` data group holds one atom per line, ``x y z
  symbol`` in Bohr; other data groups (``$periodic``, ``$user-defined
  bonds``, ``$redundant``, ...) and ``$end`` are skipped;
* xyz: atom count, comment line, then ``symbol x y z`` in Angstrom;
* ``.CHRG``: one number, the molecular charge.

Positions come back in Bohr (1 Bohr = 0.529177210903 Angstrom, CODATA 2018:
the reference's NO2 ``coord`` and ``mol.xyz`` of the same structure agree
to 5e-15 Bohr with this value).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .param import ATOMIC_NUMBER

BOHR_IN_ANGSTROM = 0.529177210903
_Z = {s.lower(): z for s, z in ATOMIC_NUMBER.items() if z > 0}


def _atomic_number(symbol: str, path: Path, line: int) -> int:
    z = _Z.get(symbol.strip().lower())
    if z is None:
        raise ValueError(f"{path}:{line}: unknown element symbol '{symbol}'")
    return z


def read_coord(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """Atomic numbers (nat,) and positions (nat, 3) in Bohr from the
    ``$coord`` data group of a Turbomole file."""
    path = Path(path)
    nums, pos = [], []
    inside = found = False
    for k, line in enumerate(path.read_text().splitlines(), 1):
        s = line.strip()
        if s.startswith("$"):
            inside = s.split()[0] == "$coord"
            if inside and "frac" in s.split()[1:]:
                raise ValueError(f"{path}:{k}: fractional coordinates ('$coord frac') are not supported")
            found = found or inside
            continue
        if not inside or not s or s.startswith("#"):
            continue
        t = s.split()
        if len(t) < 4:
            raise ValueError(f"{path}:{k}: expected 'x y z symbol', got '{s}'")
        pos.append([float(x) for x in t[:3]])
        nums.append(_atomic_number(t[3], path, k))
    if not found:
        raise ValueError(f"{path}: no $coord data group")
    if not nums:
        raise ValueError(f"{path}: $coord data group holds no atoms")
    return np.asarray(nums, dtype=np.int64), np.asarray(pos, dtype=np.float64)


def read_xyz(path: str | Path) -> tuple[np.ndarray, np.ndarray]:
    """Atomic numbers (nat,) and positions (nat, 3) in Bohr from an xyz file
    (first structure)."""
    path = Path(path)
    lines = path.read_text().splitlines()
    if not lines:
        raise ValueError(f"{path}: empty file")
    try:
        nat = int(lines[0].split()[0])
    except (ValueError, IndexError):
        raise ValueError(f"{path}:1: expected the number of atoms") from None
    if nat <= 0 or len(lines) < 2 + nat:
        raise ValueError(f"{path}: announces {nat} atoms, holds {max(len(lines) - 2, 0)} lines")
    nums, pos = [], []
    for k in range(2, 2 + nat):
        t = lines[k].split()
        if len(t) < 4:
            raise ValueError(f"{path}:{k + 1}: expected 'symbol x y z', got '{lines[k].strip()}'")
        nums.append(_atomic_number(t[0], path, k + 1))
        pos.append([float(x) / BOHR_IN_ANGSTROM for x in t[1:4]])
    return np.asarray(nums, dtype=np.int64), np.asarray(pos, dtype=np.float64)


def read(path: str | Path, device=None, dtype=torch.float64) -> tuple[torch.Tensor, torch.Tensor]:
    """Atomic numbers and positions (Bohr) as tensors, by file name: ``coord``
    (or a ``.coord``/``.tmol`` suffix) is Turbomole, ``.xyz`` is xyz."""
    path = Path(path)
    if path.name == "coord" or path.suffix in (".coord", ".tmol"):
        nums, pos = read_coord(path)
    elif path.suffix == ".xyz":
        nums, pos = read_xyz(path)
    else:
        raise ValueError(f"{path}: unknown structure format (expected 'coord', '.coord', '.tmol' or '.xyz')")
    return (torch.as_tensor(nums, dtype=torch.int64, device=device),
            torch.as_tensor(pos, dtype=dtype, device=device))


def read_chrg(path: str | Path, device=None, dtype=torch.float64) -> torch.Tensor:
    """Molecular charge from a ``.CHRG`` file; 0 if the file does not exist
    or is empty."""
    path = Path(path)
    text = path.read_text().strip() if path.is_file() else ""
    if not text:
        return torch.tensor(0.0, dtype=dtype, device=device)
    try:
        value = float(text.split()[0])
    except ValueError:
        raise ValueError(f"{path}: expected a number, got '{text.split()[0]}'") from None
    return torch.tensor(value, dtype=dtype, device=device)
