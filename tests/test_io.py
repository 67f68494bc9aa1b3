"""Structure and charge readers against the reference's own structure files
(test/test_singlepoint/mols/*: Turbomole coord, xyz and .CHRG files, kept as
data under tests/golden/mols)."""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from dxtb_amd import io

from conftest import GOLDEN

MOLS = GOLDEN / "mols"


def test_no2_coord_and_xyz_agree():
    """The same NO2 structure written in Bohr (coord) and Angstrom (mol.xyz):
    one Bohr radius converts one into the other to the files' precision."""
    zc, pc = io.read_coord(MOLS / "NO2" / "coord")
    zx, px = io.read_xyz(MOLS / "NO2" / "mol.xyz")
    assert zc.tolist() == zx.tolist() == [7, 8, 8]
    assert np.abs(pc - px).max() < 1e-13
    # a plausible wrong constant (the 2014 CODATA radius) already misses by 1e-9
    assert np.abs(px * io.BOHR_IN_ANGSTROM / 0.52917721067 - pc).max() > 1e-10


@pytest.mark.parametrize("name", ["AD7en+", "H2", "H", "LYS_xao", "NO2"])
def test_coord_matches_committed_geometries(name):
    """coord files with extra data groups ($periodic, $user-defined bonds,
    $redundant) parse to the geometries the fixture script read."""
    geo = json.loads((GOLDEN / "geometries.json").read_text())[name]
    z, p = io.read(MOLS / name / "coord")
    assert z.dtype == torch.int64 and p.dtype == torch.float64
    assert z.tolist() == geo["numbers"]
    torch.testing.assert_close(p, torch.tensor(geo["positions"], dtype=torch.float64), atol=0, rtol=0)


def test_charge_files():
    assert float(io.read_chrg(MOLS / "AD7en+" / ".CHRG")) == 1.0
    assert float(io.read_chrg(MOLS / "H2" / ".CHRG")) == 0.0
    assert float(io.read_chrg(MOLS / "LYS_xao" / ".CHRG")) == 0.0  # no file: neutral


def test_reader_errors(tmp_path):
    with pytest.raises(ValueError, match="unknown structure format"):
        io.read(tmp_path / "mol.pdb")
    bad = tmp_path / "coord"
    bad.write_text("$coord\n 0.0 0.0 0.0 xx\n$end\n")
    with pytest.raises(ValueError, match="unknown element symbol"):
        io.read_coord(bad)
    bad.write_text("$title\nfoo\n$end\n")
    with pytest.raises(ValueError, match="no \\$coord"):
        io.read_coord(bad)
    bad.write_text("$coord frac\n 0.0 0.0 0.0 h\n$end\n")
    with pytest.raises(ValueError, match="fractional"):
        io.read_coord(bad)
    xyz = tmp_path / "m.xyz"
    xyz.write_text("3\ncomment\nH 0 0 0\n")
    with pytest.raises(ValueError, match="announces 3 atoms"):
        io.read_xyz(xyz)
    chrg = tmp_path / ".CHRG"
    chrg.write_text("one\n")
    with pytest.raises(ValueError, match="expected a number"):
        io.read_chrg(chrg)
