"""Parametrization records (CPU): file formats, model family from the meta
name, registration of custom parametrizations, and the oracle with a custom
record (a new name for unchanged parameters changes nothing; a change of the
repulsion parameters changes the energy by the closed-form repulsion
difference alone)."""

from __future__ import annotations

import copy
import math
from pathlib import Path

import pytest
import torch

from dxtb_amd import param
from oracle import params as oparams
from oracle import scf as oscf

REF_PARAM = Path("/root/reference/src/dxtb/_src/param")


def _shipped(name):
    return param.GFN1_XTB if name == "gfn1" else param.GFN2_XTB


@pytest.mark.parametrize("suffix", [".json", ".toml", ".yaml"])
@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
def test_file_round_trip(tmp_path, name, suffix):
    p = _shipped(name)
    f = tmp_path / f"par{suffix}"
    p.to_file(f)
    q = param.Param.from_file(f)
    assert q == p
    assert param.register(q) == name
    assert param.register(f) == name


def test_unknown_suffix(tmp_path):
    with pytest.raises(ValueError, match="Unsupported file format"):
        param.GFN2_XTB.to_file(tmp_path / "par.ini")
    with pytest.raises(ValueError, match="Unsupported file format"):
        param.Param.from_file(tmp_path / "par.ini")


@pytest.mark.skipif(not REF_PARAM.exists(), reason="reference checkout not present")
@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
def test_reference_toml_files(name):
    """The reference's own TOML files read as the shipped records."""
    p = param.Param.from_file(REF_PARAM / name / f"{name}-xtb.toml")
    assert p == _shipped(name)
    assert param.register(p) == name


def test_model_family_from_meta_name():
    rec = copy.deepcopy(param.GFN2_XTB.data)
    for meta, fam in (("GFN2-xTB", "gfn2"), ("gfn2", "gfn2"), ("GFN1-XTB", "gfn1"), ("gfn1", "gfn1")):
        rec["meta"]["name"] = meta
        assert param.model_kind(rec) == fam
    rec["meta"]["name"] = "GFN0-xTB"
    with pytest.raises(NotImplementedError):
        param.model_kind(rec)
    rec["meta"]["name"] = "GFN-FF"
    with pytest.raises(ValueError, match="Unsupported Hamiltonian"):
        param.model_kind(rec)
    del rec["meta"]["name"]
    with pytest.raises(ValueError, match="name"):
        param.model_kind(rec)


def test_register_names():
    rec = copy.deepcopy(param.GFN2_XTB.data)
    rec["element"]["C"]["gam"] *= 1.01
    a = param.register(rec)
    assert a.startswith("gfn2-custom-") and param.kind(a) == "gfn2"
    assert param.register(param.Param(rec)) == a  # content decides the name
    assert param.register(a) == a
    rec["element"]["C"]["gam"] *= 1.01
    b = param.register(rec)
    assert b != a
    # the registered record is a copy: later changes to the caller's record
    # do not reach it
    assert param.get(a).element(6)["gam"] == pytest.approx(param.GFN2_XTB.data["element"]["C"]["gam"] * 1.01)
    with pytest.raises(ValueError):
        param.get("gfn2-custom-unknown")
    with pytest.raises(ValueError):
        param.kind("gfn0")


def test_register_needs_the_computed_sections():
    rec = copy.deepcopy(param.GFN2_XTB.data)
    del rec["multipole"]
    with pytest.raises(NotImplementedError, match="multipole"):
        param.register(rec)
    rec = copy.deepcopy(param.GFN1_XTB.data)
    rec["thirdorder"] = None
    with pytest.raises(NotImplementedError, match="thirdorder"):
        param.register(rec)


def test_element_defaults():
    """Element fields a record leaves out read as the reference's defaults."""
    rec = copy.deepcopy(param.GFN1_XTB.data)
    for k in ("dkernel", "qkernel", "mprad", "mpvcn", "xbond"):
        del rec["element"]["O"][k]
    name = param.register(rec)
    e = param.get(name).element(8)
    assert all(e[k] == 0.0 for k in ("dkernel", "qkernel", "mprad", "mpvcn", "xbond"))
    assert e["gam"] == param.GFN1_XTB.data["element"]["O"]["gam"]


def _oracle_energy(rec, numbers, positions):
    name = oparams.register(param.register(rec), rec)
    return oscf.singlepoint(numbers, positions, name, tol=1e-11)["energy"].item()


@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
def test_oracle_renamed_record_is_unchanged(geometries, name):
    """The same parameters under a custom name give the same oracle energy:
    every model branch follows the family of the name (GFN1: orthogonalised
    H 2s and exponential CN; GFN2: multipoles and the GFN2 CN)."""
    m = geometries["H2O"]
    rec = copy.deepcopy(_shipped(name).data)
    rec["meta"]["reference"] = "renamed copy"
    pos = torch.tensor(m["positions"], dtype=torch.float64)
    e_custom = _oracle_energy(rec, m["numbers"], pos)
    e_shipped = oscf.singlepoint(m["numbers"], pos, name, tol=1e-11)["energy"].item()
    assert e_custom == e_shipped


@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
def test_oracle_repulsion_closed_form(name):
    """N2: a new arep of N and a new kexp change only the repulsion
    E = Zeff_A Zeff_B / R exp(-sqrt(a_A a_B) R^kexp), the SCF not at all."""
    R = 2.1
    numbers = [7, 7]
    pos = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, R]], dtype=torch.float64)
    base = _shipped(name).data
    rec = copy.deepcopy(base)
    rec["element"]["N"]["arep"] *= 1.15
    rec["repulsion"]["effective"]["kexp"] += 0.1

    def erep(r):
        zeff, a, k = r["element"]["N"]["zeff"], r["element"]["N"]["arep"], r["repulsion"]["effective"]["kexp"]
        return zeff * zeff / R * math.exp(-a * R**k)

    e0 = oscf.singlepoint(numbers, pos, name, tol=1e-11)["energy"].item()
    e1 = _oracle_energy(rec, numbers, pos)
    assert abs((e1 - e0) - (erep(rec) - erep(base))) < 1e-10
    assert abs(erep(rec) - erep(base)) > 1e-3
