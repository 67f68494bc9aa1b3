"""Custom parametrizations on the GPU: ``Calculator(numbers, par)`` with a
record, a Param or a file, against the GFN1/GFN2 calculators (same
parameters under a new name: bitwise equal) and against the oracle with the
same perturbed record (energies and forces)."""

from __future__ import annotations

import copy

import numpy as np
import pytest
import torch

from dxtb_amd import param
from dxtb_amd.calculators import Calculator, GFN1Calculator, GFN2Calculator
from dxtb_amd.serving import Engine, Request
from oracle import params as oparams
from oracle import scf as oscf

pytestmark = pytest.mark.gpu

_CLS = {"gfn1": GFN1Calculator, "gfn2": GFN2Calculator}


def _shipped(name):
    return param.GFN1_XTB if name == "gfn1" else param.GFN2_XTB


def perturbed(name: str) -> dict:
    """Every section this path reads, changed by a few per cent."""
    rec = copy.deepcopy(_shipped(name).data)
    h = rec["hamiltonian"]["xtb"]
    h["kpol"] *= 0.97
    h["enscale"] *= 1.2
    h["wexp"] += 0.05
    for k in h["shell"]:
        h["shell"][k] *= 1.02
    if "kpair" in h:
        h["kpair"]["H-H"] = 0.9
    rec["charge"]["effective"]["gexp"] = 2.1
    rec["repulsion"]["effective"]["kexp"] += 0.05
    if isinstance(rec["thirdorder"].get("shell"), dict):
        for k in rec["thirdorder"]["shell"]:
            rec["thirdorder"]["shell"][k] *= 1.1
    if "multipole" in rec:
        d = rec["multipole"]["damped"]
        d["dmp3"] += 0.2
        d["kexp"] -= 0.2
    for sym in ("H", "C", "N", "O"):
        e = rec["element"][sym]
        e["levels"] = [v * 1.03 for v in e["levels"]]
        e["kcn"] = [v * 1.2 for v in e["kcn"]]
        e["shpoly"] = [v * 0.9 for v in e["shpoly"]]
        e["gam"] *= 0.95
        e["gam3"] *= 1.1
        e["arep"] *= 1.05
        e["dkernel"] *= 1.3
        e["qkernel"] *= 1.3
        e["mprad"] *= 1.1
    rec["meta"]["reference"] = "perturbed for a test"
    return rec


@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
def test_renamed_parameters_bitwise(cuda, geometries, name):
    """The shipped parameters under a custom name: the same kernels with the
    same values, so bitwise the same energy and forces."""
    m = geometries["H2O"]
    nums = torch.tensor(m["numbers"], device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda, requires_grad=True)
    rec = copy.deepcopy(_shipped(name).data)
    rec["meta"]["reference"] = "renamed copy"
    calc = Calculator(nums, rec, device=cuda)
    assert calc.gfn.startswith(f"{name}-custom-")
    ref = _CLS[name](nums, device=cuda)
    assert calc.get_energy(pos).item() == ref.get_energy(pos).item()
    assert torch.equal(calc.get_forces(pos), ref.get_forces(pos))
    # the shipped record itself keeps its name
    assert Calculator(nums, _shipped(name), device=cuda).gfn == name


@pytest.mark.parametrize("name", ["gfn1", "gfn2"])
@pytest.mark.parametrize("mol", ["H2O", "CH4"])
def test_perturbed_vs_oracle(cuda, geometries, name, mol):
    m = geometries[mol]
    rec = perturbed(name)
    nums = torch.tensor(m["numbers"], device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda, requires_grad=True)
    calc = Calculator(nums, param.Param(rec), device=cuda)
    e = calc.get_energy(pos)
    f = calc.get_forces(pos)
    oname = oparams.register(calc.gfn, rec)
    ref = oscf.singlepoint(m["numbers"], torch.tensor(m["positions"], dtype=torch.float64), oname, tol=1e-11,
                           forces=True)
    assert abs(e.item() - ref["energy"].item()) < 1e-8
    torch.testing.assert_close(f.cpu(), ref["forces"], atol=1e-7, rtol=0)
    # the perturbation matters
    e0 = _CLS[name](nums, device=cuda).get_energy(pos.detach())
    assert abs(e.item() - e0.item()) > 1e-3


def test_parameter_file(cuda, geometries, tmp_path):
    """A TOML file of the perturbed record gives the record's results."""
    m = geometries["H2O"]
    rec = perturbed("gfn2")
    f = tmp_path / "custom.toml"
    param.Param(rec).to_file(f)
    nums = torch.tensor(m["numbers"], device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda)
    a = Calculator(nums, f, device=cuda)
    b = Calculator(nums, rec, device=cuda)
    assert a.gfn == b.gfn
    assert a.get_energy(pos).item() == b.get_energy(pos).item()


def test_engine_with_custom_parameters(cuda, geometries):
    """The serving engine takes a Param: energy and forces of the oracle with
    the same record."""
    rec = perturbed("gfn1")
    m = geometries["CH4"]
    eng = Engine(param.Param(rec), cuda)
    assert eng.gfn.startswith("gfn1-custom-")
    req = Request(0, np.asarray(m["numbers"]), np.asarray(m["positions"], dtype=np.float64))
    out = eng.run([req], realtime=False)[0]
    oname = oparams.register(eng.gfn, rec)
    ref = oscf.singlepoint(m["numbers"], torch.tensor(m["positions"], dtype=torch.float64), oname, tol=1e-10,
                           forces=True)
    assert abs(out.energy - float(ref["energy"])) < 1e-8
    np.testing.assert_allclose(out.forces, ref["forces"].numpy(), atol=1e-6, rtol=0)
