"""Calculator facade (GFN1Calculator/GFN2Calculator) on the GPU: energies and
forces against the oracle, batched inputs, error behaviour of the reference API."""

from __future__ import annotations

import pytest
import torch

from dxtb_amd.calculators import DtypeError, GFN1Calculator, GFN2Calculator
from oracle import scf as oscf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cls,gfn", [(GFN1Calculator, "gfn1"), (GFN2Calculator, "gfn2")])
def test_energy_forces_single(cuda, geometries, cls, gfn):
    m = geometries["H2O"]
    nums = torch.tensor(m["numbers"], device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda, requires_grad=True)
    calc = cls(nums, device=cuda)
    e = calc.get_energy(pos)
    f = calc.get_forces(pos)
    ref = oscf.singlepoint(m["numbers"], torch.tensor(m["positions"], dtype=torch.float64), gfn, tol=1e-11, forces=True)
    assert e.shape == ()
    assert abs(e.item() - ref["energy"].item()) < 1e-8
    torch.testing.assert_close(f.cpu(), ref["forces"], atol=1e-7, rtol=0)


def test_batched_padded(cuda, geometries):
    a, b = geometries["H2O"], geometries["CH4"]
    nums = torch.zeros((2, 5), dtype=torch.long)
    pos = torch.zeros((2, 5, 3), dtype=torch.float64)
    nums[0, :3] = torch.tensor(a["numbers"])
    pos[0, :3] = torch.tensor(a["positions"])
    nums[1, :5] = torch.tensor(b["numbers"])
    pos[1, :5] = torch.tensor(b["positions"])
    calc = GFN2Calculator(nums.to(cuda), device=cuda)
    e = calc.get_energy(pos.to(cuda), chrg=torch.tensor([0.0, 0.0]))
    for i, g in enumerate((a, b)):
        ref = oscf.singlepoint(g["numbers"], torch.tensor(g["positions"], dtype=torch.float64), "gfn2", tol=1e-11)
        assert abs(e[i].item() - ref["energy"].item()) < 1e-8


def test_api_errors(cuda, geometries):
    m = geometries["H2O"]
    nums = torch.tensor(m["numbers"], device=cuda)
    with pytest.raises(DtypeError):
        GFN2Calculator(nums.double(), device=cuda)
    calc = GFN2Calculator(nums, device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda)
    with pytest.raises(RuntimeError):
        calc.get_forces(pos)  # requires_grad is False
    with pytest.raises(DtypeError):
        calc.get_energy(pos.float())


def test_full_mode_option(cuda, geometries):
    """opts={"scf_mode": "full"}: the reference's default SCF (Anderson
    mixing) gives the implicit mode's energy and forces."""
    m = geometries["MB16_43_01"]
    nums = torch.tensor(m["numbers"], device=cuda)
    pos = torch.tensor(m["positions"], dtype=torch.float64, device=cuda, requires_grad=True)
    full = GFN2Calculator(nums, opts={"scf_mode": "full", "x_atol": 1e-10}, device=cuda)
    impl = GFN2Calculator(nums, device=cuda)
    assert abs(full.get_energy(pos).item() - impl.get_energy(pos).item()) < 1e-8
    torch.testing.assert_close(full.get_forces(pos), impl.get_forces(pos), atol=1e-7, rtol=0)
