"""
Generalized Born implicit solvation on the HIP path.

* Born radii against the reference's fixtures (test/test_solvation/samples.py,
  MB16_43_01 and SiH4) at 1e-9 relative.
* SCF energy and forces with GBSA (P16, eps 78.4), ALPB and the Still kernel
  against the oracle (autograd forces through the GB matrix and the Born radii)
  at the parity bars 1e-8 Eh / 1e-6 Eh/Bohr, neutral and charged.
* The reference's GFN1 + ALPB (eps 78.9) fixtures for MB16_43_01: solvation
  energy (sum of the per-atom values) and SCF-energy gradient at 1e-5, the
  reference's own bar.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from dxtb_amd import basis as hbasis
from dxtb_amd import scf as hscf
from dxtb_amd.solvation import GeneralizedBorn
from oracle import scf as oscf

pytestmark = pytest.mark.gpu


def _run(nums, pos, gfn, dev, charge=0.0, solvation=None):
    N = np.asarray(nums, dtype=np.int64)[None]
    bb = hbasis.build_batch(N, gfn, dev)
    st = hscf.GeometryState(bb, torch.tensor(pos, dtype=torch.float64, device=dev)[None], charges=[charge],
                            solvation=solvation)
    res = hscf.run_scf(st)
    return st, res, hscf.forces(res)[0].cpu()


@pytest.mark.parametrize("name", ["MB16_43_01", "SiH4"])
def test_born_radii_fixture(cuda, geometries, refvals, name):
    m = geometries[name]
    N = np.asarray(m["numbers"], dtype=np.int64)[None]
    bb = hbasis.build_batch(N, "gfn2", cuda)
    st = hscf.GeometryState(bb, torch.tensor(m["positions"], dtype=torch.float64, device=cuda)[None],
                            solvation=GeneralizedBorn(None, 78.4, alpb=False))
    ref = torch.tensor(refvals["solvation"][name]["born"], dtype=torch.float64)
    torch.testing.assert_close(st.born[0, : len(ref)].cpu(), ref, rtol=1e-9, atol=0)


CASES = [
    ("H2O", "gfn2", 0.0, dict(eps=78.4, alpb=False, kernel="p16")),
    ("MB16_43_01", "gfn2", 0.0, dict(eps=78.4, alpb=False, kernel="p16")),
    ("H2O", "gfn2", 1.0, dict(eps=78.4, alpb=False, kernel="p16")),
    ("SiH4", "gfn1", 0.0, dict(eps=78.9, alpb=True, kernel="p16")),
    ("H2O", "gfn2", 0.0, dict(eps=78.9, alpb=False, kernel="still")),
    ("NO2", "gfn2", -1.0, dict(eps=78.9, alpb=True, kernel="p16")),
]


@pytest.mark.parametrize("name,gfn,charge,gb", CASES)
def test_scf_energy_forces_vs_oracle(cuda, geometries, name, gfn, charge, gb):
    m = geometries[name]
    print(f"case {name} {gfn} charge {charge} {gb}", flush=True)
    settings = GeneralizedBorn(None, gb["eps"], alpb=gb["alpb"], kernel=gb["kernel"])
    _, res, f = _run(m["numbers"], m["positions"], gfn, cuda, charge, settings)
    ref = oscf.singlepoint(m["numbers"], torch.tensor(m["positions"], dtype=torch.float64), gfn, charge=charge,
                           tol=1e-10, forces=True, solvation=gb)
    assert res.converged[0]
    assert abs(float(res.energy[0]) - float(ref["energy"])) < 1e-8
    torch.testing.assert_close(f[: len(m["numbers"])], ref["forces"], atol=1e-6, rtol=0)


def test_reference_alpb_fixture(cuda, geometries, refvals):
    """GFN1 + ALPB (P16, eps 78.9) on MB16_43_01: solvation energy and the
    gradient of the SCF energy (electronic part) against the reference."""
    m = geometries["MB16_43_01"]
    gb = GeneralizedBorn(None, 78.9, alpb=True, kernel="p16")
    _, wet, f = _run(m["numbers"], m["positions"], "gfn1", cuda, 0.0, gb)
    _, vac, _ = _run(m["numbers"], m["positions"], "gfn1", cuda, 0.0, None)
    gsolv = float(wet.eel[0] + wet.ts[0] - vac.eel[0] - vac.ts[0])
    assert abs(gsolv - sum(refvals["solvation"]["MB16_43_01"]["gsolv"])) < 1e-5
    from oracle import coulomb

    p = torch.tensor(m["positions"], dtype=torch.float64, requires_grad=True)
    (grep,) = torch.autograd.grad(coulomb.repulsion_energy(m["numbers"], p, "gfn1").sum(), p)
    grad = -f[: len(m["numbers"])] - grep
    torch.testing.assert_close(grad, torch.tensor(refvals["solvation"]["MB16_43_01"]["gradient"],
                                                  dtype=torch.float64), atol=1e-5, rtol=0)


def test_batch_mixed_solvation(cuda, geometries):
    """One padded batch with a solvated and a vacuum molecule equals the two
    single runs."""
    a, b = geometries["H2O"], geometries["SiH4"]
    gb = GeneralizedBorn(None, 78.4, alpb=False)
    from dxtb_amd import workloads

    N, P = workloads.pad([np.array(a["numbers"]), np.array(b["numbers"])],
                         [np.array(a["positions"]), np.array(b["positions"])])
    bb = hbasis.build_batch(N, "gfn2", cuda)
    st = hscf.GeometryState(bb, torch.as_tensor(P, device=cuda), solvation=[gb, None])
    res = hscf.run_scf(st)
    f = hscf.forces(res).cpu()
    _, ra, fa = _run(a["numbers"], a["positions"], "gfn2", cuda, 0.0, gb)
    _, rb, fb = _run(b["numbers"], b["positions"], "gfn2", cuda, 0.0, None)
    assert abs(float(res.energy[0] - ra.energy[0])) < 1e-9 and abs(float(res.energy[1] - rb.energy[0])) < 1e-9
    torch.testing.assert_close(f[0, :3], fa[:3], atol=1e-8, rtol=0)
    torch.testing.assert_close(f[1, :5], fb[:5], atol=1e-8, rtol=0)
