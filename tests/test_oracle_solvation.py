"""
Oracle generalized Born solvation against the reference's fixtures
(test/test_solvation/samples.py, tblite values): descreening integrals and
Born radii (MB16_43_01, SiH4), per-atom GB energies for given charges with
the P16 kernel + ALPB and with the Still kernel (eps = 78.9), and the GFN1 SCF
solvation energy and SCF-energy gradient with GB (MB16_43_01).  Closed forms:
an isolated ion has B = R_vdw and the Born energy 1/2 keps q^2 / R.

The D3 van der Waals radii restated in the oracle are pinned here for H, B,
N, O, F, Na, Al, Si and Cl; C and S come from the same table and no
reference fixture covers them.
"""

from __future__ import annotations

import pytest
import torch

from oracle import scf as oscf
from oracle import solvation as sol


@pytest.fixture(scope="module")
def gbref(refvals):
    return refvals["solvation"]


def _mol(geometries, name):
    m = geometries[name]
    return m["numbers"], torch.tensor(m["positions"], dtype=torch.float64)


@pytest.mark.parametrize("name", ["MB16_43_01", "SiH4"])
def test_psi_and_born_radii(geometries, gbref, name):
    z, pos = _mol(geometries, name)
    psi = sol.compute_psi(pos, sol.vdw_radii(z))
    born = sol.born_radii(z, pos)
    torch.testing.assert_close(psi, torch.tensor(gbref[name]["psi"], dtype=torch.float64), rtol=1e-9, atol=1e-12)
    torch.testing.assert_close(born, torch.tensor(gbref[name]["born"], dtype=torch.float64), rtol=1e-9, atol=0)


@pytest.mark.parametrize("key,alpb,kernel", [("energies", True, "p16"), ("energies_still", False, "still")])
def test_gb_energies_given_charges(geometries, gbref, key, alpb, kernel):
    z, pos = _mol(geometries, "MB16_43_01")
    q = torch.tensor(gbref["MB16_43_01"]["charges"], dtype=torch.float64)
    mat = sol.gb_matrix(z, pos, 78.9, alpb=alpb, kernel=kernel)
    e = 0.5 * q * (mat @ q)
    torch.testing.assert_close(e, torch.tensor(gbref["MB16_43_01"][key], dtype=torch.float64), atol=1e-8, rtol=0)


def test_born_ion_closed_form():
    pos = torch.zeros(1, 3, dtype=torch.float64)
    r = sol.vdw_radii([17])
    torch.testing.assert_close(sol.born_radii([17], pos), r)
    eps = 78.4
    mat = sol.gb_matrix([17], pos, eps)
    torch.testing.assert_close(0.5 * mat[0, 0], 0.5 * (1 / eps - 1) / r[0])


def test_gfn1_scf_solvation_energy_and_gradient(geometries, gbref):
    """GFN1 SCF with GB (P16, ALPB, eps 78.9): the solvation energy (sum of
    the per-atom fixture) and the gradient of the SCF energy (electronic part,
    no repulsion) at 1e-5, the reference's own bar."""
    z, pos = _mol(geometries, "MB16_43_01")
    gb = {"eps": 78.9, "alpb": True, "kernel": "p16"}
    vac = oscf.singlepoint(z, pos, "gfn1", tol=1e-10)
    wet = oscf.singlepoint(z, pos, "gfn1", tol=1e-10, forces=True, solvation=gb)
    gsolv = float(wet["eel"] + wet["ts"] - vac["eel"] - vac["ts"])
    ref = sum(gbref["MB16_43_01"]["gsolv"])
    assert abs(gsolv - ref) < 1e-5
    # SCF-energy gradient = -forces - d(repulsion)/dR
    p = pos.clone().requires_grad_(True)
    from oracle import coulomb

    (grep,) = torch.autograd.grad(coulomb.repulsion_energy(z, p, "gfn1").sum(), p)
    grad = -wet["forces"] - grep
    torch.testing.assert_close(grad, torch.tensor(gbref["MB16_43_01"]["gradient"], dtype=torch.float64), atol=1e-5,
                               rtol=0)
