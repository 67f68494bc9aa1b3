"""
Pins of the oracle's Hamiltonian, electrostatics, repulsion, SCF and forces (CPU).

Reference fixtures (extracted by tests/golden/make_fixtures.py):
* GFN2 H0 without overlap for H2/LiH/SiH4 (tblite 0.3.0 FP64,
  test/test_hamiltonian/test_gfn2.py), checked at the reference's own
  tolerance sqrt(eps)*10; GFN1 H0 (float32 npz, test/test_hamiltonian/h0.npz).
* Repulsion energies GFN1/GFN2 (test/test_classical/test_repulsion/samples.py).
* AES2 energies and potentials for LiH (test/test_coulomb/test_aes2.py): these
  also pin the GFN2 counting function and the D3 covalent radii of H and Li.
* ES2 (shell resolved) for SiH4 and ES3 (atom resolved) for MB16_43_01
  (test/test_coulomb/samples.py).
* GFN1 SCF electronic energies (tblite 0.2.1, test/test_scf/samples.py) of
  H2, LiH, H2O, CH4, SiH4, MB16_43_01, LYS_xao, C60 and vancoh2, and GFN1/GFN2
  energies of the free H, C and Rn atoms (D4SC vanishes for one atom), at the
  reference's tolerance (abs and rel sqrt(eps)*10), -TS included.
Closed forms / brute force:
* Fermi occupations sum to the electron count; the free energy vanishes for a
  large gap.
* Forces against central finite differences of the SCF energy (h = 1e-4).
"""

from __future__ import annotations

import math

import pytest
import torch

from oracle import basis as obasis
from oracle import coulomb, hamiltonian, integrals, scf

REFTOL = math.sqrt(torch.finfo(torch.float64).eps) * 10
T = lambda x: torch.tensor(x, dtype=torch.float64)  # noqa: E731


@pytest.mark.parametrize("name", ["H2", "LiH", "SiH4"])
def test_h0_gfn2_noovlp(geometries, refvals, name):
    m = geometries[name]
    b = obasis.build(m["numbers"], "gfn2")
    h = hamiltonian.h0_matrix(b, T(m["positions"]), None)
    torch.testing.assert_close(h, T(refvals["h0_gfn2_noovlp"][name]), atol=REFTOL, rtol=0)


@pytest.mark.parametrize("name", ["H2", "LiH", "H2O", "CH4", "SiH4", "LYS_xao"])
def test_h0_gfn1_npz(geometries, tblite_arrays, name):
    m = geometries[name]
    b = obasis.build(m["numbers"], "gfn1")
    pos = T(m["positions"])
    s, _, _ = integrals.multipole_matrices(b, pos)
    h = hamiltonian.h0_matrix(b, pos, s)
    torch.testing.assert_close(h, T(tblite_arrays["h0_gfn1_" + name.lower()]), atol=5e-7, rtol=0)


@pytest.mark.parametrize("name", ["H2", "H2O", "SiH4", "LYS_xao", "MB16_43_01"])
@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_repulsion(geometries, refvals, name, gfn):
    m = geometries[name]
    e = coulomb.repulsion_energy(m["numbers"], T(m["positions"]), gfn).sum()
    assert abs(e.item() - refvals["repulsion"][name][gfn]) < 1e-12


def test_aes2_lih(refvals):
    a = refvals["aes2_energy"]
    b = obasis.build(a["numbers"], "gfn2")
    c = coulomb.aes2_cache(b, T(a["positions"]))
    e = coulomb.aes2_energy(c, T(a["qat"]), T(a["qdp"]), T(a["qqp"]))
    torch.testing.assert_close(e, T(a["ref_eat_dp"]) + T(a["ref_eat_qp"]), atol=1e-12, rtol=0)
    p = refvals["aes2_potential"]
    vat, vdp, vqp = coulomb.aes2_potential(c, T(p["qat"]), T(p["qdp"]), T(p["qqp"]))
    torch.testing.assert_close(vat, T(p["ref_vat_mo"]), atol=1e-12, rtol=0)
    torch.testing.assert_close(vdp, T(p["ref_vat_dp"]), atol=1e-12, rtol=0)
    torch.testing.assert_close(vqp, T(p["ref_vat_qp"]), atol=1e-12, rtol=0)


def test_es2_shell_sih4(geometries, refvals):
    m = geometries["SiH4"]
    b = obasis.build(m["numbers"], "gfn1")
    gam = coulomb.es2_matrix(b, T(m["positions"]))
    q = T(refvals["coulomb"]["SiH4"]["q"])
    assert abs((0.5 * q @ gam @ q).item() - refvals["coulomb"]["SiH4"]["es2"]) < 1e-12


def test_es3_atom_mb16(geometries, refvals):
    m = geometries["MB16_43_01"]
    b = obasis.build(m["numbers"], "gfn1")
    hd, shell = coulomb.es3_hubbard_derivs(b)
    assert not shell
    q = T(refvals["coulomb"]["MB16_43_01"]["q"])
    assert abs((hd * q**3 / 3).sum().item() - refvals["coulomb"]["MB16_43_01"]["es3"]) < 1e-8


def _scf_vs_tblite(numbers, positions, gfn, ref):
    """The reference compares ``result.scf`` (electronic energy including the
    electronic free-energy term -TS) with tblite at abs = rel = sqrt(eps)*10
    (test/test_scf/test_scf.py:48-67, 92-121)."""
    r = scf.singlepoint(numbers, T(positions), gfn)
    assert r["converged"]
    e = (r["eel"] + r["ts"]).item()
    assert abs(e - ref) <= max(REFTOL, REFTOL * abs(ref)), (e, ref)


@pytest.mark.parametrize("name", ["H2", "LiH", "H2O", "CH4", "SiH4"])
def test_scf_gfn1_vs_tblite(geometries, refvals, name):
    m = geometries[name]
    _scf_vs_tblite(m["numbers"], m["positions"], "gfn1", refvals["scf"][name]["egfn1"])


@pytest.mark.parametrize("name", ["MB16_43_01", "LYS_xao", "C60", "vancoh2"])
def test_scf_gfn1_medium_vs_tblite(geometries, refvals, name):
    """GFN1 electronic energies of larger molecules (test/test_scf/samples.py:
    582-609, test_scf.py:92-121).  These pin the CN and H0 radii of N and O
    (a 1 % change of either fails), the H0 radius of F, the CN radius of Cl,
    and B, Na, Al of MB16_43_01."""
    m = geometries[name]
    _scf_vs_tblite(m["numbers"], m["positions"], "gfn1", refvals["scf"][name]["egfn1"])


@pytest.mark.parametrize("name,z", [("H", 1), ("C", 6), ("Rn", 86)])
@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_scf_atoms_vs_tblite(refvals, name, z, gfn):
    """Free atoms (test/test_scf/samples.py:55-72): for one atom the GFN2
    D4SC term vanishes, so tblite's GFN2 energy is the energy of this path.
    The open-shell C atom pins the GFN2 ES2 arithmetic average of its s and p
    Hubbard parameters (the harmonic average is off by 1.4e-3 Eh) and the
    shell scaling of the GFN2 third-order term (scale 1 for every shell: off
    by 2.5e-2 Eh); its -TS term is 3.6e-3 Eh."""
    _scf_vs_tblite([z], [[0.0, 0.0, 0.0]], gfn, refvals["scf"][name]["e" + gfn])


def test_fermi_counts(geometries):
    m = geometries["H2O"]
    s = scf.System(m["numbers"], T(m["positions"]), "gfn2")
    e = torch.linspace(-1.0, 0.5, s.nao, dtype=torch.float64)
    occ = s.occupation(e)
    torch.testing.assert_close(occ.sum(-1), s.nab, atol=1e-12, rtol=0)
    assert abs(s.free_energy(occ).item()) < 1e-12  # large gaps: no entropy


@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_forces_vs_finite_differences(geometries, gfn):
    m = geometries["H2O"]
    pos = T(m["positions"])
    r = scf.singlepoint(m["numbers"], pos, gfn, forces=True, tol=1e-12)
    h = 1e-4
    fd = torch.zeros_like(pos)
    for a in range(pos.shape[0]):
        for x in range(3):
            pp, pm = pos.clone(), pos.clone()
            pp[a, x] += h
            pm[a, x] -= h
            ep = scf.singlepoint(m["numbers"], pp, gfn, tol=1e-12)["energy"]
            em = scf.singlepoint(m["numbers"], pm, gfn, tol=1e-12)["energy"]
            fd[a, x] = -(ep - em) / (2 * h)
    torch.testing.assert_close(r["forces"], fd, atol=1e-7, rtol=0)
    assert r["forces"].sum(0).abs().max() < 1e-10  # translational invariance
