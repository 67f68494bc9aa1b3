"""
Pins of the oracle's integral code (CPU).

* Every 1D Obara-Saika table entry is checked against brute-force numerical
  quadrature of the defining integral (independent of the recursion).
* One-centre moments against the closed form (n-1)!!/(2p)^(n/2)
  (the reference's test/test_integrals/test_multipole_kernels.py:93-111).
* <s|r|s> = S (aA + bB)/p (test_multipole_kernels.py:128-157).
* Full GFN1 overlap matrices against the reference's tblite float32 fixtures
  (test/test_overlap/overlap.npz).
* Invariants of the final D/Q: translation invariance (catches a wrong
  sign/index in the r0->rj shift), tracelessness, symmetry of S, unit diagonal.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import basis as obasis
from oracle import integrals as oint


def _quad_1d(i, j, e, a, b, A, B, C=0.0):
    # dense trapezoid rule: exponentially convergent for Gaussian integrands
    p = a + b
    P = (a * A + b * B) / p
    w = 12.0 / math.sqrt(p)
    x = np.linspace(P - w, P + w, 20001)
    f = (x - A) ** i * (x - B) ** j * (x - C) ** e * np.exp(-a * (x - A) ** 2 - b * (x - B) ** 2)
    return np.trapezoid(f, x)


@pytest.mark.parametrize("seed", range(6))
def test_os_table_against_quadrature(seed):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.2, 3.0, size=2)
    A, B = rng.uniform(-1.5, 1.5, size=2)
    C = 0.0
    p = a + b
    t = oint.os_table(
        2, 2, 2,
        torch.tensor(0.5 / p),
        torch.tensor((a * A + b * B) / p - A),
        torch.tensor((a * A + b * B) / p - B),
        torch.tensor((a * A + b * B) / p - C),
    )
    pref = math.sqrt(math.pi / p) * math.exp(-a * b / p * (A - B) ** 2)
    for (i, j, e), v in t.items():
        ref = _quad_1d(i, j, e, a, b, A, B, C)
        assert abs(float(v) * pref - ref) < 1e-11 * max(1.0, abs(ref)), (i, j, e)


@pytest.mark.parametrize("la,lb", [(0, 0), (1, 0), (1, 1), (2, 1), (2, 2)])
def test_pair_blocks_against_quadrature(la, lb):
    """Contracted spherical blocks of all 10 raw moments vs 3D product quadrature."""
    rng = np.random.default_rng(10 * la + lb)
    aa, ab = rng.uniform(0.3, 2.0, size=2), rng.uniform(0.3, 2.0, size=3)
    ca, cb = rng.uniform(0.2, 1.0, size=2), rng.uniform(0.2, 1.0, size=3)
    RA = rng.uniform(-1.0, 1.0, size=3)
    RB = rng.uniform(-1.0, 1.0, size=3)
    blk = oint.pair_blocks(
        la, lb, torch.tensor(aa), torch.tensor(ca), torch.tensor(ab), torch.tensor(cb),
        torch.tensor(RA)[None], torch.tensor(RB)[None],
    )[0]
    cart = np.zeros((10, len(oint.NLM[la]), len(oint.NLM[lb])))
    for t, comp in enumerate(oint.COMPS):
        for x, na in enumerate(oint.NLM[la]):
            for y, nb in enumerate(oint.NLM[lb]):
                acc = 0.0
                for i in range(2):
                    for j in range(3):
                        v = ca[i] * cb[j]
                        for ax in range(3):
                            v *= _quad_1d(na[ax], nb[ax], comp[ax], aa[i], ab[j], RA[ax], RB[ax])
                        acc += v
                cart[t, x, y] = acc
    ref = oint.TRAFO[la].numpy() @ cart @ oint.TRAFO[lb].numpy().T
    np.testing.assert_allclose(blk.numpy(), ref, atol=1e-11, rtol=0)


@pytest.mark.parametrize("n", [0, 1, 2])
def test_one_centre_moments_closed_form(n):
    a, b = 0.7, 1.3
    p = a + b
    t = oint.os_table(0, 0, n, torch.tensor(0.5 / p), torch.tensor(0.0), torch.tensor(0.0), torch.tensor(0.0))
    dfact = {0: 1.0, 1: 0.0, 2: 1.0}[n]
    ref = 0.0 if n % 2 else dfact / (2 * p) ** (n // 2)
    assert abs(float(t[(0, 0, n)]) - ref) < 1e-15


def test_dipole_of_s_pair_closed_form():
    a, b = 0.9, 1.7
    A = torch.tensor([[0.3, -0.2, 0.5]], dtype=torch.float64)
    B = torch.tensor([[-0.4, 0.8, 0.1]], dtype=torch.float64)
    f64 = dict(dtype=torch.float64)
    blk = oint.pair_blocks(0, 0, torch.tensor([a], **f64), torch.tensor([1.0], **f64), torch.tensor([b], **f64),
                           torch.tensor([1.0], **f64), A, B)[0]
    S = blk[0, 0, 0]
    P = (a * A[0] + b * B[0]) / (a + b)
    torch.testing.assert_close(blk[1:4, 0, 0], S * P, atol=1e-15, rtol=1e-13)


@pytest.mark.parametrize("name", ["H2", "LiH", "H2O", "CH4", "SiH4", "LYS_xao"])
def test_overlap_vs_tblite_fixture(geometries, tblite_arrays, name):
    m = geometries[name]
    b = obasis.build(m["numbers"], "gfn1")
    s, _, _ = oint.multipole_matrices(b, torch.tensor(m["positions"], dtype=torch.float64))
    ref = torch.tensor(tblite_arrays["ovlp_gfn1_" + name.lower()], dtype=torch.float64).reshape(s.shape)
    # float32 fixture (tblite 0.3.0)
    torch.testing.assert_close(s, ref, atol=5e-7, rtol=0)


@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_final_moments_invariants(geometries, gfn):
    m = geometries["SiH4"]
    pos = torch.tensor(m["positions"], dtype=torch.float64)
    b = obasis.build(m["numbers"], gfn)
    s, d, q = oint.multipole_matrices(b, pos)
    s2, d2, q2 = oint.multipole_matrices(b, pos + torch.tensor([3.1, -2.2, 1.7], dtype=torch.float64))
    torch.testing.assert_close(s2, s, atol=1e-12, rtol=0)
    torch.testing.assert_close(d2, d, atol=1e-11, rtol=0)
    torch.testing.assert_close(q2, q, atol=1e-11, rtol=0)
    torch.testing.assert_close(s, s.T, atol=1e-15, rtol=0)
    torch.testing.assert_close(s.diagonal(), torch.ones(b.nao, dtype=torch.float64), atol=1e-14, rtol=0)
    assert (q[0] + q[2] + q[5]).abs().max() < 1e-12
    # D is centred on the ket atom: D_ij - D_ji = (R_i - R_j) S_ij
    pj = pos[b.ao_atom()]
    lhs = d - d.transpose(-1, -2)
    rhs = (pj.T.unsqueeze(-1) - pj.T.unsqueeze(-2)) * s
    torch.testing.assert_close(lhs, rhs, atol=1e-12, rtol=0)


def test_gfn1_h2s_orthogonal_to_h1s():
    b = obasis.build([1], "gfn1")
    s, _, _ = oint.multipole_matrices(b, torch.zeros(1, 3, dtype=torch.float64))
    assert abs(float(s[0, 1])) < 1e-14
