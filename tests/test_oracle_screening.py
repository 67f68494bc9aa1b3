"""
Oracle shell-pair screening (CPU): the bound of ``oracle.integrals.pair_bound``
(restating the reference's ``_bounds``, integral/driver/pytorch/impls/
pairs.py:411-494) and the selection of ``screen_keep`` (``_screen``,
pairs.py:497-541), pinned by

* what the bound must be: never below the largest |element| of a block of
  the raw (origin-centred) S, D or Q, for s, p and d shells, scaled
  exponents and a molecule far from the origin (the reference's test_bound,
  test/test_integrals/test_screening.py:66-104);
* a closed form: for an s-s pair with positive contraction coefficients the
  order-0 bound is the overlap itself times 1 + 1e-10 (2 pi Gamma(3/2)
  p^(-3/2) = (pi/p)^(3/2), the primitive overlap);
* the bound's definition evaluated independently: the integral of
  (rho + |PA|)^la (rho + |PB|)^lb (rho + |PC|)^e e^(-p rho^2) by radial
  quadrature with P from the centres, for s, p, d in both orders (pins the
  binomial expansion, |PA| = b R / p against |PB| = a R / p, and the
  transforms' row sums, which the domination tests cannot see);
* what screening must do: dropped elements are below the threshold and
  something is dropped (test_error, :145-160); the overlap of a diffuse basis
  stays positive definite (test_positive_definite, :163-175); bounds of
  several orders at once equal the single-order ones (:107-127).
"""

from __future__ import annotations

import json
import pathlib

import pytest
import torch

from oracle import basis as obasis
from oracle import integrals as oint

GOLDEN = pathlib.Path(__file__).parent / "golden" / "geometries.json"
COMP = {"S": (slice(0, 1), 0), "D": (slice(1, 4), 1), "Q": (slice(4, 10), 2)}


def _geom(name):
    g = json.loads(GOLDEN.read_text())[name]
    return g["numbers"], torch.tensor(g["positions"], dtype=torch.float64)


def _basis(numbers, ascale=1.0):
    b = obasis.build(numbers, "gfn2")
    for sh in b.shells:
        sh.alpha = sh.alpha * ascale
    return b


def two_molecules(gap: float):
    """SiH4 and a smaller NH3-like fragment ``gap`` Bohr away along x (the
    reference's screening test system, test_screening.py:130-142)."""
    numbers = [14, 1, 1, 1, 1, 7, 1, 1, 1]
    base = torch.tensor([[0.0, 0.0, 0.0], [1.6, 1.6, 1.6], [-1.6, -1.6, 1.6], [-1.6, 1.6, -1.6],
                         [1.6, -1.6, -1.6]], dtype=torch.float64)
    shift = torch.tensor([gap, 0.0, 0.0], dtype=torch.float64)
    return numbers, torch.cat([base, base[:4] * 0.6 + shift])


@pytest.mark.parametrize("name", ["SiH4", "LiH"])
@pytest.mark.parametrize("ascale", [1.0, 0.3])
@pytest.mark.parametrize("shift", [0.0, 50.0])
@pytest.mark.parametrize("op", ["S", "D", "Q"])
def test_bound_dominates_blocks(name, ascale, shift, op):
    numbers, pos = _geom(name)
    pos = pos + shift
    b = _basis(numbers, ascale)
    raw = oint.raw_matrices(b, pos)
    sl, e = COMP[op]
    bd = oint.pair_bounds(b, pos, (e,))[0]
    n = 0
    for i, si in enumerate(b.shells):
        for j, sj in enumerate(b.shells):
            if si.atom == sj.atom:
                continue
            blk = raw[sl, si.ao : si.ao + 2 * si.l + 1, sj.ao : sj.ao + 2 * sj.l + 1]
            assert blk.abs().max() <= bd[i, j], (i, j)
            n += 1
    assert n > 0


def test_ss_bound_is_the_overlap():
    numbers, pos = _geom("CH4")
    b = _basis(numbers)
    raw = oint.raw_matrices(b, pos)
    bd = oint.pair_bounds(b, pos, (0,))[0]
    hs = [i for i, sh in enumerate(b.shells) if sh.z == 1]
    assert len(hs) == 4
    for i in hs:
        for j in hs:
            if i == j:
                continue
            si, sj = b.shells[i], b.shells[j]
            assert (si.coeff > 0).all() and (sj.coeff > 0).all()
            s = raw[0, si.ao, sj.ao]
            torch.testing.assert_close(bd[i, j], s * (1.0 + 1e-10), rtol=1e-13, atol=0.0)


def test_transform_row_sums():
    """Largest |row sum| of the spherical transforms: 1 for s and p, 2 for d
    (the d_z2 row -xx/2 - yy/2 + zz)."""
    assert [oint.sph_norm(l) for l in (0, 1, 2)] == [1.0, 1.0, 2.0]


def _bound_by_quadrature(la, lb, aa, ca, ab, cb, ra, rb, e, origin):
    """The bound from its definition: per primitive pair |c_i c_j| e^{-mu R^2}
    times the integral over all space of (rho + |PA|)^la (rho + |PB|)^lb
    (rho + max(|AC|, |BC|))^e e^{-p rho^2}, rho = |r - P|, with P = (a A +
    b B) / p from the centres, by radial quadrature; times the transforms'
    largest row sums 1, 1, 2 (s, p, d) and 1 + 1e-10."""
    import math

    from scipy.integrate import quad

    rs = {0: 1.0, 1: 1.0, 2: 2.0}
    ra, rb, c = ra.numpy(), rb.numpy(), origin.numpy()
    R = float(((rb - ra) ** 2).sum() ** 0.5)
    pc = max(float(((ra - c) ** 2).sum() ** 0.5), float(((rb - c) ** 2).sum() ** 0.5))
    tot = 0.0
    for i in range(len(aa)):
        for j in range(len(ab)):
            a, b = float(aa[i]), float(ab[j])
            p = a + b
            P = (a * ra + b * rb) / p
            pa = float(((P - ra) ** 2).sum() ** 0.5)
            pb = float(((P - rb) ** 2).sum() ** 0.5)
            f = lambda x: (x + pa) ** la * (x + pb) ** lb * (x + pc) ** e * x * x * math.exp(-p * x * x)  # noqa: E731
            val = 4.0 * math.pi * quad(f, 0.0, math.inf, epsabs=0.0, epsrel=1e-13, limit=200)[0]
            tot += abs(float(ca[i]) * float(cb[j])) * math.exp(-a * b / p * R * R) * val
    return tot * rs[la] * rs[lb] * (1.0 + 1e-10)


@pytest.mark.parametrize("shift", [0.0, 7.0])
def test_bound_equals_its_definition(shift):
    numbers, pos = _geom("SiH4")
    pos = pos.clone()
    pos[1] += torch.tensor([0.3, -0.2, 0.4], dtype=torch.float64)  # an unsymmetric pair geometry
    b = _basis(numbers)
    origin = torch.tensor([shift, -0.5 * shift, 0.25 * shift], dtype=torch.float64)
    checked = set()
    for si in b.shells:
        for sj in b.shells:
            if si.atom == sj.atom or (si.l, sj.l) in checked:
                continue
            checked.add((si.l, sj.l))
            ca = oint._normalized_coeff(si.l, si.alpha, si.coeff)
            cb = oint._normalized_coeff(sj.l, sj.alpha, sj.coeff)
            got = oint.pair_bound(si.l, sj.l, si.alpha, ca, sj.alpha, cb, pos[si.atom], pos[sj.atom],
                                  (0, 1, 2), origin)
            for e in (0, 1, 2):
                ref = _bound_by_quadrature(si.l, sj.l, si.alpha, ca, sj.alpha, cb, pos[si.atom], pos[sj.atom],
                                           e, origin)
                assert abs(float(got[e]) - ref) <= 1e-9 * ref, (si.l, sj.l, e)
    assert checked == {(0, 0), (0, 1), (1, 0), (0, 2), (2, 0)}  # Si s/p/d with H s, both orders, H-H


def test_orders_consistent():
    numbers, pos = _geom("SiH4")
    b = _basis(numbers)
    both = oint.pair_bounds(b, pos, (0, 1, 2))
    for i, e in enumerate((0, 1, 2)):
        assert torch.equal(both[i], oint.pair_bounds(b, pos, (e,))[0])


@pytest.mark.parametrize("op", ["S", "D", "Q"])
@pytest.mark.parametrize("threshold", [1e-4, 1e-8, 1e-12])
def test_dropped_elements_below_threshold(op, threshold):
    numbers, pos = two_molecules(16.0)
    b = _basis(numbers)
    sl, _ = COMP[op]
    ref = oint.raw_matrices(b, pos)[sl]
    scr = oint.raw_matrices(b, pos, screening_threshold=threshold)[sl]
    assert ((scr == 0) & (ref != 0)).any()
    assert (scr - ref).abs().max() <= threshold


@pytest.mark.parametrize("threshold", [1e-4, 1e-3])
def test_diffuse_overlap_stays_positive_definite(threshold):
    numbers, pos = two_molecules(32.0)
    b = _basis(numbers, 0.3)
    ref = oint.raw_matrices(b, pos)[0]
    scr = oint.raw_matrices(b, pos, screening_threshold=threshold)[0]
    assert ((scr == 0) & (ref != 0)).any()
    assert torch.linalg.eigvalsh(scr).min() > 0


def test_same_position_pairs_always_kept():
    numbers, pos = two_molecules(16.0)
    b = _basis(numbers)
    keep = oint.screen_keep(b, pos, 1e30)
    at = b.shell_atom()
    assert torch.equal(keep, at[:, None] == at[None, :])
