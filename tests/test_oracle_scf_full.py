"""Full-mode SCF (the reference's default scf_mode="full": the Anderson or
the simple mixer, scf/unrolling/default.py:71-140, scf/mixer/anderson.py:163-317,
scf/mixer/simple.py) in the oracle, and the batched mixer of the device path
(host logic, run here on CPU tensors).

Pins:
* the fixed point: GFN1 SCF energies of the tblite fixtures at the
  reference's tolerance, and the implicit (broyden1) result of the same
  molecule to 1e-9 Eh (GFN1 and GFN2);
* simple mixing: the closed form x_k = x* + (I - b (I - M))^k (x_0 - x*) on a
  linear map g(x) = M x + c, with b = damp (simple mixer) or b = damp_init
  (the Anderson mixer's soft-start steps);
* Anderson mixing: its step minimises the residual over the affine span of
  the history, so on a linear map of dimension d (no diagonal offset) it is
  GMRES on the residual's Krylov space and lands on the fixed point after
  d Anderson steps; a dropped term, a wrong sign or a transposed history
  misses it;
* the device path's batched mixer equals the oracle's system by system on a
  batch of different linear maps, with each vector stored shell-resolved
  (weights 2l+1) against the oracle's orbital-resolved copy, padding included.
"""

from __future__ import annotations

import math
import warnings

import pytest
import torch

from dxtb_amd import calculators
from dxtb_amd import scf as hscf
from oracle import scf

REFTOL = math.sqrt(torch.finfo(torch.float64).eps) * 10
T = lambda x: torch.tensor(x, dtype=torch.float64)  # noqa: E731


def _linear(d, seed, rho=0.6):
    g = torch.Generator().manual_seed(seed)
    Q, _ = torch.linalg.qr(torch.randn(d, d, generator=g, dtype=torch.float64))
    M = Q @ torch.diag(torch.linspace(-rho, rho, d, dtype=torch.float64)) @ Q.T + 0.1 * torch.randn(
        d, d, generator=g, dtype=torch.float64)
    c = torch.randn(d, generator=g, dtype=torch.float64)
    xs = torch.linalg.solve(torch.eye(d, dtype=torch.float64) - M, c)
    return M, c, xs


@pytest.mark.parametrize("kind,steps", [("simple", 12), ("anderson", 5)])
def test_simple_mixing_closed_form(kind, steps):
    """Simple mixer with damp, and the Anderson mixer's soft-start steps (the
    first `generations` steps) with damp_init."""
    M, c, xs = _linear(4, 1)
    x0 = torch.zeros(4, dtype=torch.float64)
    mixer = scf.FullModeMixer(kind, damp=0.5, damp_init=0.1, generations=5, soft_start=True)
    b = 0.5 if kind == "simple" else 0.1
    step = torch.eye(4, dtype=torch.float64) - b * (torch.eye(4, dtype=torch.float64) - M)
    x = x0
    for k in range(1, steps + 1):
        x = mixer.iter(M @ x + c, x)
        ref = xs + torch.linalg.matrix_power(step, k) @ (x0 - xs)
        torch.testing.assert_close(x, ref, atol=1e-13, rtol=0)


def test_anderson_terminates_on_linear_maps():
    d = 3
    M, c, xs = _linear(d, 2, rho=0.9)
    mixer = scf.FullModeMixer("anderson", damp=0.5, damp_init=0.1, generations=5, soft_start=False, offset=None)
    x = torch.zeros(d, dtype=torch.float64)
    errs = []
    for _ in range(d + 1):  # one simple step, then d Anderson steps
        x = mixer.iter(M @ x + c, x)
        errs.append((x - xs).norm().item())
    assert errs[-2] > 0.1 and errs[-1] < 1e-10, errs
    # with the default diagonal offset the same steps converge, though not
    # exactly (iterated until the residual vanishes, as the SCF loop stops)
    mixer = scf.FullModeMixer("anderson", soft_start=False)
    x = torch.zeros(d, dtype=torch.float64)
    for k in range(30):
        gx = M @ x + c
        if (gx - x).norm() < 1e-12:
            break
        x = mixer.iter(gx, x)
    assert 4 < k < 12 and (x - xs).norm().item() < 1e-11


@pytest.mark.parametrize("kind,soft", [("anderson", True), ("anderson", False), ("simple", True)])
def test_batched_mixer_matches_oracle(kind, soft):
    """Three systems of 3, 5 and 4 shells (l = 0/1/2: weights 1/3/5), padded to
    5; the oracle mixes the orbital-resolved vectors with unit weights."""
    ls = [[0, 1, 0], [1, 2, 0, 1, 0], [2, 0, 1, 1]]
    L = 5
    B = len(ls)
    w = torch.zeros(B, L, dtype=torch.float64)
    maps = []
    for b, l in enumerate(ls):
        w[b, : len(l)] = T([2 * x + 1 for x in l])
        maps.append(_linear(len(l), 10 + b, rho=0.8)[:2])
    o = hscf.SCFOptions(mixer=kind, damp_soft_start=soft)
    mix = hscf.BatchedFullMixer(o, B, L, torch.device("cpu"), w)
    oracles = [scf.FullModeMixer(kind, soft_start=soft) for _ in range(B)]

    def g_batch(x):
        out = torch.zeros_like(x)
        for b, (M, c) in enumerate(maps):
            n = M.shape[0]
            out[b, :n] = M @ x[b, :n] + c
        return out

    def expand(b, v):  # shell-resolved -> orbital-resolved (each entry 2l+1 times)
        return torch.repeat_interleave(v[: len(ls[b])], w[b, : len(ls[b])].long())

    x = torch.zeros(B, L, dtype=torch.float64)
    xo = [torch.zeros(int(w[b].sum()), dtype=torch.float64) for b in range(B)]
    active = torch.ones(B, dtype=torch.bool)
    for it in range(15):
        x, d = mix.iter(g_batch(x), x, active)
        for b, (M, c) in enumerate(maps):
            gx = expand(b, M @ xo[b][_firsts(w[b], len(ls[b]))] + c)
            xo[b] = oracles[b].iter(gx, xo[b])
            # rounding only: the small systems are solved in another order
            torch.testing.assert_close(expand(b, x[b]), xo[b], atol=1e-10, rtol=1e-10)
            torch.testing.assert_close(hscf.wnorm(d[b], w[b]), oracles[b].delta.norm(), atol=1e-10, rtol=1e-10)
        assert (x[:, L - 1][[0, 2]] == 0).all()  # padding stays untouched
        if min(m.delta.norm().item() for m in oracles) < 1e-8:
            break  # converged (as the SCF loop stops; the history then becomes singular)
    assert it >= 7  # soft start and then Anderson steps


def _firsts(wb, n):
    """Index of the first orbital of each shell in the orbital-resolved vector."""
    starts = torch.cumsum(wb[:n].long(), 0) - wb[:n].long()
    return starts


def test_oracle_full_mode_fixed_point(geometries, refvals):
    for name in ("H2", "LiH", "H2O", "CH4", "SiH4"):
        m = geometries[name]
        r = scf.singlepoint(m["numbers"], T(m["positions"]), "gfn1", scf_mode="full", mixer="anderson")
        assert r["converged"]
        ref = refvals["scf"][name]["egfn1"]
        assert abs(r["eel"].item() - ref) <= max(REFTOL, REFTOL * abs(ref)), name
    for gfn in ("gfn1", "gfn2"):
        m = geometries["LYS_xao"]
        ri = scf.singlepoint(m["numbers"], T(m["positions"]), gfn)
        for mixer in ("anderson", "simple"):
            rf = scf.singlepoint(m["numbers"], T(m["positions"]), gfn, scf_mode="full", mixer=mixer, maxiter=400)
            assert rf["converged"]
            assert abs(rf["energy"].item() - ri["energy"].item()) < 1e-9, (gfn, mixer)


def test_calculator_scf_options():
    o = calculators._opts_to_scf({"scf_mode": "full"})
    assert (o.scf_mode, o.mixer) == ("full", "anderson")
    o = calculators._opts_to_scf({"mixer": "simple", "scf_mode": 0, "x_atol_max": 1e-7, "damp_init": 0.2})
    assert (o.scf_mode, o.mixer, o.x_tol_max, o.damp_init) == ("full", "simple", 1e-7, 0.2)
    assert calculators._opts_to_scf({"scf_mode": "default"}).scf_mode == "implicit"
    assert calculators._opts_to_scf({"scf_mode": 1}).scf_mode == "implicit"
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        o = calculators._opts_to_scf({"scf_mode": "full", "mixer": "broyden"})
    assert o.mixer == "anderson" and any("Anderson" in str(r.message) for r in rec)
    with pytest.raises(ValueError, match="removed"):
        calculators._opts_to_scf({"scf_mode": "nonpure"})
    with pytest.raises(ValueError, match="Unknown SCF mode"):
        calculators._opts_to_scf({"scf_mode": "bogus"})
    with pytest.raises(ValueError, match="Unknown mixer"):
        calculators._opts_to_scf({"mixer": "diis"})
    with pytest.raises(NotImplementedError):
        calculators._opts_to_scf({"scf_mode": "single-shot"})


def test_batched_mixer_singular_history_raises():
    """Identical residuals make the Anderson system singular: an error, as
    torch.linalg.solve raises in the reference mixer."""
    o = hscf.SCFOptions(mixer="anderson", damp_soft_start=False)
    w = torch.ones(1, 3, dtype=torch.float64)
    mix = hscf.BatchedFullMixer(o, 1, 3, torch.device("cpu"), w)
    x = torch.zeros(1, 3, dtype=torch.float64)
    f = torch.tensor([[1.0, 2.0, 3.0]], dtype=torch.float64)
    active = torch.ones(1, dtype=torch.bool)
    mix.iter(x + f, x, active)
    with pytest.raises(RuntimeError, match="singular"):
        mix.iter(x + f, x, active)  # the same residual again
