"""
Serving engine (continuous batching): every result is the same alone and in
the mix.  Molecules of several pool classes go through the engine with two
slots per class, so slots are retired and refilled while other molecules
iterate.  Each result is compared with the same molecule served alone by the
engine and with the one-molecule path (run_scf + forces at the molecule's own
sizes).  Padding and batch shapes change the rounding of the batched
products, so trajectories are not bitwise equal and iteration counts may
differ by a few near the 1e-8 stopping bar; the results agree to 1e-9 Eh and
1e-7 Eh/Bohr, inside the 1e-8 Eh / 1e-6 Eh/Bohr parity bars.  One molecule
is also checked against the oracle at the parity bars.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from dxtb_amd import serving as sv
from dxtb_amd import workloads

pytestmark = pytest.mark.gpu


def _requests(k=8, maxnat=30, solvated=()):
    """Seeded serving requests; the ones at positions ``solvated`` use GBSA."""
    from dxtb_amd.solvation import GeneralizedBorn

    nums, poss, chg = workloads.config_batch(3, count=60)
    sel = [i for i in range(len(nums)) if len(nums[i]) <= maxnat][:k]
    gb = GeneralizedBorn(None, 78.4, alpb=False)
    return [sv.Request(i, nums[i], poss[i], float(chg[i]), solvation=gb if k_ in solvated else None)
            for k_, i in enumerate(sel)]


def test_alone_vs_mix(cuda):
    reqs = _requests(solvated=(1, 4))
    assert len({sv.pool_class(*sv.sizes(r.numbers, "gfn2")) for r in reqs}) >= 2
    print("case serving mix of", [len(r.numbers) for r in reqs], flush=True)
    eng = sv.Engine("gfn2", cuda, capacity={c: 2 for c in sv.POOL_NAO})
    out = eng.run([sv.Request(r.rid, r.numbers, r.positions, r.charge, solvation=r.solvation) for r in reqs],
                  realtime=False)
    assert len(out) == len(reqs)
    got = {r.rid: r for r in out}
    for r in reqs:
        m = got[r.rid]
        alone = sv.Engine("gfn2", cuda).run([sv.Request(r.rid, r.numbers, r.positions, r.charge,
                                                        solvation=r.solvation)], realtime=False)[0]
        ref = sv.plain_path(sv.Request(r.rid, r.numbers, r.positions, r.charge, solvation=r.solvation), "gfn2",
                            cuda)
        print(f"case {r.rid}: iterations mix {m.iterations} alone {alone.iterations} plain {ref.iterations}, "
              f"dE {m.energy - alone.energy:.2e} {m.energy - ref.energy:.2e}", flush=True)
        assert m.converged and alone.converged and ref.converged
        # a molecule's SCF arithmetic does not depend on its batch mates or on
        # when it joined its pool: the same iterations alone and in the mix
        assert m.iterations == alone.iterations
        for other in (alone, ref):
            assert abs(m.energy - other.energy) < 1e-9
            np.testing.assert_allclose(m.forces, other.forces, atol=1e-7, rtol=0)


def test_engine_against_oracle(cuda):
    from oracle import scf as oscf

    r = _requests(k=1, maxnat=16)[0]
    print("case serving vs oracle", len(r.numbers), r.charge, flush=True)
    out = sv.Engine("gfn2", cuda).run([r], realtime=False)[0]
    ref = oscf.singlepoint(r.numbers.tolist(), torch.tensor(r.positions), "gfn2", charge=r.charge, tol=1e-10,
                           forces=True)
    assert abs(out.energy - float(ref["energy"])) < 1e-8
    np.testing.assert_allclose(out.forces, ref["forces"].numpy(), atol=1e-6, rtol=0)


def test_poisson_stream_small(cuda):
    """Requests arriving while others run are admitted and all complete."""
    reqs = _requests(k=6)
    rng = np.random.default_rng(0)
    t = np.cumsum(rng.exponential(0.02, size=len(reqs)))
    for r, a in zip(reqs, t):
        r.arrival = float(a)
    out = sv.Engine("gfn2", cuda).run(reqs, realtime=True)
    assert len(out) == len(reqs)
    for r in out:
        assert r.finished >= r.arrival and r.converged


def test_graph_replays_match_eager(cuda):
    """Graph mode (each SCF iteration a replay of a captured hipGraph, with
    slots retired and refilled between replays) gives the results of the
    eager engine: energies within 1e-9 Eh and forces within 1e-7 Eh/Bohr.
    The graphs always sweep the full Broyden history, whose zero rows change
    only the rounding; on slowly converging molecules (small gaps) that
    changes the iteration count by up to ~20 % (as padding does), so only the
    results are compared (the counts are printed)."""
    reqs = _requests(k=10, solvated=(2, 7))
    cap = {c: 3 for c in sv.POOL_NAO}
    mk = lambda: [sv.Request(r.rid, r.numbers, r.positions, r.charge, solvation=r.solvation) for r in reqs]  # noqa
    g = {r.rid: r for r in sv.Engine("gfn2", cuda, capacity=cap, graphs=True).run(mk(), realtime=False)}
    e = {r.rid: r for r in sv.Engine("gfn2", cuda, capacity=cap, graphs=False).run(mk(), realtime=False)}
    for rid in g:
        print(f"case {rid}: iterations graph {g[rid].iterations} eager {e[rid].iterations}, "
              f"dE {g[rid].energy - e[rid].energy:.1e}", flush=True)
        assert g[rid].converged and e[rid].converged
        assert abs(g[rid].energy - e[rid].energy) < 1e-9
        np.testing.assert_allclose(g[rid].forces, e[rid].forces, atol=1e-7, rtol=0)


def test_compaction_and_admission_into_busy_pool(cuda):
    """One pool class with 8 slots and 14 molecules: molecules are admitted
    while others are mid-SCF, and when retirements leave holes below a
    long-running slot the pool compacts (moves slots down).  Every result
    equals the one-molecule path's."""
    nums, poss, chg = workloads.config_batch(3, count=400)
    cls = [sv.pool_class(*sv.sizes(z, "gfn2")) for z in nums]
    target = 96
    sel = [i for i in range(len(nums)) if cls[i] == target][:14]
    assert len(sel) == 14
    print("case compaction: class", target, "molecules", sel, flush=True)
    reqs = [sv.Request(i, nums[i], poss[i], float(chg[i])) for i in sel]
    eng = sv.Engine("gfn2", cuda, capacity={target: 8})
    out = eng.run([sv.Request(r.rid, r.numbers, r.positions, r.charge) for r in reqs], realtime=False)
    pool = eng.pools[target]
    print(f"case compaction: moves {pool.moves}, admitted while busy {pool.admitted_busy}", flush=True)
    assert len(out) == len(reqs)
    assert pool.admitted_busy > 0
    got = {r.rid: r for r in out}
    for r in reqs:
        m = got[r.rid]
        ref = sv.plain_path(sv.Request(r.rid, r.numbers, r.positions, r.charge), "gfn2", cuda)
        assert m.converged == ref.converged
        if m.converged:
            assert abs(m.energy - ref.energy) < 1e-9, (r.rid, m.energy - ref.energy)
            np.testing.assert_allclose(m.forces, ref.forces, atol=1e-7, rtol=0)


def test_compact_moves_slots_exactly(cuda):
    """Pool compaction: after three iterations of 8 molecules the five lowest
    slots are dropped; compact() moves the three survivors into slots 0-2
    bit for bit (potential vectors, Broyden history in use, control words,
    geometry state, orbitals, basis maps), and the moved molecules then
    converge to the one-molecule path's results."""
    nums, poss, chg = workloads.config_batch(3, count=400)
    cls = [sv.pool_class(*sv.sizes(z, "gfn2")) for z in nums]
    sel = [i for i in range(len(nums)) if cls[i] == 96][:8]
    reqs = [sv.Request(i, nums[i], poss[i], float(chg[i])) for i in sel]
    eng = sv.Engine("gfn2", cuda, capacity={96: 8})
    eng.t0 = 0.0
    for r in reqs:
        eng.submit(r)
    eng.admit()  # setup on the default stream, as the engine does it
    pool = eng.pools[96]
    torch.cuda.synchronize()
    with torch.cuda.stream(pool.stream):
        for _ in range(3):
            pend = pool.launch()
            torch.cuda.synchronize()
            pool.finish(pend, lambda: 0.0)
    torch.cuda.synchronize()
    assert pool.occupied[:8].all(), "no molecule may converge in three iterations"
    keep = [5, 6, 7]
    m = pool.mix
    vec = ("xe", "x", "y", "ctl", "flag", "dxn")
    snap = {name: getattr(m, name)[keep].clone() for name in vec}
    hist = [int(k) for k in m.ctl[keep, 2].tolist()]  # history entries in use (the rest is never read)
    assert min(hist) >= 1
    snap_h = [(m.hc[s, :k].clone(), m.hd[s, :k].clone()) for s, k in zip(keep, hist)]
    snap.update({name: getattr(pool.state, name)[keep].clone() for name in ("S", "H0", "U", "Corb", "X", "pos")})
    host = {k: getattr(pool.state.bb, k)[keep].copy() for k in ("numbers", "nao", "sh_l", "ao_shell")}
    pool.occupied[:5] = False
    m.ctl[:5, 0] = 0
    for r in range(5):
        pool.req[r] = None
    with torch.cuda.stream(pool.stream):
        assert pool.compact()
    torch.cuda.synchronize()
    assert pool.occupied.tolist() == [True] * 3 + [False] * 5 and pool.hi == 3 and pool.moves == 3
    for name in vec:
        assert torch.equal(getattr(m, name)[:3], snap[name]), name
    for d, (k, (hc, hd)) in enumerate(zip(hist, snap_h)):
        assert torch.equal(m.hc[d, :k], hc) and torch.equal(m.hd[d, :k], hd)
    assert (m.ctl[3:, 0] == 0).all()  # the source slots are free on the device too
    for name in ("S", "H0", "U", "Corb", "X", "pos"):
        assert torch.equal(getattr(pool.state, name)[:3], snap[name]), name
    for k, v in host.items():
        assert np.array_equal(getattr(pool.state.bb, k)[:3], v), k
    assert [r.rid for r in pool.req[:3]] == [sel[i] for i in keep]
    out = []
    with torch.cuda.stream(pool.stream):
        for _ in range(300):
            if not pool.occupied.any():
                break
            pend = pool.launch()
            torch.cuda.synchronize()
            out += pool.finish(pend, lambda: 0.0)
        out += pool.collect(lambda: 0.0, wait=True)
    assert sorted(r.rid for r in out) == sorted(sel[i] for i in keep)
    for m in out:
        r = next(q for q in reqs if q.rid == m.rid)
        ref = sv.plain_path(sv.Request(r.rid, r.numbers, r.positions, r.charge), "gfn2", cuda)
        assert m.converged and ref.converged
        assert abs(m.energy - ref.energy) < 1e-9, (m.rid, m.energy - ref.energy)
        np.testing.assert_allclose(m.forces, ref.forces, atol=1e-7, rtol=0)


def test_spin_requests(cuda, geometries):
    """Requests with and without a spin share a pool: the triplet water and
    the quartet NO2 match the calculators' results for the same spin, the
    closed-shell water its singlet."""
    from dxtb_amd.calculators import GFN2Calculator

    reqs = []
    for i, (name, spin) in enumerate((("H2O", 2), ("H2O", None), ("NO2", 3))):
        g = geometries[name]
        reqs.append(sv.Request(i, np.array(g["numbers"]), np.array(g["positions"], dtype=np.float64), spin=spin))
    print("case serving spins", flush=True)
    out = {r.rid: r for r in sv.Engine("gfn2", cuda).run(reqs, realtime=False)}
    for i, (name, spin) in enumerate((("H2O", 2), ("H2O", None), ("NO2", 3))):
        g = geometries[name]
        calc = GFN2Calculator(torch.tensor(g["numbers"], device=cuda), device=cuda)
        e = calc.get_energy(torch.tensor(g["positions"], dtype=torch.float64, device=cuda), spin=spin)
        assert out[i].converged and abs(out[i].energy - e.item()) < 1e-9, (name, spin)


def test_gbsa_admitted_into_vacuum_pool(cuda):
    """A pool iterates vacuum molecules over a window of two slots; when the
    small one (H2) has retired, a GBSA request is admitted into its slot
    while the other is mid-SCF, so the pool keeps iterating the same window.
    The GBSA molecule is iterated with its solvation term (the window follows
    the pool's solvation flags): its energy and forces equal the
    one-molecule path's, and so do the vacuum molecule's."""
    from dxtb_amd.solvation import GeneralizedBorn

    nums, poss, chg = workloads.config_batch(3, count=200)
    # neutral molecules: the first charged one of the class (a cation with Cl
    # and F) has two self-consistent solutions 2.8 Eh apart, and which one an
    # SCF reaches depends on its rounding (AO order, padding), not on admission
    c48 = [i for i in range(len(nums))
           if sv.pool_class(*sv.sizes(nums[i], "gfn2")) == 48 and len(nums[i]) >= 8 and chg[i] == 0]
    assert len(c48) >= 2
    gb = GeneralizedBorn(None, 78.4, alpb=False)
    h2 = sv.Request(-1, np.array([1, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]))
    slow = sv.Request(c48[0], nums[c48[0]], poss[c48[0]], float(chg[c48[0]]))
    late = sv.Request(c48[1], nums[c48[1]], poss[c48[1]], float(chg[c48[1]]), solvation=gb)
    print("case gbsa admission into a vacuum pool", c48[:2], flush=True)
    eng = sv.Engine("gfn2", cuda, capacity={48: 4})
    eng.submit(h2)
    eng.submit(slow)
    out = []
    while not any(r.rid == -1 for r in out):
        out += eng.tick()
    p = eng.pools[48]
    assert p.occupied.tolist() == [False, True, False, False] and not p.state.any_gb
    assert 2 in p._wins  # the two-slot window was made while the pool was in vacuum
    eng.submit(late)
    out += eng.tick()
    assert p.occupied[0] and p.req[0] is late and p.hi == 2 and p.admitted_busy == 1
    while eng.busy:
        out += eng.tick()
    got = {r.rid: r for r in out}
    for r in (slow, late):
        ref = sv.plain_path(sv.Request(r.rid, r.numbers, r.positions, r.charge, solvation=r.solvation), "gfn2",
                            cuda)
        m = got[r.rid]
        print(f"case {r.rid} gb={r.solvation is not None}: dE {m.energy - ref.energy:.2e}", flush=True)
        assert m.converged and ref.converged
        assert abs(m.energy - ref.energy) < 1e-9
        np.testing.assert_allclose(m.forces, ref.forces, atol=1e-7, rtol=0)


def test_float32_request_positions(cuda):
    """Positions given in float32 are converted to FP64 before any kernel
    reads them (the plain path and the engine agree with the FP64 request)."""
    r = _requests(k=1, maxnat=16)[0]
    r32 = sv.Request(r.rid, r.numbers, np.asarray(r.positions, dtype=np.float32), r.charge)
    r64 = sv.Request(r.rid, r.numbers, np.asarray(r32.positions, dtype=np.float64), r.charge)
    a = sv.plain_path(r32, "gfn2", cuda)
    b = sv.plain_path(r64, "gfn2", cuda)
    assert abs(a.energy - b.energy) < 1e-12
    c = sv.Engine("gfn2", cuda).run([sv.Request(r.rid, r.numbers, r32.positions, r.charge)], realtime=False)[0]
    assert abs(c.energy - b.energy) < 1e-9
