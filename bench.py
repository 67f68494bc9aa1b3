"""
Benchmark of the MI355X GFN1/GFN2-xTB single-point path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads:
  gfn2-batch -- (default) GFN2 energy + forces of 1024 seeded molecules of
                10-120 atoms through the continuous-batching engine;
                molecules/s with p50/p99 completion latency.
  serving    -- serving stream: log-normal molecule sizes (median 40 atoms),
                Poisson arrivals at three offered loads; molecules/s and
                p50/p99 latency, against the plain one-molecule-per-call path.
  eigen      -- eigensolver microbenchmark: one sweep of the blocked Jacobi
                (one workgroup per matrix, or --eig-variant 3 the split
                form) on 256 random symmetric matrices of order 240.
  integrals  -- integral-driver microbenchmark: overlap + dipole + quadrupole
                and their d/dR VJP for a padded batch of 512 GFN2 molecules of
                60 atoms (H:C:N:O:S = .45:.3:.1:.1:.05); rate in shell pairs/s.

One JSON line is printed (rank 0).  Inputs are resident in HBM before the timed
region; every step recomputes the whole workload.  A second object
"roofline" reports the dominant kernel's achieved rate (algorithmic flops or
bytes per launch divided by its HIP-event time) against the MI355X peak: for
the SCF workloads the eigensolver counted with the canonical eigensolve flops
(4/3 n^3 + 2 n^3 per matrix, whatever the method), with its HBM view beside
it; for the eigensolver microbenchmark the HBM bytes of one sweep.
"cpu_baseline" times the FP64 CPU oracle on a bounded sample of the same
workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def ncart(l):
    return (l + 1) * (l + 2) // 2


def integral_flops(la, lb, na, nb):
    """Algorithmic FP64 operations of one shell pair forward (fixed formula)."""
    return na * nb * (20 + 3 * (la + 1) * (lb + 1) * 3 * 6 + ncart(la) * ncart(lb) * 10 * 3) + 10 * (
        (2 * la + 1) * ncart(la) * ncart(lb) * 2 + (2 * la + 1) * (2 * lb + 1) * ncart(lb) * 2
    )


# measured HBM traffic of the integral forward (all nine class launches) / its
# algorithmic output bytes: rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE x 1 (scales
# from a copy of known size), config-1 workload, by AO layout.  A thread
# writes a 8-40 byte row segment per component; in the reference layout the
# columns of one class are interleaved with the other classes', so most lines
# are written in pieces by several launches; grouped by l, the lanes of a
# wavefront write adjacent segments and only the lines where two classes'
# column ranges meet are shared.
INTEGRAL_FWD_TRAFFIC_RATIO = {"l": 1.450, "reference": 2.607}


def bench_integrals(args, dev):
    from dxtb_amd import basis as hbasis
    from dxtb_amd import integrals as hint
    from dxtb_amd import workloads

    nmol = args.nmol
    nums, poss, _ = workloads.config_batch(1, count=nmol)
    N, P = workloads.pad(nums, poss)
    bb = hbasis.build_batch(N, "gfn2", dev, ao_order=args.ao_order)
    pos = torch.as_tensor(P, dtype=torch.float64, device=dev)
    # algorithmic counts over the lower triangle incl. diagonal (unique shell pairs)
    tab = hbasis.shell_table("gfn2")
    pairs = 0
    flops = 0.0
    out_bytes = 0.0
    for b in range(bb.nbatch):
        n = int(bb.nsh[b])
        u = bb.sh_ush[b, :n]
        l = tab.l[u]
        npr = tab.np[u]
        ii, jj = np.tril_indices(n)
        pairs += len(ii)
        flops += float(sum(integral_flops(int(l[i]), int(l[j]), int(npr[i]), int(npr[j])) for i, j in zip(ii, jj)))
        out_bytes += 10.0 * int(bb.nao[b]) ** 2 * 8
    n = bb.nao_max
    g = torch.Generator(device="cpu").manual_seed(3)
    aS = torch.randn((bb.nbatch, n, n), generator=g, dtype=torch.float64).to(dev)
    aD = torch.randn((bb.nbatch, 3, n, n), generator=g, dtype=torch.float64).to(dev)
    aQ = torch.randn((bb.nbatch, 6, n, n), generator=g, dtype=torch.float64).to(dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        S, D, Q, _ = hint.multipole_forward(bb, pos)
        gr = hint.multipole_vjp(bb, pos, aS, aD, aQ)
        return S, gr

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    # kernel-level timing with events on the launch stream
    fwd_ms, bwd_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        S, D, Q, _ = hint.multipole_forward(bb, pos)
        e1.record(stream)
        hint.multipole_vjp(bb, pos, aS, aD, aQ)
        e2.record(stream)
        e2.synchronize()
        fwd_ms.append(e0.elapsed_time(e1))
        bwd_ms.append(e1.elapsed_time(e2))
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    fwd = float(np.mean(fwd_ms))
    # fwd includes zeroing the outputs' padding (dxtb_zero_padding of S, D, Q)
    achieved = out_bytes / (fwd * 1e-3) / 1e9
    res = {
        "metric": "integral-driver shell pairs/s (S+D+Q forward + d/dR VJP)",
        "value": pairs / (ms_step * 1e-3),
        "unit": "shell pairs/s",
        "ms_per_step": ms_step,
        "dtype": "f64",
        "config": {"workload": f"config-1 integral microbenchmark: {nmol} GFN2 molecules x 60 atoms",
                   "shell_pairs": pairs, "nao_max": n,
                   "ao_order": ("shells grouped by angular momentum (the serving pools' layout)" if args.ao_order == "l"
                                else "reference (atoms in order; the integral driver's layout)")},
        "roofline": {
            "kernel": "multipole_fwd (9 class launches + padding zero fill)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": INTEGRAL_FWD_TRAFFIC_RATIO[args.ao_order] * out_bytes,
            "algorithmic_bytes_per_launch": out_bytes,
            "fwd_ms": fwd,
            "bwd_ms": float(np.mean(bwd_ms)),
            "fp64_tflops_fwd": flops / (fwd * 1e-3) / 1e12,
            "fp64_peak_tflops": FP64_PEAK_TFLOPS,
        },
    }
    return res, (nums, poss, pairs / nmol)


def cpu_baseline_integrals(nums, poss, budget_s=15.0):
    from oracle import basis as obasis
    from oracle import integrals as oint

    torch.set_num_threads(os.cpu_count() or 1)
    pairs = 0
    t0 = time.perf_counter()
    k = 0
    while k < len(nums) and time.perf_counter() - t0 < budget_s:
        b = obasis.build(nums[k].tolist(), "gfn2")
        p = torch.tensor(poss[k], dtype=torch.float64, requires_grad=True)
        s, d, q = oint.multipole_matrices(b, p)
        (s.sum() + d.sum() + q.sum()).backward()
        pairs += b.nsh * (b.nsh + 1) // 2
        k += 1
    dt = time.perf_counter() - t0
    return {
        "value": pairs / dt,
        "unit": "shell pairs/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"{k} of the config-1 molecules, forward + autograd backward, FP64 torch CPU oracle",
    }


def jacobi_flops_per_sweep(n: int, variant: int) -> float:
    """Algorithmic FP64 operations of one sweep of the Jacobi solver on an
    order-n matrix (every rotation applied).  Scalar variant: n(n-1)/2
    rotations, each updating two rows and two columns of A (24 operations per
    2x2 block touched) and two rows of V (6 per element pair).  Blocked
    variant: per outer step the 32x32 subproblem sweeps, the products
    Q_k^T A_kk' Q_k' of the off-diagonal tiles and Q_k^T VT of the panels."""
    if n < 2:
        return 0.0
    if variant == 1:
        m = n + (n & 1)
        return (m - 1) * ((m // 2) ** 2 * 24.0 + (m // 2) * n * 6.0)
    npairs = -(-n // 32)
    nb = 2 * npairs
    ntile = npairs * (npairs - 1) // 2
    npanel = -(-n // 32)
    gemm = 2.0 * 32 ** 3
    inner = (256 * 24.0 + 512 * 6.0)  # one subproblem step
    steps_inner = 31 + 16 * (nb - 2)
    return (nb - 1) * (ntile * 2 * gemm + npairs * npanel * gemm) + npairs * steps_inner * inner


# measured HBM traffic of one dxtb_jacobi_eigh call / its algorithmic bytes,
# by form: rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch (summed over the
# call's kernels), scaled by the same counters on a copy of known size (eigen
# workload, order 240, one sweep; FETCH_SIZE x 2, WRITE_SIZE x 1).
# 2: one workgroup per matrix; 3 (and 0, which picks 3 or 2): the split form.
# Per-step eigenvector update: 16 matrices (split form; the product reads only
# the upper tiles of A, hence below 1); the one-workgroup value is from 256
# matrices before the deferred update existed.  The scalar form (1, A in LDS)
# is not measured and counted at its algorithmic bytes.
JACOBI_TRAFFIC_RATIO = {0: 0.943, 1: 1.0, 2: 1.154, 3: 0.943}
# the same with the deferred eigenvector update (256 matrices), against its
# algorithmic bytes
JACOBI_TRAFFIC_RATIO_DEFERRED = {0: 1.041, 1: 1.0, 2: 1.542, 3: 1.041}
# the library's choice of the deferred eigenvector update (eigen.hip:
# defer_call): order <= 512 and a batch capacity above 16
DEFER_MAX, DEFER_MIN_BATCH = 512, 16


def batch_capacity(nbatch: int) -> int:
    """The eigensolver's batch capacity (eigen.hip: batch_capacity)."""
    if nbatch <= 8:
        c = 1
        while c < nbatch:
            c <<= 1
        return c
    if nbatch <= 64:
        return (nbatch + 7) // 8 * 8
    return (nbatch + 31) // 32 * 32


def deferred(nmax: int, nbatch: int) -> bool:
    return nmax <= DEFER_MAX and batch_capacity(nbatch) > DEFER_MIN_BATCH


def jacobi_bytes(n: int, variant: int, defer: bool = False) -> float:
    """Algorithmic HBM bytes of one sweep: blocked variant, A read and written
    once per outer step and the eigenvector rows VT once per step, or once per
    sweep with the deferred update; scalar variant (A in LDS), the two VT rows
    of every rotation read and written once per step."""
    if variant != 1:
        return jacobi_bytes_per_sweep(n, defer)
    m = n + (n & 1)
    return (m - 1) * 2.0 * n * n * 8


def canonical_eigh_flops(n: int) -> float:
    """FP64 operations the reference's eigensolve of an order-n matrix is
    counted with, whatever the method: tridiagonalisation 4/3 n^3 plus the
    Householder back-transform 2 n^3.  Counting the Jacobi solver by its own
    sweeps instead would let a slower method (more sweeps) report a higher
    rate."""
    return (4.0 / 3.0 + 2.0) * float(n) ** 3


def _union_ms(intervals) -> float:
    """Total length of the union of (start, end) intervals."""
    tot, cs, ce = 0.0, None, None
    for a, b in sorted(intervals):
        if cs is None:
            cs, ce = a, b
        elif a <= ce:
            ce = max(ce, b)
        else:
            tot += ce - cs
            cs, ce = a, b
    if cs is not None:
        tot += ce - cs
    return tot


def jacobi_roofline(records, base) -> dict:
    """Roofline of the eigensolver calls recorded in the timed region.  Each
    call is bracketed by HIP events on its launch stream; ``base`` is an event
    recorded at the start of the timed region, so every call becomes an
    interval on one clock, and the eigensolver's busy time is the union of
    those intervals (it cannot exceed the step, however the pools' streams
    overlap).  achieved = the canonical eigensolve flops of every matrix
    solved (canonical_eigh_flops) / that busy time, against the FP64 matrix
    peak; per launch: the flops of one call / (busy time / calls).  The HBM
    view of the same calls (the bytes of the sweeps each matrix ran, and the
    PMC-measured traffic) is reported beside it.  Kernels of other pools that
    run while an eigensolver call is in flight count as eigensolver time, so
    the rates are lower bounds."""
    if not records:
        return None
    iv = []
    canon = 0.0
    flops = 0.0
    nbytes = 0.0
    traffic = 0.0
    nsweep, nmat = 0, 0
    for e0, e1, n, nao, sw, variant, _split in records:
        iv.append((base.elapsed_time(e0), base.elapsed_time(e1)))
        sweeps = sw.cpu().numpy()
        nsweep += int(sweeps.sum())
        nmat += len(sweeps)
        dfr = deferred(n, len(sweeps))
        for i in range(len(sweeps)):
            canon += canonical_eigh_flops(int(nao[i]))
            k = int(sweeps[i])
            if k:
                flops += k * jacobi_flops_per_sweep(int(nao[i]), variant)
                nb_i = k * jacobi_bytes(int(nao[i]), variant, dfr)
                nbytes += nb_i
                traffic += (JACOBI_TRAFFIC_RATIO_DEFERRED if dfr else JACOBI_TRAFFIC_RATIO)[variant] * nb_i
    busy_ms = _union_ms(iv)
    total_s = busy_ms * 1e-3
    ncall = len(iv)
    achieved = canon / total_s / 1e12
    hbm = nbytes / total_s / 1e9
    return {
        "kernel": "dxtb_jacobi_eigh (split form k_bj_* + k_jacobi_blocked above order 96; k_jacobi_blocked, one workgroup per matrix, up to 96)",
        "bound": "mfma",
        "achieved": achieved,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS,
        "traffic": traffic / ncall,
        "canonical_flops_per_launch": canon / ncall,
        "avg_launch_ms": busy_ms / ncall,
        "launches": ncall,
        "busy_ms": busy_ms,
        "summed_event_ms": float(sum(b - a for a, b in iv)),
        "hbm_algorithmic_bytes_per_launch": nbytes / ncall,
        "hbm_achieved_gbs": hbm,
        "hbm_frac": hbm / HBM_PEAK_GBS,
        "jacobi_sweep_fp64_tflops": flops / total_s / 1e12,
        "mean_sweeps": nsweep / max(nmat, 1),
    }


def jacobi_bytes_per_sweep(n: int, defer: bool = False) -> float:
    """HBM bytes one sweep of the blocked Jacobi variant moves for an order-n
    matrix: every outer step reads and writes A and the eigenvector rows VT
    once (4 n^2 doubles; the 32x32 rotations are L2-resident and small), or,
    with the deferred update, A per step, VT once per sweep and the sweep's
    32x32 rotations (n x 32 doubles per step) written and read once."""
    nb = 2 * (-(-n // 32))
    if defer:  # A per step, VT per sweep, the sweep's rotation log written and read once
        return (nb - 1) * 2.0 * n * n * 8 + 2.0 * n * n * 8 + 2.0 * (nb - 1) * n * 32 * 8
    return (nb - 1) * 4.0 * n * n * 8


def bench_eigen(args, dev):
    """Batched eigensolver microbenchmark: the blocked Jacobi kernel on
    nmol random symmetric matrices of order n (the size class of a 60-80 atom
    GFN2 molecule), one sweep per launch so every launch does the same work;
    a device copy of known size runs first (calibration of the HBM
    counters)."""
    from dxtb_amd import _lib

    lib = _lib.load()
    B, n = args.nmol, args.eig_n
    g = torch.Generator(device="cpu").manual_seed(11)
    M = torch.randn((B, n, n), generator=g, dtype=torch.float64)
    A0 = (0.5 * (M + M.transpose(1, 2))).to(dev)
    A = torch.empty_like(A0)
    VT = torch.empty_like(A0)
    ev = torch.empty((B, n), dtype=torch.float64, device=dev)
    ns = torch.full((B,), n, dtype=torch.int32, device=dev)
    sw = torch.zeros(B, dtype=torch.int32, device=dev)
    work = torch.empty(lib.dxtb_jacobi_work_size(B, n) // 8 + 1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    # calibration copy: 512 MiB read + 512 MiB written
    src = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize(dev)

    variant = args.eig_variant  # 2: one workgroup per matrix; 3: the split form (its one sweep split)
    split = 1 if variant == 3 else 0

    def launch():
        A.copy_(A0)  # the blocked variant works in place
        rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(ns), _lib.ptr(A), _lib.ptr(VT), _lib.ptr(ev), _lib.ptr(work),
                                  variant, 1, split, 0.0, None, _lib.ptr(sw), _lib.stream_ptr(dev))
        _lib.check(rc, "dxtb_jacobi_eigh")

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize(dev)
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        A.copy_(A0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(ns), _lib.ptr(A), _lib.ptr(VT), _lib.ptr(ev), _lib.ptr(work),
                                  variant, 1, split, 0.0, None, _lib.ptr(sw), _lib.stream_ptr(dev))
        e1.record(stream)
        _lib.check(rc, "dxtb_jacobi_eigh")
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    wall = (time.perf_counter() - t0) / args.steps
    assert int(sw.min()) == 1 and int(sw.max()) == 1
    t = float(np.mean(ms)) * 1e-3
    nbytes = B * jacobi_bytes_per_sweep(n, deferred(n, B))
    flops = B * jacobi_flops_per_sweep(n, 2)
    res = {
        "metric": f"batched Jacobi sweeps/s ({'split' if variant == 3 else 'one-workgroup'} blocked form, FP64)",
        "value": B / t,
        "unit": "matrix-sweeps/s",
        "ms_per_step": wall * 1e3,
        "dtype": "f64",
        "config": {"workload": f"eigensolver microbenchmark: {B} random symmetric matrices of order {n}, one sweep "
                               "per launch", "kernel_ms": t * 1e3,
                   "fp64_tflops": flops / t / 1e12, "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        "roofline": {"kernel": "k_bj_norms + k_bj_check + k_bj_inner_full + k_bj_step + k_bj_product (+ k_bj_apply) + k_jacobi_blocked (one dxtb_jacobi_eigh call)"
                               if variant == 3 else "k_jacobi_blocked", "bound": "hbm", "achieved": nbytes / t / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": nbytes / t / 1e9 / HBM_PEAK_GBS,
                     "traffic": (JACOBI_TRAFFIC_RATIO_DEFERRED if deferred(n, B) else JACOBI_TRAFFIC_RATIO)[variant]
                     * nbytes, "bytes_per_launch": nbytes},
    }
    return res, None


def _progress(label):
    return lambda done, total: print(f"{label}: {done}/{total} done", file=sys.stderr, flush=True)


def _requests(nums, poss, chg, idx, solv=None):
    """Requests for molecules ``idx``; where ``solv`` is set, with GBSA
    (generalized Born, P16 kernel, eps 78.4, no ALPB term)."""
    from dxtb_amd import serving as sv
    from dxtb_amd.solvation import GeneralizedBorn

    gb = GeneralizedBorn(None, 78.4, alpb=False)
    return [sv.Request(int(i), nums[i], poss[i], float(chg[i]),
                       solvation=gb if solv is not None and solv[i] else None) for i in idx]


def _stratified(nums, gfn, per_class):
    """Indices of up to per_class molecules of every pool class (warm-up:
    creates every pool the workload needs)."""
    from dxtb_amd import serving as sv

    seen = {}
    for i, z in enumerate(nums):
        c = sv.pool_class(*sv.sizes(z, gfn))
        seen.setdefault(c, [])
        if len(seen[c]) < per_class:
            seen[c].append(i)
    return sorted(i for v in seen.values() for i in v)


def bench_gfn2_batch(args, dev):
    """Config 2: GFN2 energy + forces of 1024 seeded molecules of 10-120 atoms
    (H/C/N/O/F/S), all submitted at once to the continuous-batching engine
    (slot pools by basis size; converged molecules leave, waiting ones enter
    between SCF iterations)."""
    from dxtb_amd import scf as hscf
    from dxtb_amd import serving as sv
    from dxtb_amd import workloads

    nums, poss, chg = workloads.config_batch(2, count=args.nmol)

    def engine():
        return sv.Engine("gfn2", dev, opts=hscf.SCFOptions(jacobi_variant=args.jacobi_variant),
                         graphs=args.engine_graphs, graph_window_max=args.graph_window_max)

    # A step serves the whole batch with a new engine (pool creation is in
    # the timed region).  A warm-up step is the same pass: the device
    # allocator's blocks and the library kernels of every window shape are
    # then in place.  (With a small stratified warm-up the timed pass ran
    # ~10 % slower than later passes of the same process; reusing the
    # warmed-up engine instead of a new one, ~5 % slower.)
    for _ in range(args.warmup):
        engine().run(_requests(nums, poss, chg, range(len(nums))), realtime=False, progress=_progress("warm-up"))
    torch.cuda.synchronize(dev)
    walls, out = [], []
    hscf.JACOBI_PROFILE = []
    torch.cuda.synchronize(dev)
    base = torch.cuda.Event(enable_timing=True)
    base.record(torch.cuda.current_stream(dev))
    for _ in range(args.steps):
        eng = engine()
        reqs = _requests(nums, poss, chg, range(len(nums)))
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        out = eng.run(reqs, realtime=False, progress=_progress("config-2 batch"))
        torch.cuda.synchronize(dev)
        walls.append(time.perf_counter() - t0)
        del eng
    roof = jacobi_roofline(hscf.JACOBI_PROFILE, base)
    roof["busy_fraction_of_steps"] = roof["busy_ms"] * 1e-3 / float(np.sum(walls))
    hscf.JACOBI_PROFILE = None
    wall = float(np.mean(walls))
    lat = np.array([r.latency for r in out])
    its = np.array([r.iterations for r in out])
    nconv = int(sum(r.converged for r in out))
    res = {
        "metric": "GFN2 energy+forces molecules/s, p50/p99 latency at |dE|<1e-8 Eh",
        "value": nconv / wall,  # molecules that met the SCF bar (the others stop at the iteration limit)
        "unit": "molecules/s",
        "ms_per_step": wall * 1e3,
        "dtype": "f64",
        "config": {"workload": f"config-2 batched GFN2 energy+forces: {args.nmol} molecules, 10-120 atoms, "
                               "H/C/N/O/F/S, all submitted at t=0 to the continuous-batching engine",
                   "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                   "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                   "scf_iterations_mean": float(its.mean()),
                   "scf_iterations_max": int(its.max()),
                   "converged": nconv,
                   "molecules": len(out),
                   "all_molecules_per_s": len(out) / wall},
        "roofline": roof,
    }
    return res, (nums, poss, chg, "config-2")


def serving_stream(args, dev, nreq: int) -> tuple[dict, dict | None, tuple]:
    """Config 3 on ``nreq`` seeded requests of the stream (log-normal sizes,
    median 40 atoms, 5-400; H/C/N/O/F/S/Cl; charges 0/+1/-1 at 80/10/10; 20 %
    with GBSA, eps 78.4): the offline capacity of the engine on the first
    third, then Poisson arrivals at 30/60/90 % of that capacity on three
    further slices, and the plain one-molecule-per-call GPU path on a sample.
    Returns (results, eigensolver roofline of the streams, workload)."""
    from dxtb_amd import scf as hscf
    from dxtb_amd import serving as sv
    from dxtb_amd import workloads

    nums, poss, chg = workloads.config_batch(3, count=nreq)
    solv = workloads.config3_solvation(len(nums))
    n = len(nums)
    eng = sv.Engine("gfn2", dev, opts=hscf.SCFOptions(jacobi_variant=args.jacobi_variant),
                    graphs=args.engine_graphs, graph_window_max=args.graph_window_max)
    # warm-up: a few molecules of every pool class the stream uses
    eng.run(_requests(nums, poss, chg, _stratified(nums, "gfn2", 4), solv), realtime=False,
            progress=_progress("config-3 warm-up"))
    cap_idx = range(0, n // 3)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    cap_out = eng.run(_requests(nums, poss, chg, cap_idx, solv), realtime=False,
                      progress=_progress("config-3 capacity batch"))
    capacity = len(cap_out) / (time.perf_counter() - t0)
    rng = np.random.default_rng(workloads.seed(3) + 1)
    per = (n - n // 3) // 3
    loads = []
    hscf.JACOBI_PROFILE = []
    torch.cuda.synchronize(dev)
    base = torch.cuda.Event(enable_timing=True)
    base.record(torch.cuda.current_stream(dev))
    for j, frac in enumerate((0.3, 0.6, 0.9)):
        idx = range(n // 3 + j * per, n // 3 + (j + 1) * per)
        reqs = _requests(nums, poss, chg, idx, solv)
        rate = frac * capacity
        t = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
        for r, a in zip(reqs, t):
            r.arrival = float(a)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        out = eng.run(reqs, realtime=True, progress=_progress(f"config-3 stream at {frac:.0%} load"))
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in out])
        loads.append({"offered_mol_s": rate, "achieved_mol_s": len(out) / wall, "requests": len(out),
                      "wall_s": wall, "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                      "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                      "converged": int(sum(r.converged for r in out))})
    roof = jacobi_roofline(hscf.JACOBI_PROFILE, base)
    hscf.JACOBI_PROFILE = None
    sample = _requests(nums, poss, chg, range(n // 3, n // 3 + args.plain_sample), solv)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for k, r in enumerate(sample):
        sv.plain_path(r, "gfn2", dev)
        if k % 8 == 7:
            print(f"plain path: {k + 1}/{len(sample)}", file=sys.stderr, flush=True)
    plain = len(sample) / (time.perf_counter() - t0)
    res = {
        "workload": f"config-3 serving stream: {n} seeded requests of the 100000-request stream (log-normal sizes, "
                    "median 40 atoms, 5-400, H/C/N/O/F/S/Cl, charges 0/+-1, 20% GBSA eps 78.4); offline capacity "
                    f"on the first {n // 3}, Poisson arrivals at 30/60/90% of it on three slices of {per}",
        "requests": n,
        "requests_note": "a prefix of the seeded stream: p50/p99 of a slice of ~400 requests are stable to a few "
                         "percent between slices, while the full 100000 requests would take ~30 min per load",
        "offline_capacity_mol_s": capacity,
        "loads": loads,
        "p99_latency_ms_at_90pct": loads[-1]["p99_latency_ms"],
        "gbsa_requests": int(solv.sum()),
        "plain_path_mol_s": plain,
        "plain_sample": len(sample),
        "serving_over_plain": capacity / plain,
    }
    return res, roof, (nums, poss, chg, "config-3", solv)


def bench_serving(args, dev):
    """Config 3 on its own (--workload serving): the serving stream's numbers
    as the line's value (throughput at the 90 % load)."""
    res3, roof, extra = serving_stream(args, dev, args.nmol)
    top = res3["loads"][-1]
    res = {
        "metric": "GFN2 energy+forces molecules/s, p50/p99 latency at |dE|<1e-8 Eh",
        "value": top["achieved_mol_s"],
        "unit": "molecules/s",
        "ms_per_step": top["wall_s"] * 1e3,
        "dtype": "f64",
        "config": dict(res3, p50_latency_ms=top["p50_latency_ms"], p99_latency_ms=top["p99_latency_ms"]),
        "roofline": roof,
    }
    return res, extra


def _spread_order(nums) -> list[int]:
    """Molecule indices ordered so that every prefix spreads over the size
    range: rank quantiles in van der Corput order (1/2, 1/4, 3/4, 1/8, ...)
    of the molecules sorted by atom count."""
    by_size = np.argsort([len(z) for z in nums], kind="stable")
    m = len(by_size)
    out, seen = [], set()
    k = 1
    while len(out) < m:
        q, f, x = k, 0.5, 0.0
        while q:
            if q & 1:
                x += f
            q >>= 1
            f *= 0.5
        i = int(by_size[min(int(x * m), m - 1)])
        if i not in seen:
            seen.add(i)
            out.append(i)
        k += 1
        if k > 64 * m:
            out += [int(j) for j in by_size if int(j) not in seen]
            break
    return out


def cpu_baseline_gfn2(nums, poss, chg, label, solv=None, budget_s=20.0):
    """The FP64 CPU oracle (kind "port": torch on the host cores, LAPACK eigh,
    xitorch-style broyden1 at the GPU path's tolerance 1e-8, variational
    autograd forces) on a size-spread sample of the workload for about
    ``budget_s`` seconds."""
    from dxtb_amd import scf as hscf
    from oracle import scf as oscf

    nthreads = os.cpu_count() or 1
    torch.set_num_threads(nthreads)
    order = _spread_order(nums)
    t0 = time.perf_counter()
    per, atoms = [], []
    for i in order:
        if time.perf_counter() - t0 >= budget_s:
            break
        sv = {"eps": 78.4, "alpb": False, "kernel": "p16"} if solv is not None and solv[i] else None
        t1 = time.perf_counter()
        oscf.singlepoint(nums[i].tolist(), torch.tensor(poss[i], dtype=torch.float64), "gfn2", charge=float(chg[i]),
                         tol=1e-8, maxiter=hscf.SCFOptions().maxiter, forces=True, solvation=sv)
        per.append(time.perf_counter() - t1)
        atoms.append(len(nums[i]))
    dt = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    k = len(per)
    return {
        "value": k / dt,
        "unit": "molecules/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"{k} {label} molecules spread over the size range (atom counts {min(atoms)}-{max(atoms)}, "
                  f"mean {np.mean(atoms):.0f}; the workload's mean {np.mean([len(z) for z in nums]):.0f}), "
                  "GFN2 energy+forces, FP64 torch CPU oracle (xitorch-style broyden1 at f_tol = x_tol = 1e-8 as "
                  "on the GPU, LAPACK eigh, variational autograd forces)",
        "p50_ms_per_molecule": float(np.percentile(per, 50) * 1e3),
        "p99_ms_per_molecule": float(np.percentile(per, 99) * 1e3),
        "torch_threads": torch.get_num_threads(),
        "nproc": nthreads,
        "cpu_model": model,
    }


EIGEN_KERNELS = ("k_bj_", "k_jacobi_")


def trace_summary(path: str, steps: int = 1) -> None:
    """Summary of a rocprofv3 kernel trace (``--kernel-trace --output-format
    csv``, the ``*_kernel_trace.csv`` file): per-kernel calls, total and mean
    duration, the device busy time (union of all kernel intervals) and the
    eigensolver's busy time (union of its kernels' intervals), per step of the
    traced run; no GPU needed."""
    import csv

    rows, streams = [], []
    with open(path) as fh:
        for d in csv.DictReader(fh):
            rows.append((int(d["Start_Timestamp"]), int(d["End_Timestamp"]),
                         d["Kernel_Name"].replace("(anonymous namespace)::", "")))
            streams.append(d.get("Stream_Id") or d.get("Queue_Id") or "0")
    if not rows:
        print("no kernels")
        return
    tot: dict = {}
    for a, b, name in rows:
        k = name.split("(")[0].replace("void ", "")[:80]
        c, t = tot.get(k, (0, 0))
        tot[k] = (c + 1, t + b - a)
    span = (max(b for _, b, _ in rows) - min(a for a, _, _ in rows)) * 1e-6
    busy = _union_ms([(a * 1e-6, b * 1e-6) for a, b, _ in rows])
    eig = [(a * 1e-6, b * 1e-6) for a, b, n in rows if any(k in n for k in EIGEN_KERNELS)]
    eig_busy = _union_ms(eig)
    # whole eigensolver calls, as bench.py times them with events: from a
    # call's argument-block kernel (k_bj_args) to its finishing kernel
    # (k_jacobi_blocked) on the same stream, or one k_jacobi_scalar launch
    calls, open_at = [], {}
    for (a, b, name), st in sorted(zip(rows, streams)):
        if "k_bj_args" in name:
            open_at[st] = a
        elif "k_jacobi_blocked" in name and st in open_at:
            calls.append((open_at.pop(st) * 1e-6, b * 1e-6))
        elif "k_jacobi_scalar" in name:
            calls.append((a * 1e-6, b * 1e-6))
    call_busy = _union_ms(calls)
    allsum = sum(t for _, t in tot.values()) * 1e-6
    print(f"kernels {len(rows)}; trace span {span:.1f} ms; device busy (union) {busy:.1f} ms; "
          f"summed kernel time {allsum:.1f} ms; eigensolver busy (union of its kernels) {eig_busy:.1f} ms "
          f"({100 * eig_busy / busy:.1f} % of busy); per step ({steps}): busy {busy / steps:.1f} ms, "
          f"eigensolver {eig_busy / steps:.1f} ms; eigensolver calls {len(calls)}, busy (union of whole calls, "
          f"as bench.py's events) {call_busy:.1f} ms")
    print("| kernel | calls | total ms | % of summed | mean us |")
    print("|---|---:|---:|---:|---:|")
    for k, (c, t) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:30]:
        print(f"| `{k}` | {c} | {t * 1e-6:.1f} | {100 * t * 1e-6 / allsum:.1f} | {t * 1e-3 / c:.1f} |")


def pmc_summary(path: str) -> None:
    """Summary of a rocprofv3 counter collection (``--pmc ... --output-format
    csv``, the ``*_counter_collection.csv`` file): per kernel and counter, the
    dispatches and the counter's sum and mean per dispatch; no GPU needed."""
    import csv

    acc: dict = {}
    with open(path) as fh:
        for d in csv.DictReader(fh):
            k = d["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")[:70]
            key = (k, d["Counter_Name"])
            n, v, ids = acc.get(key, (0, 0.0, set()))
            ids.add(d.get("Dispatch_Id", ""))
            acc[key] = (n + 1, v + float(d["Counter_Value"]), ids)
    print("| kernel | counter | dispatches | sum | mean per dispatch |")
    print("|---|---|---:|---:|---:|")
    for (k, cn), (n, v, ids) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        nd = max(len(ids), 1)
        print(f"| `{k}` | {cn} | {nd} | {v:.0f} | {v / nd:.1f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pmc-summary", default=None, metavar="COUNTER_COLLECTION_CSV",
                    help="summarise a rocprofv3 counter collection instead of benchmarking (no GPU)")
    ap.add_argument("--trace-summary", default=None, metavar="KERNEL_TRACE_CSV",
                    help="summarise a rocprofv3 kernel trace instead of benchmarking (no GPU)")
    ap.add_argument("--trace-steps", type=int, default=1)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="gfn2-batch", choices=["gfn2-batch", "serving", "integrals", "eigen"])
    ap.add_argument("--eig-n", type=int, default=240)
    ap.add_argument("--eig-variant", type=int, default=2, choices=[2, 3])
    ap.add_argument("--nmol", type=int, default=None)
    ap.add_argument("--serving-requests", type=int, default=1200,
                    help="requests of the config-3 stream run once after the config-2 steps (0: skip)")
    ap.add_argument("--plain-sample", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--engine-graphs", action="store_true",
                    help="serving engine iterates each pool by replaying a captured hipGraph")
    ap.add_argument("--graph-window-max", type=int, default=None,
                    help="with --engine-graphs: replay graphs only for windows of at most this many slots")
    ap.add_argument("--ao-order", default="l", choices=["l", "reference"],
                    help="integrals workload: AO layout of the matrices")
    ap.add_argument("--jacobi-variant", type=int, default=3, choices=[0, 2, 3],
                    help="eigensolver form above order 96: 3 split (default), 2 one workgroup per matrix, 0 auto")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.trace_steps)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (one process, one device); --gpus must be 1"}))
        sys.exit(2)
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU available"}))
        sys.exit(1)
    dev = torch.device("cuda:0")
    extra3 = None
    if args.workload == "integrals":
        args.nmol = args.nmol or 512
        res, extra = bench_integrals(args, dev)
    elif args.workload == "eigen":
        args.nmol = args.nmol or 256
        args.no_cpu_baseline = True
        res, extra = bench_eigen(args, dev)
    elif args.workload == "serving":
        args.nmol = args.nmol or 1200
        res, extra = bench_serving(args, dev)
    else:
        args.nmol = args.nmol or 1024
        res, extra = bench_gfn2_batch(args, dev)
        if args.serving_requests > 0:
            res3, roof3, extra3 = serving_stream(args, dev, args.serving_requests)
            res3["roofline_eigensolver"] = roof3
            res["config"]["config3_serving"] = res3
    res.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "data": "synthetic (seeded)"})
    if not args.no_cpu_baseline:
        if args.workload == "integrals":
            res["cpu_baseline"] = cpu_baseline_integrals(extra[0], extra[1])
        else:
            res["cpu_baseline"] = cpu_baseline_gfn2(*extra)
        if extra3 is not None:
            cpu3 = cpu_baseline_gfn2(*extra3, budget_s=15.0)
            c3 = res["config"]["config3_serving"]
            c3["cpu_baseline"] = cpu3
            c3["serving_over_cpu"] = c3["offline_capacity_mol_s"] / cpu3["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
