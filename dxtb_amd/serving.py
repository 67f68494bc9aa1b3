"""
Continuous-batching GFN1/GFN2 energy + forces engine (the serving loop).

Requests (atomic numbers, positions, charge) are admitted into *slot pools*,
one pool per basis-size class with fixed padded sizes.  An engine tick
advances every pool that holds work by one SCF iteration over its occupied
slots (one batched launch sequence per pool), retires the molecules that
converged (energy and forces of all of them in one batched call) and admits
waiting requests into the freed slots.  Requests that arrive while others
are on the GPU join at the next tick; nothing waits for a batch to drain.

A molecule's iteration sequence is exactly the one ``scf.run_scf`` runs for
it alone: its potential vector, Broyden history (xitorch broyden1 as in the
reference's scf/implicit path, exlibs/xitorch/_impls/optimize/root/
rootsolver.py:42-300), iteration count and warm-start orbitals live in its
slot, and every batched kernel treats each molecule independently.  So a
result does not depend on which molecules shared the pool.

Device memory of a pool is allocated once, when the pool is created (the
arrays of a slot are overwritten on admission); each pool iterates only the
window of rows up to its highest occupied slot.
"""

from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import param
from . import scf as hscf
from .basis import build_batch, shell_table

# static window sizes of the captured iteration graphs
GRAPH_WINDOWS = (8, 16, 32, 64, 128, 256)
# iterations a pool keeps in flight: the next one is already on the pool's
# stream while the host reads the flags of the last one and retires and
# admits molecules, so the stream does not wait for the host
PIPELINE = 2
# basis-size classes (padded nao) of the slot pools
POOL_NAO = (48, 96, 144, 192, 256, 320, 384, 448, 512, 640, 768, 1024, 1280, 1536)
# padded atom and shell counts of a class, as fractions of its nao; a molecule
# goes to the smallest class that holds all three of its sizes
NAT_FRAC, NSH_FRAC = 0.5, 0.75
# AO layout of the pools' matrices: shells grouped by angular momentum, so the
# integral kernels of one class (l_A, l_B) write and read contiguous column
# ranges (the requests' energies and forces do not depend on the layout)
AO_ORDER = "l"


@dataclass
class Request:
    rid: int
    numbers: np.ndarray  # (nat,) atomic numbers
    positions: np.ndarray  # (nat, 3) Bohr
    charge: float = 0.0
    arrival: float = 0.0  # seconds on the engine clock
    admitted: float | None = None
    finished: float | None = None
    energy: float | None = None
    forces: np.ndarray | None = None  # (nat, 3) -dE/dR, Eh/Bohr
    iterations: int = 0
    converged: bool = False
    solvation: object = None  # GeneralizedBorn settings, or None (vacuum)
    spin: int | None = None  # unpaired electrons (None: from the parity of the electron count)

    @property
    def latency(self) -> float:
        return float(self.finished - self.arrival)


def _to_device(a: np.ndarray, dev) -> torch.Tensor:
    """Host array to the device through pinned memory, without waiting: a
    copy from pageable memory can wait for the whole device, i.e. for the
    kernels of the other pools' streams."""
    return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)


def sizes(numbers: np.ndarray, gfn: str) -> tuple[int, int, int]:
    """(nat, nsh, nao) of a molecule."""
    tab = shell_table(gfn)
    z = np.asarray(numbers, dtype=np.int64)
    z = z[z != 0]
    if (z < 0).any() or (z >= len(tab.first)).any() or (tab.first[np.clip(z, 0, len(tab.first) - 1)] < 0).any():
        raise ValueError(f"atomic numbers {sorted(set(z.tolist()))} include elements not parametrized in {gfn}")
    if (tab.first[z] < 0).any():
        raise ValueError(f"elements {sorted(set(z[tab.first[z] < 0].tolist()))} are not parametrized in {gfn}")
    nsh = int(tab.count[z].sum())
    nao = 0
    for zz in z:
        f = tab.first[zz]
        nao += int((2 * tab.l[f : f + tab.count[zz]] + 1).sum())
    return len(z), nsh, nao


def class_dims(nao_max: int) -> tuple[int, int, int]:
    """Padded (nat_max, nsh_max, nao_max) of the pool class ``nao_max``."""
    return max(2, int(np.ceil(NAT_FRAC * nao_max))), max(2, int(np.ceil(NSH_FRAC * nao_max))), nao_max


def pool_class(nat: int, nsh: int, nao: int) -> int:
    """Smallest class holding a molecule of these sizes."""
    for c in POOL_NAO:
        na, ns, n = class_dims(c)
        if nat <= na and nsh <= ns and nao <= n:
            return c
    raise ValueError(f"molecule too large for the serving pools (nao {nao} > {POOL_NAO[-1]})")


def default_capacity(nao_max: int, history: int = 601, budget_bytes: float = 12e9) -> int:
    """Slots of a pool: as many as fill the device (256 workgroups of the
    per-molecule kernels) within a memory budget per pool; ``history`` is
    the Broyden history length per slot (maxiter + 1)."""
    na, ns, n = class_dims(nao_max)
    per_slot = 8 * (24 * n * n + 2 * history * (ns + 9 * na))  # matrices + Broyden history (upper bound)
    return int(max(8, min(256, budget_bytes // per_slot)))


class SlotPool:
    """Fixed-capacity slots of one basis-size class."""

    def __init__(self, gfn: str, nao_max: int, capacity: int, device, opts: hscf.SCFOptions, graphs: bool = False,
                 graph_window_max: int | None = None):
        if opts.scf_mode not in ("implicit", "default"):
            raise ValueError(f"the serving pools iterate the implicit mode (broyden1), not scf_mode "
                             f"'{opts.scf_mode}'; the full mode runs through scf.run_scf and the calculators")
        self.gfn = gfn
        self.dims = class_dims(nao_max)
        self.cap = capacity
        self.dev = torch.device(device)
        self.opts = opts
        # all work of the pool runs on its own stream (default priority), so
        # the iterations of different pools overlap on the device
        self.stream = torch.cuda.Stream(self.dev)
        # every slot starts with a placeholder molecule (H2); empty slots are
        # never iterated past the window and their results are never read
        h2 = np.zeros((capacity, 2), dtype=np.int64) + 1
        pos = np.zeros((capacity, self.dims[0], 3))
        pos[:, 1, 2] = 1.4
        bb = build_batch(h2, gfn, self.dev, self.dims, ao_order=AO_ORDER)
        self.state = hscf.GeometryState(bb, torch.as_tensor(pos, device=self.dev), opts=opts)
        self.state.U = torch.eye(nao_max, dtype=torch.float64, device=self.dev).expand(capacity, nao_max,
                                                                                      nao_max).contiguous()
        self.state.Corb = self.state.X.clone()
        nbytes = self.state.lib.dxtb_jacobi_work_size(capacity, nao_max)
        self.state._jwork = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=self.dev)  # shared by windows
        L = self.state.nvec
        # per-slot Broyden state on the device (potential vectors, history,
        # control words, stop flags): an iteration needs no host data
        self.mix = hscf.DeviceBroyden(capacity, L, opts.maxiter + 1, self.dev)
        self.occupied = np.zeros(capacity, dtype=bool)
        self.req: list[Request | None] = [None] * capacity
        # pool iterations launched, and the count at each slot's admission: a
        # slot's Broyden history holds at most the iterations since then
        self.iters_run = 0
        self.adm_it = np.zeros(capacity, dtype=np.int64)
        # pinned staging of the stop flags and iteration counts, one pair per
        # iteration in flight (PIPELINE of them) plus one
        self._stage = [(torch.zeros(capacity, dtype=torch.int32, pin_memory=True),
                        torch.zeros(capacity, dtype=torch.int32, pin_memory=True)) for _ in range(PIPELINE + 1)]
        self._wins: dict = {}
        # graph mode needs a capturable eigensolver (the Jacobi kernels)
        self.graphs = graphs and nao_max <= self.state.lib.dxtb_jacobi_max_n()
        # graphs only for windows up to this many rows (None: every window)
        self.graph_wmax = capacity if graph_window_max is None else graph_window_max
        self._graphs: dict = {}
        self._gpool = torch.cuda.graph_pool_handle() if self.graphs else None
        self.inflight: collections.deque = collections.deque()  # launched iterations awaiting finish()
        self.want_compact = False  # compact once no iteration is in flight
        self.retiring: list = []  # (event, requests, host energy, host forces, ...) in flight
        self.moves = 0  # slots moved by compaction
        self.admitted_busy = 0  # requests admitted while other slots were mid-SCF

    @property
    def nfree(self) -> int:
        return int((~self.occupied).sum())

    @property
    def hi(self) -> int:
        occ = np.nonzero(self.occupied)[0]
        return int(occ[-1]) + 1 if len(occ) else 0

    # ---- admission ---------------------------------------------------------
    def admit(self, reqs: list[Request], now: float) -> None:
        m = len(reqs)
        if m == 0:
            return
        rows = np.nonzero(~self.occupied)[0][:m]
        if len(rows) < m:
            raise RuntimeError("not enough free slots")
        if self.occupied.any():
            self.admitted_busy += m
        na = self.dims[0]
        N = np.zeros((m, na), dtype=np.int64)
        P = np.zeros((m, na, 3))
        for i, r in enumerate(reqs):
            N[i, : len(r.numbers)] = r.numbers
            P[i, : len(r.numbers)] = r.positions
        bb = build_batch(N, self.gfn, self.dev, self.dims, ao_order=AO_ORDER)
        chg = np.array([r.charge for r in reqs], dtype=np.float64)
        spins = [r.spin for r in reqs]
        st = hscf.GeometryState(bb, _to_device(P, self.dev), charges=chg, opts=self.opts,
                                solvation=[r.solvation for r in reqs],
                                spins=None if all(x is None for x in spins) else spins)
        rows64 = np.asarray(rows, dtype=np.int64)
        rt = _to_device(rows64, self.dev)
        self.state.scatter(rows64, st)
        self.mix.start(rt)
        self.occupied[rows] = True
        self.adm_it[rows] = self.iters_run
        for r, row in zip(reqs, rows):
            r.admitted = now
            self.req[row] = r

    def compact(self) -> bool:
        """Move the occupied slots above the first free ones into those free
        slots (state, Broyden vectors and history, bookkeeping), so that the
        window, which runs up to the highest occupied slot, covers only live
        molecules.  Between iterations, on the pool's stream.  Returns whether
        anything moved."""
        occ = np.nonzero(self.occupied)[0]
        n = len(occ)
        if n == 0 or occ[-1] < n:
            self.want_compact = False
            return False
        src = occ[occ >= n]
        dst = np.nonzero(~self.occupied[:n])[0]
        self.state.move(src, dst)
        st, dt = _to_device(src, self.dev), _to_device(dst, self.dev)
        self.mix.move(st, dt)
        self.mix.ctl[:, 0].index_fill_(0, st, 0)  # the source slots are free
        self.occupied[dst] = True
        self.occupied[src] = False
        self.adm_it[dst] = self.adm_it[src]
        for s_, d_ in zip(src, dst):
            self.req[d_], self.req[s_] = self.req[s_], None
        self.moves += len(src)
        self.want_compact = False
        return True

    # ---- one SCF iteration of the occupied window --------------------------
    def _graphed(self, hi: int) -> bool:
        return self.graphs and hi <= self.graph_wmax

    def _window(self, hi: int) -> int:
        """Rows iterated: the occupied prefix, rounded up to a static size in
        graph mode (one captured graph per size)."""
        if not self._graphed(hi):
            return hi
        for w in GRAPH_WINDOWS:
            if hi <= w <= self.cap:
                return w
        return self.cap

    def _iterate(self, W: int, kbound: int = 0) -> None:
        """One iteration over rows [0, W): the SCF map at every row's next
        point, then the device Broyden step (dxtb_broyden_step), which skips
        free and stopped slots.  No host data, fixed shapes (capturable);
        ``kbound`` bounds the rows' history lengths (0: the capacity)."""
        o = self.opts
        win = self._wins.get(W)
        if win is None:
            win = self._wins[W] = self.state.window(W)
        # the solvation flags of the pool change when a GBSA request is admitted
        # after the window was made (its tensors are views of the pool's rows)
        win.any_gb, win._alpb, win.gb_par = self.state.any_gb, self.state._alpb, self.state.gb_par
        m = self.mix
        if o.adaptive_eigensolver:  # fresh slots: a loose bar for the cold first solve
            started = m.ctl[:W, 1] != 0
            win._jtol = torch.where(started, hscf.jacobi_tolerance(hscf.wnorm(m.y[:W], win.wvec)),
                                    hscf.JACOBI_TOL_LOOSE).contiguous()
        gx = win.g(m.xe[:W])
        m.step(gx, win.wvec, o, nb=W, kbound=kbound)

    def launch(self) -> Launched | None:
        """Enqueue one iteration over the occupied window on the current
        (pool) stream, and the copies of its stop flags and iteration counts
        to the host; returns what ``finish`` needs.  In graph mode the
        iteration is one replay of a captured graph (captured after an eager
        iteration of the same window)."""
        hi = self.hi
        if hi == 0:
            return None
        W = self._window(hi)
        if self._graphed(hi):
            key = (W, self.state.any_gb, self.state._alpb)
            g = self._graphs.get(key)
            if g is None:
                self._iterate(W)  # eager: warms up, and is this iteration
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, pool=self._gpool, stream=torch.cuda.current_stream(self.dev)):
                    self._iterate(W)
                self._graphs[key] = g
            else:
                g.replay()
        else:
            live = self.occupied[:W]
            self._iterate(W, kbound=int(self.iters_run - self.adm_it[:W][live].min()) + 1)
        t = self.iters_run
        self.iters_run += 1
        busy = {id(e.h_flag) for e in self.inflight}
        h_flag, h_iter = next(b for b in self._stage if id(b[0]) not in busy)
        h_flag[:hi].copy_(self.mix.flag[:hi], non_blocking=True)
        h_iter[:hi].copy_(self.mix.ctl[:hi, 3], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.dev))
        return Launched(t, hi, ev, h_flag, h_iter)

    def finish(self, it: Launched | None, now_fn) -> list[Request]:
        """Host bookkeeping of a completed iteration: retires the slots it
        stopped (converged, or at the iteration limit).  Another iteration
        may be in flight: a molecule that stopped is frozen on the device
        (the Broyden step skips it), and slots admitted after this iteration
        was launched are not its molecules."""
        if it is None:
            return []
        hi = it.hi
        flag = it.h_flag[:hi].numpy().copy()
        its = it.h_iter[:hi].numpy().astype(np.int64)
        out = np.nonzero(self.occupied[:hi] & (flag != 0) & (self.adm_it[:hi] <= it.t))[0]
        if len(out) == 0:
            return []
        res = self.retire(out, flag[out] == 1, its[out], now_fn)
        # a long-running molecule in a high slot would keep the window wide
        n = int(self.occupied.sum())
        if self.hi - n >= max(4, n // 4):
            self.want_compact = True
        return res

    def retire(self, rows: np.ndarray, conv: np.ndarray, iters: np.ndarray, now_fn) -> list[Request]:
        """Enqueue energy and forces of the finished slots ``rows`` on the
        pool's stream (on copies of their rows, so the slots are free at once)
        and copy the results to pinned host memory without waiting; the
        requests come back from ``collect`` once that work has completed."""
        reqs = [self.req[r] for r in rows]
        na = self.dims[0]
        N = np.zeros((len(rows), na), dtype=np.int64)
        for i, r in enumerate(reqs):
            N[i, : len(r.numbers)] = r.numbers
        bb = build_batch(N, self.gfn, self.dev, self.dims, ao_order=AO_ORDER)
        rows64 = np.asarray(rows, dtype=np.int64)
        sub = self.state.subset(rows64, bb=bb)
        res = hscf.finalize(sub, self.mix.x.index_select(0, _to_device(rows64, self.dev)), iters.copy(),
                            conv.copy())
        f = hscf.forces(res)
        e_h = torch.empty(res.energy.shape, dtype=torch.float64, pin_memory=True)
        f_h = torch.empty(f.shape, dtype=torch.float64, pin_memory=True)
        e_h.copy_(res.energy, non_blocking=True)
        f_h.copy_(f, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.dev))
        self.retiring.append((ev, reqs, e_h, f_h, iters.copy(), conv.copy(), (res, f)))
        self.occupied[rows] = False
        for r in rows:
            self.req[r] = None
        return self.collect(now_fn)

    def collect(self, now_fn, wait: bool = False) -> list[Request]:
        """Requests whose energy and forces have reached the host."""
        out, keep = [], []
        for item in self.retiring:
            ev, reqs, e_h, f_h, its, conv, _ = item
            if not (wait or ev.query()):
                keep.append(item)
                continue
            ev.synchronize()
            t = now_fn()
            energy, frc = e_h.numpy(), f_h.numpy()
            for i, r in enumerate(reqs):
                r.energy = float(energy[i])
                r.forces = frc[i, : len(r.numbers)].copy()
                r.iterations = int(its[i])
                r.converged = bool(conv[i])
                r.finished = t
            out += reqs
        self.retiring = keep
        return out


@dataclass
class Launched:
    """An iteration of a pool in flight: its number, window, completion
    event and the pinned copies of its stop flags and iteration counts."""

    t: int
    hi: int
    event: torch.cuda.Event
    h_flag: torch.Tensor
    h_iter: torch.Tensor


@dataclass
class Engine:
    """Continuous-batching engine over the slot pools of one parametrization."""

    gfn: str = "gfn2"  # a parametrization name, Param, record or file
    device: str | torch.device = "cuda:0"
    opts: hscf.SCFOptions = field(default_factory=hscf.SCFOptions)
    capacity: dict | None = None  # class nao -> slots (default: default_capacity)
    # one captured hipGraph per pool iteration (pools within the Jacobi
    # solver's order limit): cuts the host work per iteration, but the
    # static windows and the full Broyden history cost more device time than
    # that saves on the measured workloads
    graphs: bool = False
    # with graphs: only windows of at most this many rows replay a graph (the
    # latency-bound tail of a pool); None: every window
    graph_window_max: int | None = None
    pools: dict = field(default_factory=dict)
    queue: dict = field(default_factory=dict)  # class -> FIFO of requests
    t0: float = field(default_factory=time.perf_counter)

    def __post_init__(self) -> None:
        # a Param, its record or a file: registered under its own name
        self.gfn = param.register(self.gfn)

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def pool(self, c: int) -> SlotPool:
        if c not in self.pools:
            cap = (self.capacity or {}).get(c, default_capacity(c, self.opts.maxiter + 1))
            p = SlotPool(self.gfn, c, cap, self.device, self.opts, graphs=self.graphs,
                         graph_window_max=self.graph_window_max)
            torch.cuda.synchronize(p.dev)  # pool arrays were filled on the creating stream
            self.pools[c] = p
        return self.pools[c]

    def submit(self, r: Request) -> None:
        c = pool_class(*sizes(r.numbers, self.gfn))
        self.queue.setdefault(c, []).append(r)

    @property
    def busy(self) -> bool:
        return (any(p.occupied.any() or p.inflight or p.retiring for p in self.pools.values())
                or any(self.queue.values()))

    def _admit_into(self, c: int, p: "SlotPool") -> None:
        """Admit waiting requests of class ``c`` into the free slots of pool
        ``p``.  The set-up (integrals, H0, orthogonaliser by the Jacobi
        solver, solvation) runs on the pool's own stream after the pool's
        iterations in flight and the copies of retiring rows, so other pools
        are not held up."""
        q = self.queue.get(c)
        if not q:
            return
        m = min(p.nfree, len(q))
        if m:
            with torch.cuda.stream(p.stream):
                p.admit(q[:m], self.now())
            del q[:m]

    def admit(self) -> None:
        """Admit waiting requests into every pool with no iteration in flight
        (creating the pools of new classes)."""
        for c, q in self.queue.items():
            if q:
                self._admit_into(c, self.pool(c))

    def step(self) -> list[Request]:
        """One pass of the event loop: admit into new pools, then for every
        pool: finish its completed iterations (retiring converged molecules),
        admit waiting requests into the freed slots and keep PIPELINE
        iterations in flight, so that the pool's stream already holds the
        next iteration while the host reads the flags of the last one.
        Compaction waits until the pool has no iteration in flight.  Pools
        advance independently: a pool of small molecules does not wait for a
        pool of large ones, and a busy pool takes new molecules between two
        of its iterations.  Pools of larger molecules are served first: their
        iterations are the longest, so they carry a batch's critical path (in
        insertion order the config-2 batch ran at 145-172 molecules/s
        depending on how the pools happened to interleave, largest first at
        170-171)."""
        self.admit()
        done = []
        for c in sorted(self.pools, reverse=True):
            p = self.pools[c]
            if p.retiring:
                done += p.collect(self.now)
            while p.inflight and p.inflight[0].event.query():
                it = p.inflight.popleft()
                with torch.cuda.stream(p.stream):
                    done += p.finish(it, self.now)
            if p.want_compact and not p.inflight:
                with torch.cuda.stream(p.stream):
                    p.compact()
            self._admit_into(c, p)
            while len(p.inflight) < PIPELINE and p.occupied.any() and not p.want_compact:
                with torch.cuda.stream(p.stream):
                    it = p.launch()
                if it is None:
                    break
                p.inflight.append(it)
        return done

    def tick(self) -> list[Request]:
        """One synchronous iteration of every pool (admit, launch all, finish
        all)."""
        out = self.step()
        for p in self.pools.values():
            for it in p.inflight:
                it.event.synchronize()
        return out + self.step()

    def run(self, requests: list[Request], realtime: bool = True, progress=None) -> list[Request]:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` a request
        is submitted once the engine clock passes its arrival time; otherwise
        all are submitted at once (offline batch).  ``progress(done, total)``
        is called whenever another 10 % of the requests has completed."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        i = 0
        out = []
        mark = max(1, len(reqs) // 10)
        while i < len(reqs) or self.busy:
            now = self.now()
            while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                if not realtime:
                    reqs[i].arrival = 0.0
                self.submit(reqs[i])
                i += 1
            if not self.busy:
                if i < len(reqs):
                    time.sleep(max(0.0, min(reqs[i].arrival - self.now(), 0.01)))
                continue
            n0 = len(out)
            out += self.step()
            if progress is not None and len(out) // mark > n0 // mark:
                progress(len(out), len(reqs))
        return out


def plain_path(req: Request, gfn: str, device, opts: hscf.SCFOptions | None = None) -> Request:
    """One molecule per call: setup, SCF and forces of a single request (the
    baseline the serving form is measured against)."""
    from .basis import build_batch as bb_

    dev = torch.device(device)
    N = np.asarray(req.numbers, dtype=np.int64)[None]
    bb = bb_(N, param.register(gfn), dev)
    st = hscf.GeometryState(bb, torch.as_tensor(np.asarray(req.positions, dtype=np.float64)[None], device=dev),
                            charges=np.array([req.charge]), opts=opts, solvation=[req.solvation],
                            spins=None if req.spin is None else [req.spin])
    res = hscf.run_scf(st)
    f = hscf.forces(res)
    req.energy = float(res.energy[0])
    req.forces = f[0, : len(req.numbers)].cpu().numpy()
    req.iterations = int(res.iterations[0])
    req.converged = bool(res.converged[0])
    return req
