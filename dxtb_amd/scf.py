"""
Batched GFN1/GFN2-xTB self-consistent field on the GPU.

Per geometry (once):  CN, H0 pair factors, repulsion, AES2 radii
(``dxtb_setup``); S, D, Q and H0 (``dxtb_multipole_fwd``); the orthogonaliser
X = V diag(l)^-1/2 from the Jacobi eigendecomposition of S.

Per SCF iteration, for all active molecules of the batch:
  Fock build (``dxtb_fock``) -> F' = X^T F X -> eigensolve -> C = X C'
  -> Fermi occupations (``dxtb_fermi``) -> P = C f C^T
  -> Mulliken charges and atomic multipoles (``dxtb_charges``)
  -> ES2/ES3/AES2 potentials (``dxtb_potential``) -> Broyden update.

The self-consistency variable is the potential (the reference's
``scp_mode="potential"``, ``scf/pure/iterations.py:74-110``): the shell
potential (monopole, incl. the atom part), the atomic dipole and quadrupole
potentials.  The mixer is the first ("good") Broyden method of xitorch's
``broyden1`` (``exlibs/xitorch/_impls/optimize/root/rootsolver.py:42-261``,
alpha=-0.5, damp=0.5, no line search) applied per molecule, so that a
molecule's iterates do not depend on the other molecules of the batch.
A molecule is converged when |dx| < x_tol and |y| < f_tol for its own
vector; converged molecules are frozen while the rest keeps iterating.
"""

from __future__ import annotations

import collections

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, param
from . import solvation as gbsolv
from .basis import BatchBasis, to_device
from .integrals import multipole_forward
from .params_dev import device_params

PAD_DIAG = 1.0e3  # Fock diagonal of padding AOs: above every real level
# Jacobi stops when max|offdiag| <= JACOBI_TOL * max(1, max|diag|): orbital
# errors are then O(1e-12 Eh / gap), far below what the 1e-8 Eh energy and
# 1e-6 Eh/Bohr force bars can see (the energy is variational in the orbitals).
JACOBI_TOL = 1.0e-12
# optional inexact eigensolves while the SCF is far from convergence: the
# Jacobi bar of a molecule is JACOBI_RES_FACTOR x its current residual |y|,
# within [JACOBI_TOL, JACOBI_TOL_LOOSE]; at convergence (|y| < 1e-8) the bar is
# JACOBI_TOL again, and the final evaluation always uses JACOBI_TOL.
JACOBI_RES_FACTOR = 1.0e-4
JACOBI_TOL_LOOSE = 1.0e-6


def jacobi_tolerance(res: torch.Tensor) -> torch.Tensor:
    return (JACOBI_RES_FACTOR * res).clamp(JACOBI_TOL, JACOBI_TOL_LOOSE).contiguous()
# orders up to this use the blocked Jacobi solver with one workgroup per
# matrix (variant 2), larger ones SCFOptions.jacobi_variant.  Measured on
# warm-start-like batches (4 sweeps): order 48, 0.31-0.36 ms against 0.63-0.67
# with the scalar LDS variant and 0.33-0.44 split; order 96, 0.53-0.65 ms
# against 3.7 scalar and 0.53-0.75 split.  The blocked forms agree bit for bit,
# so the choice does not change a result.
JACOBI_ONE_WG_MAX = 96
# when a list, solve() appends (start event, end event, order, basis sizes,
# sweeps tensor, variant) for every Jacobi launch (kernel timing on the
# launch stream; used by the benchmark)
JACOBI_PROFILE: list | None = None


@dataclass
class SCFOptions:
    etemp: float = 300.0
    f_tol: float = 1e-8
    x_tol: float = 1e-8
    # broyden1 iterations per molecule.  On the config-2 molecules (random
    # 10-120 atom clusters with 10-14 fractionally occupied levels at 300 K)
    # the oracle needs 94 on average and up to 503; the count of such a
    # molecule follows rounding (a different summation order moves it by
    # hundreds), so the limit leaves room: 1023 of 1024 converge within 600
    # on the device, 1014 within 250
    maxiter: int = 600
    damp: float = 0.5
    alpha: float = -0.5
    eigensolver: str = "jacobi"  # "jacobi": hand-written Jacobi for nao <= 4096, else device eigh
    # Jacobi bar tied to the SCF residual (jacobi_tolerance).  Off: on the
    # config-2 molecules (small gaps) the inexact early solves cost more
    # Broyden iterations than they save in sweeps (750 instead of 1011 of 1024
    # converged within 250 iterations)
    adaptive_eigensolver: bool = False
    # dxtb_jacobi_eigh form above order JACOBI_ONE_WG_MAX: 3 the split form
    # (sweeps spread over many workgroups per matrix), 2 one workgroup per
    # matrix, 0 the library's choice for a lone batch (split below half the
    # basis size), 1 the scalar LDS form (order <= 128).
    # The split form's small workgroups share the device with the other slot
    # pools' kernels: config-2 batch 132.7 molecules/s against 117.6 with 2
    # and 116.7 with 0
    jacobi_variant: int = 3
    # fixed-point form (scf/unrolling/base.py, scf/implicit/default.py):
    # "implicit", broyden1 on the residual; "full", the full-mode loop with
    # the Anderson or simple mixer (the reference's default mode: its
    # Broyden is not available there and falls back to Anderson)
    scf_mode: str = "implicit"
    mixer: str = "anderson"
    # full mode: converged when the unmixed residual has L2 norm < x_tol and
    # max norm < x_tol_max; mixer parameters (constants/defaults.py:148-172)
    x_tol_max: float = 1e-5
    damp_init: float = 0.1
    damp_generations: int = 5
    damp_soft_start: bool = True
    damp_diagonal_offset: float = 0.01


FULL_MODES = ("full", "full_tracking", "full-tracking")


class _SolveHistory:
    """Per-state history of the Jacobi solves: the sweeps the next solve is
    expected to need.

    dxtb_jacobi_eigh runs that many sweeps in its split form (grid-wide
    launches over many workgroups per matrix, for a few large matrices) and
    finishes what is left one workgroup per matrix; a guess above the need
    costs a few empty launches, one below it whole one-workgroup sweeps.  The
    guess is the largest count over the last HISTORY solves of windows of
    the same size class (up to 8, up to 64, more matrices) whose counts have
    reached the host (copied back without blocking), plus a margin; a class
    without history yet takes the largest count of any class, and FRESH
    applies where orbitals restart from X (first solve, admitted molecules).
    Measured on the config-2 batch: with the last solve's count alone 40 % of
    the large windows fell short (10 % slower); with one history for all
    sizes the small windows of a pool's tail inherited the large windows'
    counts and ran four to six empty sweeps per solve (10 % slower, and twice
    the run-to-run spread).  Above order 512 a missing split sweep costs
    20-140 ms (one workgroup per matrix) and a spare one about a millisecond,
    so the margin and FRESH grow by two."""

    FRESH = 12
    HISTORY = 8
    PENDING = 8  # sweep counts in flight (their copies to the host not yet complete)

    def __init__(self, n: int = 0):
        # a missing split sweep costs a one-workgroup sweep: 20-140 ms above
        # order 512, seconds above 1536
        self.margin = 1 if n <= 512 else (3 if n <= 1536 else 4)
        self.fresh_sweeps = self.FRESH + self.margin - 1 + (6 if n > 1536 else 0)
        self.recent: dict[int, list[int]] = {}  # size class -> last counts
        self.bump = False
        self.sw = None
        # counts copied back without blocking: (event, pinned buffer, rows,
        # size class) in launch order; a solve's count is used once its copy
        # has completed, however far the host runs ahead of the device
        self.pending: collections.deque = collections.deque()
        self.free: list = []  # pinned buffers not in flight

    @staticmethod
    def size_class(nb: int) -> int:
        return 0 if nb <= 8 else (1 if nb <= 64 else 2)

    def _drain(self) -> None:
        while self.pending and self.pending[0][0].query():
            _, buf, n, cls = self.pending.popleft()
            h = self.recent.get(cls, [])
            self.recent[cls] = (h + [int(buf[:n].max())])[-self.HISTORY:]
            self.free.append(buf)

    def predict(self, nb: int) -> int:
        if not torch.cuda.is_current_stream_capturing():  # no event queries inside a capture
            self._drain()
        hist = self.recent.get(self.size_class(nb)) or [v for h in self.recent.values() for v in h]
        pred = max(hist) + self.margin if hist else self.fresh_sweeps
        if self.bump:
            self.bump = False
            return max(pred, self.fresh_sweeps)
        return pred

    def fresh(self) -> None:
        self.bump = True

    def counts(self, nb: int, dev) -> torch.Tensor:
        if self.sw is None or self.sw.numel() < nb:
            self.sw = torch.zeros(nb, dtype=torch.int32, device=dev)
        return self.sw[:nb]

    def record(self, nb: int, dev) -> None:
        if torch.cuda.is_current_stream_capturing():
            return
        if len(self.pending) >= self.PENDING:
            return  # the host is far ahead: this count is not kept
        i = next((i for i, b in enumerate(self.free) if b.numel() >= nb), None)
        buf = torch.zeros(max(nb, 8), dtype=torch.int32, pin_memory=True) if i is None else self.free.pop(i)
        buf[:nb].copy_(self.sw[:nb], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(dev))
        self.pending.append((ev, buf, nb, self.size_class(nb)))


# eigensolves of the overlap matrix (orthogonaliser): from the unit matrix,
# to the rounding floor (off-diagonal 1e-15), sweeps in the split form
S_JACOBI_TOL = 1.0e-15
S_SPLIT_SWEEPS = 16


def _orthogonaliser(spad: torch.Tensor, nsize: torch.Tensor, variant: int = 3) -> torch.Tensor:
    """X with X^T S X = 1 from the eigendecomposition S = V diag(l) V^T by the
    hand-written Jacobi solver (canonical orthogonalisation X = V diag(l)^-1/2).
    The reference orthogonalises with the Cholesky factor (X = L^-T,
    exlibs/xitorch/_impls/linalg/symeig.py:38-75); any X with X^T S X = 1
    spans the same generalized eigenproblem, so orbitals, density and
    energies are the same (to rounding; degenerate orbitals may mix).
    Padding rows/columns of S carry 1 on the diagonal and keep X = 1 there."""
    lib = _lib.load()
    B, n, _ = spad.shape
    dev = spad.device
    A = spad.clone()
    VT = torch.empty_like(A)
    ev = torch.empty((B, n), dtype=torch.float64, device=dev)
    work = torch.empty(max(lib.dxtb_jacobi_work_size(B, n) // 8, 1), dtype=torch.float64, device=dev)
    var = 2 if n <= JACOBI_ONE_WG_MAX else variant
    split = S_SPLIT_SWEEPS + (6 if n > 1536 else 0)
    max_sweeps = 60 if n <= 1536 else split  # one-workgroup sweeps take seconds above order 1536
    rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(nsize), _lib.ptr(A), _lib.ptr(VT), _lib.ptr(ev), _lib.ptr(work), var,
                              max_sweeps, split, S_JACOBI_TOL, None, None, _lib.stream_ptr(dev))
    _lib.check(rc, "dxtb_jacobi_eigh (overlap)")
    live = torch.arange(n, device=dev)[None, :] < nsize[:, None]
    scale = torch.where(live, ev.clamp(min=torch.finfo(torch.float64).tiny).rsqrt(), 1.0)
    # X[i, j] = V[i, j] / sqrt(l_j) = VT[j, i] / sqrt(l_j)
    return (VT.transpose(1, 2) * scale.unsqueeze(1)).contiguous()


def alpha_beta(nel: np.ndarray, uhf=None) -> np.ndarray:
    """Alpha and beta electron counts (B, 2) of molecules with nel electrons
    and uhf unpaired electrons (``wavefunction/filling.py:180-259``): uhf must
    not exceed the rounded electron count and must have its parity; without
    it (or where it has the other parity) the parity alone decides; the paired
    electrons are split equally and the unpaired ones are alpha."""
    nel = np.atleast_1d(np.asarray(nel, dtype=np.float64))
    nint = np.floor(nel + 0.5)
    if uhf is None:
        uhf = np.zeros_like(nel)
    else:
        if not torch.is_tensor(uhf) and not np.isscalar(uhf):
            # a sequence may leave single molecules unset (None): their parity decides
            uhf = [np.nan if u is None else u for u in uhf]
        uhf = np.asarray(uhf.cpu() if torch.is_tensor(uhf) else uhf, dtype=np.float64)
        uhf = np.broadcast_to(uhf, nel.shape) if uhf.ndim == 0 else uhf
        if uhf.shape == nel.shape:
            uhf = np.where(np.isnan(uhf), np.remainder(nint, 2), uhf)
        if uhf.shape != nel.shape:
            raise RuntimeError(f"Shape mismatch for unpaired electrons ({uhf.shape}) and number of electrons "
                               f"({nel.shape}).")
        if (uhf > nint).any():
            raise ValueError(f"Number of unpaired electrons ({uhf}) larger than number of electrons ({nel}).")
        if (np.remainder(uhf, 2) != np.remainder(nint, 2)).any():
            raise ValueError(f"Odd (even) number of unpaired electrons ({uhf}) but even (odd) number of "
                             f"electrons ({nel}) given.")
    nuhf = np.where(np.remainder(uhf, 2) == np.remainder(nint, 2), uhf, np.remainder(nint, 2))
    diff = np.minimum(nuhf, nel)
    nb = (nel - diff) / 2.0
    return np.stack([nb + diff, nb], 1).astype(np.float64)


class GeometryState:
    """Everything that depends on the geometry but not on the density."""

    def __init__(self, bb: BatchBasis, positions: torch.Tensor, charges=None, opts: SCFOptions | None = None,
                 solvation=None, spins=None):
        self.bb = bb
        self.opts = opts or SCFOptions()
        self.gfn = bb.gfn
        self.mp = param.kind(bb.gfn) == "gfn2"
        dev = positions.device
        if dev.type != "cuda":
            raise RuntimeError(f"positions must be on a GPU device, not {dev} (there is no CPU path)")
        if positions.numel() != bb.nbatch * bb.nat_max * 3:
            raise ValueError(f"positions hold {positions.numel()} values, the basis needs "
                             f"{bb.nbatch} x {bb.nat_max} x 3")
        self.dev = dev
        self.lib = _lib.load()
        self.dp = device_params(bb.gfn, dev)
        self.bs = _lib.basis_struct(bb)
        # the kernels read FP64 coordinates
        self.pos = positions.detach().to(torch.float64).reshape(bb.nbatch, bb.nat_max, 3).contiguous()
        d = bb.dev
        B, na, ns, n = bb.nbatch, bb.nat_max, bb.nsh_max, bb.nao_max
        f64 = dict(dtype=torch.float64, device=dev)
        self.cn = torch.empty((B, na), **f64)
        self.fsh = torch.empty((B, ns, ns), **f64)
        self.dfsh = torch.empty((B, ns, ns), **f64)
        self.hfsh = torch.empty((B, ns, ns), **f64)
        self.erep = torch.empty((B, na), **f64)
        self.mrad = torch.empty((B, na), **f64) if self.mp else None
        st = _lib.stream_ptr(dev)
        rc = self.lib.dxtb_setup(self.bs, self.dp.struct, _lib.ptr(d["numbers_i"]), _lib.ptr(self.pos),
                                 _lib.ptr(d["nsh"]), _lib.ptr(self.cn), _lib.ptr(self.fsh), _lib.ptr(self.dfsh),
                                 _lib.ptr(self.hfsh), _lib.ptr(self.erep), _lib.ptr(self.mrad), st)
        _lib.check(rc, "dxtb_setup")
        self.S, self.D, self.Q, self.H0 = multipole_forward(bb, self.pos, self.mp, fsh=self.fsh)
        self._setup_solvation(solvation, st)
        # orthogonaliser (X^T S X = 1) with unit diagonal on padding
        spad = self.S.clone()
        pad = torch.arange(n, device=dev)[None, :] >= d["nao"][:, None]
        spad.diagonal(dim1=-2, dim2=-1).copy_(torch.where(pad, 1.0, spad.diagonal(dim1=-2, dim2=-1)))
        if n <= self.lib.dxtb_jacobi_max_n():
            self.X = _orthogonaliser(spad, d["nao"], self.opts.jacobi_variant)
        else:
            L = torch.linalg.cholesky(spad)
            eye = torch.eye(n, dtype=spad.dtype, device=dev)
            self.X = torch.linalg.solve_triangular(L, eye.expand(B, n, n), upper=False).transpose(-1, -2).contiguous()
        # reference occupations and electron counts (scf/iterator.py:147-195)
        tab_occ = _refocc_table(bb.gfn)
        n0 = np.zeros((B, n))
        nel = np.zeros(B)
        for b in range(B):
            u = bb.sh_ush[b, : bb.nsh[b]]
            occ = tab_occ[u]
            nel[b] = occ.sum()
            # per AO through the AO -> shell map (any AO order)
            n0[b, : bb.nao[b]] = (occ / (2 * bb.sh_l[b, : bb.nsh[b]] + 1))[bb.ao_shell[b, : bb.nao[b]]]
        chg = np.zeros(B) if charges is None else np.asarray(
            charges.cpu() if torch.is_tensor(charges) else charges, dtype=np.float64).reshape(B)
        nel = nel - chg
        self.nab = to_device(alpha_beta(nel, spins), dev)
        self.n0 = to_device(n0.astype(np.float64), dev)
        self.kt = self.opts.etemp * param.KELVIN2AU
        self.nvec = ns + (9 * na if self.mp else 0)
        # inner-product weights of the potential vector: the reference iterates
        # the orbital-resolved monopole potential (one entry per AO, a shell's
        # value repeated 2l+1 times; components/interactions/container.py:
        # 110-141), so in the shell-resolved vector every dot product and norm
        # weights a shell by 2l+1; the atomic dipole/quadrupole entries weigh 1
        w = np.zeros((B, self.nvec))
        for b in range(B):
            k = int(bb.nsh[b])
            w[b, :k] = 2 * bb.sh_l[b, :k] + 1
            if self.mp:
                w[b, ns : ns + 3 * int(bb.nat[b])] = 1.0
                w[b, ns + 3 * na : ns + 3 * na + 6 * int(bb.nat[b])] = 1.0
        self.wvec = to_device(w, dev)
        self.valid_sh = (torch.arange(ns, device=dev)[None, :] < d["nsh"][:, None]).double()
        # warm start of the Jacobi solver: the previous orbitals are X U with U
        # orthogonal (None: identity), U <- U V^T after every solve.  Each V is
        # orthogonal to ~1e-13 only, and the products drift linearly (3e-11
        # after 170 iterations of a 628-function molecule: 2e-8 surplus
        # electrons, 1e-8 Eh), so U is re-orthogonalised at every solve
        self.U = None
        # the previous iteration's orbitals C = X U (None: X), kept so that
        # the next solve starts from them without recomputing X U
        self.Corb = None
        self._jwork = None  # Jacobi workspace (a subset keeps the larger buffer)
        self._jtol = None  # per-molecule Jacobi tolerance of the next solve (None: JACOBI_TOL)
        self._jhist = _SolveHistory(bb.nao_max)  # shared by windows and subsets of this state

    def _setup_solvation(self, solvation, stream) -> None:
        """Generalized Born solvation per molecule (keps = 0: none): Born
        radii and dB/dpsi, and the ALPB shape descriptor."""
        bb, d = self.bb, self.bb.dev
        B, na = bb.nbatch, bb.nat_max
        f64 = dict(dtype=torch.float64, device=self.dev)
        gbs = gbsolv.per_molecule(solvation, B)
        self.any_gb = any(g is not None for g in gbs)
        self.gb_par = gbsolv.gb_params(gbs)
        self.gb_keps = to_device(np.array([g.keps if g else 0.0 for g in gbs]), self.dev)
        self.gb_kind = to_device(np.array([gbsolv.KERNELS[g.kernel] if g else 0 for g in gbs], dtype=np.int32),
                                 self.dev)
        self.gb_alpbet = to_device(np.array([g.alpbet if g else 0.0 for g in gbs]), self.dev)
        self._alpb = any(g is not None and g.alpb for g in gbs)
        self.born = torch.ones((B, na), **f64)
        self.dborn = torch.zeros((B, na), **f64)
        self.gb_adet = torch.zeros(B, **f64)
        if not self.any_gb:
            return
        gbsolv.check_elements(bb.numbers, gbs)
        rc = self.lib.dxtb_gb_born(B, na, _lib.ptr(d["nat"]), _lib.ptr(d["numbers_i"]), _lib.ptr(self.pos),
                                   _lib.ptr(self.gb_keps), _lib.ptr(gbsolv.vdw_device(self.dev)),
                                   ctypes.byref(self.gb_par), _lib.ptr(self.born), _lib.ptr(self.dborn), stream)
        _lib.check(rc, "dxtb_gb_born")
        if self._alpb:
            self.gb_adet = torch.nan_to_num(self._adet(self.pos))

    def _adet(self, pos):
        z = self.bb.dev["numbers_i"].long()
        return gbsolv.adet(pos, gbsolv.vdw_device(self.dev)[z], z != 0)

    # ---- culling: a state restricted to some molecules of the batch --------
    _PER_MOLECULE = ("pos", "cn", "fsh", "dfsh", "hfsh", "erep", "mrad", "S", "D", "Q", "H0", "X", "nab", "n0",
                     "valid_sh", "U", "Corb", "gb_keps", "gb_kind", "gb_alpbet", "gb_adet", "born", "dborn", "wvec")
    _GLOBAL_DEV = ("ush_l", "ush_np", "ush_alpha", "ush_coeff", "item_mol", "item_pair")

    _BB_HOST = ("numbers", "nat", "nsh", "nao", "sh_atom", "sh_ush", "sh_ao", "sh_l", "ao_atom", "ao_shell")

    def subset(self, keep: torch.Tensor, bb: BatchBasis | None = None) -> "GeometryState":
        """Copy holding only the molecules ``keep`` (culling of converged
        systems, as the reference's batched SCF does in
        scf/unrolling/default.py:241-322).  The integral work lists are not
        carried over (the SCF iteration does not use them) unless ``bb``, the
        basis of exactly these molecules built with the same padded sizes, is
        given (the force kernels need them)."""
        import copy

        new = copy.copy(self)
        if torch.is_tensor(keep):
            kc = keep.cpu().numpy()
        else:
            kc = np.asarray(keep, dtype=np.int64)
            keep = to_device(kc, self.dev)
        for name in self._PER_MOLECULE:
            t = getattr(self, name)
            if t is not None:
                setattr(new, name, t.index_select(0, keep).contiguous())
        if bb is None:
            ob = self.bb
            dev = {k: (v if k in self._GLOBAL_DEV else v.index_select(0, keep).contiguous())
                   for k, v in ob.dev.items()}
            bb = BatchBasis(
                ob.gfn, ob.numbers[kc], ob.nat[kc], ob.nsh[kc], ob.nao[kc], ob.nat_max, ob.nsh_max, ob.nao_max,
                ob.sh_atom[kc], ob.sh_ush[kc], ob.sh_ao[kc], ob.sh_l[kc], ob.ao_atom[kc], ob.ao_shell[kc],
                ob.item_mol[:0], ob.item_pair[:0], np.zeros(10, dtype=np.int32), dev, ob.ao_order,
            )
        elif (bb.nat_max, bb.nsh_max, bb.nao_max) != (self.bb.nat_max, self.bb.nsh_max, self.bb.nao_max):
            raise ValueError("subset basis must have the padded sizes of the state")
        new.bb = bb
        new.bs = _lib.basis_struct(bb)
        new._jtol = None
        return new

    def window(self, hi: int) -> "GeometryState":
        """The first ``hi`` molecules as views (no copy): work done on the
        window, including the orbital update of the Jacobi warm start, lands
        in this state's storage."""
        import copy

        new = copy.copy(self)
        for name in self._PER_MOLECULE:
            t = getattr(self, name)
            if t is not None:
                setattr(new, name, t[:hi])
        ob = self.bb
        dev = {k: (v if k in self._GLOBAL_DEV else v[:hi]) for k, v in ob.dev.items()}
        host = {k: getattr(ob, k)[:hi] for k in self._BB_HOST}
        new.bb = BatchBasis(ob.gfn, host["numbers"], host["nat"], host["nsh"], host["nao"], ob.nat_max, ob.nsh_max,
                            ob.nao_max, host["sh_atom"], host["sh_ush"], host["sh_ao"], host["sh_l"], host["ao_atom"],
                            host["ao_shell"], ob.item_mol[:0], ob.item_pair[:0], np.zeros(10, dtype=np.int32), dev,
                            ob.ao_order)
        new.bs = _lib.basis_struct(new.bb)
        new._jtol = None
        return new

    def scatter(self, rows: torch.Tensor, other: "GeometryState") -> None:
        """Write the molecules of ``other`` (same padded sizes) into the rows
        ``rows`` of this state, in place (admission into serving slots).  The
        warm-start orbitals of those rows restart from X."""
        ob, nb = self.bb, other.bb
        if (ob.nat_max, ob.nsh_max, ob.nao_max) != (nb.nat_max, nb.nsh_max, nb.nao_max) or ob.gfn != nb.gfn:
            raise ValueError("scatter needs states of the same parametrization and padded sizes")
        if torch.is_tensor(rows):
            rc = rows.cpu().numpy()
            rows = rows.to(self.dev)
        else:
            rc = np.asarray(rows, dtype=np.int64)
            rows = to_device(rc, self.dev)
        for name in self._PER_MOLECULE:
            if name in ("U", "Corb"):
                continue
            t = getattr(self, name)
            if t is not None:
                t.index_copy_(0, rows, getattr(other, name))
        n = ob.nao_max
        if self.U is None:
            self.U = torch.eye(n, dtype=torch.float64, device=self.dev).expand(ob.nbatch, n, n).contiguous()
        self.U.index_copy_(0, rows, torch.eye(n, dtype=torch.float64, device=self.dev).expand(len(rc), n, n))
        if self.Corb is None:
            self.Corb = self.X.clone()
        self.Corb.index_copy_(0, rows, other.X)
        self._jhist.fresh()
        self.any_gb = self.any_gb or other.any_gb
        self._alpb = self._alpb or other._alpb
        if other.any_gb:
            self.gb_par = other.gb_par
        for k, v in ob.dev.items():
            if k not in self._GLOBAL_DEV:
                v.index_copy_(0, rows, nb.dev[k])
        for k in self._BB_HOST:
            getattr(ob, k)[rc] = getattr(nb, k)

    def move(self, src, dst) -> None:
        """Copy the molecules of rows ``src`` into rows ``dst`` (disjoint), in
        place: compaction of serving slots, so the iterated window shrinks to
        the occupied rows.  Views of this state (windows) see the moved rows."""
        sc = np.asarray(src, dtype=np.int64)
        dc = np.asarray(dst, dtype=np.int64)
        if len(sc) != len(dc) or len(np.intersect1d(sc, dc)):
            raise ValueError("move needs disjoint row sets of equal length")
        if len(sc) == 0:
            return
        st, dt = to_device(sc, self.dev), to_device(dc, self.dev)
        for name in self._PER_MOLECULE:
            t = getattr(self, name)
            if t is not None:
                t.index_copy_(0, dt, t.index_select(0, st))
        for k, v in self.bb.dev.items():
            if k not in self._GLOBAL_DEV:
                v.index_copy_(0, dt, v.index_select(0, st))
        for k in self._BB_HOST:
            a = getattr(self.bb, k)
            a[dc] = a[sc]
        if self._jtol is not None:
            self._jtol.index_copy_(0, dt, self._jtol.index_select(0, st))

    # ---- one application of the SCF map -----------------------------------
    def unpack(self, x):
        B, na, ns = self.bb.nbatch, self.bb.nat_max, self.bb.nsh_max
        vsh = x[:, :ns].contiguous()
        if not self.mp:
            return vsh, None, None
        vdp = x[:, ns : ns + 3 * na].reshape(B, na, 3).contiguous()
        vqp = x[:, ns + 3 * na :].reshape(B, na, 6).contiguous()
        return vsh, vdp, vqp

    def fock(self, x):
        vsh, vdp, vqp = self.unpack(x)
        bb, d = self.bb, self.bb.dev
        zero_at = torch.zeros((bb.nbatch, bb.nat_max), dtype=torch.float64, device=self.dev)
        F = torch.empty_like(self.S)
        rc = self.lib.dxtb_fock(self.bs, _lib.ptr(d["nao"]), _lib.ptr(d["ao_shell_i"]), _lib.ptr(d["ao_atom_i"]),
                                _lib.ptr(self.pos), _lib.ptr(self.H0), _lib.ptr(self.S), _lib.ptr(self.D),
                                _lib.ptr(self.Q), _lib.ptr(vsh), _lib.ptr(zero_at), _lib.ptr(vdp), _lib.ptr(vqp),
                                PAD_DIAG, _lib.ptr(F), _lib.stream_ptr(self.dev))
        _lib.check(rc, "dxtb_fock")
        return F

    def solve(self, F):
        """Orbitals, occupations and density of a Fock matrix.

        nao <= 4096: the hand-written Jacobi solver on A = C^T F C with
        C = X U the previous iteration's orbitals (S-orthonormal, so A is the
        Fock matrix in the old orbital basis and nearly diagonal; the first
        iteration uses U = 1, C = X = L^-T); the new orbitals are C V and U
        becomes U V.  The products run on the matrix cores
        (dxtb_gemm_batched), each over its molecule's own basis size.  Larger
        bases (and eigensolver="device"): X^T F X and the device eigh
        (rocSOLVER)."""
        n = self.bb.nao_max
        nsz = self.bb.dev["nao"]
        gemm = _lib.gemm
        if self.opts.eigensolver == "jacobi" and n <= self.lib.dxtb_jacobi_max_n():
            Cb = self.X if self.Corb is None else self.Corb
            A = gemm(Cb, gemm(F, Cb, nsize=nsz), transa=True, nsize=nsz, sym=True)
            V = torch.empty_like(A)
            e = torch.empty((self.bb.nbatch, n), dtype=torch.float64, device=self.dev)
            nbytes = self.lib.dxtb_jacobi_work_size(self.bb.nbatch, n)
            if self._jwork is None or self._jwork.numel() * 8 < nbytes:
                self._jwork = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=self.dev)
            # one workgroup per matrix up to JACOBI_ONE_WG_MAX, else SCFOptions.jacobi_variant
            variant = 2 if n <= JACOBI_ONE_WG_MAX else self.opts.jacobi_variant
            nb = self.bb.nbatch
            # a captured graph replays over many iterations and admissions: FRESH
            fresh = self.U is None or torch.cuda.is_current_stream_capturing()
            split = self._jhist.fresh_sweeps if fresh else self._jhist.predict(nb)
            sw = self._jhist.counts(nb, self.dev)
            prof = JACOBI_PROFILE
            if prof is not None and torch.cuda.is_current_stream_capturing():
                prof = None  # timing events cannot be read back from a captured graph
            if prof is not None:
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record(torch.cuda.current_stream(self.dev))
            # above order 1536 a one-workgroup sweep takes seconds: every sweep runs split
            max_sweeps = 50 if n <= 1536 else split
            rc = self.lib.dxtb_jacobi_eigh(nb, n, _lib.ptr(nsz), _lib.ptr(A), _lib.ptr(V),
                                           _lib.ptr(e), _lib.ptr(self._jwork), variant, max_sweeps, split, JACOBI_TOL,
                                           _lib.ptr(self._jtol), _lib.ptr(sw), _lib.stream_ptr(self.dev))
            _lib.check(rc, "dxtb_jacobi_eigh")
            self._jhist.record(nb, self.dev)
            if prof is not None:
                ev[1].record(torch.cuda.current_stream(self.dev))
                prof.append((ev[0], ev[1], n, np.array(self.bb.nao), sw.clone(), variant, split))
            # V holds the eigenvectors as rows: C V^T and U V^T
            Un = V.transpose(1, 2).contiguous() if self.U is None else gemm(self.U, V, transb=True, nsize=nsz,
                                                                             pad=1.0)
            # one Newton-Schulz step every solve, U (3 - U^T U) / 2: U^T U = 1 + E
            # becomes 1 + O(E^2); the orbitals are built from the re-orthogonalised
            # U.  (Every solve, not every k-th: a molecule's arithmetic then does
            # not depend on when it joined a batch.)
            G = gemm(Un, Un, transa=True, nsize=nsz, sym=True, pad=1.0)
            Un = gemm(Un, G, alpha=-0.5, beta=1.5, cin=Un, nsize=nsz, pad=1.0)
            C = gemm(self.X, Un, nsize=nsz, pad=1.0)
            if self.U is None:
                self.U = Un
                self.Corb = C
            else:
                self.U.copy_(Un)  # in place: a window writes through to its parent state
                self.Corb.copy_(C)
        else:
            Fp = self.X.transpose(-1, -2) @ F @ self.X
            side = torch.cuda.current_stream(self.dev) != torch.cuda.default_stream(self.dev)
            if side:  # the device solver is not reliably ordered with a side stream
                torch.cuda.synchronize(self.dev)
            e, Cp = torch.linalg.eigh(Fp)
            if side:
                torch.cuda.synchronize(self.dev)
            C = gemm(self.X, Cp.contiguous(), nsize=nsz, pad=1.0)
        occ = torch.empty((self.bb.nbatch, 2, self.bb.nao_max), dtype=torch.float64, device=self.dev)
        ts = torch.empty(self.bb.nbatch, dtype=torch.float64, device=self.dev)
        e = e.contiguous()
        rc = self.lib.dxtb_fermi(self.bb.nbatch, self.bb.nao_max, _lib.ptr(nsz), _lib.ptr(e),
                                 _lib.ptr(self.nab), self.kt, _lib.ptr(occ), _lib.ptr(ts),
                                 _lib.stream_ptr(self.dev))
        _lib.check(rc, "dxtb_fermi")
        f = occ.sum(1)
        P = gemm(C, C, transb=True, scale=f, nsize=nsz, sym=True, pad=0.0)
        return P, e, C, occ, ts

    def charges(self, P):
        bb, d = self.bb, self.bb.dev
        B, na, ns, n = bb.nbatch, bb.nat_max, bb.nsh_max, bb.nao_max
        f64 = dict(dtype=torch.float64, device=self.dev)
        qsh = torch.empty((B, ns), **f64)
        qat = torch.empty((B, na), **f64)
        eh0 = torch.empty((B, na), **f64)
        qdp = torch.empty((B, na, 3), **f64) if self.mp else None
        qqp = torch.empty((B, na, 6), **f64) if self.mp else None
        work = torch.empty(max(self.lib.dxtb_charges_work_size(B, n) // 8, 1), **f64)
        rc = self.lib.dxtb_charges(self.bs, _lib.ptr(d["nao"]), _lib.ptr(d["nat"]), _lib.ptr(d["at_sh0"]),
                                   _lib.ptr(d["at_nsh"]), _lib.ptr(self.n0), _lib.ptr(P), _lib.ptr(self.S),
                                   _lib.ptr(self.H0), _lib.ptr(self.D), _lib.ptr(self.Q), _lib.ptr(qsh),
                                   _lib.ptr(qat), _lib.ptr(qdp), _lib.ptr(qqp), _lib.ptr(eh0), _lib.ptr(work),
                                   _lib.stream_ptr(self.dev))
        _lib.check(rc, "dxtb_charges")
        return qsh, qat, qdp, qqp, eh0

    def potential(self, qsh, qat, qdp, qqp):
        bb, d = self.bb, self.bb.dev
        B, na, ns = bb.nbatch, bb.nat_max, bb.nsh_max
        f64 = dict(dtype=torch.float64, device=self.dev)
        vsh = torch.empty((B, ns), **f64)
        vat = torch.empty((B, na), **f64)
        esh = torch.empty((B, ns), **f64)
        eat = torch.empty((B, na), **f64)
        vdp = torch.empty((B, na, 3), **f64) if self.mp else None
        vqp = torch.empty((B, na, 6), **f64) if self.mp else None
        rc = self.lib.dxtb_potential(self.bs, self.dp.struct, _lib.ptr(d["numbers_i"]), _lib.ptr(self.pos),
                                     _lib.ptr(d["nsh"]), _lib.ptr(self.mrad), _lib.ptr(qsh), _lib.ptr(qat),
                                     _lib.ptr(qdp), _lib.ptr(qqp), _lib.ptr(vsh), _lib.ptr(vat), _lib.ptr(vdp),
                                     _lib.ptr(vqp), _lib.ptr(esh), _lib.ptr(eat), _lib.stream_ptr(self.dev))
        _lib.check(rc, "dxtb_potential")
        if self.any_gb:
            rc = self.lib.dxtb_gb_potential(B, na, _lib.ptr(d["nat"]), _lib.ptr(self.pos), _lib.ptr(self.gb_keps),
                                            _lib.ptr(self.gb_kind), _lib.ptr(self.born), _lib.ptr(qat), _lib.ptr(vat),
                                            _lib.ptr(eat), _lib.stream_ptr(self.dev))
            _lib.check(rc, "dxtb_gb_potential")
            if self._alpb:  # ALPB shape term: keps alpbet adet Q_tot on every atom
                c = (self.gb_keps * self.gb_alpbet * self.gb_adet * qat.sum(-1)).unsqueeze(-1)
                add = c * (d["numbers_i"] != 0)
                vat = vat + add
                eat = eat + 0.5 * qat * add
        # total shell potential (atom part spread to its shells)
        vtot = (vsh + torch.gather(vat, 1, d["sh_atom_l"])) * self.valid_sh
        return vtot, vdp, vqp, esh, eat

    def pack(self, vtot, vdp, vqp):
        if not self.mp:
            return vtot
        B = self.bb.nbatch
        return torch.cat([vtot, vdp.reshape(B, -1), vqp.reshape(B, -1)], 1)

    def g(self, x):
        P, *_ = self.solve(self.fock(x))
        qsh, qat, qdp, qqp, _ = self.charges(P)
        vtot, vdp, vqp, _, _ = self.potential(qsh, qat, qdp, qqp)
        return self.pack(vtot, vdp, vqp)


def wnorm(v: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """Norm of the potential vectors in the reference's (orbital-resolved)
    space: sqrt(sum_i w_i v_i^2)."""
    return torch.sqrt((w * v * v).sum(-1))


class DeviceBroyden:
    """Per-molecule broyden1 state of a batch on the device, advanced one
    iteration at a time by ``dxtb_broyden_step`` (xitorch's broyden1 with the
    weighted inner product of the orbital-resolved potential vector, see
    GeometryState.wvec): the point to evaluate next (xe), the current point
    and residual, the update history, and per molecule the control words
    (occupied, started, history length, iterations), the stop flag (1
    converged, 2 iteration limit) and the norm of the last step.  A molecule
    starts at xe = 0 (the SAD guess: zero potential)."""

    def __init__(self, B: int, L: int, H: int, device):
        f64 = dict(dtype=torch.float64, device=device)
        self.L, self.H = L, H
        self.xe = torch.zeros((B, L), **f64)
        self.x = torch.zeros((B, L), **f64)
        self.y = torch.zeros((B, L), **f64)
        # entries beyond a molecule's history length are never read
        self.hc = torch.empty((B, H, L), **f64)
        self.hd = torch.empty((B, H, L), **f64)
        self.ctl = torch.zeros((B, 4), dtype=torch.int32, device=device)
        self.flag = torch.zeros(B, dtype=torch.int32, device=device)
        self.dxn = torch.zeros(B, **f64)
        self.wstride = _lib.load().dxtb_broyden_work_size(1, L, H) // 8  # doubles per molecule
        self.work = torch.empty(B * self.wstride, **f64)

    _ROWS = ("xe", "x", "y", "hc", "hd", "ctl", "flag", "dxn")

    def start(self, rows=None) -> None:
        """Rows (None: all) restart as fresh molecules at the zero potential."""
        if rows is None:
            for t in (self.xe, self.x, self.y, self.dxn, self.flag):
                t.zero_()
            self.ctl.zero_()
            self.ctl[:, 0] = 1
            return
        for t in (self.xe, self.x, self.y, self.dxn, self.flag, self.ctl):
            t.index_fill_(0, rows, 0)
        self.ctl[:, 0].index_fill_(0, rows, 1)

    def subset(self, keep: torch.Tensor) -> None:
        for name in self._ROWS:
            setattr(self, name, getattr(self, name).index_select(0, keep).contiguous())
        self.work = self.work[: keep.numel() * self.wstride]

    def move(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """Copy rows src into rows dst (disjoint), in place."""
        for name in self._ROWS:
            t = getattr(self, name)
            t.index_copy_(0, dst, t.index_select(0, src))

    def step(self, gx: torch.Tensor, w: torch.Tensor, o: "SCFOptions", nb: int | None = None,
             kbound: int = 0) -> None:
        """One iteration of the first ``nb`` rows (default: all) with g(xe) = gx.
        ``kbound`` (0: the history capacity) bounds the rows' history lengths
        (e.g. the iterations they have run); it only sizes the launch."""
        nb = self.xe.shape[0] if nb is None else nb
        lib = _lib.load()
        rc = lib.dxtb_broyden_step(nb, self.L, self.H, _lib.ptr(w), _lib.ptr(gx.contiguous()), _lib.ptr(self.xe),
                                   _lib.ptr(self.x), _lib.ptr(self.y), _lib.ptr(self.hc), _lib.ptr(self.hd),
                                   _lib.ptr(self.ctl), _lib.ptr(self.flag), _lib.ptr(self.dxn), _lib.ptr(self.work),
                                   o.alpha, o.damp, o.x_tol, o.f_tol, o.maxiter, int(kbound),
                                   _lib.stream_ptr(self.xe.device))
        _lib.check(rc, "dxtb_broyden_step")


@dataclass
class SCFResult:
    energy: torch.Tensor  # (B,) total energy
    eel: torch.Tensor  # electronic energy (H0 + interactions)
    ts: torch.Tensor  # -TS term
    erep: torch.Tensor
    iterations: np.ndarray
    converged: np.ndarray
    density: torch.Tensor
    emo: torch.Tensor
    coeffs: torch.Tensor
    occupation: torch.Tensor
    potential: tuple
    charges: tuple
    state: GeometryState


class BatchedFullMixer:
    """The reference's full-mode mixers, per molecule of a batch, on the
    potential vectors with the weighted inner product of the orbital-resolved
    vector (as DeviceBroyden).

    "anderson" (``scf/mixer/anderson.py:163-317``): the history holds the
    previous inputs x_j and residuals F_j = g(x_j) - x_j (newest first, up to
    ``generations``); with n of them, theta solves a theta = b,
    a_ij = <F0 - Fi, F0 - Fj> (diagonal times 1 + offset^2),
    b_i = <F0 - Fi, F0>, and the new input is x_bar + damp F_bar with
    x_bar = x0 + sum_j theta_j (x_j - x0), F_bar = F0 + sum_j theta_j (F_j - F0).
    The first step (the first ``generations`` with soft start) is simple
    mixing x0 + damp_init F0.  "simple" (``scf/mixer/simple.py``):
    x0 + damp F0.  Every molecule starts at the same step (one batch), so
    the step count is shared; rows a caller no longer iterates get theta = 0."""

    def __init__(self, o: SCFOptions, B: int, L: int, device, w: torch.Tensor):
        if o.mixer not in ("anderson", "simple"):
            raise ValueError(f"Unknown mixer '{o.mixer}' for the full-mode SCF.")
        self.kind = o.mixer
        self.damp, self.damp_init = o.damp, o.damp_init
        self.G, self.soft, self.offset = o.damp_generations, o.damp_soft_start, o.damp_diagonal_offset
        self.w = w
        self.step = 0
        self.X = torch.zeros((B, self.G, L), dtype=torch.float64, device=device)  # previous inputs
        self.F = torch.zeros((B, self.G, L), dtype=torch.float64, device=device)  # their residuals
        self.nh = 0  # history entries held

    def subset(self, keep: torch.Tensor, w: torch.Tensor) -> None:
        self.X, self.F, self.w = self.X[keep], self.F[keep], w

    @staticmethod
    def _solve(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        """a theta = b for a batch of small symmetric positive definite
        systems (B, n, n): Gaussian elimination without pivoting, in batched
        elementwise operations on the device."""
        a, b = a.clone(), b.clone()
        n = a.shape[-1]
        for k in range(n):
            f = a[:, k + 1 :, k] / a[:, k, k].unsqueeze(-1)
            a[:, k + 1 :, :] -= f.unsqueeze(-1) * a[:, k, :].unsqueeze(1)
            b[:, k + 1 :] -= f * b[:, k].unsqueeze(-1)
        t = torch.empty_like(b)
        for k in range(n - 1, -1, -1):
            t[:, k] = (b[:, k] - (a[:, k, k + 1 :] * t[:, k + 1 :]).sum(-1)) / a[:, k, k]
        return t

    def iter(self, x_new: torch.Tensor, x_old: torch.Tensor, active: torch.Tensor):
        """Mixed inputs and the residuals x_new - x_old (B, L)."""
        self.step += 1
        f0 = x_new - x_old
        if self.kind == "simple":
            return x_old + self.damp * f0, f0
        if self.step > self.G or (self.step > 1 and not self.soft):
            n = min(self.step - 1, self.G, self.nh)
            Xp, Fp = self.X[:, :n], self.F[:, :n]
            df = f0.unsqueeze(1) - Fp  # (B, n, L)
            dfw = df * self.w.unsqueeze(1)
            a = torch.bmm(dfw, df.transpose(1, 2))
            b = torch.bmm(dfw, f0.unsqueeze(-1)).squeeze(-1)
            if self.offset is not None:
                a = a + torch.diag_embed(torch.diagonal(a, dim1=-2, dim2=-1)) * self.offset**2
            eye = torch.eye(n, dtype=a.dtype, device=a.device).expand_as(a)
            act = active.view(-1, 1, 1)
            theta = self._solve(torch.where(act, a, eye), torch.where(act[:, :, 0], b, 0.0))
            if not bool(torch.isfinite(theta).all()):
                # as torch.linalg.solve in the reference mixer
                raise RuntimeError("Anderson mixing: singular history (linearly dependent residuals)")
            th = theta.unsqueeze(-1)
            xbar = x_old + (th * (Xp - x_old.unsqueeze(1))).sum(1)
            fbar = f0 + (th * (Fp - f0.unsqueeze(1))).sum(1)
            xmix = xbar + self.damp * fbar
        else:
            xmix = x_old + self.damp_init * f0
        # shift the history by one and put (x_old, F0) in front
        self.X = torch.cat([x_old.unsqueeze(1), self.X[:, : self.G - 1]], 1)
        self.F = torch.cat([f0.unsqueeze(1), self.F[:, : self.G - 1]], 1)
        self.nh = min(self.nh + 1, self.G)
        return xmix, f0


def run_scf(state: GeometryState, cull: float = 0.75) -> SCFResult:
    """Iterate all molecules of ``state`` to self-consistency.  The mixer runs
    on the device (DeviceBroyden); the host reads the stop flags one
    iteration behind the device (pinned copies, double buffered), so it never
    waits for the iteration it has just launched.  When the running
    molecules drop below ``cull`` of the current batch, the batch is
    compacted to them (stopped molecules are frozen on the device).
    ``state.opts.scf_mode`` picks the fixed-point form: implicit (broyden1)
    or full (Anderson / simple mixing)."""
    o = state.opts
    if o.scf_mode in FULL_MODES:
        return _run_scf_full(state, cull)
    if o.scf_mode not in ("implicit", "default"):
        raise ValueError(f"Unsupported scf_mode '{o.scf_mode}'.")
    B = state.bb.nbatch
    dev = state.dev
    f64 = dict(dtype=torch.float64, device=dev)
    x_full = torch.zeros((B, state.nvec), **f64)
    it_full = torch.zeros(B, dtype=torch.int32, device=dev)
    flag_full = torch.zeros(B, dtype=torch.int32, device=dev)
    idx = torch.arange(B, device=dev)  # full-batch index of every current row
    cur = state
    mix = DeviceBroyden(B, state.nvec, o.maxiter + 1, dev)
    mix.start()
    hflags = [torch.zeros(B, dtype=torch.int32, pin_memory=True) for _ in range(2)]
    events: list = [None, None]
    for it in range(o.maxiter + 1):
        if o.adaptive_eigensolver:
            started = mix.ctl[:, 1] != 0
            cur._jtol = torch.where(started, jacobi_tolerance(wnorm(mix.y, cur.wvec)), JACOBI_TOL_LOOSE).contiguous()
        mix.step(cur.g(mix.xe), cur.wvec, o, kbound=it + 1)  # a history grows by one entry per iteration
        nb = cur.bb.nbatch
        hflags[it % 2][:nb].copy_(mix.flag, non_blocking=True)
        events[it % 2] = torch.cuda.Event()
        events[it % 2].record(torch.cuda.current_stream(dev))
        prev = events[(it - 1) % 2]
        if prev is None:
            continue
        prev.synchronize()  # flags after the previous iteration
        fl = hflags[(it - 1) % 2][:nb].numpy()
        running = fl == 0
        n_run = int(running.sum())
        if n_run == 0:
            break
        if n_run < cull * nb:
            keep = torch.as_tensor(np.nonzero(running)[0], device=dev)
            drop = torch.as_tensor(np.nonzero(~running)[0], device=dev)
            # stopped rows are frozen: their state is final
            x_full.index_copy_(0, idx[drop], mix.x.index_select(0, drop))
            it_full.index_copy_(0, idx[drop], mix.ctl[:, 3].index_select(0, drop))
            flag_full.index_copy_(0, idx[drop], mix.flag.index_select(0, drop))
            cur = cur.subset(keep)
            mix.subset(keep)
            idx = idx[keep]
            events = [None, None]  # the staged flags refer to the old rows
    x_full.index_copy_(0, idx, mix.x)
    it_full.index_copy_(0, idx, mix.ctl[:, 3].contiguous())
    flag_full.index_copy_(0, idx, mix.flag)
    iters = it_full.cpu().numpy().astype(np.int64)
    conv = flag_full.cpu().numpy() == 1
    state._jtol = None  # the final solve of every molecule at the tight bar
    return finalize(state, x_full, iters, conv)


def _run_scf_full(state: GeometryState, cull: float) -> SCFResult:
    """Full-mode loop of the reference (``scf/unrolling/default.py:71-140``
    and its batched form ``:142-403``): the map's output for the guess is
    mixed with the guess, then each output with its input; a molecule is
    converged when its unmixed residual has (weighted) L2 norm < x_tol and
    max norm < x_tol_max, and its result is that unmixed output (the last
    one if it does not converge)."""
    o = state.opts
    B = state.bb.nbatch
    dev = state.dev
    x_full = torch.zeros((B, state.nvec), dtype=torch.float64, device=dev)
    idx = torch.arange(B, device=dev)
    cur = state
    active = torch.ones(B, dtype=torch.bool, device=dev)
    mix = BatchedFullMixer(o, B, state.nvec, dev, state.wvec)
    x0 = torch.zeros((B, state.nvec), dtype=torch.float64, device=dev)
    q, _ = mix.iter(cur.g(x0), x0, active)
    iters = np.zeros(B, dtype=np.int64)
    conv = np.zeros(B, dtype=bool)
    q_new = q
    for it in range(o.maxiter):
        q_new = cur.g(q)
        qmix, d = mix.iter(q_new, q, active)
        live = cur.wvec > 0
        dmax = torch.where(live, d.abs(), 0.0).amax(-1)
        done = active & (wnorm(d, cur.wvec) < o.x_tol) & (dmax < o.x_tol_max)
        a_host, d_host, rows = active.cpu().numpy(), done.cpu().numpy(), idx.cpu().numpy()
        iters[rows[a_host]] = it + 1
        conv[rows[d_host]] = True
        if d_host.any():
            x_full[idx[done]] = q_new[done]
        active = active & ~done
        q = torch.where(active.unsqueeze(-1), qmix, q)
        n_act = int(active.sum())
        if n_act == 0:
            break
        if n_act < cull * cur.bb.nbatch:
            keep = torch.nonzero(active).squeeze(-1)
            cur = cur.subset(keep)
            idx = idx[keep]
            q, q_new, active = q[keep], q_new[keep], active[keep]
            mix.subset(keep, cur.wvec)
    if bool(active.any()):
        x_full[idx[active]] = q_new[active]
    return finalize(state, x_full, iters, conv)


def finalize(state: GeometryState, x: torch.Tensor, iters: np.ndarray, conv: np.ndarray) -> SCFResult:
    """Energies and the density of every molecule of ``state`` at its final
    potential vector ``x`` (one more application of the SCF map)."""
    F = state.fock(x)
    P, e, C, occ, ts = state.solve(F)
    qsh, qat, qdp, qqp, eh0 = state.charges(P)
    vtot, vdp, vqp, esh, eat = state.potential(qsh, qat, qdp, qqp)
    eel = eh0.sum(-1) + esh.sum(-1) + eat.sum(-1)
    erep = state.erep.sum(-1)
    energy = erep + eel + ts
    return SCFResult(energy, eel, ts, erep, iters, conv, P, e, C, occ, (vtot, vdp, vqp), (qsh, qat, qdp, qqp), state)


def forces(res: SCFResult) -> torch.Tensor:
    """-dE/dR (nbatch, nat_max, 3) of the converged SCF (hand-written kernels).

    At self-consistency the free energy is stationary in the density, so the
    gradient is the explicit derivative at fixed P with the orthonormality
    constraint carried by the energy-weighted density W = C diag(f e) C^T."""
    st = res.state
    bb, d = st.bb, st.bb.dev
    lib = st.lib
    f = res.occupation.sum(1)
    C = res.coeffs.contiguous()
    W = _lib.gemm(C, C, transb=True, scale=(f * res.emo).contiguous(), nsize=d["nao"], sym=True, pad=0.0)
    grad = torch.zeros((bb.nbatch, bb.nat_max, 3), dtype=torch.float64, device=st.dev)
    vtot, vdp, vqp = res.potential
    vtot = vtot.contiguous()
    stream = _lib.stream_ptr(st.dev)
    rc = lib.dxtb_forces_multipole(st.bs, _lib.ptr(st.pos), _lib.ptr(res.density), _lib.ptr(W),
                                   _lib.ptr(st.fsh), _lib.ptr(vtot), _lib.ptr(vdp), _lib.ptr(vqp),
                                   _lib.ptr(grad), stream)
    _lib.check(rc, "dxtb_forces_multipole")
    qsh, qat, qdp, qqp = res.charges
    dEdh = torch.empty((bb.nbatch, bb.nsh_max), dtype=torch.float64, device=st.dev)
    dEdcn = torch.empty((bb.nbatch, bb.nat_max), dtype=torch.float64, device=st.dev)
    rc = lib.dxtb_forces_classical(st.bs, st.dp.struct, _lib.ptr(d["numbers_i"]), _lib.ptr(st.pos),
                                   _lib.ptr(d["nsh"]), _lib.ptr(d["at_sh0"]), _lib.ptr(d["at_nsh"]),
                                   _lib.ptr(res.density), _lib.ptr(st.S), _lib.ptr(st.cn), _lib.ptr(st.mrad),
                                   _lib.ptr(st.dfsh), _lib.ptr(st.hfsh), _lib.ptr(qsh), _lib.ptr(qat),
                                   _lib.ptr(qdp), _lib.ptr(qqp), _lib.ptr(dEdh), _lib.ptr(dEdcn), _lib.ptr(grad),
                                   stream)
    _lib.check(rc, "dxtb_forces_classical")
    if st.any_gb:
        rc = lib.dxtb_gb_forces(bb.nbatch, bb.nat_max, _lib.ptr(d["nat"]), _lib.ptr(d["numbers_i"]), _lib.ptr(st.pos),
                                _lib.ptr(st.gb_keps), _lib.ptr(st.gb_kind), _lib.ptr(gbsolv.vdw_device(st.dev)),
                                ctypes.byref(st.gb_par), _lib.ptr(st.born), _lib.ptr(st.dborn), _lib.ptr(qat),
                                _lib.ptr(grad), stream)
        _lib.check(rc, "dxtb_gb_forces")
        if st._alpb:  # d/dR of 1/2 keps alpbet adet Q_tot^2 (a 3x3 determinant per molecule)
            with torch.enable_grad():
                p = st.pos.detach().requires_grad_(True)
                e = (0.5 * st.gb_keps * st.gb_alpbet * qat.sum(-1) ** 2 * torch.nan_to_num(st._adet(p))).sum()
                (ga,) = torch.autograd.grad(e, p)
            grad = grad + ga
    return -grad


_REFOCC: dict = {}


def _refocc_table(gfn: str) -> np.ndarray:
    if gfn not in _REFOCC:
        from .basis import shell_table

        tab = shell_table(gfn)
        par = param.get(gfn)
        occ = np.zeros(len(tab.l))
        for u in range(len(tab.l)):
            z = int(tab.z[u])
            occ[u] = par.element(z)["refocc"][u - int(tab.first[z])]
        _REFOCC[gfn] = occ
    return _REFOCC[gfn]
