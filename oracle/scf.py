"""
Oracle: SCF single point, energy and variational forces (test infrastructure only).

Follows (paths relative to ``src/dxtb/_src``):

* reference occupation and electron count -- ``scf/iterator.py:147-195``,
  ``wavefunction/filling.py:180-259``
* Fock build -- ``scf/pure/conversions.py:255-297``
* generalized eigensolve (Cholesky, L^-1 F L^-T, eigh, back-transform) --
  ``exlibs/xitorch/_impls/linalg/symeig.py:38-75`` via ``scf/pure/ovlp_diag.py:36-62``
* Fermi smearing -- ``wavefunction/filling.py:428-742``: the occupations are
  the Fermi-Dirac distribution whose chemical potential reproduces the
  electron count of each spin channel; the oracle finds that root by
  bisection to machine precision
* density -- ``scf/utils.py:29-52``; charges and atomic multipoles --
  ``scf/pure/conversions.py:203-252``
* potentials -- ``components/interactions/base.py:135-181`` and the ES2/ES3/AES2
  modules (see ``oracle.coulomb``)
* fixed point -- ``scf/implicit/fixed_point.py:476-542``: root of
  ``y - g(y)`` on the potential vector with xitorch ``broyden1``
  (``exlibs/xitorch/_impls/optimize/root/rootsolver.py:42-261``,
  ``_jacobian.py:51-120, 250-316``), alpha = -0.5, damp = 0.5, stop when
  ``|dx| < x_tol`` and ``|y| < f_tol``
* full-mode fixed point (the reference's default ``scf_mode="full"``) --
  ``scf/unrolling/default.py:71-140``: the map's first output is mixed with
  the guess, then every output with the current input by the Anderson mixer
  (``scf/mixer/anderson.py:163-317``, the default: Broyden is not available
  there, ``scf/unrolling/base.py:81-98``) or the simple mixer
  (``scf/mixer/simple.py``); converged when the unmixed residual has
  L2 norm < x_tol and max norm < x_tol_max (``scf/mixer/base.py``), and the
  result is the unmixed output
* energies -- ``scf/base.py:510-619`` (equal partition of -TS) and
  ``calculators/types/energy.py:336-337``
* forces -- the reference differentiates the converged SCF with autograd
  (``calculators/types/autograd.py:82-193``).  The oracle uses the
  equivalent variational expression: at self-consistency the free energy is
  stationary, so ``dE/dR = d/dR [E(R; P) - Tr(W S(R))]`` at fixed density P
  and energy-weighted density W (the structure of the analytic GFN1
  gradient, ``xtb/gfn1.py:311-313``).  Pinned against central finite
  differences of the SCF energy.
"""

from __future__ import annotations

import torch

from . import coulomb, hamiltonian, integrals, params
from . import solvation as gbsolv
from .basis import Basis, build


class System:
    def __init__(self, numbers, positions: torch.Tensor, gfn: str, charge: float = 0.0, etemp: float = 300.0,
                 solvation: dict | None = None, spin: int | None = None):
        # solvation: generalized Born, {"eps": dielectric constant, "alpb": bool, "kernel": "p16" | "still"}
        self.solvation = solvation
        self.numbers = [int(z) for z in numbers if int(z) != 0]
        self.gfn = gfn
        self.positions = positions
        self.basis: Basis = build(self.numbers, gfn)
        self.charge = float(charge)
        self.kt = etemp * params.KELVIN2AU
        b = self.basis
        self.ao_shell = b.ao_shell()
        self.ao_atom = b.ao_atom()
        self.sh_atom = b.shell_atom()
        self.nat = len(self.numbers)
        self.nsh = b.nsh
        self.nao = b.nao
        ref = []
        for sh in b.shells:
            ref.append(params.elem(gfn, sh.z)["refocc"][sh.ush[1]])
        refocc = torch.tensor(ref, dtype=torch.float64)
        nsph = torch.tensor([2 * sh.l + 1 for sh in b.shells], dtype=torch.float64)
        self.n0 = (refocc / nsph)[self.ao_shell]
        nel = refocc.sum() - self.charge
        # alpha/beta split (filling.py:180-259): spin unpaired electrons if it
        # has the parity of the electron count, else the parity alone
        nint = torch.floor(nel + 0.5)
        par = torch.remainder(nint, 2)
        nuhf = par if spin is None or float(spin) % 2 != float(par) else torch.tensor(float(spin), dtype=torch.float64)
        diff = torch.minimum(nuhf, nel)
        nb = (nel - diff) / 2.0
        self.nab = torch.stack([nb + diff, nb])
        self.is_gfn2 = params.kind(gfn) == "gfn2"

    # ---- geometry-dependent parts (differentiable in positions) -----------
    def build(self, positions: torch.Tensor):
        b = self.basis
        s, d, q = integrals.multipole_matrices(b, positions)
        cn = hamiltonian.coordination_number(self.numbers, positions, self.gfn)
        h0 = hamiltonian.h0_matrix(b, positions, s, cn)
        gam = coulomb.es2_matrix(b, positions)
        hd, shell_es3 = coulomb.es3_hubbard_derivs(b)
        out = {"s": s, "h0": h0, "gam": gam, "hd": hd, "shell_es3": shell_es3}
        if self.is_gfn2:
            out["d"] = d
            out["q"] = q
            out["aes"] = coulomb.aes2_cache(b, positions)
        out["erep"] = coulomb.repulsion_energy(self.numbers, positions, self.gfn)
        if self.solvation is not None:
            sv = self.solvation
            out["gb"] = gbsolv.gb_matrix(self.numbers, positions, sv["eps"], alpb=sv.get("alpb", False),
                                            kernel=sv.get("kernel", "p16"))
        return out

    # ---- SCF maps ----------------------------------------------------------
    def charges(self, m, p):
        qorb = self.n0 - (p * m["s"]).sum(-1)
        qdp = qqp = None
        if self.is_gfn2:
            # -sum_k P_ik D[:, k, i] over i of atom A
            dorb = -(p.unsqueeze(0) * m["d"]).sum(-2).T  # (nao, 3)
            qorb_q = -(p.unsqueeze(0) * m["q"]).sum(-2).T  # (nao, 6)
            qdp = torch.zeros(self.nat, 3, dtype=p.dtype).index_add(0, self.ao_atom, dorb)
            qqp = torch.zeros(self.nat, 6, dtype=p.dtype).index_add(0, self.ao_atom, qorb_q)
        return qorb, qdp, qqp

    def reduce_sh(self, qorb):
        return torch.zeros(self.nsh, dtype=qorb.dtype).index_add(0, self.ao_shell, qorb)

    def reduce_at(self, qsh):
        return torch.zeros(self.nat, dtype=qsh.dtype).index_add(0, self.sh_atom, qsh)

    def potential(self, m, qorb, qdp, qqp):
        qsh = self.reduce_sh(qorb)
        qat = self.reduce_at(qsh)
        vsh = m["gam"] @ qsh
        vat = torch.zeros(self.nat, dtype=qorb.dtype)
        if m["shell_es3"]:
            vsh = vsh + m["hd"] * qsh * qsh
        else:
            vat = vat + m["hd"] * qat * qat
        vdp = vqp = None
        if self.is_gfn2:
            va, vdp, vqp = coulomb.aes2_potential(m["aes"], qat, qdp, qqp)
            vat = vat + va
        if "gb" in m:
            vat = vat + m["gb"] @ qat
        vsh = vsh + vat[self.sh_atom]
        return vsh[self.ao_shell], vdp, vqp

    def fock(self, m, vorb, vdp, vqp):
        f = m["h0"] - 0.5 * m["s"] * (vorb.unsqueeze(-1) + vorb.unsqueeze(-2))
        if self.is_gfn2:
            for mat, v in ((m["d"], vdp), (m["q"], vqp)):
                vo = v[self.ao_atom]  # (nao, k) ket-atom potential
                tmp = 0.5 * torch.einsum("kij,jk->ij", mat, vo)
                f = f - (tmp + tmp.T)
        return f

    def diagonalize(self, m, f):
        l = torch.linalg.cholesky(m["s"])
        linv = torch.linalg.inv(l)
        fp = linv @ f @ linv.T
        e, c = torch.linalg.eigh(0.5 * (fp + fp.T))
        return e, linv.T @ c

    def occupation(self, e):
        """Fermi-Dirac occupations (2, nao) for the alpha/beta counts: the
        chemical potential of each spin channel by bisection, both channels
        at once, until the bracket cannot shrink any further (adjacent
        doubles: the midpoint equals an end)."""
        kt = self.kt
        nab = self.nab.to(e.dtype)
        lo = torch.full_like(nab, float(e.min()) - 1.0)
        hi = torch.full_like(nab, float(e.max()) + 1.0)
        for _ in range(200):
            mu = 0.5 * (lo + hi)
            if bool(((mu == lo) | (mu == hi)).all()):
                break
            nn = torch.special.expit(-(e.unsqueeze(0) - mu.unsqueeze(1)) / kt).sum(-1)
            gt = nn > nab
            hi = torch.where(gt, mu, hi)
            lo = torch.where(gt, lo, mu)
        mu = 0.5 * (lo + hi)
        return torch.special.expit(-(e.unsqueeze(0) - mu.unsqueeze(1)) / kt)

    def solve_density(self, m, f):
        e, c = self.diagonalize(m, f)
        occ = self.occupation(e)
        ftot = occ.sum(0)
        p = (c * ftot) @ c.T
        return p, e, c, occ

    def pack(self, vorb, vdp, vqp):
        if self.is_gfn2:
            return torch.cat([vorb, vdp.reshape(-1), vqp.reshape(-1)])
        return vorb

    def unpack(self, x):
        if self.is_gfn2:
            n, a = self.nao, self.nat
            return x[:n], x[n : n + 3 * a].view(a, 3), x[n + 3 * a :].view(a, 6)
        return x, None, None

    def g(self, m, x):
        vorb, vdp, vqp = self.unpack(x)
        p, *_ = self.solve_density(m, self.fock(m, vorb, vdp, vqp))
        return self.pack(*self.potential(m, *self.charges(m, p)))

    def electronic_energy(self, m, p, qorb, qdp, qqp):
        e_h0 = (p * m["h0"]).sum()
        qsh = self.reduce_sh(qorb)
        qat = self.reduce_at(qsh)
        e = e_h0 + 0.5 * qsh @ m["gam"] @ qsh
        if m["shell_es3"]:
            e = e + (m["hd"] * qsh**3 / 3.0).sum()
        else:
            e = e + (m["hd"] * qat**3 / 3.0).sum()
        if self.is_gfn2:
            e = e + coulomb.aes2_energy(m["aes"], qat, qdp, qqp).sum()
        if "gb" in m:
            e = e + 0.5 * qat @ m["gb"] @ qat
        return e

    def free_energy(self, occ):
        eps = torch.finfo(torch.float64).eps
        o = torch.clamp(occ, min=eps)
        o1 = torch.clamp(1 - occ, min=eps)
        return (torch.log(o**o * o1**o1)).sum() * self.kt


def broyden1(fcn, x0, f_tol=1e-10, x_tol=1e-10, maxiter=300, alpha=-0.5, damp=0.5):
    """xitorch broyden1 without line search (rootsolver.py:42-261) for the root of fcn."""
    x = x0.clone()
    y = fcn(x)
    ynorm = y.norm()
    if ynorm <= f_tol * torch.finfo(y.dtype).eps:
        return x, 0, True
    cns, dns = [], []
    g0 = -alpha
    xprev, yprev = x, y

    def mv(v):
        r = g0 * v
        for c, d in zip(cns, dns):
            r = r + c * torch.dot(d, v)
        return r

    def rmv(v):
        r = g0 * v
        for c, d in zip(cns, dns):
            r = r + d * torch.dot(c, v)
        return r

    for it in range(maxiter):
        dx = -mv(y)
        dxn = dx.norm()
        xnew = x + damp * dx
        ynew = fcn(xnew)
        ynorm_new = ynew.norm()
        ddx, ddy = xnew - xprev, ynew - yprev
        v = rmv(ddx)
        cns.append(ddx - mv(ddy))
        dns.append(v / torch.dot(ddy, v))
        xprev, yprev = xnew, ynew
        if dxn < x_tol and ynorm_new < f_tol:
            return xnew, it + 1, True
        x, y = xnew, ynew
    return x, maxiter, False


class FullModeMixer:
    """The reference's full-mode mixers on one system's vector.

    ``kind="anderson"``: Eyert's Anderson mixing as written in
    ``scf/mixer/anderson.py:163-317`` -- history of the mixed inputs x_j and
    their residuals F_j = g(x_j) - x_j (newest first); with n = min(step - 1,
    generations) previous entries, theta solves a theta = b with
    a_ij = <F0 - Fi | F0 - Fj> (diagonal times 1 + offset^2) and
    b_i = <F0 - Fi | F0>; then x_bar = x0 + sum_j theta_j (x_j - x0),
    F_bar = F0 + sum_j theta_j (F_j - F0) and the new input is
    x_bar + damp F_bar.  Simple mixing x0 + damp_init F0 for the first step,
    or for the first ``generations`` steps with ``soft_start``.
    ``kind="simple"``: x_old + damp (x_new - x_old) (``scf/mixer/simple.py``).
    Defaults: ``constants/defaults.py:148-172``."""

    def __init__(self, kind="anderson", damp=0.5, damp_init=0.1, generations=5, soft_start=True, offset=0.01):
        self.kind, self.damp, self.damp_init = kind, damp, damp_init
        self.generations, self.soft_start, self.offset = generations, soft_start, offset
        self.step = 0
        self.hist = []  # previous (input, residual) pairs, newest first, at most `generations`
        self.delta = None

    def iter(self, x_new, x_old):
        self.step += 1
        f0 = x_new - x_old
        self.delta = f0
        if self.kind == "simple":
            return x_old + self.damp * f0
        if self.step > self.generations or (self.step > 1 and not self.soft_start):
            prev = self.hist[: min(self.step - 1, self.generations)]
            df = torch.stack([f0 - fj for _, fj in prev])
            a = df @ df.T
            b = df @ f0
            if self.offset is not None:
                a = a + torch.diag(torch.diagonal(a)) * self.offset**2
            theta = torch.linalg.solve(a, b)
            xbar = x_old + sum(t * (xj - x_old) for t, (xj, _) in zip(theta, prev))
            fbar = f0 + sum(t * (fj - f0) for t, (_, fj) in zip(theta, prev))
            xmix = xbar + self.damp * fbar
        else:
            xmix = x_old + self.damp_init * f0
        self.hist = [(x_old, f0)] + self.hist[: self.generations - 1]
        return xmix


def mix_full(g, x0, kind="anderson", x_tol=1e-10, x_tol_max=1e-5, maxiter=300, **mixer_opts):
    """Full-mode SCF loop (``scf/unrolling/default.py:71-140``) on the map g:
    returns (unmixed output at convergence, loop iterations, converged)."""
    mixer = FullModeMixer(kind, **mixer_opts)
    q = mixer.iter(g(x0), x0)  # the guess is mixed too (MIX_GUESS)
    q_new = q
    for it in range(maxiter):
        q_new = g(q)
        q = mixer.iter(q_new, q)
        d = mixer.delta
        if d.norm() < x_tol and d.abs().max() < x_tol_max:
            return q_new, it + 1, True
    return q_new, maxiter, False


class Result(dict):
    pass


def singlepoint(numbers, positions, gfn, charge=0.0, tol=1e-10, maxiter=300, forces=False, solvation=None,
                scf_mode="implicit", mixer="anderson", x_tol_max=1e-5, spin=None):
    """Total energy (repulsion + electronic + -TS) and optionally forces;
    ``solvation`` adds generalized Born implicit solvation (see System).
    ``scf_mode="implicit"``: broyden1 on the residual; ``"full"``: the
    full-mode loop with ``mixer`` "anderson" or "simple" (x_tol = tol)."""
    sysm = System(numbers, positions, gfn, charge, solvation=solvation, spin=spin)
    with torch.no_grad():
        m = sysm.build(positions)
        n = sysm.nao + (9 * sysm.nat if sysm.is_gfn2 else 0)
        x0 = torch.zeros(n, dtype=torch.float64)
        if scf_mode == "full":
            xs, niter, conv = mix_full(lambda x: sysm.g(m, x), x0, mixer, x_tol=tol, x_tol_max=x_tol_max,
                                       maxiter=maxiter)
        else:
            xs, niter, conv = broyden1(lambda x: x - sysm.g(m, x), x0, f_tol=tol, x_tol=tol, maxiter=maxiter)
        vorb, vdp, vqp = sysm.unpack(xs)
        f = sysm.fock(m, vorb, vdp, vqp)
        p, e, c, occ = sysm.solve_density(m, f)
        q = sysm.charges(m, p)
        eel = sysm.electronic_energy(m, p, *q)
        ts = sysm.free_energy(occ)
        etot = m["erep"].sum() + eel + ts
    res = Result(energy=etot, eel=eel, ts=ts, erep=m["erep"].sum(), iterations=niter, converged=conv,
                 density=p, emo=e, occupation=occ, charges=q, potential=(vorb, vdp, vqp), coeffs=c)
    if forces:
        w = (c * (occ.sum(0) * e)) @ c.T
        pos = positions.detach().clone().requires_grad_(True)
        m2 = sysm.build(pos)
        qq = sysm.charges(m2, p)
        lag = m2["erep"].sum() + sysm.electronic_energy(m2, p, *qq) - (w * m2["s"]).sum()
        (g,) = torch.autograd.grad(lag, pos)
        res["forces"] = -g
    return res
