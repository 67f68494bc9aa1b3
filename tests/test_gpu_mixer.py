"""Device Broyden mixer (dxtb_broyden_step) against the oracle's xitorch
broyden1 (exlibs/xitorch/_impls/optimize/root/rootsolver.py:42-261) on
linear maps g(x) = M x + c of different sizes in one padded batch (weight 0
on the padding): the same iteration counts and the same points to 1e-10,
and molecules that stop are frozen while the others go on."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from dxtb_amd import scf as hscf
from oracle import scf as oscf

pytestmark = pytest.mark.gpu


def _maps(sizes, L, seed=0):
    g = torch.Generator().manual_seed(seed)
    Ms, cs = [], []
    for n in sizes:
        Q, _ = torch.linalg.qr(torch.randn((n, n), generator=g, dtype=torch.float64))
        lam = torch.rand(n, generator=g, dtype=torch.float64) * 3.5 - 3.0  # spectrum in (-3, 0.5)
        M = (Q * lam) @ Q.T
        Mp = torch.zeros((L, L), dtype=torch.float64)
        Mp[:n, :n] = M
        c = torch.zeros(L, dtype=torch.float64)
        c[:n] = torch.randn(n, generator=g, dtype=torch.float64)
        Ms.append(Mp)
        cs.append(c)
    return torch.stack(Ms), torch.stack(cs)


@pytest.mark.parametrize("tol", [1e-8, 1e-11])
def test_device_broyden_vs_oracle(cuda, tol):
    sizes, L = [40, 25, 64, 7], 64
    M, c = _maps(sizes, L)
    w = torch.zeros((len(sizes), L), dtype=torch.float64)
    for i, n in enumerate(sizes):
        w[i, :n] = 1.0
    o = hscf.SCFOptions(f_tol=tol, x_tol=tol, maxiter=200)
    mix = hscf.DeviceBroyden(len(sizes), L, o.maxiter + 1, cuda)
    mix.start()
    Md, cd, wd = M.to(cuda), c.to(cuda), w.to(cuda)
    print(f"case device broyden tol={tol}", flush=True)
    for _ in range(o.maxiter + 1):
        gx = torch.bmm(Md, mix.xe.unsqueeze(-1)).squeeze(-1) + cd
        mix.step(gx, wd, o)
        if bool((mix.flag != 0).all()):
            break
    flags, iters, x = mix.flag.cpu(), mix.ctl[:, 3].cpu(), mix.x.cpu()
    for i, n in enumerate(sizes):
        Mi, ci = M[i, :n, :n], c[i, :n]
        xo, it, conv = oscf.broyden1(lambda v: v - (Mi @ v + ci), torch.zeros(n, dtype=torch.float64), f_tol=tol,
                                     x_tol=tol, maxiter=o.maxiter)
        assert conv and int(flags[i]) == 1
        assert int(iters[i]) == it, (i, int(iters[i]), it)
        torch.testing.assert_close(x[i, :n], xo, atol=1e-10, rtol=0)
        assert (x[i, n:] == 0).all()
        # the fixed point of the map
        xs = torch.linalg.solve(torch.eye(n, dtype=torch.float64) - Mi, ci)
        assert (x[i, :n] - xs).abs().max() < 1e3 * tol


def test_device_broyden_limit_and_skip(cuda):
    """A molecule that does not converge within maxiter stops with flag 2
    and its iteration count; an unoccupied row is left untouched."""
    sizes, L = [30, 30], 32
    M, c = _maps(sizes, L, seed=3)
    M[1] = M[1] - 40.0 * torch.eye(L, dtype=torch.float64) * (torch.arange(L) < 30)  # far out of reach
    o = hscf.SCFOptions(maxiter=5)
    mix = hscf.DeviceBroyden(3, L, o.maxiter + 1, cuda)
    mix.start(torch.tensor([0, 1], device=cuda))
    w = torch.ones((3, L), dtype=torch.float64, device=cuda)
    Md = torch.cat([M, torch.zeros((1, L, L), dtype=torch.float64)]).to(cuda)
    cd = torch.cat([c, torch.ones((1, L), dtype=torch.float64)]).to(cuda)
    for _ in range(o.maxiter + 1):
        mix.step(torch.bmm(Md, mix.xe.unsqueeze(-1)).squeeze(-1) + cd, w, o)
    assert mix.flag.tolist() == [2, 2, 0]
    assert mix.ctl[:, 3].tolist() == [5, 5, 0]
    assert (mix.x[2] == 0).all() and (mix.xe[2] == 0).all()
