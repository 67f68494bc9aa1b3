"""Hand-written batched Jacobi eigensolver (dxtb_jacobi_eigh) against the
device eigh: eigenvalues, residual |A v - e v|, orthogonality; padded batches
with odd and maximal orders and (near-)degenerate spectra."""

from __future__ import annotations

import pytest
import torch

from dxtb_amd import _lib

pytestmark = pytest.mark.gpu


def _jacobi(A, nsize, sweeps=60, tol=1e-14):
    lib = _lib.load()
    B, n, _ = A.shape
    V = torch.empty_like(A)
    e = torch.empty((B, n), dtype=A.dtype, device=A.device)
    sw = torch.zeros(B, dtype=torch.int32, device=A.device)
    ns = torch.as_tensor(nsize, dtype=torch.int32, device=A.device)
    rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(ns), _lib.ptr(A.contiguous()), _lib.ptr(V), _lib.ptr(e), sweeps, tol,
                              _lib.ptr(sw), _lib.stream_ptr(A.device))
    _lib.check(rc, "dxtb_jacobi_eigh")
    return e, V, sw


@pytest.mark.parametrize("nmax,sizes", [(7, [7, 5, 1, 2]), (64, [64, 63, 40]), (128, [128, 127, 100, 3])])
def test_random_symmetric(cuda, nmax, sizes):
    print(f"case jacobi nmax={nmax} sizes={sizes}", flush=True)
    g = torch.Generator().manual_seed(nmax)
    A = torch.zeros(len(sizes), nmax, nmax, dtype=torch.float64)
    for b, n in enumerate(sizes):
        M = torch.randn(n, n, generator=g, dtype=torch.float64)
        A[b, :n, :n] = 0.5 * (M + M.T)
    e, V, sw = _jacobi(A.to(cuda), sizes)
    e, V = e.cpu(), V.cpu()
    for b, n in enumerate(sizes):
        a = A[b, :n, :n]
        ref = torch.linalg.eigvalsh(a)
        torch.testing.assert_close(torch.sort(e[b, :n]).values, ref, atol=1e-12, rtol=0)
        v = V[b, :n, :n]
        torch.testing.assert_close(a @ v, v * e[b, :n], atol=1e-11, rtol=0)
        torch.testing.assert_close(v.T @ v, torch.eye(n, dtype=torch.float64), atol=1e-13, rtol=0)
        # padding: identity vectors, zero values
        if n < nmax:
            assert (V[b, n:, n:] - torch.eye(nmax - n, dtype=torch.float64)).abs().max().item() == 0.0
    assert (sw.cpu() < 60).all()


def test_degenerate_and_warm(cuda):
    n = 32
    g = torch.Generator().manual_seed(3)
    Qm, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    ev = torch.cat([torch.full((8,), -1.0), torch.linspace(0, 1, n - 8)]).double()
    a = Qm @ torch.diag(ev) @ Qm.T
    e, V, _ = _jacobi(a[None].to(cuda), [n])
    torch.testing.assert_close(torch.sort(e[0].cpu()).values, torch.sort(ev).values, atol=1e-12, rtol=0)
    # nearly diagonal input (warm start) converges in very few sweeps
    d = torch.diag(torch.linspace(-1, 1, n, dtype=torch.float64)) + 1e-6 * (torch.ones(n, n) - torch.eye(n)).double()
    e2, V2, sw = _jacobi(d[None].to(cuda), [n])
    assert int(sw[0]) <= 3
