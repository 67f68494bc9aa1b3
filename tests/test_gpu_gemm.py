"""Batched FP64 matrix-core products (dxtb_gemm_batched) against torch FP64
on the CPU: every transpose combination, orders that are not multiples of the
64-wide tile, per-matrix molecule sizes (leading blocks), the padding fill,
the k scaling, the beta term and the symmetric (mirrored) form.  Bar: 1e-12
absolute for O(1) entries (sums of at most 300 products: FP64 rounding of a
different summation order is ~1e-14)."""

from __future__ import annotations

import pytest
import torch

from dxtb_amd import _lib

pytestmark = pytest.mark.gpu

ATOL = 1e-12


def _ref(A, B, ta, tb, alpha, scale, beta, cin, m):
    a = A[:m, :m].T if ta else A[:m, :m]
    b = B[:m, :m].T if tb else B[:m, :m]
    if scale is not None:
        a = a * scale[:m]
    r = alpha * a @ b
    if cin is not None:
        r = r + beta * cin[:m, :m]
    return r


@pytest.mark.parametrize("n", [1, 17, 64, 100, 257])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_general(cuda, n, ta, tb):
    g = torch.Generator().manual_seed(n * 4 + 2 * ta + tb)
    nb = 5
    A = torch.randn((nb, n, n), generator=g, dtype=torch.float64)
    B = torch.randn((nb, n, n), generator=g, dtype=torch.float64)
    S = torch.rand((nb, n), generator=g, dtype=torch.float64)
    Cn = torch.randn((nb, n, n), generator=g, dtype=torch.float64)
    sizes = torch.tensor([n, max(n - 1, 0), (n + 1) // 2, 1, 0][:nb], dtype=torch.int32)
    print(f"case gemm n={n} ta={ta} tb={tb}", flush=True)
    d = lambda t: t.to(cuda)  # noqa: E731
    # padding filled with 2*identity; alpha/beta/scale
    out = _lib.gemm(d(A), d(B), transa=ta, transb=tb, alpha=-0.7, scale=d(S), beta=1.3, cin=d(Cn),
                    nsize=d(sizes), pad=2.0).cpu()
    for i in range(nb):
        m = int(sizes[i])
        ref = _ref(A[i], B[i], ta, tb, -0.7, S[i], 1.3, Cn[i], m)
        torch.testing.assert_close(out[i, :m, :m], ref, atol=ATOL, rtol=0)
        pad = 2.0 * torch.eye(n, dtype=torch.float64)
        pad[:m, :m] = out[i, :m, :m]
        torch.testing.assert_close(out[i], pad, atol=0, rtol=0)
    # no nsize, no pad: full products; output buffer aliasing cin
    o2 = d(Cn.clone())
    _lib.gemm(d(A), d(B), transa=ta, transb=tb, beta=0.5, cin=o2, out=o2)
    for i in range(nb):
        torch.testing.assert_close(o2[i].cpu(), _ref(A[i], B[i], ta, tb, 1.0, None, 0.5, Cn[i], n), atol=ATOL,
                                   rtol=0)
    # leave the padding alone
    o3 = torch.full((nb, n, n), 7.0, dtype=torch.float64, device=cuda)
    _lib.gemm(d(A), d(B), transa=ta, transb=tb, nsize=d(sizes), out=o3)
    o3 = o3.cpu()
    for i in range(nb):
        m = int(sizes[i])
        torch.testing.assert_close(o3[i, :m, :m], _ref(A[i], B[i], ta, tb, 1.0, None, 0.0, None, m), atol=ATOL,
                                   rtol=0)
        keep = torch.ones((n, n), dtype=torch.bool)
        keep[:m, :m] = False
        assert (o3[i][keep] == 7.0).all()


@pytest.mark.parametrize("n", [30, 64, 130, 200])
def test_gemm_symmetric(cuda, n):
    """C diag(f) C^T (the density) and C^T C in the mirrored form: exactly
    symmetric output, equal to the full product."""
    g = torch.Generator().manual_seed(n)
    nb = 4
    C = torch.randn((nb, n, n), generator=g, dtype=torch.float64)
    f = torch.rand((nb, n), generator=g, dtype=torch.float64)
    sizes = torch.tensor([n, n - 3, n // 2, 2], dtype=torch.int32)
    print(f"case gemm symmetric n={n}", flush=True)
    P = _lib.gemm(C.to(cuda), C.to(cuda), transb=True, scale=f.to(cuda), nsize=sizes.to(cuda), sym=True,
                  pad=0.0).cpu()
    G = _lib.gemm(C.to(cuda), C.to(cuda), transa=True, nsize=sizes.to(cuda), sym=True, pad=1.0).cpu()
    for i in range(nb):
        m = int(sizes[i])
        c = C[i, :m, :m]
        torch.testing.assert_close(P[i, :m, :m], (c * f[i, :m]) @ c.T, atol=ATOL, rtol=0)
        torch.testing.assert_close(G[i, :m, :m], c.T @ c, atol=ATOL, rtol=0)
        assert torch.equal(P[i], P[i].T) and torch.equal(G[i], G[i].T)
        assert (P[i, m:, :] == 0).all() and (P[i, :, m:] == 0).all()
        torch.testing.assert_close(G[i, m:, m:], torch.eye(n - m, dtype=torch.float64), atol=0, rtol=0)


def test_gemm_rejects_aliasing(cuda):
    A = torch.zeros((1, 8, 8), dtype=torch.float64, device=cuda)
    with pytest.raises(RuntimeError):
        _lib.gemm(A, A.clone(), out=A)
