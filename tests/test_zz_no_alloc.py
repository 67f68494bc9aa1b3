"""dxtb_jacobi_eigh_noalloc (no device allocation: fixed per-GPU scratch
array) against dxtb_jacobi_eigh with a caller workspace, through the C ABI:
two calls in a row on the same stream, the second with a smaller batch and
order, for the split, one-workgroup and scalar forms; both issue the same
kernels in the same order, so eigenvalues, eigenvectors and sweep counts are
equal bit for bit.  A request larger than the scratch array is refused before
anything is launched."""

from __future__ import annotations

import pytest
import torch

from dxtb_amd import _lib

pytestmark = pytest.mark.gpu


def _inputs(nb, n, dev, seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.randn((nb, n, n), generator=g, dtype=torch.float64)
    A = 0.5 * (M + M.transpose(1, 2))
    sizes = torch.randint(max(n // 2, 1), n + 1, (nb,), generator=g, dtype=torch.int32)
    sizes[0] = n
    return A.to(dev), sizes.to(dev)


def _solve(lib, A0, sizes, variant, split, noalloc):
    nb, n, _ = A0.shape
    A = A0.clone()
    VT = torch.empty_like(A)
    ev = torch.empty((nb, n), dtype=torch.float64, device=A.device)
    sw = torch.zeros(nb, dtype=torch.int32, device=A.device)
    st = _lib.stream_ptr(A.device)
    if noalloc:
        rc = lib.dxtb_jacobi_eigh_noalloc(nb, n, _lib.ptr(sizes), _lib.ptr(A), _lib.ptr(VT), _lib.ptr(ev), variant,
                                          30, split, 1e-12, None, _lib.ptr(sw), st)
    else:
        work = torch.empty(lib.dxtb_jacobi_work_size(nb, n) // 8 + 1, dtype=torch.float64, device=A.device)
        rc = lib.dxtb_jacobi_eigh(nb, n, _lib.ptr(sizes), _lib.ptr(A), _lib.ptr(VT), _lib.ptr(ev), _lib.ptr(work),
                                  variant, 30, split, 1e-12, None, _lib.ptr(sw), st)
    _lib.check(rc, "jacobi")
    return VT, ev, sw


@pytest.mark.parametrize("variant,split", [(3, 12), (2, 0), (1, 0)])
def test_noalloc_matches_workspace_form(cuda, variant, split):
    lib = _lib.load()
    side = torch.cuda.Stream(cuda)  # a non-null stream: the workspace form replays its recorded graphs there
    with torch.cuda.stream(side):
        for k, (nb, n) in enumerate(((40, 200), (7, 150)) if variant != 1 else ((40, 96), (7, 60))):
            print(f"case noalloc variant={variant} call {k}: {nb} x order {n}", flush=True)
            A0, sizes = _inputs(nb, n, cuda, 100 * variant + k)
            ref = _solve(lib, A0, sizes, variant, split, noalloc=False)
            got = _solve(lib, A0, sizes, variant, split, noalloc=True)
            torch.cuda.synchronize()
            for r, g in zip(ref, got):
                assert torch.equal(r, g)
            # and the result is an eigendecomposition of each leading block
            VT, ev, sw = got
            for b in range(nb):
                m = int(sizes[b])
                a = A0[b, :m, :m]
                v = VT[b, :m, :m]
                res = (v @ a - ev[b, :m, None] * v).abs().max().item()
                assert res < 1e-10 * max(1.0, a.abs().max().item()) * m, (b, res)
            assert int(sw.min()) >= 1


def test_noalloc_rejects_oversize(cuda):
    lib = _lib.load()
    nb, n = 2048, 512  # workspace far above the fixed array
    assert lib.dxtb_jacobi_work_size(nb, n) > (64 << 20)
    dummy = torch.zeros(1, dtype=torch.float64, device=cuda)
    rc = lib.dxtb_jacobi_eigh_noalloc(nb, n, _lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy), _lib.ptr(dummy), 3,
                                      5, 2, 1e-12, None, None, _lib.stream_ptr(cuda))
    assert rc == _lib.DXTB_ESCRATCH
