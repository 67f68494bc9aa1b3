"""Hand-written batched Jacobi eigensolver (dxtb_jacobi_eigh) against the
device eigh: eigenvalues, residual |A v - e v|, orthogonality; padded batches
with odd and maximal orders and (near-)degenerate spectra.  Variants: 1 the
scalar LDS form, 2 the blocked form (one workgroup per matrix), 3 the blocked
form with its first sweeps split over many workgroups per matrix."""

from __future__ import annotations

import pytest
import torch

from dxtb_amd import _lib

pytestmark = pytest.mark.gpu


def _jacobi(A, nsize, variant=0, sweeps=60, tol=1e-14, tol_b=None, split=4):
    lib = _lib.load()
    B, n, _ = A.shape
    A = A.clone().contiguous()  # the blocked variant works in place
    V = torch.empty_like(A)
    e = torch.empty((B, n), dtype=A.dtype, device=A.device)
    sw = torch.zeros(B, dtype=torch.int32, device=A.device)
    ns = torch.as_tensor(nsize, dtype=torch.int32, device=A.device)
    work = torch.empty(max(lib.dxtb_jacobi_work_size(B, n) // 8, 1), dtype=torch.float64, device=A.device)
    rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(ns), _lib.ptr(A), _lib.ptr(V), _lib.ptr(e), _lib.ptr(work), variant,
                              sweeps, split, tol, _lib.ptr(tol_b), _lib.ptr(sw), _lib.stream_ptr(A.device))
    _lib.check(rc, "dxtb_jacobi_eigh")
    return e, V, sw


# the last two batches exceed 16 matrices: the blocked forms then apply each
# sweep's rotations to the eigenvectors once per sweep (deferred update)
CASES = [(7, [7, 5, 1, 2]), (33, [33, 32, 17]), (64, [64, 63, 40]), (128, [128, 127, 100, 3]),
         (300, [300, 211, 129]), (512, [512, 333]), (1030, [1030, 700]),
         (96, [96 - 4 * k for k in range(20)]), (260, [260, 1] + [250 - 11 * k for k in range(18)])]


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("nmax,sizes", CASES)
def test_random_symmetric(cuda, nmax, sizes, variant):
    if variant == 1 and nmax > 128:
        pytest.skip("scalar variant holds A in LDS (n <= 128)")
    print(f"case jacobi variant={variant} nmax={nmax} sizes={sizes}", flush=True)
    g = torch.Generator().manual_seed(nmax)
    A = torch.zeros(len(sizes), nmax, nmax, dtype=torch.float64)
    for b, n in enumerate(sizes):
        M = torch.randn(n, n, generator=g, dtype=torch.float64)
        A[b, :n, :n] = 0.5 * (M + M.T)
    e, V, sw = _jacobi(A.to(cuda), sizes, variant)
    e, V = e.cpu(), V.cpu()
    for b, n in enumerate(sizes):
        a = A[b, :n, :n]
        ref = torch.linalg.eigvalsh(a)
        # backward-stable solvers agree to O(n eps ||A||): the bars scale with n and the spectral radius
        eps = torch.finfo(torch.float64).eps
        scale = max(1.0, ref.abs().max().item())
        torch.testing.assert_close(torch.sort(e[b, :n]).values, ref, atol=4 * n * eps * scale, rtol=0)
        v = V[b, :n, :n].T  # eigenvectors are returned as rows
        torch.testing.assert_close(a @ v, v * e[b, :n], atol=16 * n * eps * scale, rtol=0)
        torch.testing.assert_close(v.T @ v, torch.eye(n, dtype=torch.float64), atol=4 * n * eps, rtol=0)
        # padding: identity vectors, zero values
        if n < nmax:
            assert (V[b, n:, n:] - torch.eye(nmax - n, dtype=torch.float64)).abs().max().item() == 0.0
    assert (sw.cpu() < 60).all()


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("n", [32, 100])
def test_degenerate_and_warm(cuda, variant, n):
    g = torch.Generator().manual_seed(3)
    Qm, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    ev = torch.cat([torch.full((8,), -1.0), torch.linspace(0, 1, n - 8)]).double()
    a = Qm @ torch.diag(ev) @ Qm.T
    e, V, _ = _jacobi(a[None].to(cuda), [n], variant)
    torch.testing.assert_close(torch.sort(e[0].cpu()).values, torch.sort(ev).values, atol=1e-12, rtol=0)
    v = V[0].cpu().T
    torch.testing.assert_close(a @ v, v * e[0].cpu(), atol=1e-12, rtol=0)
    # nearly diagonal input (warm start) converges in very few sweeps
    d = torch.diag(torch.linspace(-1, 1, n, dtype=torch.float64)) + 1e-6 * (torch.ones(n, n) - torch.eye(n)).double()
    e2, V2, sw = _jacobi(d[None].to(cuda), [n], variant)
    assert int(sw[0]) <= 3


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_per_matrix_tolerance(cuda, variant):
    """tol_b sets the stopping bar per matrix: a loose bar stops earlier, and
    the matrix with the tight bar is solved as with the scalar tolerance."""
    n = 60
    g = torch.Generator().manual_seed(5)
    M = torch.randn(2, n, n, generator=g, dtype=torch.float64)
    A = (0.5 * (M + M.transpose(1, 2))).to(cuda)
    tb = torch.tensor([1e-4, 1e-14], dtype=torch.float64, device=cuda)
    e, V, sw = _jacobi(A, [n, n], variant, tol_b=tb)
    e1, V1, sw1 = _jacobi(A[1:], [n], variant)
    assert int(sw[0]) < int(sw[1])
    assert int(sw[1]) == int(sw1[0])
    torch.testing.assert_close(e[1], e1[0], atol=0, rtol=0)


@pytest.mark.parametrize("nmax,sizes", [(128, [128, 97]), (300, [300, 64, 0, 255]), (512, [512]),
                                        (200, [200, 0] + [199 - 9 * k for k in range(20)])])
def test_split_matches_blocked(cuda, nmax, sizes):
    """The split form runs the same rotations as the one-workgroup form: the
    same sweep counts and eigenpairs to rounding (an empty matrix included)."""
    print(f"case jacobi split nmax={nmax} sizes={sizes}", flush=True)
    g = torch.Generator().manual_seed(11)
    A = torch.zeros(len(sizes), nmax, nmax, dtype=torch.float64)
    for b, n in enumerate(sizes):
        M = torch.randn(n, n, generator=g, dtype=torch.float64)
        A[b, :n, :n] = 0.5 * (M + M.T)
    A = A.to(cuda)
    e2, V2, sw2 = _jacobi(A, sizes, 2)
    e3, V3, sw3 = _jacobi(A, sizes, 3)
    assert (sw2.cpu() - sw3.cpu()).abs().max().item() <= 1
    eps = torch.finfo(torch.float64).eps
    for b, n in enumerate(sizes):
        if n == 0:
            assert int(sw3[b]) == 0 and e3[b].abs().max().item() == 0.0
            continue
        scale = max(1.0, e2[b, :n].abs().max().item())
        torch.testing.assert_close(e3[b, :n], e2[b, :n], atol=4 * n * eps * scale, rtol=0)
        a = A[b, :n, :n]
        v = V3[b, :n, :n].T
        torch.testing.assert_close(a @ v, v * e3[b, :n], atol=16 * n * eps * scale, rtol=0)


@pytest.mark.parametrize("n,sizes", [(200, [200] * 4), (100, [100] * 20), (600, [600] * 3), (157, [157]),
                                     (200, [200, 64, 20, 157, 0, 33]), (160, [160, 64] + [150 - 5 * k for k in range(18)]),
                                     (224, [224, 64, 20, 0] + [210 - k for k in range(76)])])
def test_split_bitwise_matches_blocked(cuda, n, sizes):
    """The split form (one launch per outer step: the products of the step
    before, then the cross sweeps by four wavefronts per block pair) and the
    one-workgroup form (one wavefront per block pair) apply the same
    rotations with the same arithmetic: on warm-start-like matrices (sorted
    diagonal plus a 1e-3 perturbation) they agree bit for bit, also with the
    deferred eigenvector update (20 matrices) and with matrices of fewer
    blocks in the call (order 64: four blocks, whose only tile holds both
    pairs of a step; orders up to 32: one block pair) and in calls of 576
    block pairs or more (80 matrices of order up to 224: 96 x 7 pairs), whose
    steps sweep every pair with one wavefront, so an SCF result does not
    depend on how many of its sweeps ran split."""
    B = len(sizes)
    print(f"case jacobi bitwise split nmax={n} sizes={sizes}", flush=True)
    g = torch.Generator().manual_seed(5)
    A = torch.zeros(B, n, n, dtype=torch.float64)
    for b, m in enumerate(sizes):
        d = torch.sort(torch.randn(m, generator=g, dtype=torch.float64)).values
        M = 1e-3 * torch.randn(m, m, generator=g, dtype=torch.float64)
        A[b, :m, :m] = torch.diag(d) + 0.5 * (M + M.T)
    A = A.to(cuda)
    for split in (2, 4):
        e2, V2, s2 = _jacobi(A, sizes, 2, tol=1e-12, split=split)
        e3, V3, s3 = _jacobi(A, sizes, 3, tol=1e-12, split=split)
        assert torch.equal(s2, s3)
        assert torch.equal(e2, e3) and torch.equal(V2, V3)


def test_split_concurrent_streams_bitwise(cuda):
    """Split-form solves of two batches on two streams at once, beside a
    large product on a third, give bit for bit the results of the same
    solves run alone: the kernels' LDS exchanges do not depend on how their
    wavefronts are scheduled (a missing barrier shows up here as different
    rotations under contention)."""
    print("case jacobi split concurrent streams", flush=True)
    g = torch.Generator().manual_seed(17)
    mats = []
    for B, n in ((48, 256), (24, 320)):
        A = torch.zeros(B, n, n, dtype=torch.float64)
        for b in range(B):
            d = torch.sort(torch.randn(n, generator=g, dtype=torch.float64)).values
            M = 1e-2 * torch.randn(n, n, generator=g, dtype=torch.float64)
            A[b] = torch.diag(d) + 0.5 * (M + M.T)
        mats.append(A.to(cuda))
    solo = [_jacobi(A, [A.shape[1]] * A.shape[0], 3, tol=1e-12, split=12) for A in mats]
    big = torch.randn(4096, 4096, dtype=torch.float64, device=cuda)
    streams = [torch.cuda.Stream(cuda) for _ in range(3)]
    torch.cuda.synchronize(cuda)
    for _ in range(3):
        outs = []
        with torch.cuda.stream(streams[2]):
            c = big @ big
        for A, st in zip(mats, streams[:2]):
            with torch.cuda.stream(st):
                outs.append(_jacobi(A, [A.shape[1]] * A.shape[0], 3, tol=1e-12, split=12))
        torch.cuda.synchronize(cuda)
        del c
        for (e0, V0, s0), (e1, V1, s1) in zip(solo, outs):
            assert torch.equal(s0, s1)
            assert torch.equal(e0, e1) and torch.equal(V0, V1)


def test_split_few_sweeps(cuda):
    """max_sweeps below the split sweeps: the split form stops there too."""
    n = 200
    g = torch.Generator().manual_seed(2)
    M = torch.randn(n, n, generator=g, dtype=torch.float64)
    A = (0.5 * (M + M.T))[None].to(cuda)
    for ms in (0, 1):
        e2, _, sw2 = _jacobi(A, [n], 2, sweeps=ms)
        e3, _, sw3 = _jacobi(A, [n], 3, sweeps=ms)
        assert int(sw2[0]) == ms and int(sw3[0]) == ms
        torch.testing.assert_close(e3, e2, atol=1e-12, rtol=0)


def _jacobi_into(lib, A, nsize, work, split, sweeps=50, tol=1e-14):
    """One split-form call on the current stream with a caller-held workspace
    (A is solved in place)."""
    B, n, _ = A.shape
    V = torch.empty_like(A)
    e = torch.empty((B, n), dtype=A.dtype, device=A.device)
    sw = torch.zeros(B, dtype=torch.int32, device=A.device)
    ns = torch.as_tensor(nsize, dtype=torch.int32, device=A.device)
    rc = lib.dxtb_jacobi_eigh(B, n, _lib.ptr(ns), _lib.ptr(A), _lib.ptr(V), _lib.ptr(e), _lib.ptr(work), 3, sweeps,
                              split, tol, None, _lib.ptr(sw), _lib.stream_ptr(A.device))
    _lib.check(rc, "dxtb_jacobi_eigh")
    return e, V, sw


def test_split_graph_replay_matches_direct(cuda):
    """On a side stream the split form's launches are recorded once per
    (workspace, batch capacity, order, sweep counts) and replayed: the second
    call here replays the first call's record with other buffers and another
    batch size of the same capacity (5 and 7 matrices, capacity 8).  Both
    give the results of the same calls launched directly (default stream),
    bit for bit, including the sweep counts and a matrix that converges
    before the split sweeps end."""
    lib = _lib.load()
    assert torch.cuda.current_stream(cuda).cuda_stream == 0  # the default stream: direct launches
    SPLIT = 5
    nmax = 160
    g = torch.Generator().manual_seed(5)
    batches = []
    for sizes in ([160, 150, 97, 33, 160], [120, 160, 160, 5, 160, 64, 159]):
        A = torch.zeros(len(sizes), nmax, nmax, dtype=torch.float64)
        for b, n in enumerate(sizes):
            M = torch.randn(n, n, generator=g, dtype=torch.float64)
            d = torch.sort(torch.randn(n, generator=g, dtype=torch.float64)).values
            # near-diagonal (a warm start) except the first matrix
            A[b, :n, :n] = 0.5 * (M + M.T) if b == 0 else torch.diag(d) + 1e-3 * (M + M.T)
        batches.append((A.to(cuda), sizes))
    wsize = lib.dxtb_jacobi_work_size(8, nmax) // 8
    direct = []
    wd = torch.empty(wsize, dtype=torch.float64, device=cuda)
    for A, sizes in batches:
        direct.append(_jacobi_into(lib, A.clone(), sizes, wd, split=SPLIT))
    torch.cuda.synchronize()
    side = torch.cuda.Stream(device=cuda)
    wg = torch.empty(wsize, dtype=torch.float64, device=cuda)
    with torch.cuda.stream(side):
        assert torch.cuda.current_stream(cuda).cuda_stream != 0
        graphed = [_jacobi_into(lib, A.clone(), sizes, wg, split=SPLIT) for A, sizes in batches]
    torch.cuda.synchronize()
    for (e0, V0, s0), (e1, V1, s1), (A, sizes) in zip(direct, graphed, batches):
        assert s0.tolist() == s1.tolist()
        # sweeps left to the finisher, and stops within the split sweeps
        assert max(s0.tolist()) > SPLIT and min(s0.tolist()) <= SPLIT
        torch.testing.assert_close(e1, e0, atol=0, rtol=0)
        torch.testing.assert_close(V1, V0, atol=0, rtol=0)
        # and the eigenpairs are right
        for b, n in enumerate(sizes):
            a, v = A[b, :n, :n], V1[b, :n, :n].T
            scale = max(1.0, e1[b, :n].abs().max().item())
            torch.testing.assert_close(a @ v, v * e1[b, :n], atol=16 * n * 2.2e-16 * scale, rtol=0)
