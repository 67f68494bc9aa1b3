"""
Oracle: overlap, dipole and quadrupole integrals (test infrastructure only).

Restates the reference PyTorch integral driver:

* primitive-pair prefactors -- ``impls/pipeline.py:82-104``
* three-index Obara-Saika 1D table about a common origin -- ``impls/kernels/os.py:61-163``
* 3D assembly, contraction, Cartesian->spherical transform --
  ``impls/pipeline.py:151-235`` with the tables of ``impls/trafo.py:68-182``
* lower-triangle shell pairs, halved diagonal blocks, mirror --
  ``impls/pairs.py:92-149, 207-225, 322-399``
* unit self-overlap rescaling of the coefficients -- ``impls/pairs.py:181-204``
* shell-pair screening: the rigorous block bound of ``impls/pairs.py:411-494``
  (``_bounds``) and the selection of ``impls/pairs.py:497-541`` (``_screen``:
  pairs at one position always kept, else kept when the bound of any order
  0, 1, 2 exceeds the threshold), applied to the lower-triangle pairs
  before the mirror (``impls/pairs.py:322-399``)
* container post-processing: snorm normalisation (``integral/base.py:152-174``,
  ``integral/utils.py``), r0->rj shift of the dipole (``integral/types/dipole.py:40-86``),
  9->6 reduction, r0r0->rjrj shift and traceless quadrupole
  (``integral/types/quadrupole.py:41-389``) in the order of
  ``integral/container.py:395-535``.

(All paths relative to ``src/dxtb/_src/integral/driver/pytorch`` or
``src/dxtb/_src``.)  Integrals are differentiable with torch autograd, which
is what the oracle uses for the reference's autograd gradients.
"""

from __future__ import annotations

import math

import torch

from .basis import Basis

s3 = math.sqrt(3.0)
TRAFO = {
    0: torch.tensor([[1.0]], dtype=torch.float64),
    1: torch.eye(3, dtype=torch.float64),
    2: torch.tensor(
        [
            [0.0, s3, 0.0, 0.0, 0.0, 0.0],
            [0.0, 0.0, 0.0, 0.0, s3, 0.0],
            [-0.5, 0.0, 0.0, -0.5, 0.0, 1.0],
            [0.0, 0.0, s3, 0.0, 0.0, 0.0],
            [0.5 * s3, 0.0, 0.0, -0.5 * s3, 0.0, 0.0],
        ],
        dtype=torch.float64,
    ),
}
NLM = {
    0: [(0, 0, 0)],
    1: [(0, 1, 0), (0, 0, 1), (1, 0, 0)],  # y, z, x
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
}
# unique per-axis multipole orders: S, Dx, Dy, Dz, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz
COMPS = [
    (0, 0, 0),
    (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2),
]


def os_table(la: int, lb: int, emax: int, xij, pa, pb, pc):
    """1D Obara-Saika table I[i][j][e] (os.py:61-163): start I000 = 1,
    raise i with X_PA, j with X_PB and e with X_PC, each plus 1/(2p) times
    the index-weighted lower terms.  Filled here by raising i and j at e=0
    first, then e for every (i, j)."""
    t = {}

    def g(i, j, e):
        return t.get((i, j, e))

    def lower(i, j, e):
        acc = 0.0
        if i > 0 and g(i - 1, j, e) is not None:
            acc = acc + i * g(i - 1, j, e)
        if j > 0 and g(i, j - 1, e) is not None:
            acc = acc + j * g(i, j - 1, e)
        if e > 0 and g(i, j, e - 1) is not None:
            acc = acc + e * g(i, j, e - 1)
        return acc

    one = torch.ones_like(pa)
    t[(0, 0, 0)] = one
    for i in range(la + 1):
        for j in range(lb + 1):
            if (i, j) == (0, 0):
                continue
            if i > 0:
                v = pa * t[(i - 1, j, 0)] + xij * lower(i - 1, j, 0)
            else:
                v = pb * t[(i, j - 1, 0)] + xij * lower(i, j - 1, 0)
            t[(i, j, 0)] = v
    for e in range(1, emax + 1):
        for i in range(la + 1):
            for j in range(lb + 1):
                t[(i, j, e)] = pc * t[(i, j, e - 1)] + xij * lower(i, j, e - 1)
    return t


def pair_blocks(la, lb, alpha_a, coeff_a, alpha_b, coeff_b, ra, rb, origin=None):
    """Raw multipole blocks of a class of shell pairs about ``origin``
    (default 0), shape (npair, 10, 2la+1, 2lb+1) with COMPS order."""
    if origin is None:
        origin = torch.zeros(3, dtype=ra.dtype)
    ai = alpha_a.view(1, -1, 1)
    aj = alpha_b.view(1, 1, -1)
    ci = coeff_a.view(1, -1, 1)
    cj = coeff_b.view(1, 1, -1)
    p = ai + aj
    oij = 1.0 / p
    xij = 0.5 * oij
    vec = rb - ra  # B - A
    r2 = (vec * vec).sum(-1).view(-1, 1, 1)
    sij = torch.exp(-ai * aj * oij * r2) * math.pi**1.5 * oij**1.5 * ci * cj
    tabs = []
    for ax in range(3):
        v = vec[:, ax].view(-1, 1, 1)
        pa = v * aj * oij  # P - A
        pb = -v * ai * oij  # P - B
        pc = pa + (ra[:, ax] - origin[ax]).view(-1, 1, 1)  # P - C
        tabs.append(os_table(la, lb, 2, xij, pa, pb, pc))
    nca, ncb = len(NLM[la]), len(NLM[lb])
    out = []
    for comp in COMPS:
        rows = []
        for ca in NLM[la]:
            cols = []
            for cb in NLM[lb]:
                prod = sij
                for ax in range(3):
                    prod = prod * tabs[ax][(ca[ax], cb[ax], comp[ax])]
                cols.append(prod.sum((-2, -1)))
            rows.append(torch.stack(cols, -1))
        out.append(torch.stack(rows, -2))  # (npair, nca, ncb)
    cart = torch.stack(out, 1)  # (npair, 10, nca, ncb)
    return TRAFO[la] @ cart @ TRAFO[lb].T


def _normalized_coeff(l, a, c):
    """c / (sqrt(max(s, eps)) + eps) with the exact self overlap (pairs.py:181-204)."""
    z = torch.zeros((1, 3), dtype=torch.float64)
    s = pair_blocks(l, l, a, c, a, c, z, z)[0, 0, 0, 0]
    eps = torch.finfo(torch.float64).eps
    return c / (torch.sqrt(torch.clamp(s, min=eps)) + eps)


def sph_norm(l: int) -> float:
    """Largest absolute row sum of the Cartesian-to-spherical transform of l
    (``_SPH_NORM``, impls/pairs.py:407-408)."""
    return float(TRAFO[l].abs().sum(-1).max())


def _normalized_all(basis: Basis) -> dict:
    coeffs = {}
    for sh in basis.shells:
        if sh.ush not in coeffs:
            coeffs[sh.ush] = _normalized_coeff(sh.l, sh.alpha, sh.coeff)
    return coeffs


def pair_bound(la, lb, alpha_a, coeff_a, alpha_b, coeff_b, ra, rb, orders=(0, 1, 2), origin=None):
    """Upper bounds of the largest |element| of the spherical block of one
    shell pair (bra a at ra, ket b at rb) for each multipole order e in
    ``orders`` about ``origin`` (default 0), following impls/pairs.py:411-494:
    per primitive pair |c_i c_j| exp(-mu R^2) times
    sum_{i<=la, j<=lb, k<=e} C(la,i) C(lb,j) C(e,k) |PA|^i |PB|^j |PC|^k
    2 pi Gamma((n+3)/2) p^(-(n+3)/2), n = (la-i)+(lb-j)+(e-k), with |PA| =
    b R / p, |PB| = a R / p, |PC| <= max(|AC|, |BC|); summed over the
    primitive pairs and scaled by the two transforms' largest row sums
    (times 1 + 1e-10)."""
    c = torch.zeros(3, dtype=ra.dtype) if origin is None else torch.as_tensor(origin, dtype=ra.dtype).reshape(3)
    a, b = alpha_a[:, None], alpha_b[None, :]
    p = a + b
    cc = (coeff_a[:, None] * coeff_b[None, :]).abs()
    mu = a * b / p
    sph = sph_norm(la) * sph_norm(lb) * (1.0 + 1e-10)
    r = (rb - ra).norm()
    xc = torch.maximum((ra - c).norm(), (rb - c).norm())
    xa, xb = b / p * r, a / p * r
    pref = cc * torch.exp(-mu * r * r)
    emax = max(orders)
    mom = [2.0 * math.pi * math.gamma((n + 3) / 2) * p ** (-(n + 3) / 2) for n in range(la + lb + emax + 1)]
    out = []
    for e in orders:
        poly = torch.zeros_like(p)
        for i in range(la + 1):
            for j in range(lb + 1):
                for k in range(e + 1):
                    n = (la - i) + (lb - j) + (e - k)
                    poly = poly + (math.comb(la, i) * math.comb(lb, j) * math.comb(e, k)
                                   * xa**i * xb**j * xc**k * mom[n])
        out.append((pref * poly).sum() * sph)
    return torch.stack(out)


def pair_bounds(basis: Basis, positions: torch.Tensor, orders=(0, 1, 2), origin=None) -> torch.Tensor:
    """``pair_bound`` of every shell pair, (len(orders), nsh, nsh), computed
    for the lower triangle (bra = the larger shell index, as the reference's
    pair plan orders them) and mirrored."""
    coeffs = _normalized_all(basis)
    nsh = basis.nsh
    out = torch.zeros(len(orders), nsh, nsh, dtype=positions.dtype)
    for i, shi in enumerate(basis.shells):
        for j in range(i + 1):
            shj = basis.shells[j]
            bd = pair_bound(shi.l, shj.l, shi.alpha, coeffs[shi.ush], shj.alpha, coeffs[shj.ush],
                            positions[shi.atom], positions[shj.atom], orders, origin)
            out[:, i, j] = bd
            out[:, j, i] = bd
    return out


def screen_keep(basis: Basis, positions: torch.Tensor, threshold: float, origin=None) -> torch.Tensor:
    """(nsh, nsh) bool: the pairs ``_screen`` keeps (impls/pairs.py:497-541):
    both shells at one position, or the bound of any order 0, 1, 2 above
    ``threshold``."""
    pos = positions.detach()
    bd = pair_bounds(basis, pos, (0, 1, 2), origin)
    at = basis.shell_atom()
    same = (pos[at][:, None, :] == pos[at][None, :, :]).all(-1)
    return same | (bd > threshold).any(0)


def raw_matrices(basis: Basis, positions: torch.Tensor, screening_threshold: float | None = None,
                 origin=None) -> torch.Tensor:
    """Raw (origin 0) S, D(3), Q(6 unique) matrices, shape (10, nao, nao):
    lower-triangle shell pairs, halved diagonal blocks, mirror (pairs.py).
    With ``screening_threshold`` the pairs that ``screen_keep`` drops
    (bounds about ``origin``) are zero."""
    nao = basis.nao
    coeffs = _normalized_all(basis)
    keep = None if screening_threshold is None else screen_keep(basis, positions, screening_threshold, origin)
    # group lower-triangle pairs by ordered unique-shell class
    classes: dict = {}
    for i, shi in enumerate(basis.shells):
        for j in range(i + 1):
            shj = basis.shells[j]
            classes.setdefault((shi.ush, shj.ush), []).append((i, j))
    flat = torch.zeros(10, nao * nao, dtype=positions.dtype)
    idx_all, val_all = [], []
    for (ua, ub), pairs in classes.items():
        i0, j0 = pairs[0]
        sa, sb = basis.shells[i0], basis.shells[j0]
        ii = torch.tensor([p[0] for p in pairs])
        jj = torch.tensor([p[1] for p in pairs])
        ra = positions[torch.tensor([basis.shells[k].atom for k in ii.tolist()])]
        rb = positions[torch.tensor([basis.shells[k].atom for k in jj.tolist()])]
        blk = pair_blocks(sa.l, sb.l, sa.alpha, coeffs[ua], sb.alpha, coeffs[ub], ra, rb)
        w = torch.where(ii == jj, 0.5, 1.0).to(positions.dtype)
        if keep is not None:
            w = w * keep[ii, jj].to(positions.dtype)
        blk = blk * w.view(-1, 1, 1, 1)
        aoa = torch.tensor([basis.shells[k].ao for k in ii.tolist()])
        aob = torch.tensor([basis.shells[k].ao for k in jj.tolist()])
        rows = aoa.view(-1, 1, 1) + torch.arange(2 * sa.l + 1).view(1, -1, 1)
        cols = aob.view(-1, 1, 1) + torch.arange(2 * sb.l + 1).view(1, 1, -1)
        idx = (rows * nao + cols).reshape(-1)
        idx_all.append(idx)
        val_all.append(blk.permute(1, 0, 2, 3).reshape(10, -1))
    flat = flat.index_copy(1, torch.cat(idx_all), torch.cat(val_all, 1))
    tri = flat.view(10, nao, nao)
    return tri + tri.transpose(-1, -2)


def multipole_matrices(basis: Basis, positions: torch.Tensor, screening_threshold: float | None = None):
    """Final S (nao, nao), D (3, nao, nao) centred on the ket atom and
    traceless Q (6, nao, nao; xx, yx, yy, zx, zy, zz) centred on the ket atom,
    following integral/container.py:182-535 (screened raw integrals with
    ``screening_threshold``)."""
    raw = raw_matrices(basis, positions, screening_threshold)
    s0 = raw[0]
    d = s0.diagonal()
    norm = torch.where(d == 0.0, torch.zeros_like(d), d.pow(-0.5))
    nn = norm.unsqueeze(-1) * norm.unsqueeze(-2)
    s = s0 * nn
    dr0 = raw[1:4] * nn
    # raw Q unique comps: xx, xy, xz, yy, yz, zz -> lower 9->6: xx, yx, yy, zx, zy, zz
    q = raw[4:10] * nn
    q6 = torch.stack([q[0], q[1], q[3], q[2], q[4], q[5]])
    pos = positions[basis.ao_atom()]  # (nao, 3) ket positions
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    dx, dy, dz = dr0[0], dr0[1], dr0[2]

    def diag(c, dc):
        return -2 * c.unsqueeze(-2) * dc + (c * c).unsqueeze(-2) * s

    def off(a, b, da, db):
        return -b.unsqueeze(-2) * da - a.unsqueeze(-2) * db + (a * b).unsqueeze(-2) * s

    q6 = torch.stack(
        [
            q6[0] + diag(x, dx),
            q6[1] + off(y, x, dy, dx),
            q6[2] + diag(y, dy),
            q6[3] + off(z, x, dz, dx),
            q6[4] + off(z, y, dz, dy),
            q6[5] + diag(z, dz),
        ]
    )
    drj = dr0 - pos.T.unsqueeze(-2) * s.unsqueeze(0)
    tr = 0.5 * (q6[0] + q6[2] + q6[5])
    qt = torch.stack([1.5 * q6[0] - tr, 1.5 * q6[1], 1.5 * q6[2] - tr, 1.5 * q6[3], 1.5 * q6[4], 1.5 * q6[5] - tr])
    return s, drj, qt
