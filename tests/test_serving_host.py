"""Host-side logic of the serving engine (no GPU): basis sizes, pool classes,
capacities."""

from __future__ import annotations

import numpy as np
import pytest

from dxtb_amd import scf as hscf
from dxtb_amd import serving as sv
from dxtb_amd import workloads


def test_sizes_water_and_h2s():
    assert sv.sizes(np.array([8, 1, 1]), "gfn2") == (3, 4, 6)  # O: s,p; H: s
    assert sv.sizes(np.array([16, 1, 1]), "gfn2") == (3, 5, 11)  # S: s,p,d
    assert sv.sizes(np.array([8, 1, 1]), "gfn1") == (3, 6, 8)  # GFN1 H: 1s, 2s


def test_pool_class_holds_molecule():
    nums, _, _ = workloads.config_batch(3, count=300)
    for z in nums:
        nat, nsh, nao = sv.sizes(z, "gfn2")
        c = sv.pool_class(nat, nsh, nao)
        na, ns, n = sv.class_dims(c)
        assert nat <= na and nsh <= ns and nao <= n
        smaller = [k for k in sv.POOL_NAO if k < c]
        if smaller:  # the class is the smallest that fits
            na2, ns2, n2 = sv.class_dims(smaller[-1])
            assert not (nat <= na2 and nsh <= ns2 and nao <= n2)


def test_too_large_and_unknown_elements():
    with pytest.raises(ValueError):
        sv.pool_class(10, 10, sv.POOL_NAO[-1] + 1)
    with pytest.raises(ValueError):
        sv.sizes(np.array([118]), "gfn2")


def test_capacity_bounds():
    for c in sv.POOL_NAO:
        cap = sv.default_capacity(c)
        assert 8 <= cap <= 256
    assert sv.default_capacity(sv.POOL_NAO[-1]) < sv.default_capacity(sv.POOL_NAO[0])


def test_sweep_prediction_by_window_size(monkeypatch):
    """The split-sweep guess of the SCF's eigensolves: per window-size class,
    the largest recent count plus the margin; a class without history takes
    the largest count of any class; FRESH until anything is known, and after
    an admission; wider margins above order 512."""
    import torch

    from dxtb_amd.scf import _SolveHistory

    monkeypatch.setattr(torch.cuda, "is_current_stream_capturing", lambda: False)
    h = _SolveHistory(240)
    assert h.predict(4) == h.fresh_sweeps == 12
    h.recent = {2: [3, 7, 5]}
    assert h.predict(200) == 8          # large windows: max 7 + 1
    assert h.predict(5) == 8            # no small-window history yet: any class
    h.recent[0] = [2, 3]
    assert h.predict(5) == 4            # small windows: their own history
    assert h.predict(40) == 8           # medium class empty: largest of any class
    h.fresh()
    assert h.predict(5) == 12 and h.predict(5) == 4   # one fresh solve after an admission
    big = _SolveHistory(700)
    big.recent = {0: [3]}
    assert big.predict(2) == 6 and big.fresh_sweeps == 14
    assert [_SolveHistory.size_class(n) for n in (1, 8, 9, 64, 65)] == [0, 0, 1, 1, 2]


def test_pools_reject_full_mode():
    """The pools run broyden1 only; the full mode is refused before any
    device work."""
    with pytest.raises(ValueError, match="implicit mode"):
        sv.SlotPool("gfn2", 48, 8, "cpu", hscf.SCFOptions(scf_mode="full"))
