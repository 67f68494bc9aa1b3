"""Host-side logic of the serving engine (no GPU): basis sizes, pool classes,
capacities."""

from __future__ import annotations

import numpy as np
import pytest

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
