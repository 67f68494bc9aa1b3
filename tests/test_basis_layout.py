"""
Shell/AO maps of ``basis.build_batch`` in both AO layouts (CPU).

The reference layout puts every shell's 2l+1 functions in shell order (atoms
in order, each atom's shells in order: ``integral/driver/pytorch/base.py``
``ihelp`` orbital maps); the "l" layout groups the shells by angular
momentum.  Both must give consistent maps (each AO belongs to exactly one
shell, a shell's functions are contiguous and in component order) and the
same shells, atoms and pair lists.
"""

from __future__ import annotations

import numpy as np
import pytest

from dxtb_amd import basis as hbasis
from dxtb_amd import workloads


def _mols():
    nums, _, _ = workloads.config_batch(1, count=3)
    n2, _, _ = workloads.config_batch(2, count=4)
    return workloads.pad(list(nums) + list(n2), [np.zeros((len(z), 3)) for z in list(nums) + list(n2)])[0]


def _check_maps(bb):
    for b in range(bb.nbatch):
        ns, na = int(bb.nsh[b]), int(bb.nao[b])
        seen = np.zeros(na, dtype=int)
        for s in range(ns):
            k = 2 * int(bb.sh_l[b, s]) + 1
            r = np.arange(bb.sh_ao[b, s], bb.sh_ao[b, s] + k)
            assert r[-1] < na
            seen[r] += 1
            assert (bb.ao_shell[b, r] == s).all()
            assert (bb.ao_atom[b, r] == bb.sh_atom[b, s]).all()
        assert (seen == 1).all(), "every AO of the molecule belongs to exactly one shell"


@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_reference_layout_is_shell_order(gfn):
    N = _mols()
    bb = hbasis.build_batch(N, gfn, "cpu")
    assert bb.ao_order == "reference"
    _check_maps(bb)
    for b in range(bb.nbatch):
        ns = int(bb.nsh[b])
        k = 2 * bb.sh_l[b, :ns] + 1
        assert (bb.sh_ao[b, :ns] == np.cumsum(k) - k).all()
        # AO -> shell is the shell order expanded by 2l+1
        assert (bb.ao_shell[b, : bb.nao[b]] == np.repeat(np.arange(ns), k)).all()


@pytest.mark.parametrize("gfn", ["gfn1", "gfn2"])
def test_l_layout_groups_shells_by_angular_momentum(gfn):
    N = _mols()
    ref = hbasis.build_batch(N, gfn, "cpu")
    bl = hbasis.build_batch(N, gfn, "cpu", ao_order="l")
    assert bl.ao_order == "l"
    _check_maps(bl)
    # same shells, atoms, sizes and pair work lists
    for name in ("nat", "nsh", "nao", "sh_atom", "sh_ush", "sh_l", "item_mol", "item_pair", "class_off"):
        np.testing.assert_array_equal(getattr(bl, name), getattr(ref, name))
    for b in range(bl.nbatch):
        na, ns = int(bl.nao[b]), int(bl.nsh[b])
        lao = bl.sh_l[b, bl.ao_shell[b, :na]]
        assert (np.diff(lao) >= 0).all(), "AOs grouped by l"
        for l in range(3):
            sh = np.nonzero(bl.sh_l[b, :ns] == l)[0]
            assert (np.diff(bl.sh_ao[b, sh]) == 2 * l + 1).all(), "a group holds its shells in reference order"
        # the two layouts differ by a permutation of whole shells
        perm = np.empty(na, dtype=int)
        for s in range(ns):
            k = 2 * int(bl.sh_l[b, s]) + 1
            perm[ref.sh_ao[b, s] : ref.sh_ao[b, s] + k] = np.arange(bl.sh_ao[b, s], bl.sh_ao[b, s] + k)
        assert sorted(perm) == list(range(na))
        np.testing.assert_array_equal(bl.ao_atom[b, perm], ref.ao_atom[b, :na])


def test_l_layout_with_padded_dims():
    """Slot-pool padding (fixed nat/nsh/nao): padding shells and AOs stay
    after the molecule's own, as in the reference layout."""
    N = _mols()
    dims = (N.shape[1] + 3, 400, 700)
    bl = hbasis.build_batch(N, "gfn2", "cpu", dims=dims, ao_order="l")
    assert (bl.nat_max, bl.nsh_max, bl.nao_max) == dims
    _check_maps(bl)
    for b in range(bl.nbatch):
        assert (bl.sh_ao[b, : bl.nsh[b]] < bl.nao[b]).all()


def test_unknown_layout_rejected():
    with pytest.raises(ValueError):
        hbasis.build_batch(_mols(), "gfn2", "cpu", ao_order="atom")
