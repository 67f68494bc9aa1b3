"""Calculator options on the host (no GPU): the reference's keys this path
follows, warns about or refuses (calculators/config/*.py,
calculators/types/base.py:596-671)."""

from __future__ import annotations

import warnings

import numpy as np
import pytest

from dxtb_amd import calculators as c


def _quiet(opts, gfn="gfn2", numbers=None):
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        return c._opts_to_scf(opts, gfn, numbers)


def test_accepted_silently():
    o = _quiet({"scp_mode": "potential", "int_driver": "dxtb", "batch_mode": 1, "verbosity": 0,
                "exclude": ["disp"], "maxiter": 77})
    assert o.maxiter == 77
    _quiet({"scp_mode": 2, "int_driver": 1, "exclude": ["disp", "hal"]}, "gfn1", np.array([6, 17]))


@pytest.mark.parametrize("opts", [{"scp_mode": "charge"}, {"scp_mode": 0}, {"int_driver": "libcint"},
                                  {"exclude": []}, {"exclude": ["d4sc"]}])
def test_differs_from_path_warns(opts):
    with pytest.warns(UserWarning):
        c._opts_to_scf(opts)


def test_gfn1_halogen_warns():
    with pytest.warns(UserWarning, match="halogen"):
        c._opts_to_scf({"exclude": ["disp"]}, "gfn1", np.array([6, 1, 17]))
    _quiet({"exclude": ["disp"]}, "gfn1", np.array([6, 1, 8]))  # no halogen-bond donor


@pytest.mark.parametrize("comp", ["es2", "rep", "aes2", "es3", "all"])
def test_exclude_unsupported_raises(comp):
    with pytest.raises(NotImplementedError):
        c._opts_to_scf({"exclude": ["disp", comp]})


@pytest.mark.parametrize("opts", [{"scp_mode": "density"}, {"int_driver": "cuda"}, {"exclude": ["bogus"]},
                                  {"unknown_key": 1}])
def test_unknown_values_raise(opts):
    with pytest.raises(ValueError):
        c._opts_to_scf(opts)


def test_cache_value_key():
    k = c._Calculator._value_key
    assert k(0) == (0.0,)
    assert k([1, -1]) == (1.0, -1.0)
    assert k(np.array([2.0])) == (2.0,)
    assert k(None) is None
    assert k([None, 2]) == (None, 2.0)
    hash(k([1, 2]))
