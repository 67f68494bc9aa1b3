"""Shared pytest configuration: the ``gpu`` marker and common fixtures."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest
import torch

GOLDEN = Path(__file__).parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def geometries():
    return json.loads((GOLDEN / "geometries.json").read_text())


@pytest.fixture(scope="session")
def refvals():
    return json.loads((GOLDEN / "reference_values.json").read_text())


@pytest.fixture(scope="session")
def tblite_arrays():
    with np.load(GOLDEN / "tblite_arrays.npz", allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
