"""Analytic per-pair derivatives of the HIP kernels (dxtb_amd/csrc/pairfuncs.h)
against central finite differences, compiled and run on the host (CPU)."""

from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_pair_derivatives(tmp_path):
    exe = tmp_path / "pfc"
    src = ROOT / "tests" / "cpu" / "pairfuncs_check.cpp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(src)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout
