"""The C-ABI library builds, loads and exports every symbol of include/dxtb_hip.h (CPU only)."""

from __future__ import annotations

import re
import subprocess
from pathlib import Path

import pytest

from dxtb_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared() -> set[str]:
    text = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    return set(re.findall(r"^\s*(?:int|size_t)\s+(dxtb_\w+)\s*\(", text, flags=re.M))


def test_library_present():
    assert _lib.LIBPATH.exists(), "run `python -m dxtb_amd.build` first"


def test_exports_match_header():
    declared = _declared()
    assert declared, "no entry points declared"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIBPATH)], capture_output=True, text=True)
    exported = set(re.findall(r"\bT (dxtb_\w+)", out.stdout))
    assert declared <= exported, f"missing exports: {sorted(declared - exported)}"
    assert declared == set(_lib.exported_names()), "ctypes signatures out of sync with the header"


def test_ctypes_load_and_bind():
    lib = _lib.load()
    for name in _lib.exported_names():
        assert getattr(lib, name).restype is not None


def test_invalid_arguments_rejected_without_gpu():
    """Argument checks run on the host before any launch."""
    lib = _lib.load()
    st = _lib.DxtbBasis()  # all zero: invalid
    assert lib.dxtb_multipole_fwd(st, None, None, None, None, None, None, 1, None) == _lib.DXTB_EINVAL
    assert lib.dxtb_multipole_vjp(st, None, None, None, None, None, None) == _lib.DXTB_EINVAL


def test_jacobi_arguments_without_gpu():
    lib = _lib.load()
    # workspace: a 256-byte argument block, then per matrix of the batch
    # capacity the 32x32 rotations of the order padded to a multiple of 32 for
    # every step of a sweep (order 33: 2 * 2 - 1 = 3 steps of 64 rows; the
    # whole sweep's are kept up to order 512, two steps above), 4 doubles of
    # split-sweep state,
    # per step 128 int rotation flags (block pairs of the largest order, 4096;
    # the flags of the batch rounded up to 256 bytes), and two doubles per
    # chunk of rows of the convergence test (2048 // n rows, at least enough
    # for at most 1024 chunks)
    assert lib.dxtb_jacobi_max_n() == 4096

    def size(cap, n, steps, npad):
        flags = -(-cap * steps * 128 * 4 // 256) * 256
        chunks = -(-n // max(2048 // n, -(-n // 1024)))
        return 256 + cap * steps * npad * 32 * 8 + cap * 4 * 8 + flags + cap * chunks * 2 * 8

    assert lib.dxtb_jacobi_work_size(2, 33) == size(2, 33, 3, 64)
    # capacities: powers of two up to 8, multiples of 8 up to 64, then of 32
    for nb, cap in ((1, 1), (3, 4), (5, 8), (9, 16), (63, 64), (65, 96), (256, 256), (257, 288)):
        assert lib.dxtb_jacobi_work_size(nb, 33) == size(cap, 33, 3, 64), nb
    assert lib.dxtb_jacobi_work_size(0, 33) == 0
    # above order 512 the rotations of two steps only (by step parity): order
    # 600, 608 rows, 200 chunks of 3 rows; order 4000: 1000 chunks of 4 rows
    assert lib.dxtb_jacobi_work_size(1, 600) == size(1, 600, 2, 608)
    assert size(1, 600, 2, 608) == 256 + 2 * 608 * 32 * 8 + 4 * 8 + 1024 + 200 * 16
    assert lib.dxtb_jacobi_work_size(1, 4000) == size(1, 4000, 2, 4000)
    assert lib.dxtb_jacobi_max_n() >= 128
    args = (None,) * 5
    assert lib.dxtb_jacobi_eigh(0, 10, *args, 0, 10, 0, 1e-14, None, None, None) == _lib.DXTB_EINVAL
    assert lib.dxtb_jacobi_eigh(1, lib.dxtb_jacobi_max_n() + 1, *args, 0, 10, 0, 1e-14, None, None, None) == \
        _lib.DXTB_EINVAL
    assert lib.dxtb_jacobi_eigh(1, 10, *args, 4, 10, 0, 1e-14, None, None, None) == _lib.DXTB_EINVAL
    assert lib.dxtb_jacobi_eigh(1, 10, *args, 3, 10, -1, 1e-14, None, None, None) == _lib.DXTB_EINVAL


def test_ctypes_arity_matches_header():
    """Every ctypes signature has as many arguments as its declaration in the
    header (a missing or extra argument would shift every pointer after it)."""
    text = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for name, args in _lib._SIGNATURES.items():
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)\s*;", text, flags=re.S)
        assert m, name
        params = m.group(1).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        assert n == len(args), f"{name}: header has {n} parameters, ctypes {len(args)}"
