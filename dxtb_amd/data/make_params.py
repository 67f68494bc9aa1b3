"""
Regenerate the parameter and STO-nG data files of this package.

The GFN1-xTB and GFN2-xTB parametrizations (Apache-2.0, Grimme group, as
shipped with dxtb 0.4.0 in ``src/dxtb/_src/param/gfn{1,2}/gfn{1,2}-xtb.toml``)
and the STO-nG expansion tables (``src/dxtb/_src/basis/sto-ng/sto-{1..6}g.npy``)
are read from a checkout of the reference and written as plain JSON, the only
format this package reads at run time.  Nothing else of the reference is
copied.  Run ``python -m dxtb_amd.data.make_params [reference_root]``.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import tomli

HERE = Path(__file__).parent


def main(root: str = "/root/reference") -> None:
    base = Path(root) / "src" / "dxtb" / "_src"
    for gfn in ("gfn1", "gfn2"):
        with open(base / "param" / gfn / f"{gfn}-xtb.toml", "rb") as fh:
            par = tomli.load(fh)
        out = HERE / f"{gfn}.json"
        with open(out, "w") as fh:
            json.dump(par, fh, indent=1, sort_keys=False)
        print(f"wrote {out}")

    tables = {}
    for n in range(1, 7):
        arr = np.load(base / "basis" / "sto-ng" / f"sto-{n}g.npy", allow_pickle=False)
        # (2, ntype, n): row 0 coefficients, row 1 exponents
        tables[str(n)] = arr.astype(np.float64).tolist()
    out = HERE / "stong.json"
    with open(out, "w") as fh:
        json.dump(tables, fh)
    print(f"wrote {out}")


if __name__ == "__main__":
    main(*sys.argv[1:])
