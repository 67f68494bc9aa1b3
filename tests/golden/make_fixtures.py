"""
Extract the reference's own test vectors into small data fixtures.

Reads (never imports or runs) files of the reference test suite and writes
plain data to ``tests/golden``:

* geometries of ``test/test_singlepoint/mols/*/coord`` (Turbomole format, Bohr);
* tblite H0 / overlap / dCN arrays of ``test/test_hamiltonian/*.npz`` and
  ``test/test_overlap/overlap.npz`` (subset of keys);
* reference numbers written as ``torch.tensor(...)`` literals in
  ``test/test_hamiltonian/test_gfn2.py`` (GFN2 H0 without overlap),
  ``test/test_coulomb/samples.py`` (ES2/ES3 energies, charges, gradients),
  ``test/test_coulomb/test_aes2.py`` (AES2 energies and potentials),
  ``test/test_scf/samples.py`` (SCF electronic energies),
  ``test/test_classical/test_repulsion/samples.py`` (repulsion energies),
  ``test/test_solvation/samples.py`` (Born radii, psi, generalized Born
  energies for given charges, GFN1 solvation energies),
  ``test/test_singlepoint/refs/gfn2/*.txt`` (tblite GFN2 energies/gradients).

The literals are evaluated with ``ast`` as data only.  Run:
``python tests/golden/make_fixtures.py [/root/reference]``.
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).parent

SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co "
    "Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te "
    "I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir "
    "Pt Au Hg Tl Pb Bi Po At Rn"
).split()
ZOF = {s.lower(): i for i, s in enumerate(SYMBOLS)}


def read_coord(path: Path):
    nums, pos = [], []
    inside = False
    for line in path.read_text().splitlines():
        s = line.strip()
        if s.startswith("$coord"):
            inside = True
            continue
        if s.startswith("$"):
            inside = False
            continue
        if inside and s:
            t = s.split()
            pos.append([float(x) for x in t[:3]])
            nums.append(ZOF[t[3].lower()])
    return nums, pos


class _Lit:
    """Literal evaluator for nested dict/list/torch.tensor(...) expressions."""

    def ev(self, n):
        if isinstance(n, ast.Constant):
            return n.value
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = self.ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, (ast.List, ast.Tuple)):
            out = []
            for e in n.elts:
                if isinstance(e, ast.Starred):
                    out.extend(self.ev(e.value))
                else:
                    out.append(self.ev(e))
            return out
        if isinstance(n, ast.Dict):
            return {self.ev(k): self.ev(v) for k, v in zip(n.keys, n.values) if k is not None}
        if isinstance(n, ast.Call):
            f = n.func
            if isinstance(f, ast.Attribute) and f.attr == "tensor":
                return self.ev(n.args[0])
            if isinstance(f, ast.Attribute) and f.attr == "reshape":
                base = np.asarray(self.ev(f.value), dtype=np.float64)
                shape = [self.ev(a) for a in n.args]
                return base.reshape(shape).tolist()
        return None  # not a plain literal (e.g. a computed tensor): skipped


def module_dict(path: Path, name: str):
    tree = ast.parse(path.read_text())
    for node in tree.body:
        tgt = None
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            tgt, val = node.targets[0], node.value
        elif isinstance(node, ast.AnnAssign):
            tgt, val = node.target, node.value
        if isinstance(tgt, ast.Name) and tgt.id == name and val is not None:
            return _Lit().ev(val)
    raise KeyError(name)


def function_locals(path: Path, func: str):
    """Literal torch.tensor assignments inside one test function."""
    tree = ast.parse(path.read_text())
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == func:
            for st in node.body:
                if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name):
                    try:
                        out[st.targets[0].id] = _Lit().ev(st.value)
                    except (ValueError, IndexError):
                        pass
    return out


def main(root: str = "/root/reference") -> None:
    t = Path(root) / "test"
    geo = {}
    for d in sorted((t / "test_singlepoint" / "mols").iterdir()):
        c = d / "coord"
        if c.exists():
            nums, pos = read_coord(c)
            geo[d.name] = {"numbers": nums, "positions": pos}
    (OUT / "geometries.json").write_text(json.dumps(geo))

    arrays = {}
    for f, prefix in [
        (t / "test_hamiltonian" / "h0.npz", "h0_gfn1_"),
        (t / "test_hamiltonian" / "grad_no_overlap.npz", "h0grad_gfn1_"),
        (t / "test_overlap" / "overlap.npz", "ovlp_gfn1_"),
    ]:
        with np.load(f, allow_pickle=False) as d:
            for k in d.files:
                arrays[prefix + k] = d[k]
    np.savez_compressed(OUT / "tblite_arrays.npz", **arrays)

    data = {
        "h0_gfn2_noovlp": module_dict(t / "test_hamiltonian" / "test_gfn2.py", "ref"),
        "coulomb": module_dict(t / "test_coulomb" / "samples.py", "refs"),
        "aes2_energy": function_locals(t / "test_coulomb" / "test_aes2.py", "test_energy"),
        "aes2_potential": function_locals(t / "test_coulomb" / "test_aes2.py", "test_potential"),
        "scf": module_dict(t / "test_scf" / "samples.py", "refs"),
        "repulsion": module_dict(t / "test_classical" / "test_repulsion" / "samples.py", "refs"),
        "solvation": module_dict(t / "test_solvation" / "samples.py", "refs"),
    }
    refs = {}
    for f in sorted((t / "test_singlepoint" / "refs" / "gfn2").glob("*.txt")):
        refs[f.stem] = f.read_text()
    data["singlepoint_gfn2_txt"] = refs
    (OUT / "reference_values.json").write_text(json.dumps(data))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main(*sys.argv[1:])
