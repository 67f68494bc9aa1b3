"""
Parametrizations (GFN1-xTB, GFN2-xTB and custom ones), element data and units.

The shipped parameters are the reference's TOML files converted to JSON
(``data/make_params.py``); the element radii and unit constants, which the
reference takes from tad-mctc 0.7.0, are in ``data/elements.json``.
Mirrors the parameter access of ``src/dxtb/_src/param/module/utils.py``
(element parameters, pair parameters, valence shells) and the
parametrization record of ``src/dxtb/_src/param/base.py:60-270`` (``Param``:
loaded from and written to JSON, TOML or YAML files).

Every parametrization is known by a name: "gfn1" and "gfn2" for the shipped
ones, "gfn1-custom-<hash>" or "gfn2-custom-<hash>" for one registered from a
``Param`` (``register``).  The prefix is its model family, which selects the
GFN1 or GFN2 form of the Hamiltonian and interactions.
"""

from __future__ import annotations

import copy
import hashlib
import json
import math
from functools import lru_cache
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "data"


@lru_cache(maxsize=None)
def _json(name: str) -> dict:
    with open(DATA / f"{name}.json") as fh:
        return json.load(fh)


ELEMENTS = _json("elements")
SYMBOLS: list[str] = ["X"] + ELEMENTS["symbols"]
ATOMIC_NUMBER = {s: i for i, s in enumerate(SYMBOLS)}
MAXZ = len(SYMBOLS) - 1

BOHR_ANGSTROM = ELEMENTS["bohr_angstrom"]
AA2AU = 1.0 / BOHR_ANGSTROM
EV2AU = 1.0 / ELEMENTS["hartree_ev"]
KELVIN2AU = ELEMENTS["kb_hartree_per_kelvin"]

#: D3 covalent radii (4/3 x Pyykko 2009) in Bohr, index = Z
COV_D3 = np.array([0.0] + [4.0 / 3.0 * r * AA2AU for r in ELEMENTS["covalent_2009"]])
#: atomic radii of the H0 distance polynomial in Bohr, index = Z
ATOMIC_RADII = np.array([0.0] + [r * AA2AU for r in ELEMENTS["atomic"]])
#: D3 van der Waals radii (Angstrom, Z = 1..18) of the generalized Born
#: solvation model (the VDW_D3 table the reference takes from tad-mctc);
#: VDW_D3[Z] in Bohr, NaN where no radius is available
_VDW_D3_AA = [1.09155, 0.86735, 1.74780, 1.54910, 1.60800, 1.45515, 1.31125, 1.24085, 1.14980, 1.06870,
              1.85410, 1.74195, 2.00530, 1.89585, 1.75085, 1.65535, 1.55230, 1.45740]
VDW_D3 = np.full(MAXZ + 1, np.nan)
VDW_D3[1 : len(_VDW_D3_AA) + 1] = np.array(_VDW_D3_AA) * AA2AU

_LVAL = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}


class Param:
    """An extended tight-binding parametrization as a plain record (the
    sections ``meta``, ``hamiltonian``, ``repulsion``, ``charge``,
    ``thirdorder``, ``multipole``, ``element``, ... of the reference's TOML
    files), as ``src/dxtb/_src/param/base.py:60-270`` holds it.

    ``Param.from_file(path)`` reads ``.json``, ``.toml`` or ``.yaml``/``.yml``
    (the formats of ``Param.from_file``, param/base.py:128-156) and
    ``to_file(path)`` writes them.  ``kind`` is the model family, taken from
    ``meta.name`` as the reference selects its Hamiltonian
    (``integral/factory.py:60-83``)."""

    def __init__(self, data: dict) -> None:
        if not isinstance(data, dict) or "element" not in data:
            raise ValueError("a parametrization needs at least the 'element' section")
        self.data = copy.deepcopy(data)

    @classmethod
    def from_dict(cls, data: dict) -> "Param":
        return cls(data)

    @classmethod
    def from_file(cls, filepath) -> "Param":
        path = Path(filepath)
        if path.suffix == ".json":
            with open(path) as fh:
                return cls(json.load(fh))
        if path.suffix == ".toml":
            import tomli

            with open(path, "rb") as fh:
                return cls(tomli.load(fh))
        if path.suffix in (".yaml", ".yml"):
            import yaml

            with open(path) as fh:
                return cls(yaml.safe_load(fh))
        raise ValueError(f"Unsupported file format: {path.suffix}")

    def to_file(self, filepath, **kwargs) -> None:
        path = Path(filepath)
        if path.suffix == ".json":
            with open(path, "w") as fh:
                json.dump(self.data, fh, indent=kwargs.get("indent", 1))
        elif path.suffix == ".toml":
            with open(path, "w") as fh:
                fh.write(_toml_dumps(self.data))
        elif path.suffix in (".yaml", ".yml"):
            import yaml

            with open(path, "w") as fh:
                yaml.safe_dump(self.data, fh, sort_keys=False)
        else:
            raise ValueError(f"Unsupported file format: {path.suffix}")

    def clean_model_dump(self) -> dict:
        return copy.deepcopy(self.data)

    @property
    def kind(self) -> str:
        return model_kind(self.data)

    @property
    def name(self) -> str | None:
        return self.data.get("meta", {}).get("name")

    def __eq__(self, other) -> bool:
        return isinstance(other, Param) and self.data == other.data


# model-family names of meta.name (integral/factory.py:74-83)
_FAMILY = {"gfn1-xtb": "gfn1", "gfn1": "gfn1", "gfn2-xtb": "gfn2", "gfn2": "gfn2"}


def model_kind(data: dict) -> str:
    """Model family of a parametrization record, "gfn1" or "gfn2", from its
    ``meta.name`` (casefolded "gfn1-xtb"/"gfn1" or "gfn2-xtb"/"gfn2", as
    ``integral/factory.py:60-83``).  GFN0 is not on this path."""
    name = (data.get("meta") or {}).get("name")
    if name is None:
        raise ValueError("The `name` field of the meta information is missing in the parametrization; "
                         "the model family cannot be selected.")
    fam = _FAMILY.get(str(name).casefold())
    if fam is None:
        if str(name).casefold() in ("gfn0-xtb", "gfn0"):
            raise NotImplementedError("GFN0-xTB (no self-consistent charges) is not part of this path.")
        raise ValueError(f"Unsupported Hamiltonian type: {name}")
    return fam


def _check_sections(data: dict, fam: str) -> None:
    """The components this path computes need their sections: the reference
    drops a component whose section is absent (``new_es3``,
    ``components/interactions/coulomb/thirdorder.py:391``; ``new_aes2``,
    ``multipole.py:673``; ``new_repulsion``, ``classicals/repulsion/factory.py:85``),
    this path always computes ES2, ES3, repulsion and (GFN2) AES2."""
    need = ["hamiltonian", "charge", "thirdorder", "repulsion"] + (["multipole"] if fam == "gfn2" else [])
    missing = [k for k in need if not data.get(k)]
    if missing:
        raise NotImplementedError(f"parametrization without {missing}: this path always computes ES2, ES3, "
                                  "repulsion and, for GFN2, the anisotropic electrostatics")


def _toml_value(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isinf(v) or math.isnan(v):
            raise ValueError("TOML has no portable form for inf/nan parameters")
        return repr(v)
    if isinstance(v, str):
        return json.dumps(v)
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    if isinstance(v, dict):  # inline table
        return "{" + ", ".join(f"{_toml_key(k)} = {_toml_value(x)}" for k, x in v.items()) + "}"
    raise TypeError(f"cannot write {type(v).__name__} to TOML")


def _toml_key(k: str) -> str:
    k = str(k)
    return k if k.replace("_", "").replace("-", "").isalnum() else json.dumps(k)


def _toml_dumps(data: dict, prefix: str = "") -> str:
    """TOML text of nested dicts: scalars and lists first, then sub-tables."""
    lines, subs = [], []
    for k, v in data.items():
        if isinstance(v, dict):
            subs.append((k, v))
        else:
            lines.append(f"{_toml_key(k)} = {_toml_value(v)}")
    out = []
    if prefix and (lines or not subs):
        out.append(f"[{prefix}]")
    out += lines
    for k, v in subs:
        name = f"{prefix}.{_toml_key(k)}" if prefix else _toml_key(k)
        out.append("")
        out.append(_toml_dumps(v, name).rstrip("\n"))
    return "\n".join(out) + "\n"


#: registered parametrizations by name (the shipped ones are read from data/)
_REGISTRY: dict[str, dict] = {}


def register(par) -> str:
    """Name of a parametrization (a Param, its record as a dict, a file path or
    a shipped name), registering a custom one under "<kind>-custom-<hash of
    its content>".  The shipped GFN1/GFN2 records keep their names."""
    if isinstance(par, str) and par in ("gfn1", "gfn2"):
        return par
    if isinstance(par, str) and par in _REGISTRY:
        return par
    if isinstance(par, (str, Path)):
        par = Param.from_file(par)
    if isinstance(par, dict):
        par = Param(par)
    if not isinstance(par, Param):
        raise TypeError(f"expected a Param, a parametrization record or a file, got {type(par).__name__}")
    for name in ("gfn1", "gfn2"):
        if par.data == _json(name):
            return name
    fam = par.kind
    _check_sections(par.data, fam)
    h = hashlib.sha1(json.dumps(par.data, sort_keys=True).encode()).hexdigest()[:12]
    name = f"{fam}-custom-{h}"
    _REGISTRY.setdefault(name, copy.deepcopy(par.data))
    return name


def kind(name: str) -> str:
    """Model family ("gfn1" or "gfn2") of a parametrization name."""
    k = name.split("-", 1)[0]
    if k not in ("gfn1", "gfn2"):
        raise ValueError(f"Unknown parametrization '{name}'.")
    return k


# element fields with a default (param/element.py:135-205): the values a
# record that leaves them out stands for
ELEMENT_DEFAULTS = {"gam3": 0.0, "xbond": 0.0, "dkernel": 0.0, "qkernel": 0.0, "mprad": 0.0, "mpvcn": 0.0}


class Parametrization:
    """Read-only view of one xTB parametrization (by name, see register)."""

    def __init__(self, name: str) -> None:
        if name in ("gfn1", "gfn2"):
            self.raw = _json(name)
        elif name in _REGISTRY:
            self.raw = _REGISTRY[name]
        else:
            raise ValueError(f"Unknown parametrization '{name}' (expected 'gfn1', 'gfn2' or a registered one).")
        self.name = name
        self.kind = kind(name)
        self.h0 = self.raw["hamiltonian"]["xtb"]
        self.elements = {s: {**ELEMENT_DEFAULTS, **e} for s, e in self.raw["element"].items()}

    def has(self, z: int) -> bool:
        return 0 < z <= MAXZ and SYMBOLS[z] in self.elements

    def element(self, z: int) -> dict:
        try:
            return self.elements[SYMBOLS[z]]
        except (IndexError, KeyError):
            raise ValueError(f"Element Z={z} is not parametrized in {self.name}.") from None

    def shells(self, z: int) -> list[tuple[int, int]]:
        """(principal n, angular l) of every shell of an element."""
        return [(int(s[:-1]), _LVAL[s[-1]]) for s in self.element(z)["shells"]]

    def valence(self, z: int) -> list[bool]:
        """Only the first shell of each angular momentum is a valence shell."""
        seen: set[int] = set()
        out = []
        for _, l in self.shells(z):
            out.append(l not in seen)
            seen.add(l)
        return out

    def kpair(self, za: int, zb: int) -> float:
        kp = self.h0.get("kpair", {})
        a, b = SYMBOLS[za], SYMBOLS[zb]
        return float(kp.get(f"{a}-{b}", kp.get(f"{b}-{a}", 1.0)))


@lru_cache(maxsize=None)
def get(name: str) -> Parametrization:
    return Parametrization(name)


GFN1_XTB = Param(_json("gfn1"))
GFN2_XTB = Param(_json("gfn2"))


def stong_table(ng: int) -> np.ndarray:
    """STO-nG table (2, ntype, ng): coefficients and exponents."""
    return np.asarray(_json("stong")[str(ng)], dtype=np.float64)
