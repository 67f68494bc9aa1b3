"""
Calculator facade of the HIP path: ``Calculator(numbers, par)`` for any
GFN1/GFN2-type parametrization, ``GFN1Calculator`` / ``GFN2Calculator`` with
the shipped ones, with ``get_energy``, ``get_forces`` and the single-point
properties.

Mirrors the reference calculator API (``src/dxtb/_src/calculators/gfn1.py:38-70``,
``gfn2.py:38-69``, ``types/abc.py:88-170``, ``types/base.py:453-884``):

* ``Calculator(numbers, opts=None, device=None, dtype=torch.float64)`` with
  integer ``numbers`` of shape (nat,) or zero-padded (B, nat);
* ``get_energy(positions, chrg=0, spin=None)`` -> total energy, summed over
  atoms: (,) or (B,); ``spin`` is the number of unpaired electrons (scalar
  or one per molecule; None: from the parity of the electron count), split
  into alpha and beta counts as in ``wavefunction/filling.py:180-259``;
* ``get_forces(positions, chrg=0, spin=None, grad_mode="autograd")`` -> -dE/dR;
  like the reference, it requires ``positions.requires_grad`` (RuntimeError
  otherwise); the forces come from the hand-written backward kernels;
* ``get_energy`` is differentiable in the positions (when they require
  gradients): ``torch.autograd.grad(E, positions)`` runs the hand-written
  force kernels (``types/autograd.py:146-150`` differentiates the same way);
* the last result is cached (one entry, as ``types/decorators.py:241-289``)
  and reused for inputs of equal value; ``reset()`` clears the cache
  (``types/base.py:758-763``);
* the single-point properties of the reference's energy calculator
  (``types/energy.py:55-65``): ``get_iterations``, ``get_density``,
  ``get_mo_energies`` (ascending), ``get_occupation``, ``get_coefficients``,
  ``get_charges`` / ``get_mulliken_charges`` (orbital-resolved),
  ``get_bond_orders`` (Wiberg), ``get_potential_energy``, and
  ``get_property(name, ...)`` for each; other properties (Hessian,
  vibrations, polarizabilities, ...) raise NotImplementedError;
* ``opts`` takes the reference's SCF keys: ``scf_mode`` "implicit"
  (broyden1 on the residual, the default here) or "full" (the reference's
  default: the full-mode loop with ``mixer`` "anderson", or "simple";
  "broyden" falls back to Anderson with a warning, as in
  ``scf/unrolling/base.py:81-98``), ``maxiter``, ``x_atol``, ``f_atol``,
  ``x_atol_max``, ``damp``, ``damp_init``, ``damp_generations``,
  ``damp_soft_start``, ``fermi_etemp``; ``scp_mode``, ``int_driver`` and
  ``exclude`` warn where they differ from what this path does, or raise
  where it cannot follow them.

Dispersion (D3/D4) is not part of this path: the energy is the reference's
total energy with ``exclude=["disp"]`` (``types/base.py:612-663``).
"""

from __future__ import annotations

import warnings

import numpy as np
import torch

from . import param
from . import scf as _scf
from .basis import build_batch
from .solvation import GeneralizedBorn


class DtypeError(ValueError):
    pass


class DeviceError(ValueError):
    pass


class SCFConvergenceWarning(UserWarning):
    pass


class SCFConvergenceError(RuntimeError):
    pass


# SCF mode and mixer names and integer codes (constants/labels/scf.py)
_MODE_IMPLICIT, _MODE_FULL = ("default", "implicit"), ("full", "full_tracking", "unrolling")
_MODE_REMOVED = ("implicit_old", "implicit_nonpure", "nonpure", "non-pure", "old")
_MIXERS = {0: "simple", 1: "anderson", 2: "broyden", "linear": "simple", "l": "simple", "simple": "simple",
           "s": "simple", "anderson": "anderson", "a": "anderson", "broyden": "broyden", "b": "broyden"}


def _scf_mode(val) -> str:
    """``calculators/config/scf.py:214-264``: "implicit" or "full"."""
    if isinstance(val, bool) or not isinstance(val, (str, int)):
        raise TypeError(f"The scf_mode must be of type 'int' or 'str', but '{type(val)}' was given.")
    if isinstance(val, int):
        if val == 2:
            raise ValueError("The SCF mode with integer code 2 (non-pure implicit) was removed. Use 'implicit'.")
        if val in (0, 1):
            return "full" if val == 0 else "implicit"
    else:
        v = val.casefold()
        if v in _MODE_IMPLICIT:
            return "implicit"
        if v in _MODE_FULL:
            return "full"
        if v in _MODE_REMOVED:
            raise ValueError(f"The SCF mode '{val}' was removed. Use 'implicit'.")
        if v in ("single-shot", "single_shot", "experimental"):
            raise NotImplementedError(f"The SCF mode '{val}' is not part of this path.")
    raise ValueError(f"Unknown SCF mode '{val}'. Use one of 'implicit', 'full'.")


# components of the reference calculator (calculators/types/base.py:596-671)
# and what this path computes: ES2, ES3, AES2 (GFN2) and repulsion always;
# dispersion (D3/D4, D4SC) never; the GFN1 halogen-bond correction, the
# isotropic electrostatics and short-range basis terms (not part of GFN1/GFN2
# single points with these parametrizations' defaults) never
_EXCLUDE_ALWAYS = {"disp", "d4sc", "hal", "ies", "srb"}
_EXCLUDE_NEVER = {"es2", "es3", "aes2", "rep", "all"}
# halogen-bond donors of the GFN1 correction (classicals/halogen.py)
_HALOGEN_DONORS = (17, 35, 53, 85)


def _check_exclude(val, gfn: str, numbers: np.ndarray) -> None:
    """The reference's ``exclude`` list (``calculators/types/base.py:596-671``):
    components this path cannot switch off raise NotImplementedError; a
    list that keeps a component this path never computes warns."""
    exc = {str(v).casefold() for v in ([val] if isinstance(val, str) else list(val))}
    bad = exc & _EXCLUDE_NEVER
    if bad:
        raise NotImplementedError(f"Excluding {sorted(bad)} is not supported on the MI355X path; it always computes "
                                  "ES2, ES3, AES2 (GFN2) and repulsion.")
    unknown = exc - _EXCLUDE_ALWAYS
    if unknown:
        raise ValueError(f"Unknown component(s) {sorted(unknown)} in 'exclude'.")
    if "disp" not in exc:
        warnings.warn("Dispersion (D3/D4, D4SC) is not computed on the MI355X path: the energy is the reference's "
                      "energy with exclude=['disp'].", UserWarning, stacklevel=4)
    if param.kind(gfn) == "gfn1" and "hal" not in exc and np.isin(numbers, _HALOGEN_DONORS).any():
        warnings.warn("The GFN1 halogen-bond correction is not computed on the MI355X path: the energy is the "
                      "reference's energy with 'hal' excluded.", UserWarning, stacklevel=4)


def _opts_to_scf(opts: dict | None, gfn: str = "gfn2", numbers: np.ndarray | None = None) -> _scf.SCFOptions:
    """The reference's options (``calculators/config/*.py``) as SCF options
    of this path.  Accepted and used: the SCF tolerances, ``maxiter``,
    ``fermi_etemp``, the mixer settings, ``scf_mode``, ``mixer``; accepted
    with a warning where the value differs from what this path does:
    ``scp_mode`` (the potential is iterated: the same fixed point as
    "charge"/"fock"), ``int_driver`` (the HIP integral driver: the same
    integrals as "libcint"), ``exclude`` (see _check_exclude); accepted:
    ``verbosity``, ``batch_mode``, ``force_convergence``."""
    o = _scf.SCFOptions()
    if not opts:
        return o
    mixer = None
    for key, val in opts.items():
        if key in ("f_atol", "f_tol"):
            o.f_tol = float(val)
        elif key in ("x_atol", "x_tol"):
            o.x_tol = float(val)
        elif key in ("x_atol_max", "x_tol_max"):
            o.x_tol_max = float(val)
        elif key == "maxiter":
            o.maxiter = int(val)
        elif key in ("fermi_etemp", "etemp"):
            o.etemp = float(val)
        elif key == "damp":
            o.damp = float(val)
        elif key == "damp_init":
            o.damp_init = float(val)
        elif key == "damp_generations":
            o.damp_generations = int(val)
        elif key == "damp_soft_start":
            o.damp_soft_start = bool(val)
        elif key == "scf_mode":
            o.scf_mode = _scf_mode(val)
        elif key == "mixer":
            m = _MIXERS.get(val.casefold() if isinstance(val, str) else val)
            if m is None:
                raise ValueError(f"Unknown mixer '{val}'.")
            mixer = m
        elif key == "scp_mode":
            # constants/labels/scf.py:144-160: fock 0, charge 1, potential 2
            v = val.casefold() if isinstance(val, str) else val
            if v not in ("potential", "pot", 2):
                if v not in ("charge", "charges", "fock", "fockian", 0, 1):
                    raise ValueError(f"Unknown scp_mode '{val}'.")
                warnings.warn(f"scp_mode '{val}': the MI355X path iterates the potential (the same fixed point).",
                              UserWarning, stacklevel=3)
        elif key == "int_driver":
            # constants/labels/integrals.py:45-54: libcint 0, pytorch 1
            v = val.casefold() if isinstance(val, str) else val
            if v not in ("dxtb", "torch", "pytorch", "autograd", "hip", 1):
                if v not in ("libcint", "c", 0):
                    raise ValueError(f"Unknown integral driver '{val}'.")
                warnings.warn(f"int_driver '{val}': the MI355X path computes the integrals with its HIP driver "
                              "(the same integrals).", UserWarning, stacklevel=3)
        elif key == "exclude":
            _check_exclude(val, gfn, np.zeros(0) if numbers is None else numbers)
        elif key in ("verbosity", "batch_mode", "force_convergence"):
            continue
        else:
            raise ValueError(f"Unknown option '{key}'.")
    if o.scf_mode == "full":
        # scf/unrolling/base.py:81-98: no Broyden in the full mode
        if mixer in (None, "broyden"):
            if mixer == "broyden":
                warnings.warn("Broyden mixer is not implemented for SCF with full gradient tracking. "
                              "Using Anderson mixer instead.", UserWarning, stacklevel=3)
            mixer = "anderson"
        o.mixer = mixer
    return o


class _Calculator:
    gfn = ""

    def __init__(self, numbers: torch.Tensor, *, classical=None, interaction=None, opts: dict | None = None,
                 device=None, dtype: torch.dtype = torch.float64, **_):
        if numbers.dtype not in (torch.int16, torch.int32, torch.int64, torch.long):
            raise DtypeError(f"Atomic numbers must be an integer tensor, got {numbers.dtype}.")
        if dtype != torch.float64:
            raise DtypeError("The MI355X path computes in float64 only.")
        self.device = torch.device(device) if device is not None else numbers.device
        if self.device.type != "cuda":
            raise DeviceError("The MI355X path runs on a GPU device (cuda:N); there is no CPU fallback.")
        self.dtype = dtype
        self.numbers = numbers.to(self.device)
        self.batched = numbers.ndim == 2
        nums = np.atleast_2d(numbers.detach().cpu().numpy())
        self.opts = dict(opts or {})
        self.scf_opts = _opts_to_scf(opts, self.gfn, nums)
        self.force_convergence = bool(self.opts.get("force_convergence", False))
        self.bb = build_batch(nums, self.gfn, self.device)
        # extra interactions (calculators/types/base.py: `interaction` list):
        # generalized Born implicit solvation is supported
        self.solvation = None
        for it in (interaction or []):
            if isinstance(it, GeneralizedBorn):
                self.solvation = it
            else:
                raise NotImplementedError(f"interaction {type(it).__name__} is not supported on this path")
        self.cache: dict = {}
        self._ncalcs = 0

    def reset(self) -> None:
        self.cache.clear()

    # ------------------------------------------------------------------
    def _check(self, positions):
        if positions.dtype != self.dtype:
            raise DtypeError(f"positions have dtype {positions.dtype}, expected {self.dtype}")
        if positions.device != self.device:
            raise DeviceError(f"positions are on {positions.device}, expected {self.device}")

    def _charges(self, chrg):
        B = self.bb.nbatch
        if isinstance(chrg, (int, float)):
            if self.batched and chrg != 0:
                raise ValueError("Cannot set charge for batched calculations with a single float or integer. "
                                 "Please provide a 1D tensor of charges.")
            return np.full(B, float(chrg))
        return np.asarray(chrg.detach().cpu() if torch.is_tensor(chrg) else chrg, dtype=np.float64).reshape(B)

    def _spins(self, spin):
        """Unpaired electrons per molecule (None: from the electron count's
        parity), like the charges a scalar or one value per molecule."""
        if spin is None:
            return None
        B = self.bb.nbatch
        s = np.asarray(spin.detach().cpu() if torch.is_tensor(spin) else spin, dtype=np.float64)
        return np.broadcast_to(s, (B,)).copy() if s.ndim == 0 else s.reshape(B)

    @staticmethod
    def _value_key(v):
        """Hashable value of a charge or spin argument (scalar, sequence or tensor)."""
        if v is None:
            return None
        a = np.asarray(v.detach().cpu() if torch.is_tensor(v) else v, dtype=object)
        return tuple(None if x is None else float(x) for x in a.reshape(-1))

    def _cached(self, positions, chrg, spin):
        """The cached result if it was computed for the same inputs: one entry,
        as the reference's cache (``types/decorators.py:241-289``), validated by
        the values of the positions, charges and spins (tensor identity alone
        would match a new tensor that reuses a freed tensor's id)."""
        c = self.cache.get("singlepoint")
        if c is None:
            return None
        pos, chrg_key, spin_key, res = c
        if (chrg_key != self._value_key(chrg) or spin_key != self._value_key(spin) or pos.shape != positions.shape
                or not torch.equal(pos, positions.detach())):
            return None
        return res

    def singlepoint(self, positions, chrg=0, spin=None):
        self._check(positions)
        res = self._cached(positions, chrg, spin)
        if res is not None:
            return res
        pos = positions.detach().reshape(self.bb.nbatch, self.bb.nat_max, 3)
        state = _scf.GeometryState(self.bb, pos, self._charges(chrg), self.scf_opts, solvation=self.solvation,
                                   spins=self._spins(spin))
        res = _scf.run_scf(state)
        if not res.converged.all():
            msg = f"SCF did not converge for {int((~res.converged).sum())} molecule(s)."
            if self.force_convergence:
                raise SCFConvergenceError(msg)
            warnings.warn(msg, SCFConvergenceWarning, stacklevel=3)
        self.cache["singlepoint"] = (positions.detach().clone(), self._value_key(chrg), self._value_key(spin), res)
        self._ncalcs += 1
        return res

    def get_energy(self, positions, chrg=0, spin=None, **_):
        """Total energy.  With ``positions.requires_grad`` (and grad mode on)
        the energy is differentiable in the positions: its backward runs the
        hand-written force kernels, so ``torch.autograd.grad(E, positions)``
        is the gradient the reference gets from autograd
        (``calculators/types/autograd.py:146-150``).  First derivatives only."""
        if positions.requires_grad and torch.is_grad_enabled():
            e = _EnergyFunction.apply(positions, self, chrg, spin)
        else:
            e = self.singlepoint(positions, chrg, spin).energy
        return e if self.batched else e[0]

    def get_forces(self, positions, chrg=0, spin=None, grad_mode="autograd", **_):
        if grad_mode not in ("autograd", "backward", "functorch", "row"):
            raise ValueError(f"Unknown grad_mode '{grad_mode}'.")
        if not positions.requires_grad:
            raise RuntimeError("Position tensor needs `requires_grad=True` to compute forces.")
        res = self.singlepoint(positions, chrg, spin)
        f = _scf.forces(res)
        return f.reshape(positions.shape)

    # single-point properties (calculators/types/energy.py:55-65, abc.py:54-560)
    implemented_properties = ["energy", "forces", "bond_orders", "charges", "coefficients", "density",
                              "iterations", "mo_energies", "occupation"]

    def get_implemented_properties(self) -> list[str]:
        return list(self.implemented_properties)

    def _out(self, t):
        return t if self.batched else t[0]

    def get_potential_energy(self, positions, chrg=0, spin=None, **_):
        return self.get_energy(positions, chrg, spin)

    def get_iterations(self, positions, chrg=0, spin=None, **_):
        res = self.singlepoint(positions, chrg, spin)
        return self._out(torch.as_tensor(res.iterations, device=self.device))

    def get_density(self, positions, chrg=0, spin=None, **_):
        """Density matrix P (.., nao, nao) of the converged SCF."""
        return self._out(self.singlepoint(positions, chrg, spin).density)

    @staticmethod
    def _mo_order(res):
        # the eigensolver leaves the orbitals unsorted; the reference's eigh
        # returns them by ascending energy, padding orbitals (energy 0) included
        return torch.argsort(res.emo, dim=-1, stable=True)

    def get_mo_energies(self, positions, chrg=0, spin=None, **_):
        res = self.singlepoint(positions, chrg, spin)
        return self._out(torch.gather(res.emo, -1, self._mo_order(res)))

    def get_occupation(self, positions, chrg=0, spin=None, **_):
        """Fermi occupations (.., 2, nao) of the alpha and beta orbitals, in
        the order of get_mo_energies."""
        res = self.singlepoint(positions, chrg, spin)
        idx = self._mo_order(res).unsqueeze(1).expand_as(res.occupation)
        return self._out(torch.gather(res.occupation, -1, idx))

    def get_coefficients(self, positions, chrg=0, spin=None, **_):
        """Orbital coefficients (.., nao, nao), one orbital per column, in the
        order of get_mo_energies (signs are arbitrary, as for any eigensolver)."""
        res = self.singlepoint(positions, chrg, spin)
        idx = self._mo_order(res).unsqueeze(1).expand_as(res.coeffs)
        return self._out(torch.gather(res.coeffs, -1, idx))

    def get_charges(self, positions, chrg=0, spin=None, **_):
        """Orbital-resolved Mulliken charges n0 - diag(P S) (.., nao)
        (``scf/pure/conversions.py:203-252``)."""
        res = self.singlepoint(positions, chrg, spin)
        st = res.state
        return self._out(st.n0 - (res.density * st.S).sum(-1))

    def get_mulliken_charges(self, positions, chrg=0, spin=None, **kw):
        return self.get_charges(positions, chrg, spin, **kw)

    def get_bond_orders(self, positions, chrg=0, spin=None, **_):
        """Wiberg bond orders (.., nat, nat): the atom blocks of
        (P S) * (P S)^T summed, diagonal zero (``wavefunction/wiberg.py:33-60``)."""
        res = self.singlepoint(positions, chrg, spin)
        st = res.state
        PS = res.density @ st.S
        W = PS * PS.transpose(-1, -2)
        bb = st.bb
        live = torch.arange(bb.nao_max, device=W.device)[None, :] < bb.dev["nao"][:, None].long()
        onehot = torch.nn.functional.one_hot(bb.dev["ao_atom"], bb.nat_max).to(W.dtype) * live.unsqueeze(-1)
        wbo = onehot.transpose(1, 2) @ W @ onehot
        wbo.diagonal(dim1=-2, dim2=-1).fill_(0.0)
        return self._out(wbo)

    def get_property(self, name, positions, chrg=0, spin=None, **kw):
        getters = {"energy": self.get_energy, "forces": self.get_forces, "bond_orders": self.get_bond_orders,
                   "charges": self.get_charges, "coefficients": self.get_coefficients,
                   "density": self.get_density, "iterations": self.get_iterations,
                   "mo_energies": self.get_mo_energies, "occupation": self.get_occupation}
        if name not in getters:
            raise NotImplementedError(f"Property '{name}' is not implemented on the MI355X path.")
        return getters[name](positions, chrg, spin, **kw)


class _EnergyFunction(torch.autograd.Function):
    """SCF total energy as a function of the positions.  Backward: the
    gradient of the converged free energy from the hand-written kernels
    (integral derivatives contracted with P and W, H0/CN, electrostatics,
    repulsion and solvation terms; ``scf.forces``), times the incoming
    gradient of every molecule."""

    @staticmethod
    def forward(ctx, positions, calc, chrg, spin):
        res = calc.singlepoint(positions, chrg, spin)
        ctx.res = res
        ctx.shape = positions.shape
        return res.energy.clone()

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_e):
        g = -_scf.forces(ctx.res)  # (B, nat_max, 3) = dE/dR
        g = g * grad_e.reshape(-1, 1, 1).to(g.dtype)
        return g.reshape(ctx.shape), None, None, None


class Calculator(_Calculator):
    """Calculator for a GFN1- or GFN2-type parametrization
    (``calculators/base.py``; constructor ``types/base.py:453-575``):
    ``Calculator(numbers, par, *, interaction=None, opts=None, device=None,
    dtype=torch.float64)``.  ``par`` is a ``Param`` (``param.GFN1_XTB``,
    ``param.GFN2_XTB`` or one read with ``Param.from_file``), its record as a
    dict, or a parametrization file; its ``meta.name`` selects the GFN1 or
    GFN2 form (``integral/factory.py:60-83``), every parameter is its own."""

    def __init__(self, numbers: torch.Tensor, par, *, classical=None, interaction=None, opts: dict | None = None,
                 device=None, dtype: torch.dtype = torch.float64, **kw):
        self.gfn = param.register(par)
        self.par = param.get(self.gfn)
        super().__init__(numbers, classical=classical, interaction=interaction, opts=opts, device=device,
                         dtype=dtype, **kw)


class GFN1Calculator(Calculator):
    """GFN1-xTB on the MI355X path (``calculators/gfn1.py:38-70``)."""

    def __init__(self, numbers: torch.Tensor, *, classical=None, interaction=None, opts: dict | None = None,
                 device=None, dtype: torch.dtype = torch.float64, **kw):
        super().__init__(numbers, "gfn1", classical=classical, interaction=interaction, opts=opts, device=device,
                         dtype=dtype, **kw)


class GFN2Calculator(Calculator):
    """GFN2-xTB on the MI355X path (``calculators/gfn2.py:38-69``)."""

    def __init__(self, numbers: torch.Tensor, *, classical=None, interaction=None, opts: dict | None = None,
                 device=None, dtype: torch.dtype = torch.float64, **kw):
        super().__init__(numbers, "gfn2", classical=classical, interaction=interaction, opts=opts, device=device,
                         dtype=dtype, **kw)
