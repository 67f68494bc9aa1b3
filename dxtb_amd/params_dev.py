"""
Device tables of a parametrization for the HIP kernels (``dxtb_params_t``).

Per unique shell (row order of ``basis.ShellTable``): self-energy level and
CN shift (``xtb/base.py:146-159``, in Hartree), shell polynomial, Hubbard
parameter gam*lgam (``coulomb/secondorder.py:1054-1108``), third-order
parameter (``coulomb/thirdorder.py:183-245``), valence flag; the shell-pair
scaling K^{ll'} (``xtb/gfn1.py:66-165``, ``xtb/gfn2.py:67-156``); per element:
H0 radii, electronegativity, D3 covalent radii, pair scaling kpair,
repulsion and AES2 parameters.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, param
from .basis import shell_table

NZ = param.MAXZ + 1
_AVG = {"harmonic": 0, "arithmetic": 1, "geometric": 2}
_LAB = "spdfg"


def _hscale(gfn: str, table) -> np.ndarray:
    par = param.get(gfn)
    keys = par.h0["shell"]
    nu = len(table.l)
    lab = [_LAB[int(l)] for l in table.l]
    zs = table.z
    within = np.zeros(nu, dtype=np.int64)
    for u in range(nu):
        within[u] = u - table.first[zs[u]]
    k = np.ones((nu, nu))
    if par.kind == "gfn1":
        val = np.array([par.valence(int(zs[u]))[within[u]] for u in range(nu)])
        kii = np.array([par.h0["kpol"] if not val[u] else keys.get(lab[u] * 2, 1.0) for u in range(nu)])
        for i in range(nu):
            for j in range(nu):
                if val[i] and val[j]:
                    v = keys.get(lab[i] + lab[j], keys.get(lab[j] + lab[i], 0.5 * (kii[i] + kii[j])))
                else:
                    v = 0.5 * (kii[i] + kii[j])
                k[i, j] = v
        return k
    wexp = par.h0["wexp"]
    zeta = np.array([par.element(int(zs[u]))["slater"][within[u]] for u in range(nu)])
    for i in range(nu):
        for j in range(nu):
            a, b = lab[i], lab[j]
            kij = keys.get(a + b, keys.get(b + a, 0.5 * (keys.get(a + a, 1.0) + keys.get(b + b, 1.0))))
            k[i, j] = kij * (2.0 * math.sqrt(zeta[i] * zeta[j]) / (zeta[i] + zeta[j])) ** wexp
    return k


class DeviceParams:
    """All tables of one parametrization on one device plus the C struct."""

    def __init__(self, gfn: str, device: torch.device) -> None:
        par = param.get(gfn)
        tab = shell_table(gfn)
        self.gfn = gfn
        nu = len(tab.l)
        lev, kcn, shp, hub, hd3, val = (np.zeros(nu) for _ in range(6))
        third = par.raw["thirdorder"]
        shell_es3 = third.get("shell") not in (False, None)
        for u in range(nu):
            z = int(tab.z[u])
            w = u - int(tab.first[z])
            e = par.element(z)
            lev[u] = e["levels"][w] * param.EV2AU
            kcn[u] = e["kcn"][w] * param.EV2AU
            shp[u] = e["shpoly"][w]
            hub[u] = e["gam"] * e["lgam"][w]
            hd3[u] = e["gam3"] * third["shell"][_LAB[int(tab.l[u])]] if shell_es3 else 0.0
            val[u] = 1.0 if par.valence(z)[w] else 0.0
        zarr = lambda key: np.array(  # noqa: E731
            [par.element(z)[key] if par.has(z) else 0.0 for z in range(NZ)], dtype=np.float64
        )
        kp = np.ones((NZ, NZ))
        for a in range(1, NZ):
            for b in range(1, NZ):
                kp[a, b] = par.kpair(a, b)
        rep = par.raw["repulsion"]["effective"]
        if "enscale" in rep:
            raise NotImplementedError("repulsion EN scaling is not used by GFN1/GFN2")
        t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=device)  # noqa: E731
        self.t = {
            "ush_lev": t(lev), "ush_kcn": t(kcn), "ush_shpoly": t(shp), "ush_hub": t(hub), "ush_hd3": t(hd3),
            "ush_valence": torch.as_tensor(val.astype(np.int32), device=device),
            "hscale": t(_hscale(gfn, tab)),
            "z_rad": t(param.ATOMIC_RADII), "z_en": t(zarr("en")), "z_rcov": t(param.COV_D3),
            "kpair": t(kp), "z_gam3": t(zarr("gam3")), "z_arep": t(zarr("arep")), "z_zeff": t(zarr("zeff")),
            "z_mprad": t(zarr("mprad")), "z_mpvcn": t(zarr("mpvcn")),
            "z_dkernel": t(zarr("dkernel")), "z_qkernel": t(zarr("qkernel")),
        }
        s = _lib.DxtbParams()
        s.gfn = 1 if par.kind == "gfn1" else 2
        s.es3_shell = 1 if shell_es3 else 0
        s.es2_average = _AVG[par.raw["charge"]["effective"]["average"]]
        s.rep_light = 1 if "klight" in rep else 0
        s.enscale = float(par.h0.get("enscale", 0.0))
        s.es2_gexp = float(par.raw["charge"]["effective"]["gexp"])
        s.rep_kexp = float(rep["kexp"])
        s.rep_klight = float(rep.get("klight", 0.0))
        mp = par.raw.get("multipole", {}).get("damped", {})
        s.aes_dmp3 = float(mp.get("dmp3", 0.0))
        s.aes_dmp5 = float(mp.get("dmp5", 0.0))
        s.aes_kexp = float(mp.get("kexp", 0.0))
        s.aes_shift = float(mp.get("shift", 0.0))
        s.aes_rmax = float(mp.get("rmax", 0.0))
        for name, tensor in self.t.items():
            setattr(s, name, tensor.data_ptr())
        self.struct = s


_CACHE: dict = {}


def device_params(gfn: str, device) -> DeviceParams:
    key = (gfn, str(torch.device(device)))
    if key not in _CACHE:
        _CACHE[key] = DeviceParams(gfn, torch.device(device))
    return _CACHE[key]
