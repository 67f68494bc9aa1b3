"""
Seeded synthetic molecules for the benchmark configurations.

Geometry rule ("self-avoiding growth"): the first atom sits at the origin;
every further atom bonds to a randomly chosen existing atom at a length drawn
from U(2.0, 2.9) Bohr in a uniformly random direction and is accepted if it is
at least 3.5 Bohr from every other atom (retries, then a new parent).
Elements are drawn i.i.d. from the configuration's ratios; the hydrogen count
is adjusted by one where needed so that the electron count (valence electrons
minus the charge) is even (closed shell).
"""

from __future__ import annotations

import math

import numpy as np

VALENCE = {1: 1, 6: 4, 7: 5, 8: 6, 9: 7, 16: 6, 17: 7}
Z = {"H": 1, "C": 6, "N": 7, "O": 8, "F": 9, "S": 16, "Cl": 17}

RATIOS = {
    0: {"H": 0.5, "C": 0.3, "N": 0.1, "O": 0.1},
    1: {"H": 0.45, "C": 0.3, "N": 0.1, "O": 0.1, "S": 0.05},
    2: {"H": 0.45, "C": 0.28, "N": 0.10, "O": 0.10, "F": 0.03, "S": 0.02},
    3: {"H": 0.45, "C": 0.28, "N": 0.1, "O": 0.1, "F": 0.03, "S": 0.02, "Cl": 0.02},
}


def seed(config: int) -> int:
    return 20260000 + config


def draw_elements(rng: np.random.Generator, nat: int, ratios: dict, charge: int = 0) -> np.ndarray:
    syms = list(ratios)
    p = np.array([ratios[s] for s in syms], dtype=np.float64)
    p /= p.sum()
    z = np.array([Z[s] for s in syms])[rng.choice(len(syms), size=nat, p=p)]
    nel = sum(VALENCE[int(x)] for x in z) - charge
    if nel % 2 == 1:
        # flip one atom to/from hydrogen to fix the parity
        heavy = np.nonzero(z != 1)[0]
        hyd = np.nonzero(z == 1)[0]
        odd_heavy = [i for i in heavy if VALENCE[int(z[i])] % 2 == 0]
        if len(hyd) > 0:
            # H (1 e) -> C (4 e) changes the count by 3 (odd)
            z[hyd[rng.integers(len(hyd))]] = 6
        elif odd_heavy:
            z[odd_heavy[rng.integers(len(odd_heavy))]] = 1
        else:
            z[heavy[rng.integers(len(heavy))]] = 6 if VALENCE[int(z[heavy[0]])] % 2 else 7
    return z


def grow_geometry(rng: np.random.Generator, nat: int, bond=(2.0, 2.9), dmin: float = 3.5,
                  tries: int = 50) -> np.ndarray:
    pos = np.zeros((nat, 3))
    batch = 16  # candidates drawn together; the first acceptable one is taken
    for k in range(1, nat):
        while True:
            parent = rng.integers(k, size=batch)
            v = rng.normal(size=(batch, 3))
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            cand = pos[parent] + rng.uniform(*bond, size=(batch, 1)) * v
            d = np.linalg.norm(pos[None, :k, :] - cand[:, None, :], axis=2)
            d[np.arange(batch), parent] = np.inf
            ok = np.nonzero(d.min(1) >= dmin)[0]
            if len(ok):
                pos[k] = cand[ok[0]]
                break
    return pos


def molecule(rng, nat: int, config: int, charge: int = 0):
    z = draw_elements(rng, nat, RATIOS[config], charge)
    return z, grow_geometry(rng, nat)


def config_batch(config: int, count: int | None = None, nat: int | None = None):
    """(numbers list, positions list, charges list) of a configuration."""
    rng = np.random.default_rng(seed(config))
    nums, poss, chg = [], [], []
    if config == 0:
        count = count or 1
        sizes = [nat or 24] * count
    elif config == 1:
        count = count or 512
        sizes = [nat or 60] * count
    elif config == 2:
        count = count or 1024
        sizes = rng.integers(10, 121, size=count).tolist()
    elif config == 3:
        count = count or 100000
        sizes = np.clip(np.round(rng.lognormal(math.log(40.0), 0.7, size=count)), 5, 400).astype(int).tolist()
    else:
        raise ValueError(config)
    for n in sizes:
        c = 0
        if config == 3:
            c = int(rng.choice([0, 1, -1], p=[0.8, 0.1, 0.1]))
        z, p = molecule(rng, int(n), config, c)
        nums.append(z)
        poss.append(p)
        chg.append(c)
    return nums, poss, chg


def config3_solvation(count: int, fraction: float = 0.2) -> np.ndarray:
    """Requests of the serving stream that enable implicit solvation (GBSA,
    eps = 78.4): a seeded fraction, drawn independently of the molecules."""
    rng = np.random.default_rng(seed(3) + 7)
    return rng.random(count) < fraction


def water_cluster(nmol: int = 667, density_gcc: float = 1.0, seed_: int = 20260004):
    """Rigid waters (O-H 1.81 Bohr, 104.5 deg) at random orientations in a cube."""
    rng = np.random.default_rng(seed_)
    mass = 18.015 * 1.66053906660e-24  # g per molecule
    vol_cm3 = nmol * mass / density_gcc
    edge = (vol_cm3 ** (1.0 / 3.0)) * 1e8 / 0.529177210903  # Bohr
    th = math.radians(104.5) / 2
    local = np.array([[0, 0, 0], [1.81 * math.sin(th), 0, 1.81 * math.cos(th)],
                      [-1.81 * math.sin(th), 0, 1.81 * math.cos(th)]])
    oxy = []
    while len(oxy) < nmol:
        c = rng.uniform(0, edge, size=3)
        if oxy:
            d = np.linalg.norm(np.array(oxy) - c, axis=1)
            if d.min() < 4.9:
                continue
        oxy.append(c)
    nums, pos = [], []
    for c in oxy:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        for k, zz in enumerate((8, 1, 1)):
            nums.append(zz)
            pos.append(c + R @ local[k])
    return np.array(nums), np.array(pos)


def pad(nums_list, pos_list):
    """Zero-padded (B, nat_max) numbers and (B, nat_max, 3) positions."""
    nmax = max(len(z) for z in nums_list)
    B = len(nums_list)
    nums = np.zeros((B, nmax), dtype=np.int64)
    pos = np.zeros((B, nmax, 3))
    for i, (z, p) in enumerate(zip(nums_list, pos_list)):
        nums[i, : len(z)] = z
        pos[i, : len(z)] = p
    return nums, pos
