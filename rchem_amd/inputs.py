"""Inputs of the benchmark workloads: geometries and synthetic shells.

  * ``water_crawford()``     -- the geometry rchem runs (water_crawford.xyz,
                                 bohr; rchem.rs:29-34), STO-3G, 7 functions.
  * ``water_cluster(...)``   -- (H2O)_n: O-H 0.96 A, H-O-H 104.5 deg, uniform
                                 random orientation, on a jittered 3.1 A cubic
                                 lattice, for the 6-31G Fock-build and gradient
                                 workloads ((H2O)_64: 832 functions; (H2O)_32: 416).
  * ``random_molecules(...)``-- 512 molecules of 3-30 H/C/N/O atoms placed by a
                                 1.0-1.5 A random walk with 0.9 A minimum
                                 separation, even electron count (STO-3G).
  * ``synthetic_shells(...)``-- 2,000 shells in a 20-bohr cube, l in {0,1,2}
                                 with weights .5/.35/.15, 3 primitives each,
                                 exponents log-uniform in [0.1, 100],
                                 coefficients uniform in (0.1, 1] normalised to
                                 unit self-overlap of the x^l component.

Everything is seeded (numpy PCG64), coordinates are returned in bohr.
"""
from __future__ import annotations

import math

import numpy as np

BOHR_PER_ANGSTROM = 1.0 / 0.529177249

# water_crawford.xyz (bohr), the input of rchem's driver (rchem.rs:29).
WATER_CRAWFORD = np.array([[0.000000000000, -0.143225816552, 0.000000000000],
                           [1.638036840407, 1.136548822547, -0.000000000000],
                           [-1.638036840407, 1.136548822547, 0.000000000000]])
WATER_Z = [8, 1, 1]


def read_xyz(path, to_bohr=False):
    """Minimal xyz reader (what rchem takes from chemfiles, rchem.rs:29-34):
    element symbols and coordinates, used as given unless to_bohr."""
    sym = {"H": 1, "C": 6, "N": 7, "O": 8}
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0].split()[0])
    Z, xyz = [], []
    for ln in lines[2:2 + n]:
        parts = ln.split()
        Z.append(sym[parts[0]])
        xyz.append([float(x) for x in parts[1:4]])
    xyz = np.array(xyz)
    if to_bohr:
        xyz = xyz * BOHR_PER_ANGSTROM
    return Z, xyz


def water_crawford():
    return list(WATER_Z), WATER_CRAWFORD.copy()


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(dims=(4, 4, 4), seed=64, spacing=3.1, jitter=0.15, r_oh=0.96,
                  angle_deg=104.5):
    """(H2O)_n on a jittered cubic lattice; returns (Z, xyz in bohr)."""
    rng = np.random.default_rng(seed)
    half = math.radians(angle_deg) / 2.0
    mol = np.array([[0.0, 0.0, 0.0],
                    [r_oh * math.sin(half), 0.0, r_oh * math.cos(half)],
                    [-r_oh * math.sin(half), 0.0, r_oh * math.cos(half)]])
    Z, xyz = [], []
    for i in range(dims[0]):
        for j in range(dims[1]):
            for k in range(dims[2]):
                centre = np.array([i, j, k], dtype=float) * spacing
                centre += rng.uniform(-jitter, jitter, 3)
                R = _random_rotation(rng)
                for a in range(3):
                    xyz.append(centre + R @ mol[a])
                Z.extend([8, 1, 1])
    return Z, np.array(xyz) * BOHR_PER_ANGSTROM


def random_molecules(count=512, seed=512, nmin=3, nmax=30):
    """Random-walk H/C/N/O molecules with an even electron count."""
    rng = np.random.default_rng(seed)
    elements = [1, 6, 7, 8]
    out = []
    while len(out) < count:
        n = int(rng.integers(nmin, nmax + 1))
        Z = [int(elements[rng.integers(0, 4)]) for _ in range(n)]
        if sum(Z) % 2:
            continue
        pos = [np.zeros(3)]
        while len(pos) < n:
            for _ in range(100):
                step = rng.uniform(1.0, 1.5)
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                cand = pos[-1] + step * d
                if min(np.linalg.norm(cand - p) for p in pos) >= 0.9:
                    pos.append(cand)
                    break
            else:
                pos = [np.zeros(3)]
        out.append((Z, np.array(pos) * BOHR_PER_ANGSTROM))
    return out


def synthetic_shells(nshell=2000, seed=2000, box=20.0, nprim=3):
    """Shells (centres, l, nprim, exps, coefs) of the ERI microbenchmark."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, box, (nshell, 3))
    ls = rng.choice([0, 1, 2], size=nshell, p=[0.5, 0.35, 0.15]).astype(np.int32)
    exps = np.exp(rng.uniform(math.log(0.1), math.log(100.0), (nshell, nprim)))
    coefs = 1.1 - rng.uniform(0.1, 1.0, (nshell, nprim))  # uniform in (0.1, 1]
    # unit self-overlap of the x^l component under rchem's primitive norms:
    # <x^l|x^l> = sum_ij c_i c_j (2 sqrt(a_i a_j)/(a_i + a_j))^(l + 3/2)
    for s in range(nshell):
        a = exps[s]
        l = ls[s]
        M = (2.0 * np.sqrt(np.outer(a, a)) / np.add.outer(a, a)) ** (l + 1.5)
        coefs[s] /= math.sqrt(coefs[s] @ M @ coefs[s])
    npr = np.full(nshell, nprim, dtype=np.int32)
    return centres, ls, npr, exps.reshape(-1), coefs.reshape(-1)


def nocc_of(Z):
    ne = int(sum(Z))
    assert ne % 2 == 0, "odd electron count"
    return ne // 2
