"""Test helpers: fixture readers and finite-difference checks.

read_cfg parses the AtomEye extended CFG files the reference's tests use
(tests/aC.cfg, tests/aC_small.cfg): H0 rows are lattice vectors, positions
are fractional.
"""
import os

import numpy as np

from atomistica_amd.parameters import ATOMIC_NUMBERS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_cfg(name):
    path = os.path.join(GOLDEN, name)
    H = np.zeros((3, 3))
    nat = None
    entry = None
    sym = None
    pos, Z = [], []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            if s.startswith("Number of particles"):
                nat = int(s.split("=")[1])
            elif s.startswith("H0("):
                a = int(s[3]) - 1
                b = int(s[5]) - 1
                H[a, b] = float(s.split("=")[1].split()[0])
            elif s.startswith("entry_count"):
                entry = int(s.split("=")[1])
            elif "=" in s:
                continue
            else:
                toks = s.split()
                if len(toks) == 1:
                    try:
                        float(toks[0])
                    except ValueError:
                        sym = toks[0]
                    continue
                pos.append([float(t) for t in toks[:3]])
                Z.append(ATOMIC_NUMBERS[sym])
    pos = np.array(pos) @ H
    assert nat is None or len(pos) == nat
    return pos, H, np.array([True, True, True]), np.array(Z, np.int32)


def fd_forces(energy, pos, atoms, h=1e-5):
    """Central finite-difference forces -dE/dr for the listed atoms."""
    out = np.zeros((len(atoms), 3))
    for n, i in enumerate(atoms):
        for x in range(3):
            p = pos.copy()
            p[i, x] += h
            ep = energy(p)
            p[i, x] -= 2 * h
            em = energy(p)
            out[n, x] = -(ep - em) / (2 * h)
    return out


def fd_virial(energy, pos, cell, h=1e-6):
    """W_ab = dE/d(eps_ab) by central differences of a homogeneous strain
    (the quantity the reference returns as wpot, tests.py:83-120)."""
    W = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            F = np.eye(3) + eps
            ep = energy(pos @ F.T, cell @ F.T)
            F = np.eye(3) - eps
            em = energy(pos @ F.T, cell @ F.T)
            W[a, b] = (ep - em) / (2 * h)
    return W


from oracle.oracle import read_setfl_tables  # noqa: E402,F401
