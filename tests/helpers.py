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


def read_xyz(name):
    """Plain xyz file under tests/golden (symbol x y z per line)."""
    with open(os.path.join(GOLDEN, name)) as f:
        lines = f.read().splitlines()
    n = int(lines[0])
    pos, Z = [], []
    for line in lines[2:2 + n]:
        t = line.split()
        Z.append(ATOMIC_NUMBERS[t[0]])
        pos.append([float(x) for x in t[1:4]])
    return np.array(pos), np.array(Z, np.int32)


def small_hydrocarbon(name):
    """Idealised starting geometries of the molecules the reference builds
    with ase.build.molecule (relaxation brings them to the REBO2 minimum)."""
    s3 = 1.0 / np.sqrt(3.0)
    if name == "CH4":
        P = [[0, 0, 0]] + [[1.09 * a * s3, 1.09 * b * s3, 1.09 * a * b * s3]
                           for a in (1, -1) for b in (1, -1)]
        Z = [6, 1, 1, 1, 1]
    elif name == "C2H2":
        P = [[0, 0, -0.60], [0, 0, 0.60], [0, 0, -1.66], [0, 0, 1.66]]
        Z = [6, 6, 1, 1]
    elif name == "C2H4":
        P = [[-0.667, 0, 0], [0.667, 0, 0], [-1.23, 0.92, 0], [-1.23, -0.92, 0],
             [1.23, 0.92, 0], [1.23, -0.92, 0]]
        Z = [6, 6, 1, 1, 1, 1]
    elif name == "C2H6":
        P, Z = [[0, 0, -0.765], [0, 0, 0.765]], [6, 6]
        for k in range(3):
            t = 2 * np.pi * k / 3
            P.append([1.02 * np.cos(t), 1.02 * np.sin(t), -1.16])
            P.append([1.02 * np.cos(t + np.pi / 3), 1.02 * np.sin(t + np.pi / 3), 1.16])
            Z += [1, 1]
    elif name == "C6H6":
        P, Z = [], []
        for k in range(6):
            t = 2 * np.pi * k / 6
            P.append([1.39 * np.cos(t), 1.39 * np.sin(t), 0])
            P.append([2.48 * np.cos(t), 2.48 * np.sin(t), 0])
            Z += [6, 1]
    else:
        raise KeyError(name)
    return np.array(P, float), np.array(Z, np.int32)


def in_vacuum_box(pos, vacuum=5.0):
    """Shift a molecule into a non-periodic orthorhombic box with `vacuum`
    on every side."""
    pos = pos - pos.min(0) + vacuum
    return pos, np.diag(pos.max(0) + vacuum), np.array([False] * 3)


def bop_cutoff(params):
    """Neighbour-list range of a bond-order parameter dict: r2 unscreened; for
    the screened variants sqrt(C_dr_cut) times the largest of r2, or2, bor2
    over the pairs, C_dr_cut = Cmax^2 / (4 (Cmax - 1)) for Cmax > 2, else 1
    (default_bind_to_func.f90:44-150)."""
    r2 = np.atleast_1d(params["r2"]).astype(float)
    if "or1" not in params:
        return float(r2.max())
    cm = np.atleast_1d(params["Cmax"]).astype(float)
    x = np.where(cm > 2.0, cm ** 2 / (4 * (cm - 1)), 1.0)
    mx = max(r2.max(), np.atleast_1d(params["or2"]).max(), np.atleast_1d(params["bor2"]).max())
    return float(np.sqrt(x.max()) * mx)
