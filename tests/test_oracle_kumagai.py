"""The oracle's Kumagai forms (kumagai_func.f90 through the generic
bond-order kernel) against what the mathematics and the reference's tests
fix:

  - perfect diamond Si in closed form (4 bonds at a*sqrt(3)/4, fc = 1,
    cos = -1/3, h(0) = 1), written here from the paper's formulas;
  - Ec = 4.630 eV, a0 = 5.429 A, C11 = 166.4, C12 = 65.3 and the unrelaxed
    C44 = 120.9 GPa (tests/test_bulk_properties.py:118-121, 5 % there);
  - forces and virial are the derivatives of the energy (central FD) on a
    disordered Si solid (tests/test_forces_and_virial.py:121);
  - the dimer energy and force are continuous (tests/test_dimers.py:112-134).
"""
import math

import numpy as np
from scipy.optimize import minimize_scalar

from oracle import oracle as O
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import fd_forces, fd_virial, in_vacuum_box

K = P.Kumagai_CompMaterSci_39_457_Si
GPA = 1.0 / 160.21766208


def energy(pos, cell, pbc, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, K["r2"])
    return O.tersoff(K, pos, np.zeros(len(pos), np.int32), cell, nl, mask=mask, per_at=per_at)


def closed_form(a):
    r = a * math.sqrt(3) / 4
    t = K["h"] + 1.0 / 3.0
    go = K["c2"] * t / (K["c3"] + t * t)
    g = K["c1"] + t * go * (1 + K["c4"] * math.exp(-K["c5"] * t * t))
    zeta = 3 * g
    b = (1 + zeta ** K["eta"]) ** (-K["delta"])
    return 2 * (K["A"] * math.exp(-K["lambda1"] * r) - b * K["B"] * math.exp(-K["lambda2"] * r))


def test_diamond_closed_form():
    for a in (5.40, 5.429, 5.46):
        pos, cell, pbc, Z = W.diamond_si(2, a=a, sigma=0.0)
        e, f, w = energy(pos, cell, pbc)
        assert abs(e / len(pos) - closed_form(a)) < 1e-12
        assert np.abs(f).max() < 1e-10


def test_diamond_cohesive_energy_and_lattice_constant():
    res = minimize_scalar(lambda a: closed_form(a), bounds=(5.3, 5.6), method="bounded", options={"xatol": 1e-8})
    assert abs(-res.fun - 4.630) / 4.630 < 5e-4
    assert abs(res.x - 5.429) / 5.429 < 5e-4

    def e_of_a(a):
        pos, cell, pbc, Z = W.diamond_si(2, a=a, sigma=0.0)
        return energy(pos, cell, pbc)[0] / len(pos)
    res2 = minimize_scalar(e_of_a, bounds=(5.3, 5.6), method="bounded", options={"xatol": 1e-8})
    assert abs(res2.x - res.x) < 1e-5


def test_diamond_elastic_constants():
    a0 = minimize_scalar(lambda a: closed_form(a), bounds=(5.3, 5.6), method="bounded",
                         options={"xatol": 1e-10}).x
    pos0, cell0, pbc, Z = W.diamond_si(2, a=a0, sigma=0.0)

    def stress(F):
        pos, cell = pos0 @ F.T, cell0 @ F.T
        return energy(pos, cell, pbc)[2] / abs(np.linalg.det(cell))
    h = 1e-5
    Fp, Fm = np.eye(3), np.eye(3)
    Fp[0, 0] += h
    Fm[0, 0] -= h
    d = (stress(Fp) - stress(Fm)) / (2 * h) / GPA
    assert abs(d[0, 0] - 166.4) / 166.4 < 0.05
    assert abs(d[1, 1] - 65.3) / 65.3 < 0.05
    # unrelaxed C44: simple shear without internal relaxation
    Sp, Sm = np.eye(3), np.eye(3)
    Sp[1, 2] += h
    Sm[1, 2] -= h
    c44 = (stress(Sp) - stress(Sm))[1, 2] / (2 * h) / GPA
    assert abs(c44 - 120.9) / 120.9 < 0.05


def test_fd_forces_and_virial_disordered():
    pos, cell, pbc, Z = W.diamond_si(2, sigma=0.15, seed=21)
    e, f, w = energy(pos, cell, pbc)
    fd = fd_forces(lambda p: energy(p, cell, pbc)[0], pos, range(0, len(pos), 7))
    assert np.max(np.abs(fd - f[::7][: len(fd)])) < 1e-6 * max(1.0, np.abs(f).max())
    Wfd = fd_virial(lambda p, c: energy(p, c, pbc)[0], pos, cell)
    assert np.max(np.abs(Wfd - w)) < 1e-5 * max(1.0, np.abs(w).max())
    assert np.abs(f.sum(axis=0)).max() < 1e-9


def test_mask_decomposition_and_per_atom():
    pos, cell, pbc, Z = W.diamond_si(2, sigma=0.1, seed=5)
    e, f, w, epa, wpa = energy(pos, cell, pbc, per_at=True)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e1, f1, w1 = energy(pos, cell, pbc, mask=m)
    e2, f2, w2 = energy(pos, cell, pbc, mask=1 - m)
    assert abs(e - e1 - e2) < 1e-9 and np.abs(f - f1 - f2).max() < 1e-9
    assert abs(epa.sum() - e) < 1e-9 and np.abs(wpa.sum(axis=0) - w).max() < 1e-9


def test_dimer_continuity():
    # tests/test_dimers.py:112-134: Si2 from 1.8 to 6.2 A in 1000 steps,
    # |dE| < 0.08 eV and |dF| < 0.4 eV/A between neighbouring distances
    d = np.linspace(1.8, 6.2, 1000)
    E, F = [], []
    for x in d:
        pos, cell, pbc = in_vacuum_box(np.array([[0.0, 0.0, 0.0], [x, 0.0, 0.0]]), vacuum=4.0)
        e, f, w = energy(pos, cell, pbc)
        E.append(e)
        F.append(f[0, 0])
    E, F = np.array(E), np.array(F)
    assert np.abs(np.diff(E)).max() < 0.08
    assert np.abs(np.diff(F)).max() < 0.4
    # and the force on atom 0 is +dE/dr
    dE = np.gradient(E, d[1] - d[0])
    assert np.abs(dE[5:-5] - F[5:-5]).max() < 1e-2 * max(1.0, np.abs(F).max())
