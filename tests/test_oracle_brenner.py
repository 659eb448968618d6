"""The oracle's Brenner forms (brenner_func.f90 through the generic
bond-order kernel) against what the mathematics and the reference's tests
fix:

  - perfect diamond carbon in closed form (4 bonds, fc = 1, cos = -1/3),
    written here from the functional forms;
  - the bulk values of tests/test_bulk_properties.py:67-85 (5 % there):
    Erhart-Albe SiC for diamond C, diamond Si and B3 SiC, Brenner's potential
    II for diamond C;
  - forces and virial are the derivatives of the energy (central FD) for the
    cases of tests/test_forces_and_virial.py:91-115 (a-C, SiC, FeC with the
    Henriksson set, and Albe's PtC whose Pt-Pt pair carries the h-term);
  - mask decomposition and per-atom sums.
"""
import math

import numpy as np
from scipy.optimize import minimize_scalar

from oracle import oracle as O
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import fd_forces, fd_virial, read_cfg

GPA = 1.0 / 160.21766208
SIC = P.Erhart_PRB_71_035211_SiC

BCC_BASIS = np.array([[0, 0, 0], [.5, .5, .5]])
FCC_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
B1_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.5, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5]])


def energy(prm, pos, cell, pbc, el, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, max(prm["r2"]))
    return O.tersoff(prm, pos, np.asarray(el, np.int32), cell, nl, mask=mask, per_at=per_at)


def diamond(a, n=2, species=(0, 0)):
    pos, cell = W.lattice(W.DIAMOND_BASIS, n, a)
    el = np.where(np.arange(len(pos)) % 8 >= 4, species[1], species[0])
    return pos, cell, np.array([True] * 3), el


def closed_form_c(prm, a):
    r = a * math.sqrt(3) / 4
    S, D0, r0, beta = prm["S"][0], prm["D0"][0], prm["r0"][0], prm["beta"][0]
    g = prm["gamma"][0] * (1 + prm["c"][0] ** 2 / prm["d"][0] ** 2 -
                           prm["c"][0] ** 2 / (prm["d"][0] ** 2 + (prm["h"][0] - 1.0 / 3.0) ** 2))
    b = (1 + (3 * g) ** prm["n"][0]) ** (-0.5 / prm["n"][0])
    VR = D0 / (S - 1) * math.exp(-beta * math.sqrt(2 * S) * (r - r0))
    VA = S * D0 / (S - 1) * math.exp(-beta * math.sqrt(2 / S) * (r - r0))
    return 2 * (VR - b * VA)


def test_diamond_carbon_closed_form():
    for prm in (SIC, P.Brenner_PRB_42_9458_C_I, P.Brenner_PRB_42_9458_C_II):
        for a in (3.50, 3.566, 3.60):
            pos, cell, pbc, el = diamond(a)
            e, f, w = energy(prm, pos, cell, pbc, el)
            assert abs(e / len(pos) - closed_form_c(prm, a)) < 1e-12
            assert np.abs(f).max() < 1e-10


def _bulk(prm, species, bounds, n=2):
    def e_of_a(a):
        pos, cell, pbc, el = diamond(a, n, species)
        return energy(prm, pos, cell, pbc, el)[0] / len(pos)
    res = minimize_scalar(e_of_a, bounds=bounds, method="bounded", options={"xatol": 1e-9})
    a0 = res.x
    pos0, cell0, pbc, el = diamond(a0, n, species)

    def stress(F):
        pos, cell = pos0 @ F.T, cell0 @ F.T
        return energy(prm, pos, cell, pbc, el)[2] / abs(np.linalg.det(cell))
    h = 1e-5
    Fp, Fm, Sp, Sm = np.eye(3), np.eye(3), np.eye(3), np.eye(3)
    Fp[0, 0] += h
    Fm[0, 0] -= h
    Sp[1, 2] += h
    Sm[1, 2] -= h
    d = (stress(Fp) - stress(Fm)) / (2 * h) / GPA
    c440 = (stress(Sp) - stress(Sm))[1, 2] / (2 * h) / GPA
    return -res.fun, a0, d[0, 0], d[1, 1], c440


def _rel(x, ref, tol=0.05):
    return abs(x - ref) / abs(ref) < tol


def test_erhart_sic_bulk_properties():
    Ec, a0, C11, C12, _ = _bulk(SIC, (0, 0), (3.4, 3.7))
    assert _rel(Ec, 7.3731, 5e-4) and _rel(a0, 3.566, 5e-4)
    assert _rel(C11, 1082) and _rel(C12, 127)
    Ec, a0, C11, C12, C440 = _bulk(SIC, (1, 1), (5.2, 5.6))
    assert _rel(Ec, 4.63, 5e-3) and _rel(a0, 5.429, 5e-3)
    assert _rel(C11, 167) and _rel(C12, 65) and _rel(C440, 105)
    Ec, a0, C11, C12, C440 = _bulk(SIC, (0, 1), (4.2, 4.5))
    assert _rel(Ec, 6.340, 5e-3) and _rel(a0, 4.359, 5e-3)
    assert _rel(C11, 382) and _rel(C12, 145) and _rel(C440, 305)


def test_brenner_ii_diamond_carbon():
    Ec, a0, C11, C12, _ = _bulk(P.Brenner_PRB_42_9458_C_II, (0, 0), (3.4, 3.7))
    assert _rel(Ec, 7.376 - 0.0524, 5e-3) and _rel(a0, 3.558, 5e-3)
    assert _rel(C11, 621) and _rel(C12, 415)


def _fd_check(prm, pos, cell, pbc, el, rows):
    e, f, w = energy(prm, pos, cell, pbc, el)
    fd = fd_forces(lambda p: energy(prm, p, cell, pbc, el)[0], pos, rows)
    assert np.max(np.abs(fd - f[list(rows)])) < 1e-6 * max(1.0, np.abs(f).max())
    Wfd = fd_virial(lambda p, c: energy(prm, p, c, pbc, el)[0], pos, cell)
    assert np.max(np.abs(Wfd - w)) < 1e-5 * max(1.0, np.abs(w).max())
    assert np.abs(f.sum(axis=0)).max() < 1e-9


def test_fd_amorphous_carbon():
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    pos = pos + np.random.default_rng(7).normal(0, 0.03, pos.shape)
    _fd_check(SIC, pos, cell, pbc, np.zeros(len(pos)), range(0, 128, 13))


def test_fd_sic_disordered():
    pos, cell, pbc, el = diamond(4.36, 2, (0, 1))
    pos = pos + np.random.default_rng(3).normal(0, 0.08, pos.shape)
    _fd_check(SIC, pos, cell, pbc, el, [0, 5, 17, 40, 63])


def test_fd_iron_carbide_and_mask():
    prm = P.Henriksson_PRB_79_144107_FeC
    pos, cell = W.lattice(B1_BASIS, 2, 3.9)
    el = np.where(np.arange(len(pos)) % 8 >= 4, 1, 0)
    pos = pos + np.random.default_rng(4).normal(0, 0.08, pos.shape)
    pbc = np.array([True] * 3)
    _fd_check(prm, pos, cell, pbc, el, [0, 9, 22, 37, 60])
    e, f, w, epa, wpa = energy(prm, pos, cell, pbc, el, per_at=True)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e1, f1, w1 = energy(prm, pos, cell, pbc, el, mask=m)
    e2, f2, w2 = energy(prm, pos, cell, pbc, el, mask=1 - m)
    assert abs(e - e1 - e2) < 1e-9 and np.abs(f - f1 - f2).max() < 1e-9 and np.abs(w - w1 - w2).max() < 1e-9
    assert abs(epa.sum() - e) < 1e-9 and np.abs(wpa.sum(axis=0) - w).max() < 1e-9


def test_fd_platinum_carbon_with_h_term():
    prm = P.Albe_PRB_65_195124_PtC
    pos, cell = W.lattice(FCC_BASIS, 3, 3.92)
    el = np.zeros(len(pos), int)
    el[::7] = 1
    pos = pos + np.random.default_rng(8).normal(0, 0.1, pos.shape)
    _fd_check(prm, pos, cell, np.array([True] * 3), el, [0, 7, 30, 61, 100])
