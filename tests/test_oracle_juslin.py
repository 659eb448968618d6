"""The oracle's Juslin forms (juslin_func.f90 through the generic bond-order
kernel: ordered element pairs, per-triplet h-term) against what the
reference's tests fix:

  - the bulk values of tests/test_bulk_properties.py:96-117 (5 % there):
    bcc W Ec, a0, C11, C12; diamond C Ec, a0; B1, B2 and B3 WC Ec and a0
    (the mixed lattices pin the ordered pair and triplet indices);
  - forces and virial are the derivatives of the energy (central FD) on a
    disordered W-C-H structure (H-containing triplets carry alpha = 4 and
    omega != 1) and on the Kuopanportti Fe-C-H set;
  - mask decomposition and per-atom sums.
"""
import numpy as np
from scipy.optimize import minimize_scalar

from oracle import oracle as O
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import bop_cutoff, fd_forces, fd_virial

J = P.Juslin_JAP_98_123520_WCH
GPA = 1.0 / 160.21766208
BCC = np.array([[0, 0, 0], [.5, .5, .5]])
B1 = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.5, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5]])


def energy(prm, pos, cell, pbc, el, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, bop_cutoff(prm), avgn=200)
    return O.tersoff(prm, pos, np.asarray(el, np.int32), cell, nl, mask=mask, per_at=per_at)


def _bulk(basis, n, species, bounds, elastic=False):
    pbc = np.array([True] * 3)

    def make(a):
        pos, cell = W.lattice(basis, n, a)
        el = np.array([species[i % len(basis)] for i in range(len(pos))])
        return pos, cell, el

    def e_of_a(a):
        pos, cell, el = make(a)
        return energy(J, pos, cell, pbc, el)[0] / len(pos)
    res = minimize_scalar(e_of_a, bounds=bounds, method="bounded", options={"xatol": 1e-9})
    if not elastic:
        return -res.fun, res.x
    pos0, cell0, el = make(res.x)

    def stress(F):
        pos, cell = pos0 @ F.T, cell0 @ F.T
        return energy(J, pos, cell, pbc, el)[2] / abs(np.linalg.det(cell))
    h = 1e-5
    Fp, Fm = np.eye(3), np.eye(3)
    Fp[0, 0] += h
    Fm[0, 0] -= h
    d = (stress(Fp) - stress(Fm)) / (2 * h) / GPA
    return -res.fun, res.x, d[0, 0], d[1, 1]


def _rel(x, ref, tol=0.05):
    return abs(x - ref) / abs(ref) < tol


def test_bcc_tungsten():
    Ec, a0, C11, C12 = _bulk(BCC, 3, (0, 0), (3.0, 3.3), elastic=True)
    assert _rel(Ec, 8.89, 1e-3) and _rel(a0, 3.165, 1e-3) and _rel(C11, 542) and _rel(C12, 191)


def test_diamond_carbon_and_tungsten_carbides():
    Ec, a0 = _bulk(W.DIAMOND_BASIS, 2, (1,) * 8, (3.4, 3.7))
    assert _rel(Ec, 7.376 - 0.0524, 1e-3) and _rel(a0, 3.558, 1e-3)
    Ec, a0 = _bulk(B1, 2, (0, 0, 0, 0, 1, 1, 1, 1), (4.2, 4.6))
    assert _rel(Ec, (16.68 - 0.98) / 2, 1e-3) and _rel(a0, 4.380, 1e-3)
    Ec, a0 = _bulk(BCC, 3, (0, 1), (2.5, 2.9))
    assert _rel(Ec, (16.68 - 2.32) / 2, 1e-3) and _rel(a0, 2.704, 1e-3)
    Ec, a0 = _bulk(W.DIAMOND_BASIS, 2, (0, 0, 0, 0, 1, 1, 1, 1), (4.4, 4.9))
    assert _rel(Ec, (16.68 - 2.12) / 2, 1e-3) and _rel(a0, 4.679, 1e-3)


def _wch(seed=5):
    """A rattled B1 W-C crystal with a third of the C replaced by H."""
    pos, cell = W.lattice(B1, 2, 4.38)
    rng = np.random.default_rng(seed)
    el = np.where(np.arange(len(pos)) % 8 >= 4, 1, 0)
    el[(el == 1) & (rng.random(len(pos)) < 0.35)] = 2
    pos = pos + rng.normal(0, 0.12, pos.shape)
    return pos, cell, np.array([True] * 3), el


def _fd_check(prm, pos, cell, pbc, el, rows):
    e, f, w = energy(prm, pos, cell, pbc, el)
    fd = fd_forces(lambda p: energy(prm, p, cell, pbc, el)[0], pos, rows)
    assert np.max(np.abs(fd - f[list(rows)])) < 1e-6 * max(1.0, np.abs(f).max())
    Wfd = fd_virial(lambda p, c: energy(prm, p, c, pbc, el)[0], pos, cell)
    assert np.max(np.abs(Wfd - w)) < 1e-5 * max(1.0, np.abs(w).max())
    assert np.abs(f.sum(axis=0)).max() < 1e-9


def test_fd_wch_and_mask():
    pos, cell, pbc, el = _wch()
    assert (el == 2).any()
    _fd_check(J, pos, cell, pbc, el, [0, 4, 9, 22, 37, 60])
    e, f, w, epa, wpa = energy(J, pos, cell, pbc, el, per_at=True)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e1, f1, w1 = energy(J, pos, cell, pbc, el, mask=m)
    e2, f2, w2 = energy(J, pos, cell, pbc, el, mask=1 - m)
    assert abs(e - e1 - e2) < 1e-9 and np.abs(f - f1 - f2).max() < 1e-9 and np.abs(w - w1 - w2).max() < 1e-9
    assert abs(epa.sum() - e) < 1e-9 and np.abs(wpa.sum(axis=0) - w).max() < 1e-9


def test_fd_kuopanportti_fech():
    prm = P.Kuopanportti_CMS_111_525_FeCH
    pos, cell = W.lattice(B1, 2, 4.0)
    rng = np.random.default_rng(6)
    el = np.where(np.arange(len(pos)) % 8 >= 4, 1, 0)
    el[(el == 1) & (rng.random(len(pos)) < 0.3)] = 2
    pos = pos + rng.normal(0, 0.12, pos.shape)
    _fd_check(prm, pos, cell, np.array([True] * 3), el, [1, 8, 19, 44])


# JuslinScr (juslin_scr.f90): the same kernel with SCREENING
JS = P.Juslin_JAP_98_123520_WCH__Scr


def test_scr_dimer_equals_unscreened():
    # an isolated W-W dimer inside r1 = or1 = bor1: no third atom screens it and
    # every cutoff is 1, so the screened form reduces to Juslin exactly
    cell = np.eye(3) * 30.0
    pos = np.array([[10.0, 10.0, 10.0], [12.6, 10.0, 10.0]])
    pbc = np.array([False] * 3)
    es = energy(JS, pos, cell, pbc, [0, 0])[0]
    eu = energy(J, pos, cell, pbc, [0, 0])[0]
    assert eu < 0 and abs(es - eu) <= 1e-12 * abs(eu)


def test_scr_fd_wch():
    pos, cell, pbc, el = _wch(seed=7)
    _fd_check(JS, pos, cell, pbc, el, [0, 3, 9, 22, 37, 60])
    # second neighbours inside the cutoffs are partly screened here
    assert abs(energy(JS, pos, cell, pbc, el)[0] - energy(J, pos, cell, pbc, el)[0]) > 1e-3
