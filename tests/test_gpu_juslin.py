"""Juslin (W-C-H and Fe-C-H sets on the bond-order kernel) on the GPU
against the oracle through the C-ABI: FP64 energy, forces, virial and
per-atom properties within 1e-10 relative, the mask, the library default,
and the FP32-arithmetic / FP64-accumulate force RMS <= 1e-5 eV/A."""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import bop_cutoff

pytestmark = pytest.mark.gpu

REL = 1e-10
J = P.Juslin_JAP_98_123520_WCH
B1 = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.5, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5]])
BCC = np.array([[0, 0, 0], [.5, .5, .5]])


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def oracle(prm, pos, cell, pbc, el, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, bop_cutoff(prm), avgn=200)
    return O.tersoff(prm, pos, np.asarray(el, np.int32), cell, nl, mask=mask, per_at=per_at)


def gpu(prm, pos, cell, pbc, el, precision="fp64", default=False, **kw):
    Z = np.array([P.ATOMIC_NUMBERS[prm["el"][e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(200)
    pot = N.Juslin() if default else N.Juslin(**{k: v for k, v in prm.items() if k not in ("__ref__", "family")})
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, **kw)


def _wch(prm=J, a=4.38, seed=5, n=3):
    pos, cell = W.lattice(B1, n, a)
    rng = np.random.default_rng(seed)
    el = np.where(np.arange(len(pos)) % 8 >= 4, 1, 0)
    el[(el == 1) & (rng.random(len(pos)) < 0.35)] = 2
    pos = pos + rng.normal(0, 0.12, pos.shape)
    return prm, pos, cell, np.array([True] * 3), el


def _bccw():
    pos, cell = W.lattice(BCC, 5, 3.165)
    pos = pos + np.random.default_rng(3).normal(0, 0.08, pos.shape)
    return J, pos, cell, np.array([True] * 3), np.zeros(len(pos), int)


CASES = [("WCH_b1", lambda: _wch()), ("W_bcc", _bccw),
         ("FeCH_kuopanportti", lambda: _wch(P.Kuopanportti_CMS_111_525_FeCH, a=4.0, seed=6))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_juslin_fp64_parity(name, make):
    prm, pos, cell, pbc, el = make()
    e0, f0, w0, epa0, wpa0 = oracle(prm, pos, cell, pbc, el, per_at=True)
    e, f, w, epa, _, _, wpa, _ = gpu(prm, pos, cell, pbc, el, epot_per_at=True, wpot_per_at=True)
    print("juslin %s: E %.12f (oracle %.12f)" % (name, e, e0), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(epa, epa0) and close(wpa, wpa0)


def test_juslin_mask_and_default():
    prm, pos, cell, pbc, el = _wch(seed=9)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e0, f0, w0 = oracle(prm, pos, cell, pbc, el, mask=m)
    e, f, w, *_ = gpu(prm, pos, cell, pbc, el, mask=m)
    assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)
    # the library default is the Fortran database entry: the Python set with
    # the W-H inner cutoff 1.95 .. 2.35
    fort = dict(J, r1=[3.20, 2.60, 1.95, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
                r2=[3.80, 3.00, 2.35, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70])
    e0, f0, w0 = oracle(fort, pos, cell, pbc, el)
    e, f, w, *_ = gpu(J, pos, cell, pbc, el, default=True)
    assert abs(e - e0) <= REL * abs(e0) and close(f, f0)


def test_juslin_mixed_error():
    prm, pos, cell, pbc, el = _wch()
    e64, f64, *_ = gpu(prm, pos, cell, pbc, el)
    e32, f32, *_ = gpu(prm, pos, cell, pbc, el, precision="mixed")
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("juslin mixed: F rms %.3e" % rms, flush=True)
    assert rms <= 1e-5


JS = P.Juslin_JAP_98_123520_WCH__Scr


def gpu_scr(prm, pos, cell, pbc, el, default=False):
    Z = np.array([P.ATOMIC_NUMBERS[prm["el"][e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(300)
    pot = N.JuslinScr() if default else N.JuslinScr(**{k: v for k, v in prm.items() if k not in ("__ref__", "family")})
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl)


@pytest.mark.parametrize("case", ["WCH_b1", "W_bcc"])
def test_juslin_scr_parity(case):
    _, pos, cell, pbc, el = dict(CASES)[case]()
    print("juslin_scr", case, len(pos), flush=True)
    nl = O.nl_build(pos, cell, pbc, bop_cutoff(JS), avgn=300)
    re_, rf, rw = O.tersoff(JS, pos, np.asarray(el, np.int32), cell, nl)
    e, f, w = gpu_scr(JS, pos, cell, pbc, el)[:3]
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)
    e0, f0, w0 = gpu_scr(JS, pos, cell, pbc, el, default=True)[:3]
    assert abs(e0 - e) <= REL * abs(e) and close(f0, f)


def test_juslin_scr_screening_changes_result_and_mixed_error():
    """On rattled W-C-H the screened class differs from the unscreened one
    (the partially screened branch, the one that writes screening records,
    is taken), and its FP32-arithmetic / FP64-accumulate form keeps the
    force RMS within 1e-5 eV/A of its FP64 form."""
    _, pos, cell, pbc, el = _wch(seed=5)
    e_j = gpu(J, pos, cell, pbc, el)[0]
    Z = np.array([P.ATOMIC_NUMBERS[JS["el"][e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(300)
    pot = N.JuslinScr()
    pot.bind_to(p, nl)
    e64, f64 = pot.energy_and_forces(p, nl)[:2]
    pot.set_precision("mixed")
    e32, f32 = pot.energy_and_forces(p, nl)[:2]
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("juslin_scr: E %.8f vs Juslin %.8f; mixed F rms %.3e" % (e64, e_j, rms), flush=True)
    assert abs(e64 - e_j) > 1e-3
    assert rms <= 1e-5 and abs(e32 - e64) / len(pos) <= 1e-5


def test_juslin_scr_kuopanportti_parity():
    """The screened Fe-C-H set, whose Fe pairs have or1 = bor1 = 100 A past
    or2: those pairs keep the inner cutoff only."""
    KS = P.Kuopanportti_CMS_111_525_FeCH__Scr
    _, pos, cell, pbc, el = _wch(KS, a=4.0, seed=6)
    nl = O.nl_build(pos, cell, pbc, bop_cutoff(KS), avgn=300)
    re_, rf, rw = O.tersoff(KS, pos, np.asarray(el, np.int32), cell, nl)
    e, f, w = gpu_scr(KS, pos, cell, pbc, el)[:3]
    print("juslin_scr kuopanportti: E %.12f (oracle %.12f)" % (e, re_), flush=True)
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)
