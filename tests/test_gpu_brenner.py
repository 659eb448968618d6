"""Brenner (Albe/Erhart forms on the bond-order kernel) on the GPU against
the oracle through the C-ABI: FP64 energy, forces, virial and per-atom
properties within 1e-10 relative, the mask, and the FP32-arithmetic /
FP64-accumulate force RMS <= 1e-5 eV/A."""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import read_cfg

pytestmark = pytest.mark.gpu

REL = 1e-10
FCC = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
B1 = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
               [.5, .5, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5]])


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def oracle(prm, pos, cell, pbc, el, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, max(prm["r2"]))
    return O.tersoff(prm, pos, np.asarray(el, np.int32), cell, nl, mask=mask, per_at=per_at)


def gpu(prm, pos, cell, pbc, el, precision="fp64", **kw):
    Z = np.array([P.ATOMIC_NUMBERS[prm["el"][e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    kwargs = {k: v for k, v in prm.items() if k not in ("__ref__", "family")}
    pot = N.Brenner(**kwargs)
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, **kw)


def _ac():
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    return P.Erhart_PRB_71_035211_SiC, pos, cell, pbc, np.zeros(len(pos), int)


def _sic():
    pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.08, seed=3)
    return P.Erhart_PRB_71_035211_SiC, pos, cell, pbc, (np.arange(len(pos)) % 8 >= 4).astype(int)


def _fec():
    pos, cell = W.lattice(B1, 3, 3.9)
    pos = pos + np.random.default_rng(4).normal(0, 0.08, pos.shape)
    return P.Henriksson_PRB_79_144107_FeC, pos, cell, np.array([True] * 3), (np.arange(len(pos)) % 8 >= 4).astype(int)


def _ptc():
    pos, cell = W.lattice(FCC, 3, 3.92)
    pos = pos + np.random.default_rng(8).normal(0, 0.1, pos.shape)
    el = np.zeros(len(pos), int)
    el[::7] = 1
    return P.Albe_PRB_65_195124_PtC, pos, cell, np.array([True] * 3), el


def _c2():
    pos, cell, pbc, Z = W.diamond_si(3, a=3.558, sigma=0.05, seed=9)
    return P.Brenner_PRB_42_9458_C_II, pos, cell, pbc, np.zeros(len(pos), int)


CASES = [("aC_erhart", _ac), ("SiC_erhart", _sic), ("FeC_henriksson", _fec), ("PtC_albe", _ptc),
         ("C_brenner_II", _c2)]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_brenner_fp64_parity(name, make):
    prm, pos, cell, pbc, el = make()
    e0, f0, w0, epa0, wpa0 = oracle(prm, pos, cell, pbc, el, per_at=True)
    e, f, w, epa, _, _, wpa, _ = gpu(prm, pos, cell, pbc, el, epot_per_at=True, wpot_per_at=True)
    print("brenner %s: E %.12f (oracle %.12f)" % (name, e, e0), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(epa, epa0) and close(wpa, wpa0)


def test_brenner_mask():
    prm, pos, cell, pbc, el = _fec()
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e0, f0, w0 = oracle(prm, pos, cell, pbc, el, mask=m)
    e, f, w, *_ = gpu(prm, pos, cell, pbc, el, mask=m)
    assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)


def test_brenner_mixed_error():
    prm, pos, cell, pbc, el = _sic()
    e64, f64, *_ = gpu(prm, pos, cell, pbc, el)
    e32, f32, *_ = gpu(prm, pos, cell, pbc, el, precision="mixed")
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("brenner mixed: F rms %.3e  |dE|/N %.3e" % (rms, abs(e32 - e64) / len(pos)), flush=True)
    assert rms <= 1e-5
