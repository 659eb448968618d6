"""Screened bond-order variants (TersoffScr, KumagaiScr, BrennerScr) on the
GPU against the oracle through the C-ABI: FP64 energy, forces, virial and
per-atom properties within 1e-10 relative on structures where bonds are
partially screened, the mask, bitwise reproducibility, the library's
defaults against the reference's Python parameter sets, and the
FP32-arithmetic / FP64-accumulate force RMS <= 1e-5 eV/A."""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import bop_cutoff, read_cfg

pytestmark = pytest.mark.gpu

REL = 1e-10
TS = P.Tersoff_PRB_39_5566_Si_C__Scr
KS = P.Kumagai_CompMaterSci_39_457_Si__Scr
BS = P.Erhart_PRB_71_035211_SiC__Scr
CLS = {id(TS): N.TersoffScr, id(KS): N.KumagaiScr, id(BS): N.BrennerScr}


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def oracle(prm, pos, cell, pbc, el, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, bop_cutoff(prm), avgn=300)
    return O.tersoff(prm, pos, np.asarray(el, np.int32), cell, nl, mask=mask, per_at=per_at)


def make_pot(prm, precision="fp64", default=False):
    cls = CLS[id(prm)]
    if default:
        pot = cls()
    else:
        pot = cls(**{k: v for k, v in prm.items() if k not in ("__ref__", "family")})
    pot.set_precision(precision)
    return pot


def gpu(prm, pos, cell, pbc, el, precision="fp64", default=False, **kw):
    names = prm["el"] if isinstance(prm["el"], (list, tuple)) else [prm["el"]]
    Z = np.array([P.ATOMIC_NUMBERS[names[e]] for e in el], np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(300)
    pot = make_pot(prm, precision, default)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, **kw)


def _ac(prm, seed=7):
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    pos = pos + np.random.default_rng(seed).normal(0, 0.03, pos.shape)
    return prm, pos, cell, pbc, np.zeros(len(pos), int)


def _sic(prm):
    pos, cell = W.lattice(W.DIAMOND_BASIS, 3, 4.36)
    pos = pos + np.random.default_rng(3).normal(0, 0.15, pos.shape)
    return prm, pos, cell, np.array([True] * 3), (np.arange(len(pos)) % 8 >= 4).astype(int)


def _si(prm):
    pos, cell, pbc, Z = W.diamond_si(3, sigma=0.25, seed=4)
    el = np.zeros(len(pos), int) if prm is KS else np.ones(len(pos), int)
    return prm, pos, cell, pbc, el


CASES = [("tersoff_scr_aC", lambda: _ac(TS)), ("tersoff_scr_SiC", lambda: _sic(TS)),
         ("tersoff_scr_Si", lambda: _si(TS)), ("kumagai_scr_Si", lambda: _si(KS)),
         ("brenner_scr_aC", lambda: _ac(BS, 11)), ("brenner_scr_SiC", lambda: _sic(BS))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_screened_fp64_parity(name, make):
    prm, pos, cell, pbc, el = make()
    print("screened case", name, len(pos), flush=True)
    e0, f0, w0, epa0, wpa0 = oracle(prm, pos, cell, pbc, el, per_at=True)
    e, f, w, epa, _, _, wpa, _ = gpu(prm, pos, cell, pbc, el, epot_per_at=True, wpot_per_at=True)
    print("  E %.12f (oracle %.12f)  |dF| %.2e" % (e, e0, np.abs(f - f0).max()), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(epa, epa0) and close(wpa, wpa0)


def test_screened_mask_and_reproducible():
    prm, pos, cell, pbc, el = _sic(TS)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e0, f0, w0 = oracle(prm, pos, cell, pbc, el, mask=m)
    e, f, w, *_ = gpu(prm, pos, cell, pbc, el, mask=m)
    assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)
    e2, f2, w2, *_ = gpu(prm, pos, cell, pbc, el, mask=m)
    assert e == e2 and np.array_equal(f, f2) and np.array_equal(w, w2)


@pytest.mark.parametrize("prm", [TS, KS, BS], ids=["TersoffScr", "KumagaiScr", "BrennerScr"])
def test_library_defaults(prm):
    # the C defaults are the reference's Fortran screened database entries;
    # they agree with the Python sets where those coincide (BrennerScr's
    # Python r1 of the C-Si pair differs, so compare on pure Si / C)
    prm_, pos, cell, pbc, el = _si(prm)
    e0, f0, w0 = oracle(prm, pos, cell, pbc, el)
    e, f, w, *_ = gpu(prm, pos, cell, pbc, el, default=True)
    assert abs(e - e0) <= 1e-9 * abs(e0) and close(f, f0, 1e-9)


def test_screened_mixed_error():
    prm, pos, cell, pbc, el = _ac(TS)
    e64, f64, *_ = gpu(prm, pos, cell, pbc, el)
    e32, f32, *_ = gpu(prm, pos, cell, pbc, el, precision="mixed")
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("screened mixed: F rms %.3e  |dE|/N %.3e" % (rms, abs(e32 - e64) / len(pos)), flush=True)
    assert rms <= 1e-5
