"""Kumagai on the GPU (the bond-order kernel with the Kumagai forms) against
the oracle through the C-ABI.

Bars: FP64 energy, forces, virial and per-atom properties within 1e-10
relative; the mask decomposition; FP32-arithmetic / FP64-accumulate force RMS
<= 1e-5 eV/A; perfect diamond at the full bench size equals the closed form.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W

pytestmark = pytest.mark.gpu

REL = 1e-10
K = P.Kumagai_CompMaterSci_39_457_Si


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def oracle(pos, cell, pbc, mask=None, per_at=False):
    nl = O.nl_build(pos, cell, pbc, K["r2"])
    return O.tersoff(K, pos, np.zeros(len(pos), np.int32), cell, nl, mask=mask, per_at=per_at)


def gpu(pos, cell, pbc, precision="fp64", **kw):
    Z = np.full(len(pos), 14, np.int32)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.Kumagai()
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, **kw)


CASES = [
    ("diamond_512", lambda: W.diamond_si(4, sigma=0.05, seed=1)),
    ("disordered_64", lambda: W.diamond_si(2, sigma=0.2, seed=21)),
    ("compressed_216", lambda: W.diamond_si(3, a=5.0, sigma=0.1, seed=3)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_kumagai_fp64_parity(name, make):
    pos, cell, pbc, Z = make()
    e0, f0, w0, epa0, wpa0 = oracle(pos, cell, pbc, per_at=True)
    e, f, w, epa, _, _, wpa, _ = gpu(pos, cell, pbc, epot_per_at=True, wpot_per_at=True)
    print("kumagai %s: E %.12f (oracle %.12f)" % (name, e, e0), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(epa, epa0) and close(wpa, wpa0)


def test_kumagai_mask():
    pos, cell, pbc, Z = W.diamond_si(3, sigma=0.1, seed=7)
    m = (np.arange(len(pos)) % 3 == 0).astype(np.int32)
    e0, f0, w0 = oracle(pos, cell, pbc, mask=m)
    e, f, w, *_ = gpu(pos, cell, pbc, mask=m)
    assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)


def test_kumagai_mixed_error():
    pos, cell, pbc, Z = W.diamond_si(6, sigma=0.1, seed=9)
    e64, f64, *_ = gpu(pos, cell, pbc)
    e32, f32, *_ = gpu(pos, cell, pbc, precision="mixed")
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("kumagai mixed: F rms %.3e  |dE|/N %.3e" % (rms, abs(e32 - e64) / len(pos)), flush=True)
    assert rms <= 1e-5


def test_kumagai_diamond_million_closed_form():
    a = 5.429
    r = a * math.sqrt(3) / 4
    t = K["h"] + 1.0 / 3.0
    go = K["c2"] * t / (K["c3"] + t * t)
    g = K["c1"] + t * go * (1 + K["c4"] * math.exp(-K["c5"] * t * t))
    b = (1 + (3 * g) ** K["eta"]) ** (-K["delta"])
    e_atom = 2 * (K["A"] * math.exp(-K["lambda1"] * r) - b * K["B"] * math.exp(-K["lambda2"] * r))
    pos, cell, pbc, Z = W.diamond_si(50, a=a, sigma=0.0)
    e, f, *_ = gpu(pos, cell, pbc)
    assert abs(e / len(pos) - e_atom) <= 1e-10 * abs(e_atom)
    assert np.abs(f).max() < 1e-9
