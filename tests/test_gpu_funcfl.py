"""TabulatedEAM (single-element funcfl EAM, tabulated_eam.f90) on the GPU
against the oracle through the C-ABI, on the reference's fixture
tests/Au_u3.eam: FP64 energy, forces, virial and per-atom properties within
1e-10 relative on rattled, compressed and stretched fcc Au and a surface
slab; the FP32-arithmetic / FP64-accumulate form within 1e-5 eV/A force
RMS; the ASE calculator class; and the refusal of a mask (the reference
class has no mask feature)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import aseinterface as A
from atomistica_amd import native as N
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import GOLDEN

pytestmark = pytest.mark.gpu

REL = 1e-10
FN = os.path.join(GOLDEN, "Au_u3.eam")
FCC = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def _au(n=4, a=4.08, sigma=0.1, seed=3, slab=False):
    pos, cell = W.lattice(FCC, n, a)
    pos = pos + np.random.default_rng(seed).normal(0, sigma, pos.shape)
    pbc = np.array([True, True, not slab])
    return pos, cell, pbc, np.full(len(pos), 79, np.int32)


CASES = [("rattled", lambda: _au()), ("compressed", lambda: _au(a=3.9, sigma=0.05, seed=4)),
         ("stretched", lambda: _au(a=4.3, sigma=0.15, seed=5)), ("slab", lambda: _au(slab=True, seed=6))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_funcfl_parity(name, make):
    pos, cell, pbc, Z = make()
    T = O.read_funcfl_tables(FN)
    nl = O.nl_build(pos, cell, pbc, T["cutoff"])
    e0, f0, w0, epa0, wpa0 = O.eam(T, pos, np.zeros(len(pos), np.int32), cell, nl, per_at=True)
    p = N.from_atoms(pos, cell, pbc, Z)
    gnl = N.Neighbors(100)
    pot = N.TabulatedEAM(fn=FN)
    pot.bind_to(p, gnl)
    e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, gnl, epot_per_at=True, wpot_per_at=True)
    print("funcfl %s: E %.12f (oracle %.12f)" % (name, e, e0), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(epa, epa0) and close(wpa, wpa0)
    # the call without per-atom outputs takes the lighter kernel variant
    e2, f2, w2, *_ = pot.energy_and_forces(p, gnl)
    assert abs(e2 - e0) <= REL * abs(e0) and close(f2, f0) and close(w2, w0)
    e3, f3, *_ = pot.energy_and_forces(p, gnl)
    assert e3 == e2 and np.array_equal(f3, f2)


def test_funcfl_mixed_and_calculator():
    pos, cell, pbc, Z = _au(n=5, seed=7)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.TabulatedEAM(fn=FN)
    pot.bind_to(p, nl)
    e64, f64 = pot.energy_and_forces(p, nl)[:2]
    pot.set_precision("mixed")
    e32, f32 = pot.energy_and_forces(p, nl)[:2]
    rms = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print("funcfl mixed: F rms %.3e" % rms, flush=True)
    assert rms <= 1e-5 and abs(e32 - e64) / len(pos) <= 1e-5
    s = A.Structure(pos, cell, True, Z)
    s.calc = A.TabulatedEAM(fn=FN)
    assert abs(s.get_potential_energy() - e64) <= REL * abs(e64)
    with pytest.raises(AgpuError):
        pot.energy_and_forces(p, nl, mask=np.ones(len(pos), np.int32))
