"""REBO2 HIP path against the CPU oracle through the C-ABI (MI355X only).

Bars: energy, forces, virial and per-atom energies/virials within 1e-10
relative (forces and virials scaled by their largest entry, as in the other
parity tests); results bitwise reproducible run to run (fixed-point force
accumulation).  At the full bench size: perfect diamond at the REBO2 lattice
constant gives the cohesive energy of the small cell per atom and zero net
force.
"""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import in_vacuum_box, read_cfg, read_xyz

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.abs(b).max()) if b.size else 1.0)
    return float(np.abs(a - b).max()) <= rel * scale


def run_gpu(pos, cell, pbc, Z, per_at=False, precision="fp64", mask=None, dihedral=False):
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.Rebo2(dihedral=dihedral)
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, epot_per_at=per_at, wpot_per_at=per_at, mask=mask)


_R = {}


def run_oracle(pos, cell, pbc, Z, per_at=False, dihedral=False):
    if dihedral not in _R:
        _R[dihedral] = O.Rebo2(dihedral=dihedral)
    nl = O.nl_build(pos, cell, pbc, 2.0)
    return _R[dihedral](pos, Z, cell, nl, per_at=per_at)


def _molecule(name):
    pos, Z = read_xyz("molecules/%s.xyz" % name)
    pos, cell, pbc = in_vacuum_box(pos)
    return pos, cell, pbc, Z


CASES = [
    ("diamond_rattled", lambda: W.diamond_c(3, sigma=0.08, seed=3)),
    ("aC_small", lambda: read_cfg("aC_small.cfg")),
    ("aC", lambda: read_cfg("aC.cfg")),
    ("random_CH", lambda: W.random_ch(50, 10, 3.0, seed=7, strain=0.1)),
    ("random_CH_large", lambda: W.random_ch(1500, 500, 2.2, seed=8, strain=0.15)),
    ("naphthalene", lambda: _molecule("naphthalene")),
    ("neopentane", lambda: _molecule("neopentane")),
    ("small_cell", lambda: W.random_ch(10, 4, 2.5, seed=5, strain=0.2)),
]


@pytest.mark.parametrize("per_at", [True, False], ids=["per_atom", "totals"])
@pytest.mark.parametrize("dihedral", [False, True], ids=["plain", "torsion"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_rebo2_fp64_parity(name, make, dihedral, per_at):
    # per-atom virials take the fixed-point neighbour pass; totals only, the
    # per-centre pass with its owner gather
    pos, cell, pbc, Z = make()
    print("rebo2 case", name, len(pos), "dihedral", dihedral, "per_at", per_at, flush=True)
    e, f, w, epa, _, _, wpa, _ = run_gpu(pos, cell, pbc, Z, per_at=per_at, dihedral=dihedral)
    re_, rf, rw, repa, rwpa = run_oracle(pos, cell, pbc, Z, per_at=True, dihedral=dihedral)
    assert abs(e - re_) <= REL * max(1.0, abs(re_))
    assert close(f, rf) and close(w, rw)
    if per_at:
        assert close(epa, repa) and close(wpa, rwpa)


def test_rebo2_ignores_other_elements():
    pos, cell, pbc, Z = W.random_ch(60, 20, 2.5, seed=12, strain=0.1)
    Z = Z.copy()
    Z[::9] = 14       # atoms outside C/H are not part of any bond
    e, f, w, *_ = run_gpu(pos, cell, pbc, Z)
    re_, rf, rw = run_oracle(pos, cell, pbc, Z)
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)


def test_rebo2_bitwise_reproducible():
    pos, cell, pbc, Z = W.random_ch(2000, 600, 2.4, seed=13, strain=0.1)
    e1, f1, w1, *_ = run_gpu(pos, cell, pbc, Z, dihedral=True)
    e2, f2, w2, *_ = run_gpu(pos, cell, pbc, Z, dihedral=True)
    assert e1 == e2 and np.array_equal(f1, f2) and np.array_equal(w1, w2)


def test_rebo2_unsupported_modes():
    pos, cell, pbc, Z = W.diamond_c(2, sigma=0.02, seed=1)
    with pytest.raises(AgpuError):
        run_gpu(pos, cell, pbc, Z, mask=np.ones(len(pos), np.int32))


# (name, generator, force bar): a-C:H with 1e-5 eV/A absolute; the random
# C/H packing, whose forces are ~100 eV/A, with 2e-7 of its force RMS
MIXED_CASES = [
    ("aCH_5000", lambda: W.amorphous_ch(5000, x_h=0.2, density=2.0, seed=10), None),
    ("aC", lambda: read_cfg("aC.cfg"), None),
    ("random_CH_large", lambda: W.random_ch(1500, 500, 2.2, seed=8, strain=0.15), 2e-7),
]


@pytest.mark.parametrize("name,make,rel", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_rebo2_mixed_error(name, make, rel):
    """The mixed form (FP32 bond-record copies and FP32 arithmetic in the
    per-neighbour angular and conjugation loops; FP64 bond records, bond
    orders, radial pair terms and torsion; fixed-point accumulation) against
    the oracle with torsion: force RMS <= 1e-5 eV/A on a-C:H (the bench
    system's generator and the reference's a-C fixture), energy <= 1e-5
    eV/atom, virial <= 1e-5 eV/atom per component, bitwise reproducible."""
    pos, cell, pbc, Z = make()
    n = len(pos)
    print("rebo2 mixed case", name, n, flush=True)
    e, f, w, epa, _, _, wpa, _ = run_gpu(pos, cell, pbc, Z, per_at=True, precision="mixed", dihedral=True)
    re_, rf, rw, repa, rwpa = run_oracle(pos, cell, pbc, Z, per_at=True, dihedral=True)
    rms = float(np.sqrt(np.mean((f - rf) ** 2)))
    print("rebo2 mixed %s: F rms %.3e max %.3e (|F| rms %.3e)  |dE|/N %.3e  |dW|/N %.3e  |dEpa| max %.3e" %
          (name, rms, np.abs(f - rf).max(), float(np.sqrt(np.mean(rf ** 2))), abs(e - re_) / n,
           np.abs(w - rw).max() / n, np.abs(epa - repa).max()), flush=True)
    assert rms <= (1e-5 if rel is None else rel * float(np.sqrt(np.mean(rf ** 2))))
    assert abs(e - re_) / n <= 1e-5
    assert np.abs(w - rw).max() / n <= 1e-5
    assert np.abs(epa - repa).max() <= 1e-4
    # totals only: the per-centre neighbour pass, bitwise reproducible
    e2, f2, w2, *_ = run_gpu(pos, cell, pbc, Z, precision="mixed", dihedral=True)
    rms2 = float(np.sqrt(np.mean((f2 - rf) ** 2)))
    print("rebo2 mixed %s (per-centre pass): F rms %.3e  |dE|/N %.3e" % (name, rms2, abs(e2 - re_) / n), flush=True)
    assert rms2 <= (1e-5 if rel is None else rel * float(np.sqrt(np.mean(rf ** 2))))
    assert abs(e2 - re_) / n <= 1e-5 and np.abs(w2 - rw).max() / n <= 1e-5
    e3, f3, w3, *_ = run_gpu(pos, cell, pbc, Z, precision="mixed", dihedral=True)
    assert e3 == e2 and np.array_equal(f3, f2) and np.array_equal(w3, w2)


def test_rebo2_full_size_diamond():
    # 50^3 cells = 1,000,000 atoms at the lattice constant of the small-cell
    # energy minimum: every atom sees the same environment
    a0 = 3.5657500628
    pos, cell, pbc, Z = W.diamond_c(50, a=a0, sigma=0.0)
    e, f, w, *_ = run_gpu(pos, cell, pbc, Z)
    sp, sc, _, sZ = W.diamond_c(2, a=a0, sigma=0.0)
    re_, rf, rw = run_oracle(sp, sc, pbc, sZ)
    print("rebo2 1M diamond E/atom %.12f  small cell %.12f  max|f| %.3e" %
          (e / len(pos), re_ / len(sp), np.abs(f).max()), flush=True)
    assert abs(e / len(pos) - re_ / len(sp)) <= 1e-10 * abs(re_ / len(sp))
    assert np.abs(f).max() < 1e-9
    assert close(w / len(pos), rw / len(sp), rel=1e-9)


def test_rebo2_amorphous_hydrocarbon_sample():
    # the bench system's generator at 1/20 of its size, with torsion
    pos, cell, pbc, Z = W.amorphous_ch(5000, x_h=0.2, density=2.0, seed=10)
    e, f, w, *_ = run_gpu(pos, cell, pbc, Z, dihedral=True)
    re_, rf, rw = run_oracle(pos, cell, pbc, Z, dihedral=True)
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)
