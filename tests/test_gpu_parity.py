"""HIP path against the CPU oracle through the C-ABI (MI355X only).

Bars: neighbour lists bit-identical in content and order; FP64 energy,
forces, virial and per-atom properties within 1e-10 relative; the mixed
(FP32 arithmetic / FP64 accumulation) forms within the stated force RMS
bound; every result bitwise reproducible run to run.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import GOLDEN, read_cfg, read_setfl_tables

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.abs(b).max()) if b.size else 1.0)
    return float(np.abs(a - b).max()) <= rel * scale if b.size else True


def particles(pos, cell, pbc, Z):
    return N.from_atoms(pos, cell, pbc, Z)


# ---------------------------------------------------------------------------
# neighbour lists
# ---------------------------------------------------------------------------

NL_CASES = [
    ("diamond", lambda: W.diamond_si(4, sigma=0.05) + (3.0,)),
    ("fcc", lambda: W.fcc_cu(5, sigma=0.08) + (5.5,)),
    ("aC", lambda: read_cfg("aC.cfg") + (3.0,)),
    ("aC_5A", lambda: read_cfg("aC_small.cfg") + (5.0,)),
]


def _triclinic(seed, n, L, rc, pbc):
    rng = np.random.default_rng(seed)
    cell = np.diag([L, 1.1 * L, 0.9 * L]) + rng.uniform(-0.15, 0.15, (3, 3)) * L
    pos = rng.random((n, 3)) @ cell
    if pbc[0]:
        pos[::5] += rng.integers(-3, 4, (len(pos[::5]), 1)) * cell[0]
    return pos, cell, np.array(pbc, bool), np.full(n, 14, np.int32), rc


NL_CASES += [
    ("triclinic", lambda: _triclinic(1, 300, 14.0, 4.0, [1, 1, 1])),
    ("small_cell", lambda: _triclinic(2, 12, 5.4, 6.0, [1, 1, 1])),
    ("partial_pbc", lambda: _triclinic(3, 200, 12.0, 3.5, [1, 0, 1])),
    ("open", lambda: _triclinic(4, 150, 10.0, 3.0, [0, 0, 0])),
]


@pytest.mark.parametrize("name,make", NL_CASES, ids=[c[0] for c in NL_CASES])
def test_neighbor_list_bit_identical(name, make):
    pos, cell, pbc, Z, rc = make()
    print("nl case", name, len(pos), flush=True)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(1000)
    nl.request_interaction_range(rc)
    off, j, dc = nl.get_list(p)
    ro, rj, rdc = O.nl_build(pos, cell, pbc, rc, avgn=1000)
    assert np.array_equal(off, ro)
    assert np.array_equal(j, rj)
    assert np.array_equal(dc, rdc)


def test_neighbor_list_crowded_bin():
    # 1500 atoms inside one bin: the per-bin ordering of the counting sort
    # hands over to the radix sort, and the list stays bit-identical
    rng = np.random.default_rng(21)
    pos = 40.0 + rng.random((1500, 3)) * 2.0
    cell = np.diag([100.0, 100.0, 100.0])
    pbc = np.array([True, False, True])
    Z = np.full(len(pos), 14, np.int32)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(1600)
    nl.request_interaction_range(4.0)
    off, j, dc = nl.get_list(p)
    ro, rj, rdc = O.nl_build(pos, cell, pbc, 4.0, avgn=1600)
    assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)


def test_neighbor_list_overflow_error():
    pos, cell, pbc, Z = W.diamond_si(2, sigma=0.0)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(4)
    nl.request_interaction_range(5.0)
    with pytest.raises(AgpuError) as ei:
        nl.update(p)
    assert ei.value.code == 3


def test_neighbor_list_rebuilds_on_position_change():
    pos, cell, pbc, Z = W.diamond_si(3, sigma=0.05)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    nl.request_interaction_range(3.0)
    a = nl.get_list(p)
    p.coordinates[5] += 3 * cell[1]
    p.I_changed_positions()
    b = nl.get_list(p)
    ref = O.nl_build(p.coordinates, cell, pbc, 3.0)
    assert np.array_equal(b[2], ref[2]) and not np.array_equal(a[2], b[2])


# ---------------------------------------------------------------------------
# Tersoff
# ---------------------------------------------------------------------------

def run_tersoff(pos, cell, pbc, Z, params=None, mask=None, per_at=False, precision="fp64", pot=None):
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = pot or (N.Tersoff(**params) if params else N.Tersoff())
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    out = pot.energy_and_forces(p, nl, mask=mask, epot_per_at=per_at, wpot_per_at=per_at)
    return out, p, nl


def oracle_tersoff(pos, cell, pbc, Z, params, mask=None, per_at=False):
    el_map = {P.ATOMIC_NUMBERS[s]: k for k, s in enumerate(params["el"])}
    el = np.array([el_map.get(int(z), -1) for z in Z], np.int32)
    nl = O.nl_build(pos, cell, pbc, max(params["r2"]))
    return O.tersoff(params, pos, el, cell, nl, mask=mask, per_at=per_at)


TERSOFF_CASES = [
    ("si_4096", lambda: W.diamond_si(8, sigma=0.05, seed=1)),
    ("si_hot", lambda: W.diamond_si(3, sigma=0.25, seed=9)),
    ("aC_small", lambda: read_cfg("aC_small.cfg")),
    ("aC", lambda: read_cfg("aC.cfg")),
]


@pytest.mark.parametrize("name,make", TERSOFF_CASES, ids=[c[0] for c in TERSOFF_CASES])
def test_tersoff_fp64_parity(name, make):
    pos, cell, pbc, Z = make()
    print("tersoff case", name, len(pos), flush=True)
    prm = P.Tersoff_PRB_39_5566_Si_C
    (e, f, w, epa, _, _, wpa, _), p, nl = run_tersoff(pos, cell, pbc, Z, per_at=True)
    re_, rf, rw, repa, rwpa = oracle_tersoff(pos, cell, pbc, Z, prm, per_at=True)
    assert abs(e - re_) <= REL * abs(re_)
    assert close(f, rf) and close(w, rw)
    assert close(epa, repa) and close(wpa, rwpa)


def test_tersoff_sic_and_mask():
    pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.06, seed=11)
    Z = np.where(np.arange(len(pos)) % 8 >= 4, 14, 6).astype(np.int32)
    prm = P.Tersoff_PRB_39_5566_Si_C
    mask = (np.arange(len(pos)) % 3 != 0).astype(np.int32)
    (e, f, w, *_), p, nl = run_tersoff(pos, cell, pbc, Z, mask=mask)
    re_, rf, rw = oracle_tersoff(pos, cell, pbc, Z, prm, mask=mask)
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)


def test_tersoff_custom_parameters_and_unknown_atoms():
    pos, cell, pbc, Z = W.diamond_si(3, sigma=0.08, seed=4)
    Z = Z.copy()
    Z[::11] = 29      # atoms the potential ignores
    prm = dict(P.Tersoff_Si)
    prm["mubo"] = [0.35]
    prm["m"] = [3]
    prm["omega"] = [1.1]
    (e, f, w, *_), p, nl = run_tersoff(pos, cell, pbc, Z, params=prm)
    re_, rf, rw = oracle_tersoff(pos, cell, pbc, Z, prm)
    assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)


def test_tersoff_mixed_precision_error():
    pos, cell, pbc, Z = W.diamond_si(8, sigma=0.05, seed=1)
    prm = P.Tersoff_PRB_39_5566_Si_C
    (e, f, w, *_), p, nl = run_tersoff(pos, cell, pbc, Z, precision="mixed")
    re_, rf, rw = oracle_tersoff(pos, cell, pbc, Z, prm)
    rms = np.sqrt(np.mean((f - rf) ** 2))
    print("mixed tersoff: dE/N %.3e  F rms %.3e  F max %.3e  dW %.3e" % (
        abs(e - re_) / len(pos), rms, np.abs(f - rf).max(), np.abs(w - rw).max()))
    assert rms <= 1e-5                     # eV/A
    assert abs(e - re_) / len(pos) <= 1e-5  # eV/atom


def test_tersoff_bitwise_reproducible():
    pos, cell, pbc, Z = W.diamond_si(6, sigma=0.1, seed=2)
    (e1, f1, w1, *_), p, nl = run_tersoff(pos, cell, pbc, Z)
    (e2, f2, w2, *_), _, _ = run_tersoff(pos, cell, pbc, Z)
    assert e1 == e2 and np.array_equal(f1, f2) and np.array_equal(w1, w2)


# ---------------------------------------------------------------------------
# EAM
# ---------------------------------------------------------------------------

def fs_setfl(tmpdir, nr=10000):
    path = os.path.join(tmpdir, "fs_cu.eam.alloy")
    with open(path, "w") as f:
        f.write(W.fs_cu_setfl_text(nr=nr, nrho=nr))
    return path


def run_eam(path, pos, cell, pbc, Z, mask=None, per_at=False, precision="fp64"):
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.TabulatedAlloyEAM(fn=path)
    pot.set_precision(precision)
    pot.bind_to(p, nl)
    return pot.energy_and_forces(p, nl, mask=mask, epot_per_at=per_at, wpot_per_at=per_at)


def oracle_eam(path, pos, cell, pbc, Z, mask=None, per_at=False):
    t = read_setfl_tables(path)
    el_map = {z: k for k, z in enumerate(t["Z"])}
    el = np.array([el_map.get(int(z), -1) for z in Z], np.int32)
    nl = O.nl_build(pos, cell, pbc, t["cutoff"])
    return O.eam(t, pos, el, cell, nl, mask=mask, per_at=per_at)


@pytest.mark.parametrize("table", ["mishin", "fs"])
def test_eam_fp64_parity(table):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(GOLDEN, "Cu_mishin1.eam.alloy") if table == "mishin" else fs_setfl(d)
        pos, cell, pbc, Z = W.fcc_cu(6, sigma=0.08, seed=2)
        e, f, w, epa, _, _, wpa, _ = run_eam(path, pos, cell, pbc, Z, per_at=True)
        re_, rf, rw, repa, rwpa = oracle_eam(path, pos, cell, pbc, Z, per_at=True)
        assert abs(e - re_) <= REL * abs(re_)
        assert close(f, rf) and close(w, rw)
        assert close(epa, repa) and close(wpa, rwpa)
        mask = (np.arange(len(pos)) % 4 != 1).astype(np.int32)
        e, f, w, *_ = run_eam(path, pos, cell, pbc, Z, mask=mask)
        re_, rf, rw = oracle_eam(path, pos, cell, pbc, Z, mask=mask)
        assert abs(e - re_) <= REL * abs(re_) and close(f, rf) and close(w, rw)


def test_eam_mixed_precision_error():
    with tempfile.TemporaryDirectory() as d:
        path = fs_setfl(d)
        pos, cell, pbc, Z = W.fcc_cu(10, sigma=0.08, seed=2)
        e, f, w, *_ = run_eam(path, pos, cell, pbc, Z, precision="mixed")
        re_, rf, rw = oracle_eam(path, pos, cell, pbc, Z)
        rms = np.sqrt(np.mean((f - rf) ** 2))
        print("mixed eam: dE/N %.3e  F rms %.3e  F max %.3e" % (abs(e - re_) / len(pos), rms,
                                                               np.abs(f - rf).max()))
        assert rms <= 1e-5 and abs(e - re_) / len(pos) <= 1e-5


def test_eam_full_size_properties():
    # 256 000-atom perfect fcc: closed-form energy per atom, zero forces
    with tempfile.TemporaryDirectory() as d:
        path = fs_setfl(d)
        pos, cell, pbc, Z = W.fcc_cu(40, sigma=0.0)
        e, f, w, *_ = run_eam(path, pos, cell, pbc, Z)
        e_ref, _, _ = W.fs_cu_perfect_energy()
        assert abs(e / len(pos) - e_ref) < 1e-6
        assert np.abs(f).max() < 1e-9
        pos, cell, pbc, Z = W.fcc_cu(40, sigma=0.08, seed=2)
        e, f, w, *_ = run_eam(path, pos, cell, pbc, Z)
        assert np.abs(f.sum(axis=0)).max() < 1e-8      # Newton's third law


# ---------------------------------------------------------------------------
# Ewald
# ---------------------------------------------------------------------------

def test_ewald_parity_and_potential():
    pos, cell, pbc, Z, q = W.rocksalt(3, sigma=0.1, seed=5)
    rc = 8.0
    beta = W.ewald_beta(rc, 1e-8)
    kmax = W.ewald_kmax(beta, 1e-8)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(600)
    pot = N.Ewald(cutoff=rc, beta=beta, kmax=kmax)
    pot.bind_to(p, nl)
    e, f, w = pot.energy_and_forces(p, nl, q)
    phi = pot.potential(p, nl, q)
    onl = O.nl_build(pos, cell, pbc, rc, avgn=600)
    er, fr, wr, pr = O.ewald_real(pos, q, cell, onl, rc, beta)
    ek, fk, wk, pk = O.ewald_kspace(pos, q, cell, beta, kmax)
    assert abs(e - (er + ek)) <= REL * abs(er + ek)
    assert close(f, fr + fk) and close(w, wr + wk) and close(phi, pr + pk)


def test_ewald_madelung_full_size():
    pos, cell, pbc, Z, q = W.rocksalt(30, sigma=0.0)
    rc = 10.0
    beta = W.ewald_beta(rc, 1e-8)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(220)
    pot = N.Ewald(cutoff=rc, beta=beta, kmax=W.ewald_kmax(beta, 1e-8))
    pot.bind_to(p, nl)
    e, f, w = pot.energy_and_forces(p, nl, q)
    assert abs(e / len(pos) + 1.747564594633182 / 5.64) < 1e-7
    assert np.abs(f).max() < 1e-6


def test_ewald_requires_full_pbc():
    pos, cell, pbc, Z, q = W.rocksalt(2, sigma=0.0)
    p = particles(pos, cell, np.array([True, True, False]), Z)
    nl = N.Neighbors(600)
    pot = N.Ewald(cutoff=6.0)
    with pytest.raises(AgpuError) as ei:
        pot.bind_to(p, nl)
    assert ei.value.code == 8


def test_per_bond_request_is_unsupported_for_ewald():
    """Classes without per-bond output in the reference (no per_bond
    feature) refuse the arrays as gen_factory.py:303-304 does."""
    pos, cell, pbc, Z, q = W.rocksalt(2, sigma=0.0, seed=5)
    p = particles(pos, cell, pbc, Z)
    nl = N.Neighbors(300)
    pot = N.Ewald(cutoff=6.0)
    pot.bind_to(p, nl)
    with pytest.raises(AgpuError) as ei:
        # the potential-class argument list (Coulomb classes have their own)
        N._Potential.energy_and_forces(pot, p, nl, charges=q, epot_per_bond=True)
    assert ei.value.code == 6
