"""Staged and CSR neighbour rows.

EAM and the Ewald real-space sum read the list in bin order straight from the
staging rows of the build; the copy into CSR rows (offsets + entries in atom
order, the layout python_neighbors.f90:690-760 fills) runs only when a caller
needs it: an export, or a bond-order potential on the same list.  Both layouts
hold the same entries in the same order, so every result here is compared
bit for bit, and the exported list against the oracle's.
"""

import os

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import GOLDEN, read_setfl_tables

pytestmark = pytest.mark.gpu

MISHIN = os.path.join(GOLDEN, "Cu_mishin1.eam.alloy")


def _eam(p, nl, prebuild):
    pot = N.TabulatedAlloyEAM(fn=MISHIN)
    pot.bind_to(p, nl)
    if prebuild:
        nl.update(p)          # CSR rows first: the potential reads those
    return pot.energy_and_forces(p, nl, epot_per_at=True, wpot_per_at=True)


def test_eam_staged_rows_equal_csr_rows_and_export():
    pos, cell, pbc, Z = W.fcc_cu(6, sigma=0.08, seed=4)
    print("eam staged vs csr", len(pos), flush=True)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    staged = _eam(p, nl, prebuild=False)
    p2 = N.from_atoms(pos, cell, pbc, Z)
    csr = _eam(p2, N.Neighbors(100), prebuild=True)
    for a, b in zip(staged, csr):
        if a is not None:
            assert np.array_equal(np.asarray(a), np.asarray(b))
    # the export after a staged-only build copies the rows on demand
    off, j, dc = nl.get_list(p)
    ro, rj, rdc = O.nl_build(pos, cell, pbc, read_setfl_tables(MISHIN)["cutoff"])
    assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)
    # moved atoms: a second staged build, then the export again
    p.coordinates[7] += 2.5 * cell[2]
    p.coordinates[11] += np.array([0.3, -0.2, 0.1])
    p.I_changed_positions()
    e2 = _eam(p, nl, prebuild=False)
    assert e2[0] != staged[0]
    off, j, dc = nl.get_list(p)
    ro, rj, rdc = O.nl_build(p.coordinates, cell, pbc, read_setfl_tables(MISHIN)["cutoff"])
    assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)


def test_bond_order_after_ewald_on_one_list():
    # Ewald leaves the rows staged; Tersoff on the same list asks for CSR rows
    pos, cell, pbc, Z = W.diamond_si(4, sigma=0.05, seed=3)
    q = np.where(np.arange(len(pos)) % 2 == 0, 0.4, -0.4)
    rc = 6.0
    beta = W.ewald_beta(rc, 1e-8)
    kmax = W.ewald_kmax(beta, 1e-8)
    print("ewald + tersoff shared list", len(pos), flush=True)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(200)
    ew = N.Ewald(cutoff=rc, beta=beta, kmax=kmax)
    ters = N.Tersoff()
    ew.bind_to(p, nl)
    ters.bind_to(p, nl)
    e_ew = ew.energy_and_forces(p, nl, q)
    e_t = ters.energy_and_forces(p, nl)
    # references on lists of their own, built with CSR rows first
    pr = N.from_atoms(pos, cell, pbc, Z)
    nr = N.Neighbors(200)
    nr.request_interaction_range(rc)
    ew_r = N.Ewald(cutoff=rc, beta=beta, kmax=kmax)
    ew_r.bind_to(pr, nr)
    nr.update(pr)
    r_ew = ew_r.energy_and_forces(pr, nr, q)
    t_r = N.Tersoff()
    t_r.bind_to(pr, nr)
    r_t = t_r.energy_and_forces(pr, nr)
    for a, b in zip(e_ew, r_ew):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    for a, b in zip(e_t[:3], r_t[:3]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_capacity_overflow_without_csr_copy():
    # the reference's nat*avgn rule is applied to builds whose rows stay staged
    pos, cell, pbc, Z = W.fcc_cu(3, sigma=0.0)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(8)
    pot = N.TabulatedAlloyEAM(fn=MISHIN)
    pot.bind_to(p, nl)
    with pytest.raises(AgpuError) as ei:
        pot.energy_and_forces(p, nl)
    assert ei.value.code == 3


@pytest.mark.parametrize("shape", ["ortho", "triclinic"])
def test_pair_modes_with_wrapped_atoms(shape):
    # dense bins (a wavefront per bin) and sparse ones (a thread per atom),
    # with a third of the atoms moved by whole cell vectors so that both home
    # atoms inside the cell and wrapped ones are paired
    pos, cell, pbc, Z = W.fcc_cu(6, sigma=0.08, seed=8)
    rng = np.random.default_rng(8)
    if shape == "triclinic":
        cell = cell + np.array([[0.0, 0.0, 0.0], [2.1, 0.0, 0.0], [-1.3, 1.7, 0.0]])
        pos = (pos @ np.linalg.inv(W.fcc_cu(6, sigma=0.0)[1])) @ cell
    sel = rng.random(len(pos)) < 0.33
    pos[sel] += rng.integers(-3, 4, (int(sel.sum()), 3)) @ cell
    rc = 5.5
    ref = O.nl_build(pos, cell, pbc, rc)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    nl.request_interaction_range(rc)
    pot = N.Tersoff()            # handle for set_launch; its range is below rc
    pot.bind_to(p, nl)
    for mode in (1, 2):
        print("wrapped", shape, "pair mode", mode, flush=True)
        pot.set_launch(nl, pair_mode=mode)
        p.changed_positions()
        off, j, dc = nl.get_list(p)
        assert np.array_equal(off, ref[0]) and np.array_equal(j, ref[1]) and np.array_equal(dc, ref[2])


def _sc_lattice(n, a, jitter, seed):
    g = np.arange(n, dtype=np.float64) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = pos + np.random.default_rng(seed).normal(0.0, jitter, pos.shape)
    return pos, np.eye(3) * (n * a)


@pytest.mark.parametrize("case", ["at_cutoff", "at_cutoff_triclinic", "far_cluster", "very_far_cluster"])
def test_sparse_search_screening_band(case):
    """The thread-per-atom search screens candidates in FP32 and takes the
    FP64 test only near the cut-off: lattice distances equal to the cut-off
    up to 1e-13 A (every such pair decided by the FP64 test), a triclinic
    cell, and clusters far from the origin in an open box (a wide band, and
    one so wide that every candidate takes the FP64 test).  Rows equal the
    oracle's entry for entry in both pair modes."""
    a = 3.3                      # above the Tersoff range, so the list cut-off is rc
    if case.startswith("at_cutoff"):
        pos, cell = _sc_lattice(12, a, 1e-13, 3)
        pbc = np.array([True, True, True])
        if case == "at_cutoff_triclinic":
            shear = np.array([[1.0, 0.0, 0.0], [0.25, 1.0, 0.0], [0.0, -0.5, 1.0]])
            pos, cell = pos @ shear, cell @ shear
        rc = a
    else:
        off = 1.5e3 if case == "far_cluster" else 2.0e5
        pos, _ = _sc_lattice(8, a, 0.05, 4)
        pos = pos + off
        cell = np.eye(3) * (8 * a)
        pbc = np.array([False, False, False])
        rc = 1.05 * a
    Z = np.full(len(pos), 14, np.int32)
    ref = O.nl_build(pos, cell, pbc, rc)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    nl.request_interaction_range(rc)
    pot = N.Tersoff()
    pot.bind_to(p, nl)
    for mode in (1, 2):
        print("screening band", case, "pair mode", mode, flush=True)
        pot.set_launch(nl, pair_mode=mode)
        p.changed_positions()
        off_, j, dc = nl.get_list(p)
        assert np.array_equal(off_, ref[0]) and np.array_equal(j, ref[1]) and np.array_equal(dc, ref[2])


def test_staging_stride_follows_rows():
    """The staging stride narrows to the longest row after a build and widens
    again through the overflow rebuild when rows grow: the lists stay the
    oracle's, and a bond-order potential on them the oracle's."""
    from atomistica_amd import parameters as P
    pos, cell, pbc, Z = W.diamond_si(4, sigma=0.05, seed=8)
    prm = P.Tersoff_PRB_39_5566_Si_C
    rc = max(prm["r2"])
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.Tersoff()
    pot.bind_to(p, nl)
    el = np.ones(len(pos), np.int32)
    for scale in (1.0, 1.0, 0.76, 0.76, 1.0):
        # 0.76: the second shell (3.84 A) moves inside the 3.0 A cut-off
        print("staging stride case scale", scale, flush=True)
        ps = pos * scale
        cs = cell * scale
        p.set_cell(cs, pbc)
        p.coordinates[:] = ps
        p.I_changed_positions()
        e, f, w, *_ = pot.energy_and_forces(p, nl)
        off, j, dc = nl.get_list(p)
        ro, rj, rdc = O.nl_build(ps, cs, pbc, rc)
        assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)
        e0, f0, w0 = O.tersoff(prm, ps, el, cs, O.nl_build(ps, cs, pbc, rc))
        assert abs(e - e0) <= 1e-10 * abs(e0)
        scale_f = max(1.0, np.abs(f0).max())
        assert np.abs(f - f0).max() <= 1e-10 * scale_f


def _eam_oracle(pos, cell, pbc, Z, avgn=100):
    t = read_setfl_tables(MISHIN)
    el = np.zeros(len(pos), np.int32)
    return O.eam(t, pos, el, cell, O.nl_build(pos, cell, pbc, t["cutoff"], avgn=avgn))


def _close(a, b, rel=1e-10):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def test_eam_deferred_list_check_redo():
    """EAM checks its list build's flags with its own at the end of the call.
    A build that outgrew the staging stride, and one whose crowded bin hands
    over to the radix sort, are marked unusable on the device: the passes
    write nothing, the call rebuilds and runs again, and the results are the
    oracle's."""
    # rows outgrow the stride the previous build narrowed to
    pos, cell, pbc, Z = W.fcc_cu(4, sigma=0.05, seed=12)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(200)
    pot = N.TabulatedAlloyEAM(fn=MISHIN)
    pot.bind_to(p, nl)
    for scale in (1.0, 1.0, 0.85, 0.85):
        print("deferred check, stride case scale", scale, flush=True)
        ps, cs = pos * scale, cell * scale
        p.set_cell(cs, pbc)
        p.coordinates[:] = ps
        p.I_changed_positions()
        e, f, w, *_ = pot.energy_and_forces(p, nl)
        e0, f0, w0 = _eam_oracle(ps, cs, pbc, Z, avgn=200)
        assert abs(e - e0) <= 1e-10 * abs(e0) and _close(f, f0) and _close(w, w0)
    # a crowded bin: the counting sort hands over to the radix sort
    rng = np.random.default_rng(22)
    pos = np.concatenate([40.0 + rng.random((400, 3)) * 3.0, rng.random((200, 3)) * 100.0])
    cell = np.diag([100.0, 100.0, 100.0])
    pbc = np.array([True, True, True])
    Z = np.full(len(pos), 29, np.int32)
    print("deferred check, crowded bin", len(pos), flush=True)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(500)
    pot = N.TabulatedAlloyEAM(fn=MISHIN)
    pot.bind_to(p, nl)
    e, f, w, *_ = pot.energy_and_forces(p, nl)
    e0, f0, w0 = _eam_oracle(pos, cell, pbc, Z, avgn=500)
    assert abs(e - e0) <= 1e-10 * abs(e0) and _close(f, f0) and _close(w, w0)
    off, j, dc = nl.get_list(p)
    ro, rj, rdc = O.nl_build(pos, cell, pbc, read_setfl_tables(MISHIN)["cutoff"], avgn=500)
    assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)


def test_tersoff_deferred_list_check_redo():
    """The Tersoff row path checks a new build's flags with its own when the
    last build's rows fit it.  A row longer than the lane group the last
    build gave, and a row longer than the staging stride, each make the call
    redo itself (the first writes nothing); every call matches the oracle,
    the per-atom arrays too."""
    from atomistica_amd import parameters as P
    prm = P.Tersoff_PRB_39_5566_Si_C
    pos, cell, pbc, Z = W.diamond_si(3, sigma=0.0, seed=1)
    a0 = 5.431
    el = np.ones(len(pos), np.int32)             # Si of the (C, Si) set
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.Tersoff()                            # the (C, Si) set above
    pot.bind_to(p, nl)
    # atom 0 towards the tetrahedral site: its neighbours' rows reach 5
    # entries (frac 0.3; lanes of 4 so far) and its own 10 (frac 1.0; a
    # stride of 8 so far), then back to the lattice
    for frac, per_at in ((0.0, False), (0.3, True), (1.0, False), (0.0, True), (0.8, False)):
        ps = pos.copy()
        ps[0] = pos[0] + frac * np.full(3, a0 / 2)
        print("tersoff deferred check, frac", frac, flush=True)
        p.coordinates[:] = ps
        p.I_changed_positions()
        e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, nl, epot_per_at=per_at, wpot_per_at=per_at)
        ref = O.tersoff(prm, ps, el, cell, O.nl_build(ps, cell, pbc, max(prm["r2"])), per_at=True)
        assert abs(e - ref[0]) <= 1e-10 * abs(ref[0])
        assert _close(f, ref[1]) and _close(w, ref[2])
        if per_at:
            assert _close(epa, ref[3]) and _close(wpa, ref[4])
