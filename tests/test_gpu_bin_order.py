"""Bin-order storage.

The force kernels name atoms by their slot in the list's bin order: the
staging rows hold neighbour slots, and positions, elements and charges are
gathered from per-slot records, so a neighbour's record sits next to the
home atom's own whatever numbering the caller uses (the reference sorts
particles by cell for the same reason, neighbors.f90:464,561-564).  Each case
here evaluates a system and a random renumbering of it: both match the
oracle at 1e-10, and the renumbered results are the original ones renumbered
(to 1e-10: the reference's row order follows the caller's numbering inside a
bin, so sums run in another order).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from helpers import GOLDEN, read_cfg, read_setfl_tables

pytestmark = pytest.mark.gpu

REL = 1e-10
MISHIN = os.path.join(GOLDEN, "Cu_mishin1.eam.alloy")


def close(a, b, rel=REL):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(1.0, float(np.abs(b).max()) if b.size else 1.0)
    return float(np.abs(a - b).max()) <= rel * scale


def _perm(n, seed):
    return np.random.default_rng(seed).permutation(n)


def _check(gpu, ref, perm_gpu, e_key=0):
    """gpu/ref: (e, f, w, epa, wpa) of the original order; perm_gpu: the same
    from the renumbered system, mapped back to the original numbering."""
    assert abs(gpu[0] - ref[0]) <= REL * max(1.0, abs(ref[0]))
    assert abs(perm_gpu[0] - ref[0]) <= REL * max(1.0, abs(ref[0]))
    for a, b, c in zip(gpu[1:], ref[1:], perm_gpu[1:]):
        if b is None:
            continue
        assert close(a, b) and close(c, b) and close(c, a)


def _back(x, perm):
    """per-atom array of the renumbered system (atom k = original perm[k])
    back to the original numbering"""
    if x is None:
        return None
    out = np.empty_like(np.asarray(x))
    out[perm] = np.asarray(x)
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_eam_renumbered(precision):
    pos, cell, pbc, Z = W.fcc_cu(6, sigma=0.08, seed=21)
    perm = _perm(len(pos), 1)
    print("bin order eam", precision, len(pos), flush=True)

    def run(ps, Zs):
        p = N.from_atoms(ps, cell, pbc, Zs)
        nl = N.Neighbors(100)
        pot = N.TabulatedAlloyEAM(fn=MISHIN)
        pot.set_precision(precision)
        pot.bind_to(p, nl)
        e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, nl, epot_per_at=True, wpot_per_at=True)
        return (e, f, w, epa, wpa), p, nl

    a, _, _ = run(pos, Z)
    b, pb, nlb = run(pos[perm], Z[perm])
    b = (b[0], _back(b[1], perm), b[2], _back(b[3], perm), _back(b[4], perm))
    t = read_setfl_tables(MISHIN)
    el = np.zeros(len(pos), np.int32)
    ref = O.eam(t, pos, el, cell, O.nl_build(pos, cell, pbc, t["cutoff"]), per_at=True)
    if precision == "fp64":
        _check(a, ref, b)
        # the caller-order list exported after a build read by slot
        off, j, dc = nlb.get_list(pb)
        ro, rj, rdc = O.nl_build(pos[perm], cell, pbc, t["cutoff"])
        assert np.array_equal(off, ro) and np.array_equal(j, rj) and np.array_equal(dc, rdc)
    else:
        rms = lambda x, y: float(np.sqrt(np.mean((np.asarray(x) - np.asarray(y)) ** 2)))
        assert rms(a[1], ref[1]) <= 1e-5 and rms(b[1], ref[1]) <= 1e-5


def test_tersoff_renumbered():
    pos, cell, pbc, Z = W.diamond_si(6, sigma=0.06, seed=22)
    rng = np.random.default_rng(3)
    Z = np.where(rng.random(len(pos)) < 0.25, 6, 14).astype(np.int32)
    perm = _perm(len(pos), 2)
    mask = (np.arange(len(pos)) % 5 != 2).astype(np.int32)
    print("bin order tersoff", len(pos), flush=True)
    prm = P.Tersoff_PRB_39_5566_Si_C

    def run(ps, Zs, m):
        p = N.from_atoms(ps, cell, pbc, Zs)
        nl = N.Neighbors(100)
        pot = N.Tersoff()
        pot.bind_to(p, nl)
        e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, nl, mask=m, epot_per_at=True, wpot_per_at=True)
        return (e, f, w, epa, wpa)

    el_map = {P.ATOMIC_NUMBERS[s]: k for k, s in enumerate(prm["el"])}
    el = np.array([el_map.get(int(z), -1) for z in Z], np.int32)
    onl = O.nl_build(pos, cell, pbc, max(prm["r2"]))
    for m in (None, mask):
        a = run(pos, Z, m)
        b = run(pos[perm], Z[perm], None if m is None else m[perm])
        b = (b[0], _back(b[1], perm), b[2], _back(b[3], perm), _back(b[4], perm))
        ref = O.tersoff(prm, pos, el, cell, onl, mask=m, per_at=True)
        _check(a, ref, b)


@pytest.mark.parametrize("per_at", [True, False], ids=["per_atom", "totals"])
def test_rebo2_renumbered(per_at):
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    rng = np.random.default_rng(5)
    Z = np.where(rng.random(len(pos)) < 0.2, 1, 6).astype(np.int32)
    perm = _perm(len(pos), 3)
    print("bin order rebo2", len(pos), per_at, flush=True)

    def run(ps, Zs):
        p = N.from_atoms(ps, cell, pbc, Zs)
        nl = N.Neighbors(100)
        pot = N.Rebo2(dihedral=True)
        pot.bind_to(p, nl)
        e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, nl, epot_per_at=per_at, wpot_per_at=per_at)
        return (e, f, w, epa, wpa)

    a = run(pos, Z)
    b = run(pos[perm], Z[perm])
    b = (b[0], _back(b[1], perm), b[2], _back(b[3], perm), _back(b[4], perm))
    ref = O.Rebo2(dihedral=True)(pos, Z, cell, O.nl_build(pos, cell, pbc, 2.0), per_at=True)
    if not per_at:
        ref = ref[:3] + (None, None)
    _check(a, ref, b)


def test_ewald_renumbered():
    pos, cell, pbc, Z, q = W.rocksalt(3, sigma=0.1, seed=23)
    perm = _perm(len(pos), 4)
    rc = 8.0
    beta = W.ewald_beta(rc, 1e-8)
    kmax = W.ewald_kmax(beta, 1e-8)
    print("bin order ewald", len(pos), flush=True)

    def run(ps, Zs, qs):
        p = N.from_atoms(ps, cell, pbc, Zs)
        nl = N.Neighbors(600)
        pot = N.Ewald(cutoff=rc, beta=beta, kmax=kmax)
        pot.bind_to(p, nl)
        e, f, w = pot.energy_and_forces(p, nl, qs)
        phi = pot.potential(p, nl, qs)
        return (e, f, w, phi)

    a = run(pos, Z, q)
    b = run(pos[perm], Z[perm], q[perm])
    b = (b[0], _back(b[1], perm), b[2], _back(b[3], perm))
    onl = O.nl_build(pos, cell, pbc, rc, avgn=600)
    er, fr, wr, pr = O.ewald_real(pos, q, cell, onl, rc, beta)
    ek, fk, wk, pk = O.ewald_kspace(pos, q, cell, beta, kmax)
    ref = (er + ek, fr + fk, wr + wk, pr + pk)
    _check(a, ref, b)
