"""The LAMMPS emulation of the GPU tests, checked on the oracle alone.

LAMMPS hands a potential its local atoms plus ghost copies in an open box
and marks the local atoms with the mask (pair_atomistica.cpp:499-561);
reverse communication then returns ghost forces, energies and virials to
their owners.  For the oracle that round trip must reproduce its own
periodic evaluation; the GPU tests use this same emulation against the
oracle, so this pins the emulation itself.
"""
import os
import tempfile

import numpy as np
import pytest

from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from oracle import oracle as O


def ghosts(pos, cell, border):
    L = np.diag(cell)
    xs, owner = [], []
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                if a == b == c == 0:
                    continue
                img = pos + np.array([a, b, c]) * L
                inside = np.all((img > -border) & (img < L + border), axis=1)
                xs.append(img[inside])
                owner.append(np.nonzero(inside)[0])
    return np.concatenate(xs), np.concatenate(owner)


def reverse(arr, nlocal, gown):
    out = arr[:nlocal].copy()
    np.add.at(out, gown, arr[nlocal:])
    return out


def _wrap(pos, cell):
    L = np.diag(cell)
    return pos - np.floor(pos / L) * L


def _evaluate(kind, pos, cell, pbc, Z, mask, fn):
    if kind == "tersoff":
        prm = P.Tersoff_PRB_39_5566_Si_C
        m = {P.ATOMIC_NUMBERS[s]: k for k, s in enumerate(prm["el"])}
        el = np.array([m[int(z)] for z in Z], np.int32)
        nl = O.nl_build(pos, cell, pbc, max(prm["r2"]), avgn=200)
        return O.tersoff(prm, pos, el, cell, nl, mask=mask, per_at=True)
    t = O.read_setfl_tables(fn)
    nl = O.nl_build(pos, cell, pbc, t["cutoff"], avgn=200)
    return O.eam(t, pos, np.zeros(len(pos), np.int32), cell, nl, mask=mask, per_at=True)


@pytest.mark.parametrize("kind", ["tersoff", "eam"])
def test_masked_open_box_round_trip(kind):
    fn = None
    if kind == "tersoff":
        pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.06, seed=11)
        Z = np.where(np.arange(len(pos)) % 8 >= 4, 14, 6).astype(np.int32)
        border, rc = 6.5, 3.0
    else:
        fn = os.path.join(tempfile.mkdtemp(), "fs_cu.eam.alloy")
        with open(fn, "w") as fh:
            fh.write(W.fs_cu_setfl_text(nr=2000, nrho=2000))
        pos, cell, pbc, Z = W.fcc_cu(4, sigma=0.08, seed=3)
        border, rc = 6.0, 5.5
    pos = _wrap(pos, cell)
    gx, gown = ghosts(pos, cell, border)
    nlocal = len(pos)
    x = np.concatenate([pos, gx])
    Zall = np.concatenate([Z, Z[gown]])
    lo, hi = x.min(axis=0), x.max(axis=0)
    box = np.diag(hi - lo + 2 * (rc + 1.0))
    mask = (np.arange(len(x)) < nlocal).astype(np.int32)
    eo, fo, wo, eao, wao = _evaluate(kind, x - lo + rc + 1.0, box, [False] * 3, Zall, mask, fn)
    ep, fp, wp, eap, wap = _evaluate(kind, pos, cell, [True] * 3, Z, None, fn)
    tol = 1e-10 * max(1.0, np.abs(fp).max())
    assert abs(eo - ep) <= 1e-10 * abs(ep)
    assert np.abs(reverse(fo, nlocal, gown) - fp).max() <= tol
    assert np.abs(wo - wp).max() <= 1e-10 * max(1.0, np.abs(wp).max())
    assert np.abs(reverse(eao, nlocal, gown) - eap).max() <= 1e-10 * max(1.0, np.abs(eap).max())
    assert np.abs(reverse(wao, nlocal, gown) - wap).max() <= 1e-10 * max(1.0, np.abs(wap).max())
    # ghosts carry forces and shares of the local centres' energies (the
    # per-atom split puts half a bond on each end, bop_kernel.f90:1366-1368)
    assert np.abs(fo[nlocal:]).max() > 1e-3
    assert abs(eao.sum() - eo) <= 1e-10 * abs(eo)
