"""Launch-shape autotuner through the C-ABI.

Bars:
  - both neighbour-list pair kernels (thread per atom, wavefront per bin
    block) give the oracle's list bit for bit, on dense and sparse bins;
  - every single-element EAM launch variant (lanes per atom x entries per
    lane) gives the oracle's energy, forces and virial within 1e-10
    relative, and the variants agree with each other to 1e-12;
  - autotune() reports its timings, keeps a variant, and the results after
    it still meet the oracle bar; set_launch() rejects unknown variants.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError
from helpers import read_setfl_tables

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def _setup(d, n=6, sigma=0.08, seed=2):
    path = os.path.join(d, "fs_cu.eam.alloy")
    with open(path, "w") as f:
        f.write(W.fs_cu_setfl_text(nr=4000, nrho=4000))
    pos, cell, pbc, Z = W.fcc_cu(n, sigma=sigma, seed=seed)
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    pot = N.TabulatedAlloyEAM(fn=path)
    pot.bind_to(p, nl)
    return path, pos, cell, pbc, Z, p, nl, pot


def _oracle(path, pos, cell, pbc, Z):
    t = read_setfl_tables(path)
    el = np.zeros(len(pos), np.int32)
    nl = O.nl_build(pos, cell, pbc, t["cutoff"])
    return O.eam(t, pos, el, cell, nl), t["cutoff"]


@pytest.mark.parametrize("bin_size", [0.0, 1.2])
def test_pair_modes_identical_lists(bin_size):
    with tempfile.TemporaryDirectory() as d:
        path, pos, cell, pbc, Z, p, nl, pot = _setup(d)
        (_, _, _), rc = _oracle(path, pos, cell, pbc, Z)
        ref = O.nl_build(pos, cell, pbc, rc, bin_size=bin_size)
        if bin_size:
            nl.set_bin_size(bin_size)
        for mode in (1, 2):
            print("pair mode", mode, "bin size", bin_size, flush=True)
            pot.set_launch(nl, pair_mode=mode)
            p.changed_positions()
            off, j, dc = nl.get_list(p)
            assert np.array_equal(off, ref[0]) and np.array_equal(j, ref[1]) and np.array_equal(dc, ref[2])


def test_eam_variants_parity():
    with tempfile.TemporaryDirectory() as d:
        path, pos, cell, pbc, Z, p, nl, pot = _setup(d)
        (e0, f0, w0), _ = _oracle(path, pos, cell, pbc, Z)
        nv = pot.set_launch(nl)
        assert nv >= 2
        res = []
        for k in range(nv):
            print("eam variant", k, flush=True)
            pot.set_launch(nl, variant=k)
            e, f, w, *_ = pot.energy_and_forces(p, nl)
            assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)
            res.append((e, f.copy(), w.copy()))
        for e, f, w in res[1:]:
            assert abs(e - res[0][0]) <= 1e-12 * abs(e0)
            assert close(f, res[0][1], 1e-12) and close(w, res[0][2], 1e-12)
        with pytest.raises(AgpuError):
            pot.set_launch(nl, variant=nv)
        with pytest.raises(AgpuError):
            pot.set_launch(nl, pair_mode=3)


def test_autotune_keeps_parity():
    with tempfile.TemporaryDirectory() as d:
        path, pos, cell, pbc, Z, p, nl, pot = _setup(d, n=8)
        (e0, f0, w0), _ = _oracle(path, pos, cell, pbc, Z)
        rep = pot.autotune(p, nl, reps=3)
        print(rep, flush=True)
        assert "nl_pairs ->" in rep and "eam ->" in rep
        e, f, w, *_ = pot.energy_and_forces(p, nl)
        assert abs(e - e0) <= REL * abs(e0) and close(f, f0) and close(w, w0)
