"""NVE dynamics on the GPU through the C-ABI (velocity Verlet around the
potential, Verlet-shell list, steps replayed as HIP graphs).

Bars:
  - the first steps follow a velocity-Verlet integration written here in
    numpy over the oracle's forces, within 1e-9 A;
  - graph replay and step-by-step give the same trajectory, bitwise, also
    when the list goes stale inside a replay (rolled back and redone);
  - total energy is conserved (Tersoff Si, 512 atoms, 1 fs, 400 steps);
  - EAM runs the same way.
"""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W

pytestmark = pytest.mark.gpu

M_SI = 28.0855
DT = 0.1            # 1.018 fs in Angstrom sqrt(amu/eV)
KB = 8.617333262e-5


def _si(n=4, T=600.0, seed=1):
    pos, cell, pbc, Z = W.diamond_si(n, sigma=0.02, seed=seed)
    rng = np.random.default_rng(seed + 100)
    v = rng.normal(0.0, np.sqrt(KB * T / M_SI), pos.shape)
    v -= v.mean(axis=0)
    return pos, cell, pbc, Z, v


def _md(pot, pos, cell, pbc, Z, v, skin, dt=DT, masses=M_SI):
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(100)
    return N.MD(pot, p, nl, masses, dt, skin=skin, v0=v)


def test_md_matches_numpy_verlet_over_oracle_forces():
    pos, cell, pbc, Z, v = _si(3, T=900.0)
    md = _md(N.Tersoff(), pos, cell, pbc, Z, v, skin=0.5)
    nsteps = 10
    md.run(nsteps, graph_steps=0)
    r_gpu, v_gpu, _, _ = md.state()
    prm = P.Tersoff_PRB_39_5566_Si_C
    el = np.ones(len(pos), np.int32)

    def force(r):
        nl = O.nl_build(r, cell, pbc, 3.0)
        return O.tersoff(prm, r, el, cell, nl)[1]
    r, vv = pos.copy(), v.copy()
    f = force(r)
    for _ in range(nsteps):
        vv = vv + 0.5 * DT * f / M_SI
        r = r + DT * vv
        f = force(r)
        vv = vv + 0.5 * DT * f / M_SI
    print("md vs numpy: |dr| %.3e  |dv| %.3e" % (np.abs(r_gpu - r).max(), np.abs(v_gpu - vv).max()), flush=True)
    assert np.abs(r_gpu - r).max() < 1e-9
    assert np.abs(v_gpu - vv).max() < 1e-9


@pytest.mark.parametrize("skin", [0.6, 0.05])
def test_md_graph_equals_stepwise(skin):
    pos, cell, pbc, Z, v = _si(4, T=1500.0, seed=3)
    a = _md(N.Tersoff(), pos, cell, pbc, Z, v, skin=skin)
    b = _md(N.Tersoff(), pos, cell, pbc, Z, v, skin=skin)
    ea = a.run(60, graph_steps=10)
    eb = b.run(60, graph_steps=0)
    ra, va, fa, ca = a.state()
    rb, vb, fb, cb = b.state()
    print("md graph (skin %.2f): %s / stepwise %s" % (skin, ca, cb), flush=True)
    assert ea == eb
    assert np.array_equal(ra, rb) and np.array_equal(va, vb) and np.array_equal(fa, fb)
    assert ca["steps"] == cb["steps"] == 60 and ca["builds"] == cb["builds"]
    if skin > 0.3:
        assert ca["replays"] > 0


def test_md_energy_conservation_tersoff():
    pos, cell, pbc, Z, v = _si(4, T=600.0, seed=5)
    md = _md(N.Tersoff(), pos, cell, pbc, Z, v, skin=0.5)
    e0 = md.run(0) + md.kinetic_energy()
    drift = []
    for _ in range(8):
        ep = md.run(50, graph_steps=10)
        drift.append(ep + md.kinetic_energy() - e0)
    _, _, _, c = md.state()
    print("md energy drift per atom:", np.array(drift) / len(pos), c, flush=True)
    assert np.abs(np.array(drift)).max() / len(pos) < 2e-4


def test_md_eam_graph_equals_stepwise(tmp_path):
    fn = tmp_path / "fs_cu.eam.alloy"
    fn.write_text(W.fs_cu_setfl_text(nr=4000, nrho=4000))
    pos, cell, pbc, Z = W.fcc_cu(4, sigma=0.02, seed=2)
    rng = np.random.default_rng(9)
    v = rng.normal(0.0, np.sqrt(KB * 800.0 / 63.546), pos.shape)
    a = _md(N.TabulatedAlloyEAM(fn=str(fn)), pos, cell, pbc, Z, v, skin=0.4, masses=63.546)
    b = _md(N.TabulatedAlloyEAM(fn=str(fn)), pos, cell, pbc, Z, v, skin=0.4, masses=63.546)
    a.run(40, graph_steps=8)
    b.run(40, graph_steps=0)
    ra, va, _, ca = a.state()
    rb, vb, _, cb = b.state()
    assert np.array_equal(ra, rb) and np.array_equal(va, vb)
    assert ca["replays"] > 0 and ca["builds"] == cb["builds"]


def _bcc_w(n=3, T=900.0, seed=6):
    pos, cell = W.lattice(np.array([[0, 0, 0], [.5, .5, .5]]), n, 3.165)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0, 0.03, pos.shape)
    Z = np.full(len(pos), 74, np.int32)
    v = rng.normal(0.0, np.sqrt(KB * T / 183.84), pos.shape)
    v -= v.mean(axis=0)
    return pos, cell, np.array([True] * 3), Z, v


@pytest.mark.parametrize("kind", ["tersoff_scr", "juslin_scr"])
def test_md_screened_graph_request_equals_stepwise(kind):
    """The screened bond-order classes size their screening records on the
    host every call, so a run that asks for graph replays runs them step by
    step instead (no error), with the same trajectory bit for bit."""
    if kind == "tersoff_scr":
        pos, cell, pbc, Z, v = _si(3, T=900.0, seed=4)
        make, mass = N.TersoffScr, M_SI
    else:
        pos, cell, pbc, Z, v = _bcc_w()
        make, mass = N.JuslinScr, 183.84
    a = _md(make(), pos, cell, pbc, Z, v, skin=0.5, masses=mass)
    b = _md(make(), pos, cell, pbc, Z, v, skin=0.5, masses=mass)
    ea = a.run(40, graph_steps=32)
    eb = b.run(40, graph_steps=0)
    ra, va, fa, ca = a.state()
    rb, vb, fb, cb = b.state()
    print("md %s: %s / %s" % (kind, ca, cb), flush=True)
    assert ea == eb
    assert np.array_equal(ra, rb) and np.array_equal(va, vb) and np.array_equal(fa, fb)
    assert ca["steps"] == 40 and ca["replays"] == 0
