"""LAMMPS pair-style step through the C-ABI, with LAMMPS emulated on the host.

A periodic system is handed over the way LAMMPS does it
(pair_atomistica.cpp:172-561): the local atoms followed by ghost copies of
every atom within the border of the box, no periodicity.  After the call the
ghost forces are returned to their owners (LAMMPS' reverse communication,
newton on).  Energy, forces and virial must then equal the periodic
evaluation of the same system within 1e-10 relative.
"""
import numpy as np
import pytest

from atomistica_amd import native as N
from atomistica_amd import parameters as P
from atomistica_amd import workloads as W
from atomistica_amd._lib import AgpuError

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    scale = max(1.0, float(np.abs(b).max()))
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()) <= rel * scale


def ghosts(pos, cell, border):
    """Ghost copies of every atom whose image lies within `border` of the
    (orthorhombic) box, in LAMMPS order after the local atoms."""
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


def run_lammps(pot, pos, cell, Z, border):
    gx, gown = ghosts(pos, cell, border)
    nlocal = len(pos)
    x = np.concatenate([pos, gx])
    Zall = np.concatenate([Z, Z[gown]]).astype(np.int32)
    p = N.Particles()
    nl = N.Neighbors(200)
    eng, f, vir, _ = N.lammps_compute(pot, p, nl, nlocal, x, Zall)
    floc = f[:nlocal].copy()
    np.add.at(floc, gown, f[nlocal:])           # reverse communication
    return eng, floc, vir, len(gx)


def periodic(pot, pos, cell, Z):
    p = N.from_atoms(pos, cell, np.array([True] * 3), Z)
    nl = N.Neighbors(200)
    pot.bind_to(p, nl)
    e, f, w, *_ = pot.energy_and_forces(p, nl)
    vir = -np.array([w[0, 0], w[1, 1], w[2, 2], 0.5 * (w[0, 1] + w[1, 0]), 0.5 * (w[0, 2] + w[2, 0]),
                     0.5 * (w[1, 2] + w[2, 1])])
    return e, f, vir


def _wrap(pos, cell):
    L = np.diag(cell)
    return pos - np.floor(pos / L) * L


def test_lammps_tersoff_sic():
    pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.06, seed=11)
    Z = np.where(np.arange(len(pos)) % 8 >= 4, 14, 6).astype(np.int32)
    pos = _wrap(pos, cell)
    e0, f0, v0 = periodic(N.Tersoff(), pos, cell, Z)
    e1, f1, v1, ng = run_lammps(N.Tersoff(), pos, cell, Z, 2 * 3.0 + 0.5)
    print("lammps tersoff: %d local, %d ghosts, dE %.3e" % (len(pos), ng, e1 - e0), flush=True)
    assert abs(e1 - e0) <= REL * abs(e0)
    assert close(f1, f0) and close(v1, v0)


def test_lammps_eam(tmp_path):
    fn = tmp_path / "fs_cu.eam.alloy"
    fn.write_text(W.fs_cu_setfl_text(nr=4000, nrho=4000))
    pos, cell, pbc, Z = W.fcc_cu(4, sigma=0.08, seed=3)
    pos = _wrap(pos, cell)
    e0, f0, v0 = periodic(N.TabulatedAlloyEAM(fn=str(fn)), pos, cell, Z)
    e1, f1, v1, ng = run_lammps(N.TabulatedAlloyEAM(fn=str(fn)), pos, cell, Z, 5.5 + 0.5)
    print("lammps eam: %d local, %d ghosts, dE %.3e" % (len(pos), ng, e1 - e0), flush=True)
    assert abs(e1 - e0) <= REL * abs(e0)
    assert close(f1, f0) and close(v1, v0)


def lammps_full_list(x, rc, seed=0):
    """A LAMMPS-style full list with ghosts: every atom (local and ghost) gets
    the atoms within rc (cut-off + skin) in a shuffled order, the rows stored
    in reverse order in one array (seed = offset of each row), and some
    entries carry special-bond bits in the top two bits."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    pairs = cKDTree(x).query_ball_point(x, rc)
    rows = []
    for i, js in enumerate(pairs):
        js = np.array([j for j in js if j != i], dtype=np.int64)
        rng.shuffle(js)
        special = rng.random(len(js)) < 0.1
        js = js | (special.astype(np.int64) << 30)
        rows.append(js.astype(np.uint32).view(np.int32))
    neigh, seed_, pos = [], np.zeros(len(x), np.int64), 0
    for i in reversed(range(len(x))):
        seed_[i] = pos
        neigh.append(rows[i])
        pos += len(rows[i])
    return seed_, np.array([len(r) for r in rows], np.int32), np.concatenate(neigh)


def run_lammps_list(pot, pos, cell, Z, border, rc_list):
    gx, gown = ghosts(pos, cell, border)
    nlocal = len(pos)
    x = np.concatenate([pos, gx])
    Zall = np.concatenate([Z, Z[gown]]).astype(np.int32)
    seed, numneigh, neigh = lammps_full_list(x, rc_list)
    p = N.Particles()
    nl = N.Neighbors(200)
    eng, f, vir, _ = N.lammps_compute_list(pot, p, nl, nlocal, x, Zall, seed, numneigh, neigh)
    floc = f[:nlocal].copy()
    np.add.at(floc, gown, f[nlocal:])
    return eng, floc, vir


def test_lammps_list_tersoff_sic():
    """Tersoff over LAMMPS' own full list with ghosts (skin 1 A, shuffled
    rows, special-bond bits set) equals the periodic evaluation."""
    pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.06, seed=11)
    Z = np.where(np.arange(len(pos)) % 8 >= 4, 14, 6).astype(np.int32)
    pos = _wrap(pos, cell)
    e0, f0, v0 = periodic(N.Tersoff(), pos, cell, Z)
    e1, f1, v1 = run_lammps_list(N.Tersoff(), pos, cell, Z, 2 * 3.0 + 1.5, 3.0 + 1.0)
    print("lammps list tersoff: dE %.3e" % (e1 - e0), flush=True)
    assert abs(e1 - e0) <= REL * abs(e0)
    assert close(f1, f0) and close(v1, v0)


def test_lammps_list_eam():
    import tempfile
    import os
    d = tempfile.mkdtemp()
    fn = os.path.join(d, "fs_cu.eam.alloy")
    with open(fn, "w") as fh:
        fh.write(W.fs_cu_setfl_text(nr=4000, nrho=4000))
    pos, cell, pbc, Z = W.fcc_cu(4, sigma=0.08, seed=3)
    pos = _wrap(pos, cell)
    e0, f0, v0 = periodic(N.TabulatedAlloyEAM(fn=fn), pos, cell, Z)
    e1, f1, v1 = run_lammps_list(N.TabulatedAlloyEAM(fn=fn), pos, cell, Z, 5.5 + 1.5, 5.5 + 1.0)
    print("lammps list eam: dE %.3e" % (e1 - e0), flush=True)
    assert abs(e1 - e0) <= REL * abs(e0)
    assert close(f1, f0) and close(v1, v0)


def test_lammps_list_bad_index():
    pos, cell, pbc, Z = W.diamond_c(2, sigma=0.02, seed=1)
    n = len(pos)
    seed = np.zeros(n, np.int64)
    numneigh = np.zeros(n, np.int32)
    numneigh[0] = 1
    neigh = np.array([n + 5], np.int32)
    with pytest.raises(AgpuError):
        N.lammps_compute_list(N.Tersoff(), N.Particles(), N.Neighbors(100), n, pos, Z, seed, numneigh, neigh)


def test_lammps_unsupported_classes():
    pos, cell, pbc, Z = W.diamond_c(2, sigma=0.02, seed=1)
    with pytest.raises(AgpuError):
        N.lammps_compute(N.Rebo2(), N.Particles(), N.Neighbors(100), len(pos), pos, Z)


def _ach(n=400, seed=21):
    pos, cell, pbc, Z = W.amorphous_ch(n, x_h=0.2, density=2.0, seed=seed)
    return _wrap(pos, cell), cell, Z


def _periodic_rebo2(pos, cell, Z):
    p = N.from_atoms(pos, cell, np.array([True] * 3), Z)
    nl = N.Neighbors(50)
    pot = N.Rebo2(dihedral=True)
    pot.bind_to(p, nl)
    e, f, w, epa, _, _, wpa, _ = pot.energy_and_forces(p, nl, epot_per_at=True, wpot_per_at=True)
    vir = -np.array([w[0, 0], w[1, 1], w[2, 2], 0.5 * (w[0, 1] + w[1, 0]), 0.5 * (w[0, 2] + w[2, 0]),
                     0.5 * (w[1, 2] + w[2, 1])])
    # LAMMPS vatom of each atom: -(W_xx, W_yy, W_zz, W_yx, W_zx, W_zy)
    vat = -np.stack([wpa[:, 0, 0], wpa[:, 1, 1], wpa[:, 2, 2], wpa[:, 1, 0], wpa[:, 2, 0], wpa[:, 2, 1]], axis=1)
    return e, f, vir, epa, vat


@pytest.mark.parametrize("own_list", [False, True], ids=["list_here", "lammps_list"])
def test_lammps_rebo2_tag_half_rule(own_list):
    """REBO2 (torsion on) through the tagged step: each bond once from the
    local atom with tag(j) >= tag(i), ghosts 8 A deep (neighbours of
    neighbours of neighbours), then reverse communication of forces and
    per-atom energies; equals the periodic evaluation."""
    pos, cell, Z = _ach()
    e0, f0, v0, epa0, vat0 = _periodic_rebo2(pos, cell, Z)
    gx, gown = ghosts(pos, cell, 8.0)
    nlocal = len(pos)
    x = np.concatenate([pos, gx])
    Zall = np.concatenate([Z, Z[gown]]).astype(np.int32)
    tag = np.concatenate([np.arange(nlocal), gown]).astype(np.int32) + 1
    kw = {}
    if own_list:
        kw = dict(zip(("seed", "numneigh", "neigh"), lammps_full_list(x, 2.0 + 1.0, seed=5)))
    eng, f, vir, ea, va = N.lammps_compute_tagged(N.Rebo2(dihedral=True), N.Particles(), N.Neighbors(50), nlocal,
                                                  x, Zall, tag, eatom=True, vatom=True, **kw)
    floc = f[:nlocal].copy()
    np.add.at(floc, gown, f[nlocal:])
    eloc = ea[:nlocal].copy()
    np.add.at(eloc, gown, ea[nlocal:])
    print("lammps rebo2 (%s): %d local, %d ghosts, dE %.3e" % ("lammps list" if own_list else "list here",
                                                               nlocal, len(gx), eng - e0), flush=True)
    vloc = va[:nlocal].copy()
    np.add.at(vloc, gown, va[nlocal:])
    assert abs(eng - e0) <= REL * abs(e0)
    assert close(floc, f0) and close(vir, v0) and close(eloc, epa0) and close(vloc, vat0)


def test_lammps_tagged_tersoff_vatom():
    """The tagged step for a masked class (Tersoff SiC): per-atom virials
    after reverse communication equal the periodic per-atom virials."""
    pos, cell, pbc, Z = W.diamond_si(3, a=4.36, sigma=0.06, seed=11)
    Z = np.where(np.arange(len(pos)) % 8 >= 4, 14, 6).astype(np.int32)
    pos = _wrap(pos, cell)
    p0 = N.from_atoms(pos, cell, np.array([True] * 3), Z)
    nl0 = N.Neighbors(100)
    pot0 = N.Tersoff()
    pot0.bind_to(p0, nl0)
    e0, f0, w0, _, _, _, wpa0, _ = pot0.energy_and_forces(p0, nl0, wpot_per_at=True)
    vat0 = -np.stack([wpa0[:, 0, 0], wpa0[:, 1, 1], wpa0[:, 2, 2], wpa0[:, 1, 0], wpa0[:, 2, 0], wpa0[:, 2, 1]], axis=1)
    gx, gown = ghosts(pos, cell, 2 * 3.0 + 0.5)
    nlocal = len(pos)
    x = np.concatenate([pos, gx])
    Zall = np.concatenate([Z, Z[gown]]).astype(np.int32)
    tag = np.concatenate([np.arange(nlocal), gown]).astype(np.int32) + 1
    eng, f, vir, ea, va = N.lammps_compute_tagged(N.Tersoff(), N.Particles(), N.Neighbors(200), nlocal, x, Zall,
                                                  tag, vatom=True)
    vloc = va[:nlocal].copy()
    np.add.at(vloc, gown, va[nlocal:])
    assert abs(eng - e0) <= REL * abs(e0) and close(vloc, vat0)
