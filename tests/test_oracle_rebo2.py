"""Pins of the CPU REBO2 oracle (oracle/orc_rebo2.c).

What fixes the oracle, independently of how it is written:
  - the tables reproduce their node data (values and the Fcc derivative
    data) exactly at integer coordinates, and are C1 across box faces;
  - the carbon g(cos) splines hit the published knot values, slopes and
    curvatures (Brenner et al. 2002, Table 3; rebo2_type.f90 cc_g_*);
  - diamond carbon: Ec = 7.370 eV, a0 = 3.566 A, C11 = 1080 GPa, C12 = 130 GPa
    (tests/test_bulk_properties.py:126-129, tolerance 5 % there);
  - 26 relaxed hydrocarbons reproduce Brenner's atomization energies to the
    reference's ETOL = 0.005 eV (tests/test_rebo2_molecules.py), exercising
    P_CC, P_CH, F_CC, F_CH, F_HH, the H-centred g and the exp(lambda dr) term;
  - forces and virial are the derivatives of the energy (central FD) on
    diamond, amorphous carbon (tests/aC_small.cfg) and a random C/H solid
    (tests/test_forces_and_virial.py:137-141);
  - dimer curves are smooth (tests/test_dimers.py: C2, H2, CH limits).
"""
import numpy as np
import pytest
from scipy.optimize import minimize, minimize_scalar

from oracle import oracle as O
from atomistica_amd import workloads as W
from helpers import fd_forces, fd_virial, in_vacuum_box, read_cfg, read_xyz, small_hydrocarbon

GPA = 1.0 / 160.21766208

# Brenner et al., J. Phys. Condens. Matter 14, 783 (2002), Table 12 (the
# entries the reference's tests/test_rebo2_molecules.py keeps active)
MOLECULES_XYZ = {
    "C2H": -11.5722, "H3C2H2": -26.5601, "CH2=C=CH2": -30.2392, "propyne": -30.3076,
    "CH3CH=C=CH2": -43.1367, "1-butyne": -43.0510, "1-butene": -50.0487,
    "cis-butene": -50.2017, "i-C4H9": -52.0451, "t-C4H9": -52.3778,
    "1,3-pentadiene": -55.9025, "1,4-pentadiene": -56.5078, "cyclopentene": -57.1119,
    "cyclopentane": -63.6443, "2-pentene": -62.9456, "1-butene,2-methyl": -62.9658,
    "n-pentane": -69.0761, "isopentane": -69.0739, "neopentane": -69.0614,
    "cyclohexane": -76.4606, "naphthalene": -93.8784,
}
MOLECULES_BUILT = {"CH4": -18.1851, "C2H2": -17.5651, "C2H4": -24.4077,
                   "C2H6": -30.8457, "C6H6": -59.3096}
ETOL = 0.005


@pytest.fixture(scope="module")
def R():
    return O.Rebo2()


def energy(R, pos, cell, pbc, Z, per_at=False):
    nl = O.nl_build(pos, cell, pbc, 2.0)
    return R(pos, Z, cell, nl, per_at=per_at)


def test_tables_reproduce_node_data(R):
    F, Fi, Fj, Fk = O.Rebo2.default_table("Fcc")
    for i in range(5):
        for j in range(5):
            for k in range(10):
                v = R.table("Fcc", i, j, k)
                assert np.allclose(v, [F[i, j, k], Fi[i, j, k], Fj[i, j, k], Fk[i, j, k]],
                                   atol=1e-14), (i, j, k)
    for name in ("Fch", "Fhh"):
        T = O.Rebo2.default_table(name)
        for i in range(5):
            for j in range(5):
                for k in range(10):
                    v = R.table(name, i, j, k)
                    assert abs(v[0] - T[i, j, k]) < 1e-14 and np.all(np.abs(v[1:]) < 1e-13)
    for name in ("Pcc", "Pch"):
        T = O.Rebo2.default_table(name)
        for i in range(6):
            for j in range(6):
                v = R.table(name, i, j)
                assert abs(v[0] - T[i, j]) < 1e-14 and np.all(np.abs(v[1:3]) < 1e-13)


def test_default_tables_are_symmetric():
    F, Fi, Fj, Fk = O.Rebo2.default_table("Fcc")
    assert np.array_equal(F, F.transpose(1, 0, 2))
    assert np.array_equal(Fi, Fj.transpose(1, 0, 2))
    assert np.array_equal(Fk, Fk.transpose(1, 0, 2))
    T = O.Rebo2.default_table("Fch")
    assert np.array_equal(T, T.transpose(1, 0, 2))


def test_tables_c1_across_box_faces(R):
    rng = np.random.default_rng(3)
    for _ in range(40):
        y, z = rng.uniform(0, 3), rng.uniform(0, 8)
        x = float(rng.integers(1, 3))
        lo = R.table("Fcc", x - 1e-9, y, z)
        hi = R.table("Fcc", x + 1e-9, y, z)
        assert np.allclose(lo, hi, atol=1e-7)
        a, b = rng.uniform(0, 4), float(rng.integers(1, 4))
        assert np.allclose(R.table("Pch", a, b - 1e-9), R.table("Pch", a, b + 1e-9), atol=1e-7)


def test_carbon_g_spline_knots(R):
    theta = [-1.0, -0.5, -1.0 / 3.0, 0.0, 0.5, 1.0]
    g1 = [-0.01, 0.05280, 0.09733, 0.37545, 2.0014, 8.0]
    g2 = [0.0, 0.0, 0.09733, 0.271856, 0.416335, 1.0]
    dg = [0.104, 0.17, 0.40]
    d2g = [0.0, 0.37, 1.98]
    for t, a, b in zip(theta, g1, g2):
        assert abs(R.g(1, t, 4.0)[0] - a) < 1e-6        # high coordination: g_C
        if t >= -1.0 / 3.0:
            assert abs(R.g(1, t, 3.0)[0] - b) < 1e-6    # low coordination: gamma_C
    for t, d, dd in zip(theta[:3], dg, d2g):
        assert abs(R.g(1, t, 4.0)[1] - d) < 1e-6
        h = 1e-5
        curv = (R.g(1, t + h, 4.0)[1] - R.g(1, t - h, 4.0)[1]) / (2 * h)
        if t > -1.0:
            assert abs(curv - dd) < 1e-4
    # interpolation in N between 3.2 and 3.7, and its N derivative
    for t in (-0.7, 0.1, 0.8):
        for n in (3.3, 3.45, 3.6):
            h = 1e-6
            fd = (R.g(1, t, n + h)[0] - R.g(1, t, n - h)[0]) / (2 * h)
            assert abs(R.g(1, t, n)[2] - fd) < 1e-6


def test_hydrogen_g_continuous(R):
    # the H-centred g switches polynomial at cos = -1/2 and cos = -5/6
    for c in (-0.5, -10.0 / 12.0):
        a = R.g(3, c - 1e-9, 0.0)[0]
        b = R.g(3, c + 1e-9, 0.0)[0]
        assert abs(a - b) < 1e-3 * max(1.0, abs(a))


def test_diamond_cohesive_energy_lattice_constant(R):
    def e_of_a(a):
        pos, cell = W.lattice(W.DIAMOND_BASIS, 3, a)
        Z = np.full(len(pos), 6, np.int32)
        return energy(R, pos, cell, [True] * 3, Z)[0] / len(pos)
    res = minimize_scalar(e_of_a, bounds=(3.5, 3.65), method="bounded",
                          options={"xatol": 1e-7})
    assert abs(-res.fun - 7.370) / 7.370 < 5e-4
    assert abs(res.x - 3.566) / 3.566 < 5e-4


def test_diamond_elastic_constants(R):
    a0 = 3.5657500628
    pos0, cell0 = W.lattice(W.DIAMOND_BASIS, 2, a0)
    Z = np.full(len(pos0), 6, np.int32)

    def stress(e11):
        F = np.eye(3)
        F[0, 0] += e11
        pos, cell = pos0 @ F.T, cell0 @ F.T
        w = energy(R, pos, cell, [True] * 3, Z)[2]
        return w / abs(np.linalg.det(cell))
    h = 1e-5
    d = (stress(h) - stress(-h)) / (2 * h) / GPA
    assert abs(d[0, 0] - 1080) / 1080 < 0.05
    assert abs(d[1, 1] - 130) / 130 < 0.05


def _relax(R, pos, Z, seed=42):
    pos, cell, pbc = in_vacuum_box(pos)
    pos = pos + np.random.default_rng(seed).normal(0.0, 0.05, pos.shape)

    def fun(x):
        p = x.reshape(-1, 3)
        e, f, _ = energy(R, p, cell, pbc, Z)
        return e, -f.ravel()
    res = minimize(fun, pos.ravel(), jac=True, method="L-BFGS-B",
                   options={"gtol": 1e-5, "ftol": 1e-15, "maxiter": 5000})
    e, g = fun(res.x)
    assert np.abs(g).max() < 1e-4
    return e


@pytest.mark.parametrize("name", sorted(MOLECULES_XYZ))
def test_molecule_atomization_energy_xyz(R, name):
    pos, Z = read_xyz("molecules/%s.xyz" % name)
    assert abs(_relax(R, pos, Z) - MOLECULES_XYZ[name]) < ETOL


@pytest.mark.parametrize("name", sorted(MOLECULES_BUILT))
def test_molecule_atomization_energy_built(R, name):
    pos, Z = small_hydrocarbon(name)
    assert abs(_relax(R, pos, Z) - MOLECULES_BUILT[name]) < ETOL


def _fd_case(R, pos, cell, pbc, Z, atoms, ftol, wtol):
    e, f, w = energy(R, pos, cell, pbc, Z)
    fd = fd_forces(lambda p: energy(R, p, cell, pbc, Z)[0], pos, atoms, h=1e-5)
    assert np.max(np.abs(f[list(atoms)] - fd)) < ftol * max(1.0, np.max(np.abs(f)))
    Wfd = fd_virial(lambda p, c: energy(R, p, c, pbc, Z)[0], pos, cell, h=1e-6)
    assert np.max(np.abs(w - Wfd)) < wtol * max(1.0, np.max(np.abs(w)))


def test_fd_diamond_rattled(R):
    pos, cell, pbc, Z = W.diamond_c(n=2, sigma=0.08, seed=11)
    _fd_case(R, pos, cell, pbc, Z, [0, 7, 21, 50], 1e-5, 1e-6)


def test_fd_amorphous_carbon(R):
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    _fd_case(R, pos, cell, pbc, Z, range(0, len(pos), 17), 1e-5, 1e-6)


def test_fd_random_ch(R):
    pos, cell, pbc, Z = W.random_ch(50, 10, 3.0, seed=7, strain=0.1)
    _fd_case(R, pos, cell, pbc, Z, range(0, 60, 5), 1e-6, 1e-6)


def test_per_atom_sums_and_threads(R):
    pos, cell, pbc, Z = W.random_ch(120, 40, 2.5, seed=9, strain=0.1)
    nl = O.nl_build(pos, cell, pbc, 2.0)
    e, f, w, epa, wpa = R(pos, Z, cell, nl, per_at=True)
    assert abs(epa.sum() - e) < 1e-10 * abs(e)
    assert np.allclose(wpa.sum(0), w, atol=1e-10 * np.abs(w).max())
    assert np.abs(f.sum(0)).max() < 1e-10
    e4, f4, w4 = R(pos, Z, cell, nl, nthreads=4)
    assert abs(e4 - e) < 1e-12 * abs(e)
    assert np.allclose(f4, f, atol=1e-12)


@pytest.mark.parametrize("pair,dmin,dmax,de,df", [
    ((6, 6), 1.2, 3.1, 0.05, 0.5),
    ((1, 1), 0.6, 1.8, 0.02, 0.2),
    ((6, 1), 0.8, 1.9, 0.03, 0.3),
])
def test_dimer_curves_smooth(R, pair, dmin, dmax, de, df):
    # tests/test_dimers.py: 1000 distances, bounded jumps of E and F
    Z = np.array(pair, np.int32)
    es, fs = [], []
    for d in np.linspace(dmin, dmax, 1000):
        pos = np.array([[4.0, 4.0, 4.0], [4.0 + d, 4.0, 4.0]])
        cell = np.diag([8.0 + dmax, 8.0, 8.0])
        e, f, _ = energy(R, pos, cell, [False] * 3, Z)
        es.append(e)
        fs.append(f[0, 0])
    assert np.max(np.abs(np.diff(es))) < de
    assert np.max(np.abs(np.diff(fs))) < df


# ---------------------------------------------------------------------------
# torsion term (with_dihedral = .true.)
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def Rd():
    return O.Rebo2(dihedral=True)


def test_torsion_table_nodes(Rd):
    T = O.Rebo2.default_table("Tcc")
    assert T[2, 2, 0] == -0.070280085 and T[2, 2, 5] == -0.00809675
    for k in range(10):
        assert abs(Rd.table("Tcc", 2, 2, k)[0] - T[2, 2, k]) < 1e-14


def test_torsion_ethylene_closed_form(R, Rd):
    # Rotating one CH2 of ethylene about the C-C axis leaves every bond
    # length and bond angle unchanged, so only the torsion term changes:
    # planar -> sin^2(omega) = 0 for all four H-C-C-H; twisted by 90 deg ->
    # sin^2 = 1, and E_tors = 1/2 T_CC(2, 2, 0) V_A(r_CC) * 4.
    planar = np.array([[-0.667, 0, 0], [0.667, 0, 0], [-1.23, 0.92, 0], [-1.23, -0.92, 0],
                       [1.23, 0.92, 0], [1.23, -0.92, 0]])
    twisted = planar.copy()
    twisted[4] = [1.23, 0, 0.92]
    twisted[5] = [1.23, 0, -0.92]
    Z = np.array([6, 6, 1, 1, 1, 1], np.int32)
    r = 1.334
    VA = -(12388.79197798 * np.exp(-4.7204523127 * r) + 17.56740646509 * np.exp(-1.4332132499 * r)
           + 30.71493208065 * np.exp(-1.3826912506 * r))
    for X, n_sin2 in ((planar, 0.0), (twisted, 4.0)):
        pos, cell, pbc = in_vacuum_box(X)
        e0 = energy(R, pos, cell, pbc, Z)[0]
        e1 = energy(Rd, pos, cell, pbc, Z)[0]
        assert abs((e1 - e0) - 0.5 * (-0.070280085) * VA * n_sin2) < 1e-12


def test_torsion_fd_forces_and_virial(Rd):
    pos, cell, pbc, Z = read_cfg("aC_small.cfg")
    _fd_case(Rd, pos, cell, pbc, Z, range(0, len(pos), 17), 1e-6, 1e-6)
    pos, cell, pbc, Z = W.random_ch(50, 10, 3.0, seed=7, strain=0.1)
    _fd_case(Rd, pos, cell, pbc, Z, range(0, 60, 5), 1e-6, 1e-6)
