"""The oracle's neighbour-list restatement against closed forms.

Pins: the brute-force image set {(j, n): |r_i - r_j + A n|^2 < rc^2}
(independent enumeration), the reference's own cases from
tests/test_neighbor_list.py (:61-104 pbc counts, :106-122 lattice shifts,
:155-175 coordination 4 in the floating-point edge case) and the
minimum-image check of :39-59 on the aC fixture.
"""
import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import workloads as W
from helpers import read_cfg


def as_sets(nl, nat):
    off, j, dc = nl
    out = []
    for i in range(nat):
        s = set()
        for k in range(off[i], off[i + 1]):
            s.add((int(j[k]), int(dc[k, 0]), int(dc[k, 1]), int(dc[k, 2])))
        assert len(s) == off[i + 1] - off[i], "duplicate image in a row"
        out.append(s)
    return out


def random_cell(rng, n, L, skew):
    cell = np.diag([L, L * 1.1, L * 0.9]) + rng.uniform(-skew, skew, (3, 3)) * L
    frac = rng.random((n, 3))
    return frac @ cell, cell


@pytest.mark.parametrize("case", range(8))
def test_builder_equals_brute_force(case):
    rng = np.random.default_rng(100 + case)
    n = [20, 50, 80, 33, 64, 17, 40, 90][case]
    L = [6.0, 9.0, 12.0, 5.0, 10.0, 4.5, 8.0, 15.0][case]
    skew = [0.0, 0.1, 0.2, 0.15, 0.0, 0.2, 0.05, 0.1][case]
    pbc = [[1, 1, 1], [1, 1, 1], [1, 0, 1], [1, 1, 1], [0, 0, 0], [1, 1, 1], [0, 1, 1], [1, 1, 0]][case]
    rc = [3.0, 4.0, 3.5, 4.5, 3.0, 3.2, 5.0, 3.0][case]
    pos, cell = random_cell(rng, n, L, skew)
    pbc = np.array(pbc, bool)
    # unwrapped positions: some atoms a few cells away along a periodic axis
    # (along non-periodic axes atoms stay inside the box)
    if pbc.any():
        ax = int(np.argmax(pbc))
        pos[::7] += rng.integers(-2, 3, (len(pos[::7]), 1)) * cell[ax]
    nl = O.nl_build(pos, cell, pbc, rc, avgn=400)
    nb = O.nl_brute(pos, cell, pbc, rc)
    assert as_sets(nl, n) == as_sets(nb, n)


def test_reference_pbc_counts():
    # tests/test_neighbor_list.py:61-104: two atoms 0.2 apart through the boundary
    pos = np.array([[0.1, 0.5, 0.5], [0.9, 0.5, 0.5]])
    cell = np.eye(3)
    for pbc, expect in [([1, 1, 1], 2), ([0, 0, 0], 0), ([0, 0, 1], 0), ([1, 0, 0], 2)]:
        off, j, dc = O.nl_build(pos, cell, np.array(pbc, bool), 0.3)
        assert off[-1] == expect


def test_small_cell_multiple_images():
    # a 2.5 A cubic cell with a 6 A cutoff: many images of every atom, self images included
    pos = np.array([[0.1, 0.2, 0.3], [1.3, 1.1, 0.9]])
    cell = np.eye(3) * 2.5
    nl = O.nl_build(pos, cell, np.ones(3, bool), 6.0, avgn=2000)
    nb = O.nl_brute(pos, cell, np.ones(3, bool), 6.0, cap=100000)
    assert as_sets(nl, 2) == as_sets(nb, 2)
    assert nl[0][1] - nl[0][0] > 50


def test_floating_point_edge_coordination_four():
    # tests/test_neighbor_list.py:155-175
    pos = np.array([[-4.41173839e-52, 0.0, 0.0], [-4.41173839e-52, 2.26371743, 2.26371743],
                    [2.26371743, 0.0, 2.26371743], [2.26371743, 2.26371743, 0.0],
                    [1.13185872, 1.13185872, 1.13185872], [1.13185872, 3.39557615, 3.39557615],
                    [3.39557615, 1.13185872, 3.39557615], [3.39557615, 3.39557615, 1.13185872]])
    cell = np.eye(3) * 4.527434867899659
    off, j, dc = O.nl_build(pos, cell, np.ones(3, bool), 3.0)
    assert (np.diff(off) == 4).all()
    # scaled-positions round trip (a2 of the reference test)
    frac = pos @ np.linalg.inv(cell)
    off2, _, _ = O.nl_build(frac @ cell, cell, np.ones(3, bool), 3.0)
    assert (np.diff(off2) == 4).all()


def test_shift_by_lattice_vectors_keeps_neighbours():
    # tests/test_neighbor_list.py:106-122 on the aC fixture
    pos, cell, pbc, Z = read_cfg("aC.cfg")
    nl1 = O.nl_build(pos, cell, pbc, 3.0)
    p2 = pos.copy()
    p2[100] += 3 * cell[0]
    p2[200] += np.array([1, 3, -4]) @ cell
    nl2 = O.nl_build(p2, cell, pbc, 3.0)
    for i in range(len(pos)):
        a = sorted(nl1[1][nl1[0][i]:nl1[0][i + 1]])
        b = sorted(nl2[1][nl2[0][i]:nl2[0][i + 1]])
        assert a == b


def test_minimum_image_distances_on_aC():
    # tests/test_neighbor_list.py:39-59: NL vectors equal minimum-image vectors
    pos, cell, pbc, Z = read_cfg("aC.cfg")
    off, j, dc = O.nl_build(pos, cell, pbc, 5.0)
    i = np.repeat(np.arange(len(pos)), np.diff(off))
    dr = pos[i] - pos[j] + dc @ cell
    frac = (pos[i] - pos[j]) @ np.linalg.inv(cell)
    mic = (frac - np.round(frac)) @ cell
    assert np.max(np.abs(dr - mic)) < 1e-12


def test_threaded_builder_identical():
    pos, cell, pbc, Z = read_cfg("aC.cfg")
    a = O.nl_build(pos, cell, pbc, 3.5)
    b = O.nl_build(pos, cell, pbc, 3.5, nthreads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_overflow_is_an_error():
    pos, cell, pbc, Z = W.diamond_si(2, sigma=0.0)
    with pytest.raises(O.OracleError):
        O.nl_build(pos, cell, pbc, 5.0, avgn=4)


def test_inverse_cell():
    rng = np.random.default_rng(3)
    for _ in range(10):
        A = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        B = O.inverse_cell(A)
        # Bbox = Abox^-1 with Abox(:,k) = row k of the ASE cell
        assert np.allclose(B @ A.T, np.eye(3), atol=1e-13)
