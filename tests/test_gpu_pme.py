"""PME (smooth particle-mesh Ewald) on the GPU against the oracle through the
C-ABI: the full class (real space + reciprocal - self, pme.f90:386-478) in
FP64 within 1e-10 relative -- energy, forces, virial and the electrostatic
potential -- on the reference's test case (NaCl a = 5.64, cutoff 5, grid
10^3, rattled) and larger grids and orders; convergence to the direct
Ewald sum on the GPU; bitwise reproducibility."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from atomistica_amd import native as N
from atomistica_amd import workloads as W

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(1.0, np.abs(b).max())


def oracle(pos, q, cell, pbc, cutoff, grid, order):
    beta = math.sqrt(math.log(10.0) * 12.0) / cutoff
    nl = O.nl_build(pos, cell, pbc, cutoff, avgn=400)
    er, fr, wr, pr = O.ewald_real(pos, q, cell, nl, cutoff, beta)
    ek, fk, wk, pk = O.pme_recip(pos, q, cell, beta, grid, order)
    return er + ek, fr + fk, wr + wk, pr + pk


def gpu(pos, q, cell, pbc, Z, cutoff, grid, order, cls=None):
    p = N.from_atoms(pos, cell, pbc, Z)
    nl = N.Neighbors(400)
    pot = cls if cls is not None else N.PME(cutoff=cutoff, order=order, grid=grid)
    pot.bind_to(p, nl)
    e, f, w = pot.energy_and_forces(p, nl, q)
    phi = pot.potential(p, nl, q)
    return e, f, w, phi


def _nacl(n, seed, amp=0.5):
    pos, cell, pbc, Z, q = W.rocksalt(n, sigma=0.0)
    pos = pos + 0.1 + np.random.default_rng(seed).uniform(-amp, amp, pos.shape)
    return pos, cell, pbc, Z, q


CASES = [("reference_case", 2, 1, 5.0, (10, 10, 10), 8),
         ("grid_16_18_20_order6", 2, 2, 5.0, (16, 18, 20), 6),
         ("nacl_3_grid24_order8", 3, 3, 6.0, (24, 24, 24), 8),
         ("order4_grid12", 2, 4, 4.5, (12, 12, 12), 4)]


@pytest.mark.parametrize("name,n,seed,cutoff,grid,order", CASES, ids=[c[0] for c in CASES])
def test_pme_parity(name, n, seed, cutoff, grid, order):
    pos, cell, pbc, Z, q = _nacl(n, seed)
    print("pme case", name, len(pos), grid, order, flush=True)
    e0, f0, w0, p0 = oracle(pos, q, cell, pbc, cutoff, grid, order)
    e, f, w, phi = gpu(pos, q, cell, pbc, Z, cutoff, grid, order)
    print("  E %.12f (oracle %.12f)" % (e, e0), flush=True)
    assert abs(e - e0) <= REL * abs(e0)
    assert close(f, f0) and close(w, w0) and close(phi, p0)


def test_pme_converges_to_direct_ewald_on_gpu():
    # beta = 0.66 (cutoff 8): the B-spline interpolation error at 48^3 and
    # order 10 is then below 1e-11 relative
    pos, cell, pbc, Z, q = _nacl(2, 6, amp=0.2)
    cutoff = 8.0
    beta = math.sqrt(math.log(10.0) * 12.0) / cutoff
    ep, fp, wp, pp = gpu(pos, q, cell, pbc, Z, cutoff, (48, 48, 48), 10)
    ee, fe, we, pe = gpu(pos, q, cell, pbc, Z, cutoff, None, None,
                         cls=N.Ewald(cutoff=cutoff, beta=beta, kmax=2 * beta * math.sqrt(-math.log(1e-16))))
    print("pme vs ewald: dE %.3e  dF %.3e" % (abs(ep - ee), np.abs(fp - fe).max()), flush=True)
    assert abs(ep - ee) < 1e-9 * abs(ee) and np.abs(fp - fe).max() < 1e-8 and np.abs(pp - pe).max() < 1e-8


def test_pme_reproducible_and_charge_derivative():
    pos, cell, pbc, Z, q = _nacl(2, 7)
    a = gpu(pos, q, cell, pbc, Z, 5.0, (10, 12, 14), 8)
    b = gpu(pos, q, cell, pbc, Z, 5.0, (10, 12, 14), 8)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
    # phi = dE/dq (central FD in one charge)
    h = 1e-5
    qp, qm = q.copy(), q.copy()
    qp[3] += h
    qm[3] -= h
    d = (gpu(pos, qp, cell, pbc, Z, 5.0, (10, 12, 14), 8)[0] - gpu(pos, qm, cell, pbc, Z, 5.0, (10, 12, 14), 8)[0]) / (2 * h)
    assert abs(d - a[3][3]) < 1e-7
