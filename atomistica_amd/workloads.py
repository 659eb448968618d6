"""Synthetic, seeded workloads for the potential-evaluation path.

Each generator returns plain numpy arrays (positions ``(nat, 3)``, cell rows,
pbc, atomic numbers) so the same inputs feed the GPU path and the CPU oracle.
Atom order is (ix, iy, iz, basis) for lattices; seeds use
``numpy.random.default_rng``.
"""
import math

import numpy as np

DIAMOND_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                          [.25, .25, .25], [.25, .75, .75], [.75, .25, .75],
                          [.75, .75, .25]])
FCC_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
# rocksalt: 4 cations (fcc) + 4 anions shifted by a/2 along x
ROCKSALT_BASIS = np.vstack([FCC_BASIS, FCC_BASIS + np.array([.5, 0, 0])])
ROCKSALT_CHARGE = np.array([1.0] * 4 + [-1.0] * 4)


def lattice(basis, n, a):
    """Positions of an n x n x n (or (nx, ny, nz)) supercell of a cubic basis."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    cell = np.diag([nx * a, ny * a, nz * a]).astype(np.float64)
    return pos, cell


def diamond_si(n=8, a=5.431, sigma=0.05, seed=1):
    """Tersoff Si workload: n^3 conventional diamond cells (8 n^3 atoms)."""
    pos, cell = lattice(DIAMOND_BASIS, n, a)
    if sigma > 0:
        pos = pos + np.random.default_rng(seed).normal(0.0, sigma, pos.shape)
    Z = np.full(len(pos), 14, np.int32)
    return pos, cell, np.array([True] * 3), Z


def fcc_cu(n=40, a=3.615, sigma=0.08, seed=2):
    """EAM Cu workload: n^3 fcc cells (4 n^3 atoms)."""
    pos, cell = lattice(FCC_BASIS, n, a)
    if sigma > 0:
        pos = pos + np.random.default_rng(seed).normal(0.0, sigma, pos.shape)
    Z = np.full(len(pos), 29, np.int32)
    return pos, cell, np.array([True] * 3), Z


def rocksalt(n=30, a=5.64, sigma=0.1, seed=5):
    """Ewald workload: n^3 rocksalt cells of +/-1 e point charges (8 n^3 ions)."""
    pos, cell = lattice(ROCKSALT_BASIS, n, a)
    if sigma > 0:
        pos = pos + np.random.default_rng(seed).normal(0.0, sigma, pos.shape)
    q = np.tile(ROCKSALT_CHARGE, len(pos) // 8)
    Z = np.where(q > 0, 11, 17).astype(np.int32)
    return pos, cell, np.array([True] * 3), Z, q


def rocksalt_cells(ncells=256, nmin=2, nmax=8, a=5.64, sigma=0.1, seed=5):
    """Batched Ewald workload: ncells rocksalt cells of n^3 conventional cells
    each (8 n^3 ions, n uniform in nmin..nmax), Gaussian displacements sigma.
    Returns ([(pos, cell), ...], q concatenated)."""
    rng = np.random.default_rng(seed)
    out, qs = [], []
    for _ in range(ncells):
        n = int(rng.integers(nmin, nmax + 1))
        pos, cell, _, _, q = rocksalt(n, a=a, sigma=sigma, seed=int(rng.integers(0, 2 ** 31)))
        out.append((pos, cell))
        qs.append(q)
    return out, np.concatenate(qs)


# --------------------------------------------------------------------------
# Finnis-Sinclair-form Cu tabulated as a setfl file
# --------------------------------------------------------------------------

FS_R0 = 2.556


def fs_rho(r):
    return np.exp(-3.0 * (r / FS_R0 - 1.0))


def fs_phi(r):
    return 0.8 * np.exp(-8.0 * (r / FS_R0 - 1.0))


def fs_embed(rho):
    return -2.2 * np.sqrt(rho)


def fs_cu_setfl_text(nr=10000, nrho=10000, cutoff=5.5, rho_max=40.0):
    """setfl (DYNAMO alloy) text: F(rho), rho(r) and r*phi(r) on uniform grids.

    The pair column holds r*phi(r), the convention the setfl reader assumes.
    """
    dr = cutoff / (nr - 1)
    drho = rho_max / (nrho - 1)
    rg = np.arange(nr) * dr
    rhog = np.arange(nrho) * drho
    lines = ["Finnis-Sinclair-form Cu: rho=exp(-3(r/2.556-1)), phi=0.8exp(-8(r/2.556-1))",
             "F(rho)=-2.2 sqrt(rho); pair column is r*phi(r)",
             "synthetic table",
             "1 Cu",
             "%d %.17g %d %.17g %.17g" % (nrho, drho, nr, dr, cutoff),
             "29 63.546 3.615 fcc"]

    def block(vals):
        out = []
        for k in range(0, len(vals), 5):
            out.append(" ".join("%.17e" % v for v in vals[k:k + 5]))
        return out

    lines += block(fs_embed(rhog))
    lines += block(fs_rho(rg))
    lines += block(rg * fs_phi(rg))
    return "\n".join(lines) + "\n"


def fs_cu_perfect_energy(a=3.615, cutoff=5.5, shells=8):
    """Analytic E/atom of perfect fcc under the Finnis-Sinclair-form functions
    (exact functions, no spline), summing every neighbour inside the cutoff."""
    rng = np.arange(-shells, shells + 1)
    g = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    pts = (g[:, None, :] + FCC_BASIS[None]).reshape(-1, 3) * a
    d = np.linalg.norm(pts, axis=1)
    d = d[(d > 1e-9) & (d < cutoff)]
    rho = fs_rho(d).sum()
    return fs_embed(rho) + 0.5 * fs_phi(d).sum(), len(d), rho


# --------------------------------------------------------------------------
# Batched random Si structures for fitting
# --------------------------------------------------------------------------

SI_MASS = 28.0855
AMU_G = 1.66053906660e-24


def random_si_structures(nstruct=4096, seed=4, nmin=8, nmax=512, rho0=2.33,
                         drho=0.1, strain=0.15, dmin=2.0, max_tries=200):
    """Random triclinic periodic Si cells at density rho0*(1 +/- drho) g/cm^3
    with atom counts log-uniform in [nmin, nmax] and a minimum distance dmin
    over every periodic image.  Returns a list of (pos, cell)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nstruct):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        dens = rho0 * (1.0 + rng.uniform(-drho, drho))
        V = n * SI_MASS * AMU_G / dens * 1e24      # A^3
        E = rng.uniform(-strain, strain, (3, 3))
        cell = np.eye(3) + E
        cell *= (V / abs(np.linalg.det(cell))) ** (1.0 / 3.0)
        seed_s = int(rng.integers(0, 2 ** 63))
        out.append((random_positions(n, cell, dmin, seed_s, max_tries), cell))
    return out


def random_positions(n, cell, dmin, seed, max_tries=200):
    """n random points in a periodic cell with minimum distance dmin over all
    images (random sequential addition in the native library, splitmix64
    stream seeded by ``seed``)."""
    import ctypes
    from ._lib import call
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    pos = np.zeros((n, 3))
    call("agpu_random_positions", int(n), cell.ctypes.data_as(ctypes.c_void_p), float(dmin),
         ctypes.c_uint64(seed), int(max_tries), pos.ctypes.data_as(ctypes.c_void_p))
    return pos


def _random_positions(n, cell, dmin, rng, max_tries, block=32):
    """Random sequential addition of n points in a periodic (triclinic) cell
    with a minimum distance dmin over every periodic image.  Candidates are
    proposed in blocks: each block is tested against the accepted points in
    one vectorised step, then accepted in order against each other.  If
    max_tries blocks in a row add nothing the cell is jammed and dmin is
    relaxed by 3 %."""
    inv = np.linalg.inv(cell)
    # lattice translations covering dmin around the cell
    h = dmin * np.linalg.norm(inv, axis=0)
    rngs = [np.arange(-int(math.ceil(x)) - 1, int(math.ceil(x)) + 2) for x in h]
    shifts = np.stack(np.meshgrid(*rngs, indexing="ij"), -1).reshape(-1, 3)
    vs = shifts @ cell
    vs2 = np.einsum("kx,kx->k", vs, vs)
    frac = np.zeros((n, 3))
    m = 0
    scale = 1.0
    stall = 0
    while m < n:
        d2min = dmin * dmin * scale
        if np.any((vs2 > 0) & (vs2 < d2min)):
            scale *= 0.97             # the cell itself is shorter than dmin
            continue
        cand = rng.random((block, 3))
        ok = np.ones(block, bool)
        if m:
            d = cand[:, None, :] - frac[None, :m, :]
            d -= np.round(d)
            v = (d[:, :, None, :] + shifts[None, None]) @ cell          # (B, m, S, 3)
            ok &= np.einsum("bmsx,bmsx->bms", v, v).reshape(block, -1).min(1) >= d2min
        added = 0
        for c in np.nonzero(ok)[0]:
            if m >= n:
                break
            if added:
                d = cand[c] - frac[m - added:m]
                d -= np.round(d)
                v = (d[:, None, :] + shifts[None]) @ cell
                if np.einsum("msx,msx->ms", v, v).min() < d2min:
                    continue
            frac[m] = cand[c]
            m += 1
            added += 1
        stall = 0 if added else stall + 1
        if stall >= max_tries:
            scale *= 0.97             # relax the minimum distance if the cell is jammed
            stall = 0
    return frac @ cell


def log_uniform_counts(nstruct=4096, seed=4, nmin=8, nmax=512):
    rng = np.random.default_rng(seed)
    return np.array([int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
                     for _ in range(nstruct)])


def ewald_beta(cutoff, tol=1e-8):
    """beta with erfc(beta * cutoff) = tol (bisection)."""
    lo, hi = 0.0, 10.0 / cutoff
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid * cutoff) > tol:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def ewald_kmax(beta, tol=1e-8):
    """kmax with exp(-kmax^2 / 4 beta^2) = tol."""
    return 2.0 * beta * math.sqrt(-math.log(tol))


# --------------------------------------------------------------------------
# hydrocarbons (REBO2)
# --------------------------------------------------------------------------

C_MASS = 12.011
H_MASS = 1.008


def diamond_c(n=50, a=3.566, sigma=0.05, seed=6):
    """REBO2 carbon workload: n^3 conventional diamond cells (8 n^3 atoms)."""
    pos, cell = lattice(DIAMOND_BASIS, n, a)
    if sigma > 0:
        pos = pos + np.random.default_rng(seed).normal(0.0, sigma, pos.shape)
    Z = np.full(len(pos), 6, np.int32)
    return pos, cell, np.array([True] * 3), Z


def random_ch(nC=50, nH=10, density=3.0, seed=7, dmin=0.9, strain=0.0):
    """Random periodic C/H solid at the given mass density (g/cm^3) with a
    minimum interatomic distance dmin over every periodic image; the cell is
    cubic, or triclinic with random entries of +/-strain."""
    rng = np.random.default_rng(seed)
    V = (nC * C_MASS + nH * H_MASS) * AMU_G / density * 1e24
    cell = np.eye(3) + rng.uniform(-strain, strain, (3, 3))
    cell *= (V / abs(np.linalg.det(cell))) ** (1.0 / 3.0)
    pos = _random_positions(nC + nH, cell, dmin, rng, 200)
    Z = np.array([6] * nC + [1] * nH, np.int32)
    perm = rng.permutation(nC + nH)
    return pos[perm], cell, np.array([True] * 3), Z[perm]


def amorphous_ch(nat=100000, x_h=0.2, density=2.0, seed=10,
                 dmin=(("CC", 1.2), ("CH", 1.0), ("HH", 0.8)), batch=20000):
    """Amorphous hydrocarbon: nat atoms (fraction x_h hydrogen) uniformly at
    random in a cubic periodic box of the given mass density (g/cm^3), with
    rejection at the pair minimum distances dmin (random sequential
    addition, candidates proposed in batches and accepted in order).
    Returns (pos, cell, pbc, Z) with the atoms in insertion order."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    nH = int(round(nat * x_h))
    nC = nat - nH
    V = (nC * C_MASS + nH * H_MASS) * AMU_G / density * 1e24
    L = V ** (1.0 / 3.0)
    d = dict(dmin)
    dmax = max(d.values())
    dij = np.array([[d["CC"], d["CH"]], [d["CH"], d["HH"]]])    # species 0 = C, 1 = H
    species = np.array([0] * nC + [1] * nH)
    rng.shuffle(species)
    pos = np.zeros((nat, 3))
    n = 0
    tree = None
    while n < nat:
        m = min(batch, nat - n)
        cand = rng.random((m, 3)) * L
        sp = species[n:n + m]
        ok = np.ones(m, bool)
        if tree is not None:
            hits = tree.query_ball_point(cand, dmax)
            for c, h in enumerate(hits):
                if h:
                    dd = pos[h] - cand[c]
                    dd -= L * np.round(dd / L)
                    r = np.sqrt(np.einsum("kx,kx->k", dd, dd))
                    if np.any(r < dij[sp[c], species[h]]):
                        ok[c] = False
        # conflicts inside the batch: keep the earlier candidate
        bt = cKDTree(cand, boxsize=L)
        for a, b in sorted(bt.query_pairs(dmax)):
            if ok[a] and ok[b]:
                dd = cand[b] - cand[a]
                dd -= L * np.round(dd / L)
                if np.sqrt(dd @ dd) < dij[sp[a], sp[b]]:
                    ok[b] = False
        acc = cand[ok]
        # keep the species sequence: accepted candidates take the next slots
        k = len(acc)
        spk = sp[ok]
        # reorder the remaining species so that the accepted ones come first
        rest = np.concatenate([spk, sp[~ok], species[n + m:]])
        species[n:] = rest
        pos[n:n + k] = acc
        n += k
        tree = cKDTree(np.mod(pos[:n], L), boxsize=L)
    Z = np.where(species == 0, 6, 1).astype(np.int32)
    return pos, np.diag([L, L, L]).astype(np.float64), np.array([True] * 3), Z
