"""ctypes front end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package
``atomistica_amd``.  Every routine is a restatement of the reference Fortran
(file:line citations in the C sources next to this file).

Array conventions
-----------------
``r``       (nat, 3) float64 unwrapped positions
``cell``    (3, 3) float64, row k = lattice vector k (ASE order)
``pbc``     (3,) bool
neighbour lists are returned as ``(offsets[nat+1] int64, nbr[int32], dc[(n,3) int32])``
``wpot``    (3, 3) with ``wpot[a, b] = W(a, b)`` of the reference
``wpot_per_at`` (nat, 3, 3) with ``[i, a, b] = W_i(a, b)``
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)


class OracleError(RuntimeError):
    pass


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def use_library(path):
    """Load the oracle from another build of the same sources (the CPU
    baseline's -O3 -march=native build); must come before the first call."""
    global _lib, _LIBPATH
    if _lib is not None and _LIBPATH != path:
        raise OracleError("oracle library already loaded from " + _LIBPATH)
    _LIBPATH = path


def build_native(outdir):
    """Compile the oracle sources with -O3 -march=native (keeping
    -ffp-contract=off) into outdir; returns the library path."""
    import glob
    import subprocess
    out = os.path.join(outdir, "liboracle_native.so")
    srcs = sorted(glob.glob(os.path.join(_HERE, "orc_*.c")))
    subprocess.run(["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-std=c11",
                    "-shared", "-o", out] + srcs + ["-lm"], check=True, capture_output=True)
    return out


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise OracleError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        _lib = ctypes.CDLL(_LIBPATH)
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _check(err, what):
    if err != 0:
        raise OracleError("%s failed with code %d" % (what, err))


def inverse_cell(cell):
    cell = _c(cell, np.float64)
    out = np.zeros(9)
    _check(lib().orc_inverse_cell(_d(cell), _d(out)), "inverse_cell")
    # out holds the column-major image of Bbox = Abox^-1
    return out.reshape(3, 3).T.copy()


def nl_geometry(cell, pbc, cutoff, bin_size=0.0):
    cell = _c(cell, np.float64)
    pbc_ = _c(np.asarray(pbc, dtype=bool).astype(np.int32), np.int32)
    nc = np.zeros(3, np.int32)
    nd = np.zeros(3, np.int32)
    lib().orc_nl_geometry(_d(cell), _i(pbc_), ctypes.c_double(cutoff),
                          ctypes.c_double(bin_size), _i(nc), _i(nd))
    return nc, nd


def nl_build(r, cell, pbc, cutoff, bin_size=0.0, avgn=100, nthreads=1):
    r = _c(r, np.float64)
    nat = len(r)
    cell = _c(cell, np.float64)
    pbc_ = _c(np.asarray(pbc, dtype=bool).astype(np.int32), np.int32)
    cap = max(nat * avgn, 1)
    offsets = np.zeros(nat + 1, np.int64)
    nbr = np.zeros(cap, np.int32)
    dc = np.zeros((cap, 3), np.int32)
    if nthreads > 1:
        err = lib().orc_nl_build_mt(nat, _d(r), _d(cell), _i(pbc_), ctypes.c_double(cutoff),
                                    ctypes.c_double(bin_size), ctypes.c_int64(cap),
                                    _l(offsets), _i(nbr), _i(dc), int(nthreads))
    else:
        err = lib().orc_nl_build(nat, _d(r), _d(cell), _i(pbc_), ctypes.c_double(cutoff),
                                 ctypes.c_double(bin_size), ctypes.c_int64(cap),
                                 _l(offsets), _i(nbr), _i(dc))
    _check(err, "nl_build (neighbour list overflow?)")
    n = offsets[-1]
    return offsets, nbr[:n].copy(), dc[:n].copy()


def nl_brute(r, cell, pbc, cutoff, cap=None):
    r = _c(r, np.float64)
    nat = len(r)
    cell = _c(cell, np.float64)
    pbc_ = _c(np.asarray(pbc, dtype=bool).astype(np.int32), np.int32)
    cap = cap or max(nat * 400, 1000)
    offsets = np.zeros(nat + 1, np.int64)
    nbr = np.zeros(cap, np.int32)
    dc = np.zeros((cap, 3), np.int32)
    err = lib().orc_nl_brute(nat, _d(r), _d(cell), _i(pbc_), ctypes.c_double(cutoff),
                             ctypes.c_int64(cap), _l(offsets), _i(nbr), _i(dc))
    _check(err, "nl_brute")
    n = offsets[-1]
    return offsets, nbr[:n].copy(), dc[:n].copy()


# --------------------------------------------------------------------------
# Tersoff
# --------------------------------------------------------------------------

class _TersoffParams(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int), ("Z", ctypes.c_int * 3)] + \
        [(k, ctypes.c_double * 6) for k in ("A", "B", "xi", "lambda_", "mu", "omega", "mubo")] + \
        [("m", ctypes.c_int * 6)] + \
        [(k, ctypes.c_double * 3) for k in ("beta", "n", "c", "d", "h")] + \
        [(k, ctypes.c_double * 9) for k in ("r1", "r2")] + \
        [("fam", ctypes.c_int)] + \
        [(k, ctypes.c_double * 3) for k in ("kc1", "kc2", "kc3", "kc4", "kc5", "keta", "kdelta")] + \
        [("kalpha", ctypes.c_double * 6), ("kbeta", ctypes.c_int * 6)] + \
        [(k, ctypes.c_double * 9) for k in ("bD0", "br0", "bS", "bbeta", "bgam", "bc", "bd", "bh", "bn")] + \
        [("scr", ctypes.c_int)] + \
        [(k, ctypes.c_double * 9) for k in ("or1", "or2", "bor1", "bor2", "Cmin", "Cmax")] + \
        [("jal", ctypes.c_double * 27), ("jom", ctypes.c_double * 27), ("jm", ctypes.c_int * 27)]


_SCR_KEYS = ("or1", "or2", "bor1", "bor2", "Cmin", "Cmax")


def _screening(s, params):
    """The screened variants' parameters (or1, or2, bor1, bor2, Cmin, Cmax per
    element pair, tersoff_params.f90:82-89) when the dict has them."""
    if "or1" not in params:
        return s
    s.scr = 1
    for k in _SCR_KEYS:
        arr, vals = getattr(s, k), params[k]
        if not isinstance(vals, (list, tuple)):
            vals = [vals] * 6            # one-element sets give scalars
        for i in range(9):               # 9: the ordered pairs of JuslinScr
            arr[i] = float(vals[i]) if i < len(vals) else 0.0
    return s


def _tersoff_struct(params, Zs):
    s = _TersoffParams()
    nel = len(params["el"])
    s.nel = nel
    for i, z in enumerate(Zs):
        s.Z[i] = z
    npairs = nel * (nel + 1) // 2
    defaults = {"A": 1.0, "B": 1.0, "xi": 1.0, "lambda": 1.0, "mu": 1.0,
                "omega": 1.0, "mubo": 0.0, "r1": 1.0, "r2": 2.0}
    for k, dflt in defaults.items():
        arr = getattr(s, "lambda_" if k == "lambda" else k)
        vals = params.get(k, [dflt] * npairs)
        for i in range(6):
            arr[i] = float(vals[i]) if i < len(vals) else dflt
    mvals = params.get("m", [1] * npairs)
    for i in range(6):
        s.m[i] = int(mvals[i]) if i < len(mvals) else 1
    for k in ("beta", "n", "c", "d", "h"):
        arr = getattr(s, k)
        vals = params[k]
        for i in range(3):
            arr[i] = float(vals[i]) if i < len(vals) else 1.0
    return s


def _kumagai_struct(k):
    """Kumagai parameters (one element; names of kumagai_params.f90) in the
    shared bond-order parameter block with fam = 1."""
    s = _TersoffParams()
    s.nel = 1
    s.Z[0] = 14
    for i in range(6):
        s.A[i], s.B[i] = k["A"], k["B"]
        s.lambda_[i], s.mu[i] = k["lambda1"], k["lambda2"]
        s.xi[i] = s.omega[i] = 1.0
        s.mubo[i] = 0.0
        s.m[i] = 1
        s.r1[i], s.r2[i] = k["r1"], k["r2"]
        s.kalpha[i], s.kbeta[i] = k["alpha"], int(k["beta"])
    for i in range(3):
        s.beta[i] = s.n[i] = s.c[i] = s.d[i] = 1.0
        s.h[i] = k["h"]
        s.kc1[i], s.kc2[i], s.kc3[i], s.kc4[i], s.kc5[i] = k["c1"], k["c2"], k["c3"], k["c4"], k["c5"]
        s.keta[i], s.kdelta[i] = k["eta"], k["delta"]
    s.fam = 1
    return s


def _brenner_struct(b):
    """Brenner parameters (brenner_params.f90 names, per element pair) in the
    shared bond-order parameter block with fam = 2."""
    s = _TersoffParams()
    nel = len(b["el"])
    s.nel = nel
    npairs = nel * (nel + 1) // 2
    for i in range(6):
        inside = i < npairs
        s.A[i] = s.B[i] = s.lambda_[i] = s.mu[i] = s.xi[i] = s.omega[i] = 1.0
        s.mubo[i] = float(b["mu"][i]) if inside else 0.0
        s.m[i] = int(b["m"][i]) if inside else 1
        s.r1[i] = float(b["r1"][i]) if inside else 1.0
        s.r2[i] = float(b["r2"][i]) if inside else 2.0
        for dst, src, dflt in (("bD0", "D0", 0.0), ("br0", "r0", 0.0), ("bS", "S", 2.0), ("bbeta", "beta", 1.0),
                               ("bgam", "gamma", 0.0), ("bc", "c", 0.0), ("bd", "d", 1.0), ("bh", "h", 0.0),
                               ("bn", "n", 1.0)):
            getattr(s, dst)[i] = float(b[src][i]) if inside else dflt
    for i in range(3):
        s.beta[i] = s.n[i] = s.c[i] = s.d[i] = s.h[i] = 1.0
    s.fam = 2
    return s


def juslin_symmetrized(j):
    """Juslin parameters over ordered pairs with the entries of r0 < 0 taken
    from the reversed pair, as BIND_TO_FUNC does (juslin_module.f90:273-305)."""
    nel = len(j["el"])
    out = {k: list(v) if isinstance(v, (list, tuple)) else v for k, v in j.items()}
    keys = ("D0", "r0", "S", "beta", "gamma", "c", "d", "h", "n", "r1", "r2") + \
        tuple(k for k in _SCR_KEYS if k in out)      # screened: juslin_module.f90:286-293
    for a in range(nel):
        for b in range(nel):
            if out["r0"][b + a * nel] < 0.0:
                for k in keys:
                    out[k][b + a * nel] = out[k][a + b * nel]
    return out


def _juslin_struct(j):
    """Juslin parameters (juslin_params.f90 names) in the shared bond-order
    block with fam = 3: ordered pairs j + i nel, triplets k + nel (j + nel i)."""
    j = juslin_symmetrized(j)
    s = _TersoffParams()
    nel = len(j["el"])
    s.nel = nel
    for i in range(9):
        inside = i < nel * nel
        s.A[i % 6] = s.B[i % 6] = s.lambda_[i % 6] = s.mu[i % 6] = s.xi[i % 6] = s.omega[i % 6] = 1.0
        s.r1[i] = float(j["r1"][i]) if inside else 1.0
        s.r2[i] = float(j["r2"][i]) if inside else 2.0
        for dst, src, dflt in (("bD0", "D0", 0.0), ("br0", "r0", 0.0), ("bS", "S", 2.0), ("bbeta", "beta", 1.0),
                               ("bgam", "gamma", 0.0), ("bc", "c", 0.0), ("bd", "d", 1.0), ("bh", "h", 0.0),
                               ("bn", "n", 1.0)):
            getattr(s, dst)[i] = float(j[src][i]) if inside else dflt
    for t in range(27):
        inside = t < nel ** 3
        s.jal[t] = float(j["alpha"][t]) if inside else 0.0
        s.jom[t] = float(j["omega"][t]) if inside else 1.0
        s.jm[t] = int(j["m"][t]) if inside and "m" in j else 1
    for i in range(3):
        s.beta[i] = s.n[i] = s.c[i] = s.d[i] = s.h[i] = 1.0
    s.fam = 3
    return s


def _per_bond_arrays(nbond):
    return np.zeros(nbond), np.zeros((nbond, 3)), np.zeros(9 * nbond)


def _per_bond_out(epb, fpb, wpb):
    """(epot_per_bond, f_per_bond, wpot_per_bond) in list-entry order, the
    virial as W[e, a, b] (stored W(a, b) at 9 e + a + 3 b)."""
    return [epb, fpb, wpb.reshape(-1, 3, 3).transpose(0, 2, 1).copy()]


def tersoff(params, r, el, cell, nl, mask=None, per_at=False, nthreads=1, per_bond=False):
    """Energy, forces and virial; ``el`` holds database element indices (-1 = none).
    ``params`` is a Tersoff parameter dict, or a Kumagai or Brenner one
    (``"family": "kumagai"`` / ``"brenner"``) for those forms of the same
    kernel; a dict with the screening keys (or1, or2, bor1, bor2, Cmin, Cmax)
    selects the screened variant of its family.  ``per_bond`` appends the
    per-bond energy, pair force and virial, one entry per list entry
    (bop_kernel.f90:1370-1373,1406-1408,1511-1512,1593-1594)."""
    r = _c(r, np.float64)
    nat = len(r)
    el = _c(el, np.int32)
    cell = _c(cell, np.float64)
    offsets, nbr, dc = nl
    offsets = _c(offsets, np.int64)
    nbr = _c(nbr, np.int32)
    dc = _c(dc, np.int32)
    fam = params.get("family")
    s = (_kumagai_struct(params) if fam == "kumagai" else
         _brenner_struct(params) if fam == "brenner" else
         _juslin_struct(params) if fam == "juslin" else _tersoff_struct(params, [0, 0, 0]))
    _screening(s, juslin_symmetrized(params) if fam == "juslin" else params)
    epot = ctypes.c_double(0.0)
    f = np.zeros((nat, 3))
    w = np.zeros(9)
    epa = np.zeros(nat) if per_at else None
    wpa = np.zeros(9 * nat) if per_at else None
    m = _c(mask, np.int32) if mask is not None else None
    pb = _per_bond_arrays(len(nbr)) if per_bond else (None, None, None)
    err = lib().orc_tersoff(ctypes.byref(s), nat, _d(r), _i(el), _d(cell), _l(offsets),
                            _i(nbr), _i(dc), _i(m) if m is not None else None,
                            ctypes.byref(epot), _d(f), _d(w),
                            _d(epa) if per_at else None, _d(wpa) if per_at else None,
                            *[_d(a) if a is not None else None for a in pb],
                            int(nthreads))
    _check(err, "tersoff")
    wpot = w.reshape(3, 3).T.copy()
    out = [epot.value, f, wpot]
    if per_at:
        out += [epa, wpa.reshape(nat, 3, 3).transpose(0, 2, 1).copy()]
    if per_bond:
        out += _per_bond_out(*pb)
    return tuple(out)


# --------------------------------------------------------------------------
# EAM
# --------------------------------------------------------------------------

def spline_init(y, dx, x0=0.0):
    y = _c(y, np.float64)
    n = len(y)
    coef = np.zeros(8 * n)
    _check(lib().orc_spline_init(n, ctypes.c_double(x0), ctypes.c_double(dx), _d(y), _d(coef)),
           "spline_init")
    return coef


def eam(tables, r, el, cell, nl, mask=None, per_at=False, nthreads=1):
    """``tables``: dict from :func:`atomistica_amd.eam_tables` style data --
    keys nel, cutoff, F (list of (y, dF)), rho (list of (y, dr)), phi (nel x nel
    list of (y, dr)) with y already padded / scaled as the reader does."""
    nel = tables["nel"]
    r = _c(r, np.float64)
    nat = len(r)
    el = _c(el, np.int32)
    cell = _c(cell, np.float64)
    offsets, nbr, dc = nl
    offsets = _c(offsets, np.int64)
    nbr = _c(nbr, np.int32)
    dc = _c(dc, np.int32)
    keep = []

    def pack(items):
        ns = np.array([len(y) for y, _ in items], np.int32)
        ds = np.array([d for _, d in items], np.float64)
        coefs = [spline_init(y, d) for y, d in items]
        keep.extend(coefs)
        ptrs = (_dp * len(items))(*[_d(c) for c in coefs])
        keep.append(ptrs)
        keep.extend([ns, ds])
        return ns, ds, ptrs

    nF, dF, cF = pack(tables["F"])
    nR, dR, cR = pack(tables["rho"])
    phi_items = []
    for a in range(nel):
        for b in range(nel):
            phi_items.append(tables["phi"][a][b])
    nP, dP, cP = pack(phi_items)
    epot = ctypes.c_double(0.0)
    f = np.zeros((nat, 3))
    w = np.zeros(9)
    epa = np.zeros(nat) if per_at else None
    wpa = np.zeros(9 * nat) if per_at else None
    m = _c(mask, np.int32) if mask is not None else None
    err = lib().orc_eam(nel, nat, _d(r), _i(el), _d(cell), _l(offsets), _i(nbr), _i(dc),
                        _i(m) if m is not None else None, ctypes.c_double(tables["cutoff"]),
                        _i(nF), _d(dF), cF, _i(nR), _d(dR), cR, _i(nP), _d(dP), cP,
                        ctypes.byref(epot), _d(f), _d(w),
                        _d(epa) if per_at else None, _d(wpa) if per_at else None,
                        int(bool(tables.get("funcfl"))), int(nthreads))
    _check(err, "eam (spline argument outside its table?)")
    out = [epot.value, f, w.reshape(3, 3).T.copy()]
    if per_at:
        out += [epa, wpa.reshape(nat, 3, 3).transpose(0, 2, 1).copy()]
    return tuple(out)


# --------------------------------------------------------------------------
# Ewald
# --------------------------------------------------------------------------

def ewald_real(r, q, cell, nl, cutoff, beta, nthreads=1):
    """Real-space + self term. Returns (epot, f, wpot, phi) in e^2/A units."""
    r = _c(r, np.float64)
    nat = len(r)
    q = _c(q, np.float64)
    cell = _c(cell, np.float64)
    offsets, nbr, dc = nl
    epot = ctypes.c_double(0.0)
    f = np.zeros((nat, 3))
    w = np.zeros(9)
    phi = np.zeros(nat)
    err = lib().orc_ewald_real(nat, _d(r), _d(q), _d(cell), _l(_c(offsets, np.int64)),
                               _i(_c(nbr, np.int32)), _i(_c(dc, np.int32)),
                               ctypes.c_double(cutoff), ctypes.c_double(beta),
                               ctypes.byref(epot), _d(f), _d(w), _d(phi), int(nthreads))
    _check(err, "ewald_real")
    return epot.value, f, w.reshape(3, 3).T.copy(), phi


def ewald_kspace(r, q, cell, beta, kmax, nthreads=1):
    r = _c(r, np.float64)
    nat = len(r)
    q = _c(q, np.float64)
    cell = _c(cell, np.float64)
    epot = ctypes.c_double(0.0)
    f = np.zeros((nat, 3))
    w = np.zeros(9)
    phi = np.zeros(nat)
    err = lib().orc_ewald_kspace(nat, _d(r), _d(q), _d(cell), ctypes.c_double(beta),
                                 ctypes.c_double(kmax), ctypes.byref(epot), _d(f), _d(w),
                                 _d(phi), int(nthreads))
    _check(err, "ewald_kspace")
    return epot.value, f, w.reshape(3, 3).T.copy(), phi


# --------------------------------------------------------------------------
# smooth PME reciprocal sum (numpy restatement of pme_kernel.f90)
# --------------------------------------------------------------------------

def _bsp_one_pass(c, x, k):
    """bsp_one_pass (pme_kernel.f90), 0-based, vectorised over atoms (c has
    shape (order, n))."""
    div = 1.0 / (k - 1)
    c[k - 1] = div * x * c[k - 2]
    for j in range(1, k - 1):
        c[k - j - 1] = div * ((x + j) * c[k - j - 2] + (k - j - x) * c[k - j - 1])
    c[0] = div * (1 - x) * c[0]


def pme_bspline(w, order):
    """fill_bspline (pme_kernel.f90): cardinal B-spline weights and their
    derivatives at offset w for one dimension; returns (order, n) arrays."""
    w = np.atleast_1d(np.asarray(w, float))
    c = np.zeros((order, len(w)))
    c[1] = w
    c[0] = 1.0 - w
    for k in range(3, order):
        _bsp_one_pass(c, w, k)
    d = np.empty_like(c)
    d[0] = -c[0]
    d[1:] = c[:-1] - c[1:]
    _bsp_one_pass(c, w, order)
    return c, d


def pme_bsp_moduli(K, order):
    """load_bsp_moduli / dftmod (pme_kernel.f90): |b(m)|^-2 denominators."""
    arr, _ = pme_bspline([0.0], order)
    bsp = np.zeros(K)
    bsp[1:order + 1] = arr[:min(order, K - 1), 0]
    j = np.arange(K)
    mod = np.empty(K)
    for k in range(K):
        a = 2 * np.pi * k * j / K
        mod[k] = np.sum(bsp * np.cos(a)) ** 2 + np.sum(bsp * np.sin(a)) ** 2
    for k in range(K):
        if mod[k] < 1e-7:
            mod[k] = 0.5 * (mod[k - 1] + mod[(k + 1) % K])
    return mod


def pme_recip(r, q, cell, beta, grid, order=8):
    """Reciprocal part of smooth PME as pme_grid_potential_and_field does it
    (pme_kernel.f90:251-752): scaled fractionals with the columns of Bbox,
    B-spline charge spreading, backward FFT, energy and virial with the
    influence function exp(-pi^2 m^2/beta^2)/(B(m) m^2), forward FFT, field
    gather.  Returns (epot, f, wpot, phi) in e^2/A units; f = q E."""
    r = np.asarray(r, float)
    q = np.asarray(q, float)
    nat = len(r)
    K = [int(g) for g in grid]
    recip = np.linalg.inv(np.asarray(cell, float).T)      # Bbox(i, j)
    V = abs(np.linalg.det(cell))
    fr, th, dth, base = [], [], [], []
    for d in range(3):
        w = r[:, 0] * recip[0, d] + r[:, 1] * recip[1, d] + r[:, 2] * recip[2, d]
        an = np.where(w >= 0, np.floor(w + 0.5), np.ceil(w - 0.5))     # Fortran anint
        f = K[d] * (w - an + 0.5)
        ib = f.astype(int)
        t, dt = pme_bspline(f - ib, order)
        fr.append(f)
        th.append(t)
        dth.append(dt)
        base.append(ib)
    # grid indices (b - order + 1 + t) mod K, weights theta[t]
    idx = [np.mod(base[d][None, :] - order + 1 + np.arange(order)[:, None], K[d]) for d in range(3)]
    Q = np.zeros(K, complex)
    I = idx[0][:, None, None, :]
    J = idx[1][None, :, None, :]
    Kk = idx[2][None, None, :, :]
    W = th[0][:, None, None, :] * th[1][None, :, None, :] * th[2][None, None, :, :] * q[None, None, None, :]
    np.add.at(Q, (np.broadcast_to(I, W.shape), np.broadcast_to(J, W.shape), np.broadcast_to(Kk, W.shape)), W)
    Q = np.fft.ifftn(Q) * Q.size                              # FFTW backward, unnormalised
    mods = [pme_bsp_moduli(K[d], order) for d in range(3)]
    m = []
    for d in range(3):
        k = np.arange(K[d])
        nf = (K[d] + 1) // 2
        m.append(np.where(k < nf, k, k - K[d]).astype(float))
    M1, M2, M3 = np.meshgrid(m[0], m[1], m[2], indexing="ij")
    mh = [recip[i, 0] * M1 + recip[i, 1] * M2 + recip[i, 2] * M3 for i in range(3)]
    msq = mh[0] ** 2 + mh[1] ** 2 + mh[2] ** 2
    msq[0, 0, 0] = 1.0
    fac = np.pi ** 2 / beta ** 2
    denom = mods[0][:, None, None] * mods[1][None, :, None] * mods[2][None, None, :] * msq
    eterm = np.exp(-fac * msq) / denom
    eterm[0, 0, 0] = 0.0
    vterm = 2.0 * (fac * msq + 1.0) / msq
    struc2 = (Q * np.conj(Q)).real
    es = eterm * struc2
    vir = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            vir[a, b] = np.sum(es * (vterm * mh[a] * mh[b] - (1.0 if a == b else 0.0)))
    energy = np.sum(es)
    Q = Q * eterm
    Q = np.fft.fftn(Q)                                         # FFTW forward
    pre = 1.0 / (2 * np.pi * V)
    T = Q.real
    G = T[np.broadcast_to(I, W.shape), np.broadcast_to(J, W.shape), np.broadcast_to(Kk, W.shape)]
    t1, t2, t3 = th[0][:, None, None, :], th[1][None, :, None, :], th[2][None, None, :, :]
    d1, d2, d3 = dth[0][:, None, None, :], dth[1][None, :, None, :], dth[2][None, None, :, :]
    f0 = np.sum(G * t1 * t2 * t3, axis=(0, 1, 2))
    g1 = -K[0] * np.sum(G * d1 * t2 * t3, axis=(0, 1, 2))
    g2 = -K[1] * np.sum(G * t1 * d2 * t3, axis=(0, 1, 2))
    g3 = -K[2] * np.sum(G * t1 * t2 * d3, axis=(0, 1, 2))
    fac2 = 1.0 / (np.pi * V)
    phi = fac2 * f0
    E = np.stack([fac2 * (recip[i, 0] * g1 + recip[i, 1] * g2 + recip[i, 2] * g3) for i in range(3)], axis=1)
    return pre * energy, q[:, None] * E, pre * vir, phi


# --------------------------------------------------------------------------
# setfl tables for eam()
# --------------------------------------------------------------------------

# Units of the reference (src/support/Units.f90:74,76)
HARTREE_REF = 27.2113961
BOHR_REF = 0.529177249


def read_funcfl_tables(path):
    """Tables of a single-element funcfl file read as tabulated_eam_init does
    (src/potentials/eam/tabulated_eam.f90:141-214): a comment line, Z mass a0
    lattice, nF dF nr dr cutoff, then F(rho) (nF values from 0), Z(r) and
    rho(r) (nr values each, from 0), read list-directed; Z is scaled by
    sqrt(Hartree Bohr / 2) (:178).  The pair entry holds the scaled Z table
    and "funcfl" is set, so eam() takes r*phi = Z^2."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[1].split()
    Z = int(head[0])
    nF, dF, nr, dr, cutoff = lines[2].replace(",", " ").split()[:5]
    nF, nr = int(nF), int(nr)
    dF, dr, cutoff = float(dF), float(dr), float(cutoff)
    vals = " ".join(lines[3:]).replace(",", " ").replace("D", "e").replace("d", "e").split()
    need = nF + 2 * nr
    if len(vals) < need:
        raise OracleError("funcfl file truncated: %d of %d values" % (len(vals), need))
    v = np.array([float(x) for x in vals[:need]])
    Fy, Zy, Ry = v[:nF], v[nF:nF + nr], v[nF + nr:need]
    zs = np.sqrt(0.5 * HARTREE_REF * BOHR_REF)
    return {"nel": 1, "cutoff": cutoff, "F": [(Fy, dF)], "rho": [(Ry, dr)],
            "phi": [[(zs * Zy, dr)]], "names": None, "Z": [Z], "funcfl": True}


def read_setfl_tables(path):
    """Tables of a setfl file in the form the oracle's eam() takes, read as
    tabulated_alloy_eam_init does (:147-259): F from rho=0, rho(r) and r*phi(r)
    padded with two zeros, pair tables scaled by 1/2."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[3].split()
    nel = int(head[0])
    names = head[1:1 + nel]
    nF, dF, nr, dr, cutoff = lines[4].replace(",", " ").split()[:5]
    nF, nr = int(nF), int(nr)
    dF, dr, cutoff = float(dF), float(dr), float(cutoff)
    toks = []
    pos = 5
    F, rho, Zs = [], [], []

    def take(n):
        nonlocal pos, toks
        out = []
        while len(out) < n:
            if not toks:
                toks = lines[pos].replace(",", " ").split()
                pos += 1
                continue
            out.append(float(toks.pop(0).replace("D", "e").replace("d", "e")))
        return np.array(out)

    for e in range(nel):
        Zs.append(int(lines[pos].split()[0]))
        pos += 1
        toks = []
        F.append((take(nF), dF))
        toks = []      # one list-directed read per table: the rest of its last line is skipped
        rho.append((np.concatenate([take(nr), [0.0, 0.0]]), dr))
        toks = []
    phi = [[None] * nel for _ in range(nel)]
    for a in range(nel):
        for b in range(a + 1):
            y = 0.5 * np.concatenate([take(nr), [0.0, 0.0]])
            toks = []
            phi[a][b] = phi[b][a] = (y, dr)
    return {"nel": nel, "cutoff": cutoff, "F": F, "rho": rho, "phi": phi, "names": names, "Z": Zs}


# --------------------------------------------------------------------------
# REBO2 (C, H)
# --------------------------------------------------------------------------

class Rebo2:
    """Handle on the oracle's REBO2 parameter set and tables."""

    def __init__(self, zero_tables=False, dihedral=False):
        L = lib()
        L.orc_rebo2_new.restype = ctypes.c_void_p
        L.orc_rebo2_new.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_rebo2_free.argtypes = [ctypes.c_void_p]
        L.orc_rebo2_eval_g.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, _dp]
        L.orc_rebo2_eval_table.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, _dp]
        L.orc_rebo2_default_table.argtypes = [ctypes.c_int, _dp]
        self._h = L.orc_rebo2_new(int(bool(zero_tables)), int(bool(dihedral)))
        if not self._h:
            raise OracleError("orc_rebo2_new failed")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rebo2_free(self._h)
            self._h = None

    def g(self, ktyp, costh, n):
        """(g, dg/dcos, dg/dN) for centre type ktyp (1 = C, 3 = H)."""
        out = np.zeros(3)
        lib().orc_rebo2_eval_g(self._h, int(ktyp), float(costh), float(n), _d(out))
        return out

    TABLES = {"Fcc": 0, "Fch": 1, "Fhh": 2, "Pcc": 3, "Pch": 4, "Tcc": 5}

    def table(self, name, x, y, z=0.0):
        """Interpolated table value and its partial derivatives."""
        out = np.zeros(4)
        _check(lib().orc_rebo2_eval_table(self._h, self.TABLES[name], float(x), float(y),
                                          float(z), _d(out)), "rebo2 table")
        return out

    @staticmethod
    def default_table(name):
        """Tabulated node data: F tables (5, 5, 10) (Fcc also returns dF/di,j,k), P (6, 6)."""
        w = Rebo2.TABLES[name]
        if w == 0:
            out = np.zeros(1000)
            _check(lib().orc_rebo2_default_table(w, _d(out)), "rebo2 default table")
            return tuple(out[250 * q:250 * (q + 1)].reshape(5, 5, 10) for q in range(4))
        tri = w in (1, 2, 5)
        out = np.zeros(250 if tri else 36)
        _check(lib().orc_rebo2_default_table(w, _d(out)), "rebo2 default table")
        return out.reshape((5, 5, 10) if tri else (6, 6))

    def __call__(self, r, Z, cell, nl, per_at=False, nthreads=1, per_bond=False):
        """Energy, forces and virial (plus per-atom energies/virials, and the
        per-bond energy, pair force and virial on the half-pair entries,
        bop_kernel_rebo2.f90:2546-2547,2571-2572,2762-2763)."""
        L = lib()
        L.orc_rebo2.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _dp, _lp, _ip, _ip,
                                _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        r = _c(r, np.float64)
        nat = len(r)
        Z = _c(Z, np.int32)
        cell = _c(cell, np.float64)
        offsets, nbr, dc = nl
        offsets = _c(offsets, np.int64)
        nbr = _c(nbr, np.int32)
        dc = _c(dc, np.int32)
        epot = ctypes.c_double(0.0)
        f = np.zeros((nat, 3))
        w = np.zeros(9)
        epa = np.zeros(nat) if per_at else None
        wpa = np.zeros(9 * nat) if per_at else None
        pb = _per_bond_arrays(len(nbr)) if per_bond else (None, None, None)
        err = L.orc_rebo2(self._h, nat, _d(r), _i(Z), _d(cell), _l(offsets), _i(nbr), _i(dc),
                          ctypes.byref(epot), _d(f), _d(w), _d(epa) if per_at else None,
                          _d(wpa) if per_at else None, *[_d(a) if a is not None else None for a in pb],
                          int(nthreads))
        _check(err, "rebo2")
        wpot = w.reshape(3, 3).T.copy()
        out = [epot.value, f, wpot]
        if per_at:
            out += [epa, wpa.reshape(nat, 3, 3).transpose(0, 2, 1).copy()]
        if per_bond:
            out += _per_bond_out(*pb)
        return tuple(out)
