"""Host-side mirror of the reference's ``_atomistica`` extension types.

``Particles``, ``Neighbors`` and the potential classes keep the reference's
method names, argument meanings and error behaviour
(src/python/c/particles.c, neighbors.c, potential.c:347-666, coulomb.c),
and drive the MI355X kernels through the C-ABI in ``libatomistica_amd.so``.

Host numpy arrays are staged through HBM by the library (``AGPU_HOST``);
torch CUDA tensors (device pointers) are passed through without copies by
the ``*_device`` variants.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AgpuError, call
from .parameters import ATOMIC_NUMBERS, pair_index


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dev(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return ctypes.c_void_p(t.data_ptr())


class Particles:
    """particles_t (src/python/f90/python_particles.f90:84-170)."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("agpu_particles_create", ctypes.byref(h))
        self._h = h
        self.nat = 0
        self.coordinates = np.zeros((0, 3))
        self.Z = np.zeros(0, np.int32)
        self.cell = np.zeros((3, 3))
        self.pbc = np.array([True, True, True])
        self._dev_positions = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib.load().agpu_particles_destroy(h)
            self._h = None

    def __len__(self):
        return self.nat

    def register_data(self, *args):
        pass

    def allocate(self, nat):
        call("agpu_particles_allocate", self._h, int(nat))
        self.nat = int(nat)
        self.coordinates = np.zeros((self.nat, 3))
        self.Z = np.zeros(self.nat, np.int32)

    def set_cell(self, cell, pbc=None):
        cell = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        if pbc is None:
            pbc = [True, True, True]
        pbc_ = np.ascontiguousarray(np.asarray(pbc, dtype=bool).astype(np.int32).reshape(3))
        call("agpu_particles_set_cell", self._h, cell.ctypes.data_as(_lib.dp),
             pbc_.ctypes.data_as(_lib.ip))
        self.cell = cell.copy()
        self.pbc = pbc_.astype(bool)

    def update_elements(self):
        Z = np.ascontiguousarray(self.Z, dtype=np.int32)
        call("agpu_particles_set_atomic_numbers", self._h, _ptr(Z), _lib.HOST)

    def I_changed_positions(self):
        """Push host coordinates to HBM and mark positions changed."""
        r = np.ascontiguousarray(self.coordinates, dtype=np.float64)
        call("agpu_particles_set_positions", self._h, _ptr(r), _lib.HOST)
        self._dev_positions = None

    def set_positions_device(self, r):
        """Alias a (nat, 3) float64 device tensor as the positions."""
        if tuple(r.shape) != (self.nat, 3):
            raise ValueError("positions must have shape (nat, 3)")
        call("agpu_particles_set_position_pointer", self._h, self.nat, _dev(r))
        self._dev_positions = r

    def changed_positions(self):
        call("agpu_particles_changed_positions", self._h)

    def get_cell(self):
        return self.cell.copy()

    def get_volume(self):
        return abs(np.linalg.det(self.cell))


class Neighbors:
    """neighbors_t (src/python/f90/python_neighbors.f90:51-139)."""

    def __init__(self, avgn=100):
        h = ctypes.c_void_p()
        call("agpu_neighbors_create", ctypes.byref(h), int(avgn))
        self._h = h
        self.avgn = int(avgn)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib.load().agpu_neighbors_destroy(h)
            self._h = None

    def request_interaction_range(self, cutoff):
        call("agpu_neighbors_request_interaction_range", self._h, float(cutoff))

    def set_bin_size(self, bin_size):
        call("agpu_neighbors_set_bin_size", self._h, float(bin_size))

    def update(self, p):
        call("agpu_neighbors_update", self._h, p._h)

    def get_size(self):
        total = ctypes.c_int64()
        maxrow = ctypes.c_int()
        call("agpu_neighbors_get_size", self._h, ctypes.byref(total), ctypes.byref(maxrow))
        return total.value, maxrow.value

    def get_list(self, p):
        """CSR list in the reference's fill order: (offsets, j, dc)."""
        self.update(p)
        total, _ = self.get_size()
        offsets = np.zeros(p.nat + 1, np.int64)
        j = np.zeros(max(total, 1), np.int32)
        dc = np.zeros((max(total, 1), 3), np.int32)
        call("agpu_neighbors_get", self._h, _ptr(offsets), _ptr(j), _ptr(dc), _lib.HOST)
        return offsets, j[:total], dc[:total]

    def get_neighbors(self, p, vec=False):
        """(i, j, abs_dr) or (i, j, dr, abs_dr) -- neighbors.c:143-232, with
        dr = r_i - r_j + cell.dc (neighbors_wrap.f90:461-492)."""
        offsets, j, dc = self.get_list(p)
        i = np.repeat(np.arange(p.nat, dtype=np.int32), np.diff(offsets))
        r = p.coordinates
        dr = r[i] - r[j] + dc.astype(np.float64) @ p.cell
        absdr = np.sqrt(np.sum(dr * dr, axis=1))
        if vec:
            return i, j, dr, absdr
        return i, j, absdr

    def get_coordination_numbers(self, p, cutoff):
        i, j, absdr = self.get_neighbors(p)
        return np.bincount(i[absdr < cutoff], minlength=p.nat)

    def coordination(self, p, cutoff):
        return self.get_coordination_numbers(p, cutoff)


class _Potential:
    kind = 0

    def __init__(self):
        self._h = None
        self.precision = "fp64"

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib.load().agpu_potential_destroy(h)
            self._h = None

    def set_precision(self, precision):
        """'fp64' (default) or 'mixed' (FP32 arithmetic, FP64 accumulation)."""
        mode = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        call("agpu_potential_set_precision", self._h, mode)
        self.precision = precision

    def get_cutoff(self):
        c = ctypes.c_double()
        call("agpu_potential_get_cutoff", self._h, ctypes.byref(c))
        return c.value

    def register_data(self, p):
        pass

    def bind_to(self, p, nl):
        call("agpu_potential_bind_to", self._h, p._h, nl._h)

    def autotune(self, p, nl, reps=5):
        """Times the kernel launch variants on this configuration and keeps
        the fastest (neighbour-list pair search; EAM lanes per atom and
        entries per lane).  Returns the timing report as text."""
        buf = ctypes.create_string_buffer(4096)
        call("agpu_potential_autotune", self._h, p._h, nl._h, int(reps), buf, len(buf))
        return buf.value.decode()

    def set_launch(self, nl, pair_mode=0, variant=-1):
        """Pins the launch variants (pair_mode 0 auto, 1 thread per atom,
        2 wavefront per bin block; variant as enumerated by the autotuner,
        -1 keeps it).  Returns the number of kernel variants offered."""
        n = ctypes.c_int(0)
        call("agpu_potential_set_launch", self._h, nl._h, int(pair_mode), int(variant), ctypes.byref(n))
        return n.value

    def set_Coulomb(self, coul):
        raise AgpuError(_lib.ERR_UNSUPPORTED, "set_Coulomb: this class has no charge coupling")

    def energy_and_forces(self, p, nl, forces=None, charges=None, epot_per_at=False,
                          epot_per_bond=False, f_per_bond=False, wpot_per_at=False,
                          wpot_per_bond=False, mask=None):
        """Returns (epot, forces, wpot, epot_per_at, epot_per_bond, f_per_bond,
        wpot_per_at, wpot_per_bond) as potential.c:622-657 does.  Per-bond
        arrays have one entry per neighbour-list entry in the order of
        Neighbors.get_list (the packed order of potential.c:587-611):
        epot_per_bond (nb,), f_per_bond (nb, 3), wpot_per_bond (nb, 3, 3)."""
        nat = p.nat
        if forces is None:
            forces = np.zeros((nat, 3))
        elif forces.shape != (nat, 3):
            raise RuntimeError("Length of force array (= %d) does not equal number of atoms (= %d)."
                               % (len(forces), nat))
        if not (forces.flags.c_contiguous and forces.dtype == np.float64):
            raise ValueError("forces must be a C-contiguous float64 array")
        if mask is not None:
            mask = np.ascontiguousarray(mask, dtype=np.int32)
            if len(mask) != nat:
                raise RuntimeError("mask has %d entries but the particles hold %d atoms" % (len(mask), nat))
        nb = 0
        if epot_per_bond or f_per_bond or wpot_per_bond:
            # sized by the list the call will use (neighbors_size of the
            # reference bounds it)
            nl.update(p)
            nb = nl.get_size()[0]
        epot = ctypes.c_double(0.0)
        wpot = np.zeros(9)
        epa = np.zeros(nat) if epot_per_at else None
        wpa = np.zeros(9 * nat) if wpot_per_at else None
        epb = np.zeros(nb) if epot_per_bond else None
        fpb = np.zeros((nb, 3)) if f_per_bond else None
        wpb = np.zeros(9 * nb) if wpot_per_bond else None
        q = np.ascontiguousarray(charges, dtype=np.float64) if charges is not None else None
        call("agpu_potential_energy_and_forces", self._h, p._h, nl._h, _ptr(q),
             ctypes.byref(epot), _ptr(forces), _ptr(wpot), _ptr(mask), _ptr(epa), _ptr(epb), _ptr(fpb),
             _ptr(wpa), _ptr(wpb), _lib.HOST)
        w = wpot.reshape(3, 3, order="F").copy()
        wpa_out = wpa.reshape(nat, 3, 3).transpose(0, 2, 1).copy() if wpot_per_at else None
        wpb_out = wpb.reshape(nb, 3, 3).transpose(0, 2, 1).copy() if wpot_per_bond else None
        return epot.value, forces, w, epa, epb, fpb, wpa_out, wpb_out

    def energy_and_forces_device(self, p, nl, f, epot, wpot, charges=None, mask=None,
                                 epot_per_at=None, wpot_per_at=None, epot_per_bond=None,
                                 f_per_bond=None, wpot_per_bond=None):
        """Device variant: every array is a torch CUDA tensor (f (nat,3) and
        wpot (9,) float64 accumulated, epot (1,) float64 set; per-bond arrays
        sized by the updated list)."""
        call("agpu_potential_energy_and_forces", self._h, p._h, nl._h, _dev(charges), _dev(epot),
             _dev(f), _dev(wpot), _dev(mask), _dev(epot_per_at), _dev(epot_per_bond), _dev(f_per_bond),
             _dev(wpot_per_at), _dev(wpot_per_bond), 0)


def _tersoff_struct(params):
    s = _lib.TersoffParams()
    el = params["el"]
    nel = len(el)
    if not 1 <= nel <= 3:
        raise AgpuError(_lib.ERR_ARG, "Tersoff: 1 to 3 elements supported")
    s.nel = nel
    for k, sym in enumerate(el):
        if sym not in ATOMIC_NUMBERS:
            raise AgpuError(_lib.ERR_ELEMENT, "Unknown element '%s'." % sym)
        s.Z[k] = ATOMIC_NUMBERS[sym]
    npairs = nel * (nel + 1) // 2
    defaults = {"A": 1.0, "B": 1.0, "xi": 1.0, "lambda": 1.0, "mu": 1.0, "omega": 1.0,
                "mubo": 0.0, "r1": 1.0, "r2": 2.0}
    for k, dflt in defaults.items():
        arr = getattr(s, "lambda_" if k == "lambda" else k)
        vals = list(params.get(k, [dflt] * npairs))
        for q in range(6):
            arr[q] = float(vals[q]) if q < len(vals) else dflt
    mv = list(params.get("m", [1] * npairs))
    for q in range(6):
        s.m[q] = int(mv[q]) if q < len(mv) else 1
    for k in ("beta", "n", "c", "d", "h"):
        arr = getattr(s, k)
        vals = list(params.get(k, [1.0] * nel))
        for q in range(3):
            arr[q] = float(vals[q]) if q < len(vals) else 1.0
    return s


class Tersoff(_Potential):
    """Tersoff (tersoff.f90, classname Tersoff).  Keyword arguments are the
    ptrdict names of tersoff_registry.f90:22-115 (el, A, B, xi, lambda, mu,
    omega, mubo, m, beta, n, c, d, h, r1, r2); without arguments the
    Tersoff_PRB_39_5566_SiC database is used."""
    kind = _lib.KIND_TERSOFF

    def __init__(self, **kwargs):
        super().__init__()
        h = ctypes.c_void_p()
        if kwargs:
            from .parameters import Tersoff_PRB_39_5566_Si_C
            params = dict(Tersoff_PRB_39_5566_Si_C)
            params.update(kwargs)
            self._params = params
            s = _tersoff_struct(params)
            call("agpu_tersoff_create", ctypes.byref(h), ctypes.byref(s))
        else:
            self._params = None
            call("agpu_tersoff_create", ctypes.byref(h), None)
        self._h = h

    def set_params(self, **kwargs):
        from .parameters import Tersoff_PRB_39_5566_Si_C
        params = dict(self._params or Tersoff_PRB_39_5566_Si_C)
        params.update(kwargs)
        self._params = params
        s = _tersoff_struct(params)
        call("agpu_tersoff_set_params", self._h, ctypes.byref(s))


class Kumagai(_Potential):
    """Kumagai (kumagai.f90, classname Kumagai): the generic bond-order
    kernel with the Kumagai forms.  Keyword arguments are the names of
    kumagai_params.f90 (A, B, lambda1, lambda2, eta, delta, alpha, beta, c1
    .. c5, h, r1, r2, el); without arguments Kumagai_CompMaterSci_39_457_Si."""
    kind = _lib.KIND_KUMAGAI

    def __init__(self, **kwargs):
        super().__init__()
        from .parameters import Kumagai_CompMaterSci_39_457_Si
        params = dict(Kumagai_CompMaterSci_39_457_Si)
        params.update(kwargs)
        self._params = params
        s = _kumagai_struct(params)
        h = ctypes.c_void_p()
        call("agpu_kumagai_create", ctypes.byref(h), ctypes.byref(s))
        self._h = h


class Brenner(_Potential):
    """Brenner (brenner.f90, classname Brenner): the generic bond-order
    kernel with the Brenner (Albe/Erhart) forms.  Keyword arguments are the
    names of brenner_params.f90 (el, D0, r0, S, beta, gamma, c, d, h, mu, n,
    m, r1, r2; one value per element pair); without arguments
    Erhart_PRB_71_035211_SiC."""
    kind = _lib.KIND_BRENNER

    def __init__(self, **kwargs):
        super().__init__()
        from .parameters import Erhart_PRB_71_035211_SiC
        params = dict(Erhart_PRB_71_035211_SiC)
        params.update(kwargs)
        self._params = params
        s = _brenner_struct(params)
        h = ctypes.c_void_p()
        call("agpu_brenner_create", ctypes.byref(h), ctypes.byref(s))
        self._h = h


class Juslin(_Potential):
    """Juslin (juslin.f90, classname Juslin): the bond-order kernel with the
    Brenner forms over ordered element pairs and a per-triplet h-term.
    Keyword arguments are the names of juslin_params.f90 (el, D0, r0, S,
    beta, gamma, c, d, h, n over nel^2 ordered pairs j + i nel; alpha,
    omega, m over nel^3 ordered triplets; r1, r2); without arguments
    Juslin_J_Appl_Phys_98_123520_WCH."""
    kind = _lib.KIND_JUSLIN

    def __init__(self, **kwargs):
        super().__init__()
        h = ctypes.c_void_p()
        if kwargs:
            from .parameters import Juslin_JAP_98_123520_WCH
            params = dict(Juslin_JAP_98_123520_WCH)
            params.update(kwargs)
            self._params = params
            s = _juslin_struct(params)
            call("agpu_juslin_create", ctypes.byref(h), ctypes.byref(s))
        else:
            self._params = None
            call("agpu_juslin_create", ctypes.byref(h), None)
        self._h = h


def _juslin_struct(params):
    from .parameters import ATOMIC_NUMBERS
    nel = len(params["el"])
    npairs, ntrip = nel * nel, nel ** 3
    s = _lib.JuslinParams()
    s.nel = nel
    for i, e in enumerate(params["el"]):
        s.Z[i] = ATOMIC_NUMBERS[e]
    for k in ("D0", "r0", "S", "beta", "gamma", "c", "d", "h", "n", "r1", "r2"):
        vals = params[k]
        if len(vals) < npairs:
            raise ValueError("Juslin: %s needs one value per ordered element pair (%d)" % (k, npairs))
        arr = getattr(s, k)
        for i in range(npairs):
            arr[i] = float(vals[i])
    for k in ("alpha", "omega", "m"):
        vals = params.get(k, [1] * ntrip if k == "m" else None)
        if vals is None or len(vals) < ntrip:
            raise ValueError("Juslin: %s needs one value per ordered element triplet (%d)" % (k, ntrip))
        arr = getattr(s, k)
        for t in range(ntrip):
            arr[t] = int(vals[t]) if k == "m" else float(vals[t])
    return s


def _kumagai_struct(params):
    from .parameters import ATOMIC_NUMBERS
    s = _lib.KumagaiParams()
    el = params["el"][0] if isinstance(params["el"], (list, tuple)) else params["el"]
    s.Z = ATOMIC_NUMBERS[el]
    for k in ("A", "B", "lambda1", "lambda2", "eta", "delta", "alpha", "c1", "c2", "c3", "c4", "c5", "h",
              "r1", "r2"):
        v = params[k]
        setattr(s, k, float(v[0] if isinstance(v, (list, tuple)) else v))
    b = params["beta"]
    s.beta = int(b[0] if isinstance(b, (list, tuple)) else b)
    return s


def _brenner_struct(params):
    from .parameters import ATOMIC_NUMBERS
    nel = len(params["el"])
    npairs = nel * (nel + 1) // 2
    s = _lib.BrennerParams()
    s.nel = nel
    for i, e in enumerate(params["el"]):
        s.Z[i] = ATOMIC_NUMBERS[e]
    for k in ("D0", "r0", "S", "beta", "gamma", "c", "d", "h", "mu", "n", "r1", "r2"):
        vals = params[k]
        if len(vals) < npairs:
            raise ValueError("Brenner: %s needs one value per element pair (%d)" % (k, npairs))
        arr = getattr(s, k)
        for i in range(npairs):
            arr[i] = float(vals[i])
    for i in range(npairs):
        s.m[i] = int(params["m"][i])
    return s


def _screening_struct(params, npairs):
    """or1, or2, bor1, bor2, Cmin, Cmax per element pair (scalars for
    one-element sets)."""
    from .parameters import SCREENING_KEYS
    s = _lib.ScreeningParams()
    for k in SCREENING_KEYS:
        vals = params[k]
        if not isinstance(vals, (list, tuple)):
            vals = [vals]
        if len(vals) < npairs:
            raise ValueError("screening: %s needs one value per element pair (%d)" % (k, npairs))
        arr = getattr(s, k)
        for i in range(npairs):
            arr[i] = float(vals[i])
    return s


class _Screened(_Potential):
    """Common constructor of the screened variants: keyword arguments
    update the default screened set; without arguments the library's
    default (the reference's Fortran database entry) is used."""
    _create = None
    _default = None

    def __init__(self, **kwargs):
        super().__init__()
        from . import parameters as _P
        h = ctypes.c_void_p()
        if kwargs:
            params = dict(getattr(_P, self._default))
            params.update(kwargs)
            self._params = params
            s, npairs = self._struct(params)
            sc = _screening_struct(params, npairs)
            call(self._create, ctypes.byref(h), ctypes.byref(s), ctypes.byref(sc))
        else:
            self._params = None
            call(self._create, ctypes.byref(h), None, None)
        self._h = h


class TersoffScr(_Screened):
    """TersoffScr (tersoff_scr.f90): Tersoff with the screening of
    bop_kernel.f90 (SCREENING) and exp cutoffs.  Keyword arguments are those
    of Tersoff plus or1, or2, bor1, bor2, Cmin, Cmax per element pair;
    without arguments the screened Tersoff_PRB_39_5566_SiC."""
    kind = _lib.KIND_TERSOFF_SCR
    _create = "agpu_tersoff_scr_create"
    _default = "Tersoff_PRB_39_5566_Si_C__Scr"

    @staticmethod
    def _struct(params):
        nel = len(params["el"])
        return _tersoff_struct(params), nel * (nel + 1) // 2


class KumagaiScr(_Screened):
    """KumagaiScr (kumagai_scr.f90); keyword arguments of Kumagai plus the
    screening parameters; without arguments the screened
    Kumagai_CompMaterSci_39_457_Si."""
    kind = _lib.KIND_KUMAGAI_SCR
    _create = "agpu_kumagai_scr_create"
    _default = "Kumagai_CompMaterSci_39_457_Si__Scr"

    @staticmethod
    def _struct(params):
        return _kumagai_struct(params), 1


class BrennerScr(_Screened):
    """BrennerScr (brenner_scr.f90); keyword arguments of Brenner plus the
    screening parameters; without arguments the screened
    Erhart_PRB_71_035211_SiC."""
    kind = _lib.KIND_BRENNER_SCR
    _create = "agpu_brenner_scr_create"
    _default = "Erhart_PRB_71_035211_SiC__Scr"

    @staticmethod
    def _struct(params):
        nel = len(params["el"])
        return _brenner_struct(params), nel * (nel + 1) // 2


class JuslinScr(_Screened):
    """JuslinScr (juslin_scr.f90); keyword arguments of Juslin plus the
    screening parameters (or1, or2, bor1, bor2, Cmin, Cmax per ordered pair
    j + i nel); without arguments the screened Juslin W-C-H set."""
    kind = _lib.KIND_JUSLIN_SCR
    _create = "agpu_juslin_scr_create"
    _default = "Juslin_JAP_98_123520_WCH__Scr"

    @staticmethod
    def _struct(params):
        nel = len(params["el"])
        return _juslin_struct(params), nel * nel


class TabulatedAlloyEAM(_Potential):
    """TabulatedAlloyEAM (tabulated_alloy_eam.f90): fn = setfl file,
    elements = filter string."""
    kind = _lib.KIND_EAM_ALLOY

    def __init__(self, fn, elements="*"):
        super().__init__()
        h = ctypes.c_void_p()
        call("agpu_eam_alloy_create", ctypes.byref(h), str(fn).encode(), str(elements).encode())
        self._h = h


class TabulatedEAM(_Potential):
    """TabulatedEAM (tabulated_eam.f90): fn = single-element funcfl file
    (F(rho), effective charge Z(r), rho(r)), elements = filter string.  No
    mask, as the reference class."""
    kind = _lib.KIND_EAM

    def __init__(self, fn, elements="*"):
        super().__init__()
        h = ctypes.c_void_p()
        call("agpu_eam_create", ctypes.byref(h), str(fn).encode(), str(elements).encode())
        self._h = h


class Rebo2(_Potential):
    """Second-generation REBO for C and H (rebo2.f90, classname Rebo2) with
    the default tables of rebo2_default_tables.f90.  Keyword arguments follow
    rebo2_registry.f90: zero_tables (bool) and dihedral (bool, the torsion
    term; default False as with_dihedral in rebo2_type.f90)."""
    kind = _lib.KIND_REBO2

    def __init__(self, zero_tables=False, dihedral=False, elements="C,H"):
        super().__init__()
        if str(elements).replace(" ", "") not in ("C,H", "H,C", "*"):
            raise AgpuError(_lib.ERR_UNSUPPORTED, "Rebo2: only the elements C and H are implemented")
        flags = (_lib.REBO2_ZERO_TABLES if zero_tables else 0) | (_lib.REBO2_DIHEDRAL if dihedral else 0)
        h = ctypes.c_void_p()
        call("agpu_rebo2_create", ctypes.byref(h), int(flags))
        self._h = h


class Ewald(_Potential):
    """Ewald Coulomb solver (Coulomb flavour: energy_and_forces(p, nl, q, f),
    potential(p, nl, q, phi)).  beta <= 0: beta^2 = ln(10)*12/cutoff^2 as
    pme.f90:192; kmax <= 0: exp(-kmax^2/4beta^2) = 1e-13."""
    kind = _lib.KIND_EWALD

    def __init__(self, cutoff=10.0, beta=0.0, kmax=0.0):
        super().__init__()
        h = ctypes.c_void_p()
        call("agpu_ewald_create", ctypes.byref(h), float(cutoff), float(beta), float(kmax))
        self._h = h
        self.cutoff = cutoff

    def energy_and_forces(self, p, nl, q, f=None):
        nat = p.nat
        if f is None:
            f = np.zeros((nat, 3))
        q = np.ascontiguousarray(q, dtype=np.float64)
        epot = ctypes.c_double(0.0)
        wpot = np.zeros(9)
        call("agpu_potential_energy_and_forces", self._h, p._h, nl._h, _ptr(q), ctypes.byref(epot),
             _ptr(f), _ptr(wpot), None, None, None, None, None, None, _lib.HOST)
        return epot.value, f, wpot.reshape(3, 3, order="F").copy()

    def potential(self, p, nl, q, phi=None):
        nat = p.nat
        if phi is None:
            phi = np.zeros(nat)
        q = np.ascontiguousarray(q, dtype=np.float64)
        call("agpu_coulomb_potential", self._h, p._h, nl._h, _ptr(q), _ptr(phi), _lib.HOST)
        return phi


class PME(Ewald):
    """PME (pme.f90, classname PME): pme_init(cutoff, order, grid); real
    space and self term as Ewald with beta^2 = ln(10)*12/cutoff^2, smooth-PME
    reciprocal sum with B-splines of the given order on a K1 x K2 x K3
    FFT grid."""
    kind = _lib.KIND_PME

    def __init__(self, cutoff=10.0, order=8, grid=(64, 64, 64)):
        _Potential.__init__(self)
        h = ctypes.c_void_p()
        g = (ctypes.c_int * 3)(*[int(x) for x in grid])
        call("agpu_pme_create", ctypes.byref(h), float(cutoff), int(order), g)
        self._h = h
        self.cutoff = cutoff
        self.order = order
        self.grid = tuple(int(x) for x in grid)


def lammps_compute(pot, p, nl, nlocal, x, Z, f=None, eatom=False):
    """One LAMMPS pair-style step: local atoms then ghosts in x (nall, 3).
    Returns (eng, f (nall, 3), virial[6] in LAMMPS order and sign, eatom)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    nall = len(x)
    if f is None:
        f = np.zeros((nall, 3))
    eng = ctypes.c_double(0.0)
    vir = np.zeros(6)
    ea = np.zeros(nall) if eatom else None
    call("agpu_lammps_compute", pot._h, p._h, nl._h, int(nlocal), int(nall), _ptr(x), _ptr(Z), _ptr(f),
         ctypes.byref(eng), _ptr(vir), _ptr(ea))
    return eng.value, f, vir, ea


def lammps_compute_list(pot, p, nl, nlocal, x, Z, seed, numneigh, neigh, f=None, eatom=False):
    """The LAMMPS step over LAMMPS' own full list with ghosts: row i is
    neigh[seed[i] : seed[i] + numneigh[i]] (indices into x, special-bond bits
    in the top two bits are ignored).  Returns as :func:`lammps_compute`."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    seed = np.ascontiguousarray(seed, dtype=np.int64)
    numneigh = np.ascontiguousarray(numneigh, dtype=np.int32)
    neigh = np.ascontiguousarray(neigh, dtype=np.int32)
    nall = len(x)
    if len(seed) != nall or len(numneigh) != nall:
        raise ValueError("seed and numneigh need one entry per local and ghost atom")
    if f is None:
        f = np.zeros((nall, 3))
    eng = ctypes.c_double(0.0)
    vir = np.zeros(6)
    ea = np.zeros(nall) if eatom else None
    call("agpu_lammps_compute_list", pot._h, p._h, nl._h, int(nlocal), int(nall), _ptr(x), _ptr(Z), _ptr(seed),
         _ptr(numneigh), _ptr(neigh), _ptr(f), ctypes.byref(eng), _ptr(vir), _ptr(ea))
    return eng.value, f, vir, ea


def lammps_compute_tagged(pot, p, nl, nlocal, x, Z, tag, seed=None, numneigh=None, neigh=None, f=None,
                          eatom=False, vatom=False):
    """The LAMMPS step with the atom tags (required by REBO2, whose bonds are
    evaluated once from the local atom with tag(j) >= tag(i)); with seed,
    numneigh, neigh over LAMMPS' own list, else over a list built here."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    tag = np.ascontiguousarray(tag, dtype=np.int32)
    nall = len(x)
    if len(tag) != nall:
        raise ValueError("tag needs one entry per local and ghost atom")
    if seed is not None:
        seed = np.ascontiguousarray(seed, dtype=np.int64)
        numneigh = np.ascontiguousarray(numneigh, dtype=np.int32)
        neigh = np.ascontiguousarray(neigh, dtype=np.int32)
    if f is None:
        f = np.zeros((nall, 3))
    eng = ctypes.c_double(0.0)
    vir = np.zeros(6)
    ea = np.zeros(nall) if eatom else None
    va = np.zeros((nall, 6)) if vatom else None
    call("agpu_lammps_compute_tagged", pot._h, p._h, nl._h, int(nlocal), int(nall), _ptr(x), _ptr(Z), _ptr(tag),
         _ptr(seed), _ptr(numneigh), _ptr(neigh), _ptr(f), ctypes.byref(eng), _ptr(vir), _ptr(ea), _ptr(va))
    if vatom:
        return eng.value, f, vir, ea, va
    return eng.value, f, vir, ea


class MD:
    """NVE velocity-Verlet dynamics on the GPU around one potential (Tersoff
    family or EAM), with a Verlet shell ``skin`` on the list and the steps
    replayed as HIP graphs of ``graph_steps`` steps.  Units are the
    potential's (eV, Angstrom, amu: time unit 10.18 fs)."""

    def __init__(self, pot, p, nl, masses, dt, skin=0.3, v0=None):
        pot.bind_to(p, nl)
        self.nat = p.nat
        masses = np.ascontiguousarray(np.broadcast_to(np.asarray(masses, dtype=np.float64), (self.nat,)))
        v0 = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        h = ctypes.c_void_p()
        call("agpu_md_create", ctypes.byref(h), pot._h, p._h, nl._h, _ptr(masses), _ptr(v0), float(dt), float(skin))
        self._h = h
        self._keep = (pot, p, nl)
        self.masses = masses

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib.load().agpu_md_destroy(h)
            self._h = None

    def run(self, nsteps, graph_steps=10):
        """nsteps steps; returns the potential energy after the last one."""
        e = ctypes.c_double(0.0)
        call("agpu_md_run", self._h, int(nsteps), int(graph_steps), ctypes.byref(e))
        return e.value

    def state(self):
        """(r, v, f, counters {steps, builds, replays, rollbacks}) on the host."""
        r = np.zeros((self.nat, 3))
        v = np.zeros((self.nat, 3))
        f = np.zeros((self.nat, 3))
        c = np.zeros(4, np.int64)
        call("agpu_md_get", self._h, _ptr(r), _ptr(v), _ptr(f), None, _ptr(c), _lib.HOST)
        return r, v, f, dict(zip(("steps", "builds", "replays", "rollbacks"), c.tolist()))

    def kinetic_energy(self):
        _, v, _, _ = self.state()
        return 0.5 * float(np.sum(self.masses[:, None] * v * v))


def use_torch_stream():
    """Launch every kernel on torch's current stream of the current device,
    so library work is ordered with torch ops on the same tensors."""
    import torch
    call("agpu_set_stream", ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def profile_enable(on=True):
    call("agpu_profile_enable", 1 if on else 0)


def profile_reset():
    call("agpu_profile_reset")


def profile_list():
    """Names of every kernel timed since the last profile_reset()."""
    buf = ctypes.create_string_buffer(8192)
    call("agpu_profile_list", buf, 8192)
    v = buf.value.decode()
    return v.split(",") if v else []


def profile_query(kernel):
    """(total ms, launches) of one kernel, from HIP events on its launch stream."""
    t = ctypes.c_double()
    n = ctypes.c_int()
    call("agpu_profile_query", kernel.encode(), ctypes.byref(t), ctypes.byref(n))
    return t.value, n.value


def from_atoms(positions, cell, pbc, numbers):
    """Build a Particles object from plain arrays (native.from_atoms analogue)."""
    p = Particles()
    p.allocate(len(positions))
    p.set_cell(cell, pbc)
    p.Z[:] = numbers
    p.coordinates[:, :] = positions
    p.I_changed_positions()
    p.update_elements()
    return p


def tersoff_struct(params):
    return _tersoff_struct(params)


__all__ = ["Particles", "Neighbors", "Tersoff", "TabulatedAlloyEAM", "Ewald", "from_atoms",
           "pair_index"]
