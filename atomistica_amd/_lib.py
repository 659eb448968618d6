"""ctypes binding of ``libatomistica_amd.so`` (the C-ABI of include/atomistica_amd.h).

The library is loaded from the package directory.  There is no fallback:
if the shared object is missing or the GPU runtime is absent every call
raises, so nothing silently runs elsewhere.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("ATOMISTICA_AMD_LIB") or os.path.join(HERE, "libatomistica_amd.so")

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_NL_OVERFLOW = 3
ERR_DOMAIN = 4
ERR_SPLINE = 5
ERR_UNSUPPORTED = 6
ERR_ELEMENT = 7
ERR_PBC = 8
ERR_IO = 9
ERR_STATE = 10
ERR_NEB = 11

HOST = 1
PREC_FP64 = 0
PREC_MIXED = 1
KIND_TERSOFF = 1
KIND_EAM_ALLOY = 2
KIND_EWALD = 3
KIND_REBO2 = 4
KIND_KUMAGAI = 5
KIND_BRENNER = 6
KIND_TERSOFF_SCR = 7
KIND_KUMAGAI_SCR = 8
KIND_BRENNER_SCR = 9
KIND_PME = 10
KIND_JUSLIN = 11
KIND_JUSLIN_SCR = 12
KIND_EAM = 13
REBO2_ZERO_TABLES = 1
REBO2_DIHEDRAL = 2

_ERRNAMES = {1: "ARG", 2: "HIP", 3: "NL_OVERFLOW", 4: "DOMAIN", 5: "SPLINE",
             6: "UNSUPPORTED", 7: "ELEMENT", 8: "PBC", 9: "IO", 10: "STATE", 11: "NEB"}


class AgpuError(RuntimeError):
    """Error raised by the native library; ``code`` is the AGPU_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__("%s (AGPU_ERR_%s)" % (msg, _ERRNAMES.get(code, code)))
        self.code = code


vp = ctypes.c_void_p
dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
lp = ctypes.POINTER(ctypes.c_int64)
i = ctypes.c_int
d = ctypes.c_double


class TersoffParams(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int), ("Z", ctypes.c_int * 3)] + \
        [(k, ctypes.c_double * 6) for k in ("A", "B", "xi", "lambda_", "mu", "omega", "mubo")] + \
        [("m", ctypes.c_int * 6)] + \
        [(k, ctypes.c_double * 3) for k in ("beta", "n", "c", "d", "h")] + \
        [(k, ctypes.c_double * 6) for k in ("r1", "r2")]


class KumagaiParams(ctypes.Structure):
    _fields_ = [("Z", ctypes.c_int)] + \
        [(k, ctypes.c_double) for k in ("A", "B", "lambda1", "lambda2", "eta", "delta", "alpha")] + \
        [("beta", ctypes.c_int)] + \
        [(k, ctypes.c_double) for k in ("c1", "c2", "c3", "c4", "c5", "h", "r1", "r2")]


class BrennerParams(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int), ("Z", ctypes.c_int * 3)] + \
        [(k, ctypes.c_double * 6) for k in ("D0", "r0", "S", "beta", "gamma", "c", "d", "h", "mu", "n")] + \
        [("m", ctypes.c_int * 6)] + \
        [(k, ctypes.c_double * 6) for k in ("r1", "r2")]


class JuslinParams(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int), ("Z", ctypes.c_int * 3)] + \
        [(k, ctypes.c_double * 9) for k in ("D0", "r0", "S", "beta", "gamma", "c", "d", "h", "n")] + \
        [("alpha", ctypes.c_double * 27), ("omega", ctypes.c_double * 27), ("m", ctypes.c_int * 27)] + \
        [(k, ctypes.c_double * 9) for k in ("r1", "r2")]


class ScreeningParams(ctypes.Structure):
    # per element pair: symmetric index (6 used) or, for JuslinScr, ordered (9)
    _fields_ = [(k, ctypes.c_double * 9) for k in ("or1", "or2", "bor1", "bor2", "Cmin", "Cmax")]


# name -> argtypes (every function returns int unless listed in _RESTYPE)
PROTOTYPES = {
    "agpu_version": [],
    "agpu_get_error": [ctypes.c_char_p, i],
    "agpu_device_count": [ip],
    "agpu_set_device": [i],
    "agpu_set_stream": [vp],
    "agpu_get_stream": [ctypes.POINTER(vp)],
    "agpu_synchronize": [],
    "agpu_malloc": [ctypes.POINTER(vp), ctypes.c_size_t],
    "agpu_free": [vp],
    "agpu_memcpy": [vp, vp, ctypes.c_size_t, i],
    "agpu_profile_enable": [i],
    "agpu_profile_reset": [],
    "agpu_profile_query": [ctypes.c_char_p, dp, ip],
    "agpu_profile_list": [ctypes.c_char_p, i],
    "agpu_particles_create": [ctypes.POINTER(vp)],
    "agpu_particles_destroy": [vp],
    "agpu_particles_allocate": [vp, i],
    "agpu_particles_set_cell": [vp, dp, ip],
    "agpu_particles_set_atomic_numbers": [vp, vp, i],
    "agpu_particles_set_positions": [vp, vp, i],
    "agpu_particles_set_position_pointer": [vp, i, vp],
    "agpu_particles_changed_positions": [vp],
    "agpu_particles_get_nat": [vp, ip],
    "agpu_neighbors_create": [ctypes.POINTER(vp), i],
    "agpu_neighbors_destroy": [vp],
    "agpu_neighbors_request_interaction_range": [vp, d],
    "agpu_neighbors_set_bin_size": [vp, d],
    "agpu_neighbors_update": [vp, vp],
    "agpu_neighbors_get_size": [vp, lp, ip],
    "agpu_neighbors_get": [vp, vp, vp, vp, i],
    "agpu_tersoff_create": [ctypes.POINTER(vp), ctypes.POINTER(TersoffParams)],
    "agpu_tersoff_set_params": [vp, ctypes.POINTER(TersoffParams)],
    "agpu_kumagai_create": [ctypes.POINTER(vp), ctypes.POINTER(KumagaiParams)],
    "agpu_brenner_create": [ctypes.POINTER(vp), ctypes.POINTER(BrennerParams)],
    "agpu_juslin_create": [ctypes.POINTER(vp), ctypes.POINTER(JuslinParams)],
    "agpu_tersoff_scr_create": [ctypes.POINTER(vp), vp, vp],
    "agpu_kumagai_scr_create": [ctypes.POINTER(vp), vp, vp],
    "agpu_brenner_scr_create": [ctypes.POINTER(vp), vp, vp],
    "agpu_juslin_scr_create": [ctypes.POINTER(vp), vp, vp],
    "agpu_eam_alloy_create": [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_char_p],
    "agpu_eam_create": [ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_char_p],
    "agpu_ewald_create": [ctypes.POINTER(vp), d, d, d],
    "agpu_pme_create": [ctypes.POINTER(vp), d, i, ip],
    "agpu_rebo2_create": [ctypes.POINTER(vp), i],
    "agpu_potential_destroy": [vp],
    "agpu_potential_kind": [vp, ip],
    "agpu_potential_set_precision": [vp, i],
    "agpu_potential_get_cutoff": [vp, dp],
    "agpu_potential_bind_to": [vp, vp, vp],
    "agpu_potential_energy_and_forces": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i],
    "agpu_potential_energy_and_forces_noalloc": [vp, vp, vp, vp, vp, vp, vp],
    "agpu_coulomb_potential": [vp, vp, vp, vp, vp, i],
    "agpu_potential_autotune": [vp, vp, vp, i, ctypes.c_char_p, i],
    "agpu_potential_set_launch": [vp, vp, i, i, ip],
    "agpu_batch_create": [ctypes.POINTER(vp)],
    "agpu_batch_destroy": [vp],
    "agpu_batch_set_structures": [vp, i, vp, vp],
    "agpu_batch_set_positions": [vp, vp, i],
    "agpu_batch_get_size": [vp, ip, ip, lp],
    "agpu_tersoff_batch_forward": [vp, vp, i, vp, vp, i],
    "agpu_tersoff_batch_backward": [vp, vp, vp, vp, vp, i],
    "agpu_ewald_batch_compute": [vp, vp, d, d, d, i, vp, vp, vp, vp, i],
    "agpu_md_create": [ctypes.POINTER(vp), vp, vp, vp, vp, vp, d, d],
    "agpu_md_destroy": [vp],
    "agpu_md_run": [vp, i, i, vp],
    "agpu_md_get": [vp, vp, vp, vp, vp, vp, i],
    "agpu_random_positions": [i, vp, d, ctypes.c_uint64, i, vp],
    "agpu_lammps_compute": [vp, vp, vp, i, i, vp, vp, vp, vp, vp, vp],
    "agpu_lammps_compute_list": [vp, vp, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    "agpu_lammps_compute_tagged": [vp, vp, vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
}
_RESTYPE = {"agpu_version": ctypes.c_char_p}

_lib = None


def load():
    """Load the shared object and attach prototypes (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("atomistica_amd: %s is missing -- build it with "
                              "`python -m atomistica_amd.build` (no fallback path exists)"
                              % LIBPATH)
        lib = ctypes.CDLL(LIBPATH)
        for name, argt in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.argtypes = argt
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(4096)
    load().agpu_get_error(buf, 4096)
    return buf.value.decode(errors="replace")


def call(name, *args):
    """Call a C-ABI function and raise AgpuError on a nonzero return code."""
    rc = getattr(load(), name)(*args)
    if rc != OK:
        raise AgpuError(rc, "%s: %s" % (name, last_error()))
    return rc
