"""Published parameter sets for the bond-order potentials.

The numbers are data, kept identical to the reference's database so that the
same keyword dictionaries can be passed to either implementation:
``src/python/atomistica/parameters.py:60-77`` (Tersoff Si/C, PRB 39, 5566)
and ``src/potentials/bop/tersoff/tersoff_params.f90:99-138`` (the Fortran
default ``Tersoff_PRB_39_5566_SiC``).

Pair-indexed entries follow the symmetric pair index of ``src/macros.inc:123``
(0-based here): for elements (C, Si) the order is C-C, C-Si, Si-Si.
"""
import copy
from math import log, sqrt


def pair_index(i, j, nel):
    """0-based symmetric pair index (PAIR_INDEX of src/macros.inc:123)."""
    return min(i + j * nel, j + i * nel) - min(i * (i + 1) // 2, j * (j + 1) // 2)


Tersoff_PRB_39_5566_Si_C = {
    "__ref__": "Tersoff J., Phys. Rev. B 39, 5566 (1989)",
    "el": ["C", "Si"],
    "A": [1.3936e3, sqrt(1.3936e3 * 1.8308e3), 1.8308e3],
    "B": [3.4674e2, sqrt(3.4674e2 * 4.7118e2), 4.7118e2],
    "xi": [1.0, 0.9776e0, 1.0],
    "lambda": [3.4879e0, (3.4879e0 + 2.4799e0) / 2, 2.4799e0],
    "mu": [2.2119e0, (2.2119e0 + 1.7322e0) / 2, 1.7322e0],
    "omega": [1.0, 1.0, 1.0],
    "mubo": [0.0, 0.0, 0.0],
    "m": [1, 1, 1],
    "beta": [1.5724e-7, 1.1000e-6],
    "n": [7.2751e-1, 7.8734e-1],
    "c": [3.8049e4, 1.0039e5],
    "d": [4.3484e0, 1.6217e1],
    "h": [-5.7058e-1, -5.9825e-1],
    "r1": [1.80, sqrt(1.80 * 2.70), 2.70],
    "r2": [2.10, sqrt(2.10 * 3.00), 3.00],
}

# Single-element Si slice of the same set (the Si-Si entries, pair index 2).
Tersoff_Si = {
    "__ref__": "Tersoff J., Phys. Rev. B 39, 5566 (1989), Si",
    "el": ["Si"],
    "A": [1.8308e3],
    "B": [4.7118e2],
    "xi": [1.0],
    "lambda": [2.4799e0],
    "mu": [1.7322e0],
    "omega": [1.0],
    "mubo": [0.0],
    "m": [1],
    "beta": [1.1000e-6],
    "n": [7.8734e-1],
    "c": [1.0039e5],
    "d": [1.6217e1],
    "h": [-5.9825e-1],
    "r1": [2.70],
    "r2": [3.00],
}

# Kumagai, Izumi, Hara, Sakai, Comp. Mater. Sci. 39, 457 (2007), Si
# (kumagai_params.f90:97-132; the names there, one element).
Kumagai_CompMaterSci_39_457_Si = {
    "__ref__": "Kumagai, Izumi, Hara, Sakai, Comp. Mater. Sci. 39, 457 (2007)",
    "family": "kumagai",
    "el": ["Si"],
    "A": 3281.5905,
    "B": 121.00047,
    "lambda1": 3.2300135,
    "lambda2": 1.3457970,
    "eta": 1.0000000,
    "delta": 0.53298909,
    "alpha": 2.3890327,
    "beta": 1,
    "c1": 0.20173476,
    "c2": 730418.72,
    "c3": 1000000.0,
    "c4": 1.0000000,
    "c5": 26.000000,
    "h": -0.36500000,
    "r1": 2.70,
    "r2": 3.30,
}

# Brenner-form (Albe/Erhart) sets without screening, every parameter per
# element pair in PAIR_INDEX order (brenner_params.f90:78-185 and the
# reference's Python parameters.py for Brenner's 1990 carbon sets).
Erhart_PRB_71_035211_SiC = {
    "__ref__": "Erhart P., Albe K., Phys. Rev. B 71, 035211 (2005)",
    "family": "brenner",
    "el": ["C", "Si"],
    "D0": [6.00, 4.36, 3.24],
    "r0": [1.4276, 1.79, 2.232],
    "S": [2.167, 1.847, 1.842],
    "beta": [2.0099, 1.6991, 1.4761],
    "gamma": [0.11233, 0.011877, 0.114354],
    "c": [181.910, 273987.0, 2.00494],
    "d": [6.28433, 180.314, 0.81472],
    "h": [0.5556, 0.68, 0.259],
    "mu": [0.0, 0.0, 0.0],
    "n": [1.0, 1.0, 1.0],
    "m": [1, 1, 1],
    "r1": [1.85, 2.20, 2.68],
    "r2": [2.15, 2.60, 2.96],
}

Albe_PRB_65_195124_PtC = {
    "__ref__": "Albe K., Nordlund K., Averback R. S., Phys. Rev. B 65, 195124 (2002)",
    "family": "brenner",
    "el": ["Pt", "C"],
    "D0": [3.683, 5.3, 6.0],
    "r0": [2.384, 1.84, 1.39],
    "S": [2.24297, 1.1965, 1.22],
    "beta": [1.64249, 1.836, 2.1],
    "gamma": [0.0008542, 0.0097, 0.00020813],
    "c": [34.0, 1.23, 330.0],
    "d": [1.1, 0.36, 3.5],
    "h": [1.0, 1.0, 1.0],
    "mu": [1.335, 0.0, 0.0],
    "n": [1.0, 1.0, 1.0],
    "m": [1, 1, 1],
    "r1": [2.9, 2.5, 1.7],
    "r2": [3.3, 2.8, 2.0],
}

Henriksson_PRB_79_144107_FeC = {
    "__ref__": "Henriksson K.O.E., Nordlund K., Phys. Rev. B 79, 144107 (2009)",
    "family": "brenner",
    "el": ["Fe", "C"],
    "D0": [1.5, 4.82645134, 6.0],
    "r0": [2.29, 1.47736510, 1.39],
    "S": [2.0693109, 1.43134755, 1.22],
    "beta": [1.4, 1.63208170, 2.1],
    "gamma": [0.0115751, 0.00205862, 0.00020813],
    "c": [1.2898716, 8.95583221, 330.0],
    "d": [0.3413219, 0.72062047, 3.5],
    "h": [-0.26, 0.87099874, 1.0],
    "mu": [0.0, 0.0, 0.0],
    "n": [1.0, 1.0, 1.0],
    "m": [1, 1, 1],
    "r1": [2.95, 2.3, 1.70],
    "r2": [3.35, 2.7, 2.00],
}

Brenner_PRB_42_9458_C_I = {
    "__ref__": "Brenner D., Phys. Rev. B 42, 9458 (1990) [potential I]",
    "family": "brenner",
    "el": ["C"],
    "D0": [6.325], "r0": [1.315], "S": [1.29], "beta": [1.5], "gamma": [0.011304],
    "c": [19.0], "d": [2.5], "h": [1.0], "mu": [0.0], "n": [1.0 / (2 * 0.80469)], "m": [1],
    "r1": [1.70], "r2": [2.00],
}

Brenner_PRB_42_9458_C_II = {
    "__ref__": "Brenner D., Phys. Rev. B 42, 9458 (1990) [potential II]",
    "family": "brenner",
    "el": ["C"],
    "D0": [6.0], "r0": [1.39], "S": [1.22], "beta": [2.1], "gamma": [0.00020813],
    "c": [330.0], "d": [3.5], "h": [1.0], "mu": [0.0], "n": [1.0 / (2 * 0.5)], "m": [1],
    "r1": [1.70], "r2": [2.00],
}

# The eleven scalar Tersoff parameters trained by the parameter-gradient path,
# in the order used throughout (theta vector of a single-element set).
# Juslin-form (W-C-H, Fe-C-H) sets: every pair parameter over ordered element
# pairs j + i nel (PAIR_INDEX_NS), alpha, omega, m over ordered triplets
# k + nel (j + nel i) of centre i (TRIPLET_INDEX_NS); an r0 < 0 entry takes
# the reversed pair's values (reference: src/python/atomistica/parameters.py:
# 279-295, 312-328).
Juslin_JAP_98_123520_WCH = {
    "__ref__": "Juslin N., Erhart P., Traskelin P., Nord J., Henriksson K.O.E, Nordlund K., Salonen E., "
               "Albe K., J. Appl. Phys. 98, 123520 (2005)",
    "family": "juslin",
    "el": ["W", "C", "H"],
    "D0": [5.41861, 6.64, 2.748, 0.0, 6.0, 3.6422, 0.0, 3.642, 4.7509],
    "r0": [2.34095, 1.90547, 1.727, -1.0, 1.39, 1.1199, -1.0, 1.1199, 0.74144],
    "S": [1.92708, 2.96149, 1.2489, 0.0, 1.22, 1.69077, 0.0, 1.69077, 2.3432],
    "beta": [1.38528, 1.80370, 1.52328, 0.0, 2.1, 1.9583, 0.0, 1.9583, 1.9436],
    "gamma": [0.00188227, 0.072855, 0.0054, 0.0, 0.00020813, 0.00020813, 0.0, 12.33, 12.33],
    "c": [2.14969, 1.10304, 1.788, 0.0, 330.0, 330.0, 0.0, 0.0, 0.0],
    "d": [0.17126, 0.33018, 0.8255, 0.0, 3.5, 3.5, 0.0, 1.0, 1.0],
    "h": [-0.27780, 0.75107, 0.38912, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0],
    "n": [1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0],
    "alpha": [0.45876, 0.0, 0.0, 0.45876, 0.0, 0.0, 0.45876, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 4.0, 0.0, 4.0,
              4.0, 0.0, 0.0, 0.0, 0.0, 4.0, 4.0, 0.0, 4.0, 4.0],
    "omega": [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.94586, 4.54415,
              1.0, 1.0, 1.0, 1.0, 0.33946, 0.22006, 1.0, 1.0, 1.0],
    "m": [1] * 27,
    "r1": [3.20, 2.60, 2.68, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "r2": [3.80, 3.00, 2.96, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
}

Kuopanportti_CMS_111_525_FeCH = {
    "__ref__": "Kuopanportti P., Hayward, N., Fu C., Kuronen A., Nordlund K., Comp. Mat. Sci. 111, 525 (2016)",
    "family": "juslin",
    "el": ["Fe", "C", "H"],
    "D0": [1.5, 4.82645134, 1.630, 0.0, 6.0, 3.6422, 0.0, 3.642, 4.7509],
    "r0": [2.29, 1.47736510, 1.589, -1.0, 1.39, 1.1199, -1.0, 1.1199, 0.74144],
    "S": [2.0693, 1.43134755, 4.000, 0.0, 1.22, 1.69077, 0.0, 1.69077, 2.3432],
    "beta": [1.4, 1.63208170, 1.875, 0.0, 2.1, 1.9583, 0.0, 1.9583, 1.9436],
    "gamma": [0.01158, 0.00205862, 0.01332, 0.0, 0.00020813, 0.00020813, 0.0, 12.33, 12.33],
    "c": [1.2899, 8.95583221, 424.5, 0.0, 330.0, 330.0, 0.0, 0.0, 0.0],
    "d": [0.3413, 0.72062047, 7.282, 0.0, 3.5, 3.5, 0.0, 1.0, 1.0],
    "h": [-0.26, 0.87099874, -0.1091, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0],
    "n": [1.0, 1.0, 1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 1.0],
    "alpha": [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 4.0, 0.0, 4.0, 4.0, 0.0, 0.0,
              0.0, 0.0, 4.0, 4.0, 0.0, 4.0, 4.0],
    "omega": [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 2.94586, 4.54415,
              1.0, 1.0, 1.0, 1.0, 0.33946, 0.22006, 1.0, 1.0, 1.0],
    "m": [1] * 27,
    "r1": [2.95, 2.30, 2.2974, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "r2": [3.35, 2.70, 2.6966, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
}


# ---------------------------------------------------------------------------
# Screened variants (TersoffScr, KumagaiScr, BrennerScr): inner cutoff
# [r1, r2], pair cutoff [or1, or2], bond-order cutoff [bor1, bor2], all of the
# exp_cutoff form, and the screening parameters Cmin, Cmax (reference:
# src/python/atomistica/parameters.py:79-107, 196-218, 409-420).
# ---------------------------------------------------------------------------
SCREENING_KEYS = ("or1", "or2", "bor1", "bor2", "Cmin", "Cmax")

# JuslinScr: the screened W-C-H set of the Fortran database
# (juslin_params.f90:113-122, the SCREENING branch), per ordered pair
# j + i nel.  The reference's Python dict of the same name
# (src/python/atomistica/parameters.py:297-308) sets Cmin and Cmax to the
# r1 and r2 lists instead; this set keeps the database's Cmin = 1, Cmax = 3,
# which is what the compiled JuslinScr class uses by default.
Juslin_JAP_98_123520_WCH__Scr = copy.deepcopy(Juslin_JAP_98_123520_WCH)
Juslin_JAP_98_123520_WCH__Scr.update({
    "or1": [3.20, 2.60, 2.68, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "or2": [3.80, 3.00, 2.96, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
    "bor1": [3.20, 2.60, 2.68, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "bor2": [3.80, 3.00, 2.96, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
    "Cmin": [1.00, 1.00, 1.00, 0.0, 1.00, 1.00, 0.0, 1.00, 1.00],
    "Cmax": [3.00, 3.00, 3.00, 0.0, 3.00, 3.00, 0.0, 3.00, 3.00],
})

# screened Fe-C-H set of the Fortran database (juslin_params.f90:131-139):
# the Fe-Fe, Fe-C and Fe-H pair and bond-order cutoffs start at 100 A, past
# their upper end, so those pairs keep only the inner cutoff [r1, r2]
Kuopanportti_CMS_111_525_FeCH__Scr = copy.deepcopy(Kuopanportti_CMS_111_525_FeCH)
Kuopanportti_CMS_111_525_FeCH__Scr.update({
    "or1": [100.0, 100.0, 100.0, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "or2": [3.35, 2.70, 2.6966, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
    "bor1": [100.0, 100.0, 100.0, 0.0, 1.70, 1.30, 0.0, 1.30, 1.10],
    "bor2": [3.35, 2.70, 2.6966, 0.0, 2.00, 1.80, 0.0, 1.80, 1.70],
    "Cmin": [1.00, 1.00, 1.00, 0.0, 1.00, 1.00, 0.0, 1.00, 1.00],
    "Cmax": [3.00, 3.00, 3.00, 0.0, 3.00, 3.00, 0.0, 3.00, 3.00],
})

Tersoff_PRB_39_5566_Si_C__Scr = copy.deepcopy(Tersoff_PRB_39_5566_Si_C)
Tersoff_PRB_39_5566_Si_C__Scr.update({
    "m": [3, 3, 3],
    "r1": [2.00, sqrt(2.00 * 2.50), 2.50],
    "r2": [2.00 * 1.2, sqrt(2.00 * 2.50) * 1.2, 2.50 * 1.2],
    "or1": [2.00, sqrt(2.00 * 3.00), 3.00],
    "or2": [2.00 * 2.0, sqrt(2.00 * 3.00) * 2.0, 3.00 * 2.0],
    "bor1": [2.00, sqrt(2.00 * 3.00), 3.00],
    "bor2": [2.00 * 2.0, sqrt(2.00 * 3.00) * 2.0, 3.00 * 2.0],
    "Cmin": [1.00, 1.00, 1.00],
    "Cmax": [3.00, 3.00, 3.00],
})
# mubo is the inverse dimer length of each pair
_p = Tersoff_PRB_39_5566_Si_C__Scr
_p["mubo"] = [(_p["lambda"][i] - _p["mu"][i]) /
              log((_p["lambda"][i] * _p["A"][i]) / (_p["mu"][i] * _p["B"][i])) for i in range(3)]
del _p

Kumagai_CompMaterSci_39_457_Si__Scr = copy.deepcopy(Kumagai_CompMaterSci_39_457_Si)
Kumagai_CompMaterSci_39_457_Si__Scr.update({
    "r1": 2.50, "r2": 2.50 * 1.2,
    "or1": 3.00, "or2": 3.00 * 2.0,
    "bor1": 3.00, "bor2": 3.00 * 2.0,
    "Cmin": 1.00, "Cmax": 3.00,
})

Erhart_PRB_71_035211_SiC__Scr = copy.deepcopy(Erhart_PRB_71_035211_SiC)
Erhart_PRB_71_035211_SiC__Scr.update({
    "mu": [1.0 / 1.4276, 1.0 / 1.79, 1.0 / 1.842],
    "n": [1.0, 1.0, 1.0],
    "m": [3, 3, 3],
    "r1": [2.00, 2.40, 2.50],
    "r2": [2.00 * 1.2, 2.40 * 1.2, 2.50 * 1.2],
    "or1": [2.00, 2.40, 3.00],
    "or2": [2.00 * 2.0, 2.40 * 2.0, 3.00 * 2.0],
    "bor1": [2.00, 2.40, 3.00],
    "bor2": [2.00 * 2.0, 2.40 * 2.0, 3.00 * 2.0],
    "Cmin": [1.00, 1.00, 1.00],
    "Cmax": [3.00, 3.00, 3.00],
})


TERSOFF_THETA_NAMES = ("A", "B", "lambda", "mu", "beta", "n", "c", "d", "h", "r1", "r2")

_SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca "
            "Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr "
            "Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd "
            "Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg "
            "Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm "
            "Md No Lr").split()
ATOMIC_NUMBERS = {sym: z + 1 for z, sym in enumerate(_SYMBOLS)}
ELEMENT_NAMES = {v: k for k, v in ATOMIC_NUMBERS.items()}
