"""Published parameter sets for the bond-order potentials.

The numbers are data, kept identical to the reference's database so that the
same keyword dictionaries can be passed to either implementation:
``src/python/atomistica/parameters.py:60-77`` (Tersoff Si/C, PRB 39, 5566)
and ``src/potentials/bop/tersoff/tersoff_params.f90:99-138`` (the Fortran
default ``Tersoff_PRB_39_5566_SiC``).

Pair-indexed entries follow the symmetric pair index of ``src/macros.inc:123``
(0-based here): for elements (C, Si) the order is C-C, C-Si, Si-Si.
"""
from math import sqrt


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

# The eleven scalar Tersoff parameters trained by the parameter-gradient path,
# in the order used throughout (theta vector of a single-element set).
TERSOFF_THETA_NAMES = ("A", "B", "lambda", "mu", "beta", "n", "c", "d", "h", "r1", "r2")

ATOMIC_NUMBERS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8,
                  "F": 9, "Ne": 10, "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15,
                  "S": 16, "Cl": 17, "Ar": 18, "K": 19, "Ca": 20, "Ti": 22, "Fe": 26,
                  "Co": 27, "Ni": 28, "Cu": 29, "Zn": 30, "Ga": 31, "Ge": 32, "As": 33,
                  "Ag": 47, "Sn": 50, "Pt": 78, "Au": 79, "Pb": 82}
ELEMENT_NAMES = {v: k for k, v in ATOMIC_NUMBERS.items()}
