"""ASE-style calculator over the GPU potentials.

Mirrors the reference's ``Atomistica`` calculator
(src/python/atomistica/aseinterface.py:143-449): positions are taken minus
the cell displacement, all potentials accumulate into one force array,
per-atom energies and stresses are computed only on request, the stress is
the Voigt form of wpot/V with symmetrised off-diagonals, Coulomb results are
multiplied by Hartree*Bohr, and the particle/neighbour objects are rebuilt
when the number or the species of atoms change.

ASE itself is optional: any object with the ``ase.Atoms`` accessors used
here (get_positions, get_cell, get_pbc, get_atomic_numbers, get_celldisp,
get_volume, has, get_array) works, e.g. :class:`Structure`.
"""
import numpy as np

from . import native

# ASE's default constants (CODATA 2014): Hartree in eV, Bohr in Angstrom
HARTREE = 27.211386024367243
BOHR = 0.5291772105638411
all_changes = ["positions", "numbers", "cell", "pbc", "initial_charges", "initial_magmoms"]


class Structure:
    """Minimal stand-in for ase.Atoms (positions, cell, pbc, numbers, charges)."""

    def __init__(self, positions, cell, pbc=True, numbers=None, charges=None, celldisp=None):
        self.positions = np.array(positions, dtype=np.float64)
        self.cell = np.array(cell, dtype=np.float64).reshape(3, 3)
        self.pbc = np.array([pbc] * 3 if np.isscalar(pbc) else pbc, dtype=bool)
        self.numbers = np.array(numbers, dtype=np.int32)
        self.celldisp = np.zeros(3) if celldisp is None else np.asarray(celldisp, dtype=np.float64)
        self.arrays = {}
        if charges is not None:
            self.arrays["initial_charges"] = np.asarray(charges, dtype=np.float64)
        self.calc = None

    def __len__(self):
        return len(self.positions)

    def get_positions(self):
        return self.positions.copy()

    def get_cell(self):
        return self.cell.copy()

    def get_pbc(self):
        return self.pbc.copy()

    def get_atomic_numbers(self):
        return self.numbers.copy()

    def get_celldisp(self):
        return self.celldisp.reshape(3, 1)

    def get_volume(self):
        return abs(np.linalg.det(self.cell))

    def has(self, name):
        return name in self.arrays

    def get_array(self, name):
        return self.arrays[name].copy()

    def get_potential_energy(self):
        return self.calc.get_potential_energy(self)

    def get_forces(self):
        return self.calc.get_forces(self)

    def get_stress(self):
        return self.calc.get_stress(self)


class Atomistica:
    """Calculator driving one or more GPU potentials (and Coulomb solvers)."""

    implemented_properties = ["energy", "free_energy", "stress", "forces", "charges"]
    potential_class = None
    avgn = 100

    def __init__(self, potentials=None, avgn=None, precision="fp64", **kwargs):
        self.pots = []
        self.couls = []
        if potentials is not None:
            for pot in potentials:
                if isinstance(pot, Atomistica):
                    raise TypeError("Potential passed to Atomistica class is already an Atomistica "
                                    "object. Pass native potentials from atomistica_amd.native.")
                (self.couls if hasattr(pot, "potential") else self.pots).append(pot)
        else:
            pot = self.potential_class(**kwargs)
            (self.couls if hasattr(pot, "potential") else self.pots).append(pot)
        for pot in self.pots + self.couls:
            pot.set_precision(precision)
        if avgn:
            self.avgn = avgn
        self.particles = None
        self.nl = None
        self.charges = None
        self.initial_charges = None
        self.mask = None
        self.atoms = None
        self.results = {}
        self.properties_changed = list(all_changes)
        self._last = None
        # per-bond properties (aseinterface.py:209-216 of the reference)
        self.compute_epot_per_bond = False
        self.compute_f_per_bond = False
        self.compute_wpot_per_bond = False
        self.epot_per_bond = None
        self.f_per_bond = None
        self.wpot_per_bond = None

    # -- state -------------------------------------------------------------
    def initialize(self, atoms):
        if self.mask is not None and len(self.mask) != len(atoms):
            raise RuntimeError("mask has {0} entries but the Atoms object holds {1} atoms"
                               .format(len(self.mask), len(atoms)))
        self.particles = native.Particles()
        self.particles.allocate(len(atoms))
        self.particles.set_cell(atoms.get_cell(), atoms.get_pbc())
        self.particles.Z[:] = atoms.get_atomic_numbers()
        self.particles.coordinates[:, :] = atoms.get_positions() - np.asarray(atoms.get_celldisp()).ravel()
        self.particles.I_changed_positions()
        self.particles.update_elements()
        self.nl = native.Neighbors(self.avgn)
        for pot in self.pots:
            pot.bind_to(self.particles, self.nl)
        if self.couls:
            if atoms.has("initial_charges"):
                ic = atoms.get_array("initial_charges")
                if self.charges is None or len(self.charges) != len(atoms) or \
                        (self.initial_charges is not None and np.any(ic != self.initial_charges)):
                    self.charges = ic
                    self.initial_charges = ic.copy()
            if self.charges is None or len(self.charges) != len(atoms):
                self.charges = np.zeros(len(atoms))
        for coul in self.couls:
            coul.bind_to(self.particles, self.nl)
        self.properties_changed = list(all_changes)

    def set_mask(self, mask):
        if mask is None or self.mask is None or np.any(self.mask != mask):
            self.mask = None if mask is None else np.asarray(mask)
            self.properties_changed += ["mask"]

    def set_per_bond(self, epot=None, f=None, wpot=None):
        """Request per-bond energies, pair forces and virials on the next
        calculation (aseinterface.py:290-297); they are kept in
        epot_per_bond, f_per_bond and wpot_per_bond in the order of
        get_neighbors()."""
        if epot is not None:
            self.compute_epot_per_bond = epot
        if f is not None:
            self.compute_f_per_bond = f
        if wpot is not None:
            self.compute_wpot_per_bond = wpot
        self.properties_changed += ["per_bond"]

    def update(self, atoms):
        if atoms is None:
            return
        if self.particles is None or len(self.particles.Z) != len(atoms) or \
                (self.charges is not None and len(self.charges) != len(atoms)):
            self.initialize(atoms)
        elif np.any(self.particles.Z != atoms.get_atomic_numbers()):
            self.initialize(atoms)
        if np.any(self.particles.cell != atoms.get_cell()) or np.any(self.particles.pbc != atoms.get_pbc()):
            self.particles.set_cell(atoms.get_cell(), atoms.get_pbc())
            self.properties_changed += ["cell"]
        r = atoms.get_positions() - np.asarray(atoms.get_celldisp()).ravel()
        if np.any(self.particles.coordinates != r):
            self.particles.coordinates[:, :] = r
            self.particles.I_changed_positions()
            self.properties_changed += ["positions"]
        if atoms.has("initial_charges") and self.initial_charges is not None and \
                np.any(atoms.get_array("initial_charges") != self.initial_charges):
            self.charges = atoms.get_array("initial_charges")
            self.initial_charges = self.charges.copy()
            self.properties_changed += ["initial_charges"]

    # -- evaluation ----------------------------------------------------------
    def calculate(self, atoms, properties=("energy",), system_changes=None):
        self.atoms = atoms
        self.update(atoms)
        nat = len(self.particles)
        epot = 0.0
        forces = np.zeros((nat, 3))
        wpot = np.zeros((3, 3))
        self.results = {}
        compute_epot_per_at = "energies" in properties
        compute_wpot_per_at = "stresses" in properties
        kw = dict(epot_per_at=compute_epot_per_at, wpot_per_at=compute_wpot_per_at,
                  epot_per_bond=self.compute_epot_per_bond, f_per_bond=self.compute_f_per_bond,
                  wpot_per_bond=self.compute_wpot_per_bond)
        if self.mask is not None:
            kw["mask"] = np.asarray(self.mask, dtype=np.int32)
        epot_per_at = np.zeros(nat) if compute_epot_per_at else None
        wpot_per_at = np.zeros((nat, 3, 3)) if compute_wpot_per_at else None
        for pot in self.pots:
            # the per-bond arrays are those of the last potential, as in the
            # reference (aseinterface.py:380-392)
            _e, _f, _w, epa, self.epot_per_bond, self.f_per_bond, wpa, self.wpot_per_bond = \
                pot.energy_and_forces(self.particles, self.nl, forces=forces, **kw)
            epot += _e
            wpot += _w
            if compute_epot_per_at:
                epot_per_at += epa
            if compute_wpot_per_at:
                wpot_per_at += wpa
        if self.couls:
            conv = HARTREE * BOHR
            ec = 0.0
            fc = np.zeros((nat, 3))
            wc = np.zeros((3, 3))
            for coul in self.couls:
                _e, _f, _w = coul.energy_and_forces(self.particles, self.nl, self.charges, fc)
                ec += _e
                wc += _w
            epot += ec * conv
            wpot += wc * conv
            forces += fc * conv
            self.results["charges"] = self.charges
        self.results["energy"] = epot
        self.results["free_energy"] = epot
        volume = atoms.get_volume()
        self.results["stress"] = np.array([wpot[0, 0], wpot[1, 1], wpot[2, 2],
                                           (wpot[1, 2] + wpot[2, 1]) / 2,
                                           (wpot[0, 2] + wpot[2, 0]) / 2,
                                           (wpot[0, 1] + wpot[1, 0]) / 2]) / volume
        self.results["forces"] = forces
        if compute_epot_per_at:
            self.results["energies"] = epot_per_at
        if compute_wpot_per_at:
            w = wpot_per_at
            self.results["stresses"] = np.transpose([w[:, 0, 0], w[:, 1, 1], w[:, 2, 2],
                                                     (w[:, 1, 2] + w[:, 2, 1]) / 2,
                                                     (w[:, 0, 2] + w[:, 2, 0]) / 2,
                                                     (w[:, 0, 1] + w[:, 1, 0]) / 2])
        self.properties_changed = []
        self._last = properties

    def _get(self, atoms, name, props):
        self.calculate(atoms, props)
        return self.results[name]

    def get_potential_energy(self, atoms=None):
        return self._get(atoms or self.atoms, "energy", ("energy",))

    def get_forces(self, atoms=None):
        return self._get(atoms or self.atoms, "forces", ("energy", "forces"))

    def get_stress(self, atoms=None):
        return self._get(atoms or self.atoms, "stress", ("energy", "stress"))

    def get_potential_energies(self, atoms=None):
        return self._get(atoms or self.atoms, "energies", ("energies",))

    def get_stresses(self, atoms=None):
        return self._get(atoms or self.atoms, "stresses", ("stresses",))

    def get_electrostatic_potential(self, atoms=None):
        if atoms is not None:
            self.update(atoms)
        phi = np.zeros(len(self.particles))
        for coul in self.couls:
            coul.potential(self.particles, self.nl, self.charges, phi)
        return phi * HARTREE * BOHR

    def get_neighbors(self):
        return self.nl.get_neighbors(self.particles)


def _make(cls, name):
    return type(name, (Atomistica,), {"potential_class": cls, "name": name})


# one calculator class per GPU potential class (aseinterface.py:493-505)
Tersoff = _make(native.Tersoff, "Tersoff")
TabulatedAlloyEAM = _make(native.TabulatedAlloyEAM, "TabulatedAlloyEAM")
TabulatedEAM = _make(native.TabulatedEAM, "TabulatedEAM")
Ewald = _make(native.Ewald, "Ewald")
Rebo2 = _make(native.Rebo2, "Rebo2")
Kumagai = _make(native.Kumagai, "Kumagai")
Brenner = _make(native.Brenner, "Brenner")
TersoffScr = _make(native.TersoffScr, "TersoffScr")
KumagaiScr = _make(native.KumagaiScr, "KumagaiScr")
BrennerScr = _make(native.BrennerScr, "BrennerScr")
JuslinScr = _make(native.JuslinScr, "JuslinScr")
PME = _make(native.PME, "PME")
Juslin = _make(native.Juslin, "Juslin")
