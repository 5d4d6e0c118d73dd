"""ASE calculator over the engine (python/gfn1_rs/ase.py, GFN1RSCalculator).

The ASE boundary: positions in Angstrom, energies in eV, forces in eV/Angstrom,
dipoles in e*Angstrom, Hessians in eV/Angstrom^2, every factor taken from
``ase.units`` (ase.py:38-57); the engine underneath speaks atomic units.
Properties energy, forces, charges, dipole for molecules and for fully
periodic cells (the Gamma-point path; a cell's dipole is the reference-cell
Mulliken dipole, as in the reference; no stress); for molecules also the
native L-BFGS (``optimize_native``), the Hessian, the harmonic frequencies,
the dipole derivatives and the IR spectrum.
Importing this module needs ASE, as the reference's does.
"""
from typing import Any

import numpy as np
from ase.calculators.calculator import Calculator, all_changes
from ase.units import Bohr, Hartree, invcm

from .api import FORCE_AU_TO_EV_PER_ANGSTROM, Gfn1NativeCalculator

# Every conversion at the ASE boundary is a product of ase.units values, so
# results agree with whatever constants the installed ASE uses.
_ENERGY = Hartree                      # eV per Eh
_LENGTH = Bohr                         # Angstrom per bohr
_FORCE = Hartree / Bohr                # eV/Angstrom per Eh/bohr
_HESSIAN = Hartree / Bohr ** 2         # eV/Angstrom^2 per Eh/bohr^2
_THIRD = Hartree / Bohr ** 3
_FOURTH = Hartree / Bohr ** 4
_DIPOLE = Bohr                         # e Angstrom per e bohr
_POLARIZABILITY = Bohr ** 2 / Hartree
_WAVENUMBER = invcm                    # eV per cm^-1
# The native layer's own eV/Angstrom force unit (its CODATA constants, not
# ASE's); only used to undo it when a result carries forces but no gradient.
_NATIVE_FORCE_FACTOR = FORCE_AU_TO_EV_PER_ANGSTROM


def _forces_hartree_per_bohr(result):
    """Forces in Eh/bohr from a native result: the negated analytic gradient
    when there is one, otherwise the native eV/Angstrom forces taken back to
    atomic units with the native factor; None when the result has neither."""
    gradient = getattr(result, "gradient_hartree_per_bohr", None)
    if gradient is not None:
        return -np.asarray(gradient, dtype=float)
    native_forces = getattr(result, "forces_ev_per_angstrom", None)
    if native_forces is None:
        return None
    return np.asarray(native_forces, dtype=float) / _NATIVE_FORCE_FACTOR


def _origin_bohr(origin_angstrom):
    """A caller's origin (Angstrom, ASE's length unit) in bohr, or None."""
    if origin_angstrom is None:
        return None
    return np.asarray(origin_angstrom, dtype=float) / _LENGTH


def _to_ase_gradient_dict(raw, energy_keys=()):
    """ASE units for a dictionary of atomic-unit results: for each (source,
    target) pair in energy_keys, target = source in eV; "gradient" and
    "forces" become eV/Angstrom and keep their atomic-unit arrays under a
    "_hartree_per_bohr" suffix.  Other entries pass through unchanged."""
    out = dict(raw)
    for source, target in energy_keys:
        if source in out:
            out[target] = float(out[source]) * _ENERGY
    for key in ("gradient", "forces"):
        if out.get(key) is not None:
            atomic = np.asarray(out[key], dtype=float)
            out[key + "_hartree_per_bohr"] = atomic
            out[key] = atomic * _FORCE
    return out


class GFN1AMDCalculator(Calculator):
    """GFN1RSCalculator (ase.py:115-560): single points of molecules and of
    fully periodic cells at the Gamma point, the other methods for molecules."""

    implemented_properties = ["energy", "forces", "charges", "dipole"]
    default_parameters = {
        "param_path": None, "charge": 0.0, "multiplicity": None, "max_scc": 250, "energy_tolerance": 1.0e-6,
        "charge_tolerance": 2.0e-5, "mixing": 0.4, "scc_broyden": True, "scc_broyden_size": 250,
        "electronic_temperature": 300.0, "enable_dispersion": True, "enable_cn_hamiltonian": True,
        "level_shift": 0.0,
    }

    def __init__(self, restart=None, label=None, atoms=None, **kwargs: Any):
        Calculator.__init__(self, restart=restart, label=label, atoms=atoms, **kwargs)
        self._native = self._make_native_calculator()

    def set(self, **kwargs: Any):
        changed = Calculator.set(self, **kwargs)
        if changed:
            self._native = self._make_native_calculator()
        return changed

    def _make_native_calculator(self):
        p = self.parameters
        return Gfn1NativeCalculator(
            param_path=p.param_path, charge=float(p.charge), multiplicity=p.multiplicity, max_scc=int(p.max_scc),
            energy_tolerance=float(p.energy_tolerance), charge_tolerance=float(p.charge_tolerance),
            mixing=float(p.mixing), scc_broyden=bool(p.scc_broyden), scc_broyden_size=int(p.scc_broyden_size),
            electronic_temperature=float(p.electronic_temperature), enable_dispersion=bool(p.enable_dispersion),
            enable_cn_hamiltonian=bool(p.enable_cn_hamiltonian), level_shift=float(p.level_shift))

    @staticmethod
    def _molecule(atoms):
        if atoms is None:
            raise RuntimeError("no ASE Atoms object supplied or attached")
        if any(atoms.get_pbc()):
            raise NotImplementedError("this method is built for molecules only; periodic cells have single points")
        return atoms.get_atomic_numbers().tolist(), np.asarray(atoms.get_positions(), dtype=float)

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        """Single point in ASE units (ase.py:358-434), plus the native
        passthroughs native_energy_terms_hartree / _ev, native_dipole_au,
        native_forces_hartree_per_bohr, native_converged, native_iterations.
        Atoms periodic in every direction go to the native periodic single
        point (Gamma point, cell in Angstrom); partly periodic ones raise."""
        Calculator.calculate(self, atoms, properties, system_changes)
        if self.atoms is None:
            raise RuntimeError("no ASE Atoms object supplied or attached")
        want_forces = "forces" in set(properties)
        pbc = [bool(p) for p in self.atoms.get_pbc()]
        numbers = self.atoms.get_atomic_numbers().tolist()
        positions = np.asarray(self.atoms.get_positions(), dtype=float)
        if any(pbc):
            if not all(pbc):
                raise NotImplementedError("only cells periodic in all three directions are supported")
            result = self._native.calculate_periodic(numbers, positions, np.asarray(self.atoms.get_cell(), dtype=float),
                                                     unit="angstrom", compute_gradient=want_forces)
        else:
            result = self._native.calculate(numbers, positions, unit="angstrom", compute_gradient=want_forces)
        terms_au = dict(result.energy_terms_hartree())
        dipole_au = np.asarray(result.dipole, dtype=float)
        update = {
            "energy": float(result.energy_hartree) * _ENERGY,
            "charges": np.asarray(result.charges, dtype=float),
            "dipole": dipole_au * _DIPOLE,
            "native_dipole_au": dipole_au,
            "native_energy_terms_hartree": terms_au,
            "native_energy_terms_ev": {name: value * _ENERGY for name, value in terms_au.items()},
            "native_converged": bool(result.converged),
            "native_iterations": int(result.iterations),
        }
        forces_au = _forces_hartree_per_bohr(result)
        if forces_au is not None:
            update["native_forces_hartree_per_bohr"] = forces_au
            update["forces"] = forces_au * _FORCE
        self.results.update(update)

    def optimize_native(self, atoms=None, *, max_iterations=250, gradient_tolerance=1.0e-4, step_tolerance=1.0e-7,
                        history=12, max_atom_step=0.30):
        """The engine's L-BFGS (ase.py:436-489): thresholds in atomic units,
        atoms updated in Angstrom; returns the native OptimizationResult."""
        atoms = self.atoms if atoms is None else atoms
        numbers, positions = self._molecule(atoms)
        result = self._native.optimize(numbers, positions, unit="angstrom", max_iterations=int(max_iterations),
                                       gradient_tolerance=float(gradient_tolerance),
                                       step_tolerance=float(step_tolerance), history=int(history),
                                       max_atom_step=float(max_atom_step))
        atoms.set_positions(np.asarray(result.positions_angstrom, dtype=float))
        self.reset()
        return result

    def get_dipole_au(self, atoms=None):
        """Mulliken dipole in e*bohr (ase.py:571-580)."""
        numbers, positions = self._molecule(atoms if atoms is not None else self.atoms)
        return np.asarray(self._native.calculate(numbers, positions, unit="angstrom").dipole, dtype=float)

    def get_hessian(self, atoms=None):
        """(3N, 3N) Hessian in eV/Angstrom^2 (ase.py:641-671)."""
        numbers, positions = self._molecule(atoms if atoms is not None else self.atoms)
        return np.asarray(self._native.hessian(numbers, positions, unit="angstrom"), dtype=float) * _HESSIAN

    def get_dipole_derivatives(self, atoms=None, origin=None):
        """dipole (e*Angstrom), dipole_au (e*bohr) and ddipole_dr ([coord][alpha],
        e, the same number in either unit system) (ase.py:603-622); origin in
        Angstrom."""
        numbers, positions = self._molecule(atoms if atoms is not None else self.atoms)
        o = _origin_bohr(origin)
        out = self._native.dipole_derivatives(numbers, positions, unit="angstrom",
                                              origin=None if o is None else tuple(o))
        dipole_au = np.asarray(out["dipole"], dtype=float)
        out["dipole_au"] = dipole_au
        out["dipole"] = dipole_au * _DIPOLE
        out["ddipole_dr"] = np.asarray(out["ddipole_dr"], dtype=float)
        return out

    def get_ir_spectrum(self, atoms=None, origin=None):
        """wavenumbers (cm^-1), intensities_km_per_mol, and the atomic-unit
        intensities_au and dipole_gradients (ase.py:624-639)."""
        numbers, positions = self._molecule(atoms if atoms is not None else self.atoms)
        return self._native.ir_spectrum(numbers, positions, unit="angstrom")

    def get_vibrational_frequencies(self, atoms=None):
        """wavenumbers (cm^-1), energies_ev and modes (ase.py:673-687)."""
        numbers, positions = self._molecule(atoms if atoms is not None else self.atoms)
        out = self._native.vibrational_frequencies(numbers, positions, unit="angstrom")
        out["energies_ev"] = np.asarray(out["wavenumbers"], dtype=float) * _WAVENUMBER
        return out


# the reference's class name, for code written against it
GFN1RSCalculator = GFN1AMDCalculator
