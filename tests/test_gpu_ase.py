"""The ASE calculator (gfn1_amd.ase) over the engine: unit conversions at the
ASE boundary (python/gfn1_rs/ase.py:38-57, tests/python/test_ase_units.py) and
the native optimiser / Hessian passthroughs.

ASE is not installed in this environment.  When it is missing, a minimal
stand-in for the two ASE modules the calculator imports is put in place: the
Calculator base (parameters, results, reset) and ase.units with ASE's own
CODATA 2014 values.  With ASE installed the real package is used.  The
comparison with ASE proper is therefore unpinned here.
"""
import sys
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from ase_stand_in import ASE_BOHR, ASE_HARTREE, ASE_INVCM, install_stand_in as _install_stand_in  # noqa: F401,E402


class _Atoms:
    """The few Atoms methods the calculator calls."""

    def __init__(self, numbers, positions, pbc=False, cell=None):
        self.numbers = np.asarray(numbers)
        self.positions = np.asarray(positions, dtype=float)
        self.pbc = np.array(pbc if isinstance(pbc, (list, tuple)) else [pbc] * 3)
        self.cell = np.zeros((3, 3)) if cell is None else np.asarray(cell, dtype=float)

    def get_atomic_numbers(self):
        return self.numbers

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, p):
        self.positions = np.asarray(p, dtype=float)

    def get_pbc(self):
        return self.pbc

    def get_cell(self):
        return self.cell.copy()


WATER = ([8, 1, 1], [[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]])


@pytest.fixture(scope="module")
def ase_calc():
    stand_in = _install_stand_in()
    from gfn1_amd import ase as gase
    yield gase, stand_in
    if stand_in:
        for k in ("ase", "ase.calculators", "ase.calculators.calculator", "ase.units"):
            sys.modules.pop(k, None)


def test_single_point_units(ase_calc):
    gase, _ = ase_calc
    from gfn1_amd import ANGSTROM_TO_BOHR, analytic_gradient
    from ase.units import Bohr, Hartree
    calc = gase.GFN1RSCalculator()
    atoms = _Atoms(*WATER)
    calc.calculate(atoms, properties=("energy", "forces"))
    ref = analytic_gradient(WATER[0], np.asarray(WATER[1]) * ANGSTROM_TO_BOHR)
    assert abs(calc.results["energy"] - ref.total_energy * Hartree) < 1e-9
    assert np.allclose(calc.results["forces"], -ref.gradient * Hartree / Bohr, rtol=0, atol=1e-9)
    assert np.allclose(calc.results["native_forces_hartree_per_bohr"], -ref.gradient, rtol=0, atol=1e-12)
    assert np.allclose(calc.results["dipole"], ref.electronic_result.dipole * Bohr, rtol=0, atol=1e-12)
    assert np.allclose(calc.results["charges"], ref.electronic_result.atomic_charges, rtol=0, atol=1e-12)
    assert abs(sum(calc.results["charges"])) < 1e-8
    # a fully periodic cell: the Gamma-point single point, cell in Angstrom
    cell = np.diag([6.0, 6.5, 7.0])
    calc.calculate(_Atoms(*WATER, pbc=True, cell=cell), properties=("energy", "forces"))
    pref = analytic_gradient(WATER[0], np.asarray(WATER[1]) * ANGSTROM_TO_BOHR, lattice=cell * ANGSTROM_TO_BOHR)
    assert abs(calc.results["energy"] - pref.total_energy * Hartree) < 1e-9
    assert np.allclose(calc.results["forces"], -pref.gradient * Hartree / Bohr, rtol=0, atol=1e-9)
    assert abs(pref.total_energy - ref.total_energy) > 1e-6  # the images matter
    with pytest.raises(NotImplementedError):
        calc.calculate(_Atoms(*WATER, pbc=[True, True, False], cell=cell))
    with pytest.raises(NotImplementedError):
        calc.optimize_native(_Atoms(*WATER, pbc=True, cell=cell))


def test_optimize_hessian_frequencies(ase_calc):
    gase, _ = ase_calc
    from ase.units import Bohr, Hartree
    calc = gase.GFN1RSCalculator()
    atoms = _Atoms(*WATER)
    e0 = calc._native.calculate(*WATER).energy_hartree
    res = calc.optimize_native(atoms, gradient_tolerance=1e-6, max_iterations=200)
    assert res.converged and res.energy_hartree < e0
    assert np.allclose(atoms.get_positions(), res.positions_angstrom)
    h = calc.get_hessian(atoms)
    h_au = calc._native.hessian(atoms.get_atomic_numbers(), atoms.get_positions())
    assert np.allclose(h, h_au * Hartree / Bohr ** 2, rtol=1e-14, atol=0)
    out = calc.get_vibrational_frequencies(atoms)
    w = np.asarray(out["wavenumbers"])
    assert np.all(np.abs(w[:6]) < 30.0) and 1300.0 < w[6] < 1800.0 and 3000.0 < w[7] < 4000.0
    assert np.allclose(out["energies_ev"], w * ASE_INVCM)
    dd = calc.get_dipole_derivatives(atoms)
    assert np.allclose(dd["dipole"], dd["dipole_au"] * Bohr, rtol=1e-14, atol=0)
    ir = calc.get_ir_spectrum(atoms)
    assert np.all(np.asarray(ir["intensities_km_per_mol"])[6:] > 1.0)

