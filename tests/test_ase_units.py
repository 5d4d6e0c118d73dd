"""The ASE unit contract of gfn1_amd.ase with the native layer mocked out
(CPU only), after the reference's pure-numpy conversion tests
(tests/python/test_ase_units.py:442-600).

The fake native result advertises the native CODATA energy factor on purpose,
so a conversion that used it instead of ase.units would show.  Without ASE
installed the calculator runs against the stand-in of tests/ase_stand_in.py
(ASE's CODATA 2014 constants); with ASE the real package is used.  The
periodic stress property of the reference's calculator is out of scope here.
"""
import sys

import numpy as np
import pytest

from ase_stand_in import install_stand_in

NATIVE_HARTREE_TO_EV = 27.211386245988


class _FakeResult:
    """Duck type of the native CalculationResult (python.rs:2845-2935)."""

    def __init__(self, *, energy_hartree, gradient=None, forces_ev=None):
        self.energy_hartree = energy_hartree
        self.energy_ev = energy_hartree * NATIVE_HARTREE_TO_EV
        self.gradient_hartree_per_bohr = gradient
        self.forces_ev_per_angstrom = forces_ev
        self.charges = [0.25, -0.25]
        self.dipole = (0.1, -0.2, 0.35)
        self.iterations = 7
        self.converged = True

    def energy_terms_hartree(self):
        return {"total_internal": self.energy_hartree, "repulsion": 0.125}


@pytest.fixture(scope="module")
def ase_mod():
    stand_in = install_stand_in()
    sys.modules.pop("gfn1_amd.ase", None)
    from gfn1_amd import ase as gase
    yield gase
    sys.modules.pop("gfn1_amd.ase", None)
    if stand_in:
        for k in ("ase", "ase.calculators", "ase.calculators.calculator", "ase.units"):
            sys.modules.pop(k, None)


def units():
    from ase.units import Bohr, Hartree, invcm
    return Bohr, Hartree, invcm


def test_constants_are_exactly_the_ase_units_ones(ase_mod):
    Bohr, Hartree, invcm = units()
    assert ase_mod._ENERGY == Hartree
    assert ase_mod._LENGTH == Bohr
    assert ase_mod._FORCE == Hartree / Bohr
    assert ase_mod._HESSIAN == Hartree / Bohr ** 2
    assert ase_mod._THIRD == Hartree / Bohr ** 3
    assert ase_mod._FOURTH == Hartree / Bohr ** 4
    assert ase_mod._DIPOLE == Bohr
    assert ase_mod._POLARIZABILITY == Bohr ** 2 / Hartree
    assert ase_mod._WAVENUMBER == invcm


def test_no_hand_typed_conversion_literals_in_the_module(ase_mod):
    with open(ase_mod.__file__) as f:
        code = "\n".join(line for line in f.read().splitlines() if not line.lstrip().startswith("#"))
    for literal in ("27.211", "0.52917", "1.88972", "1822.88"):
        assert literal not in code, literal


def test_forces_helper_prefers_the_analytic_gradient(ase_mod):
    gradient = [[0.1, -0.2, 0.3], [-0.1, 0.2, -0.3]]
    result = _FakeResult(energy_hartree=-5.0, gradient=gradient, forces_ev=[[9.0, 9.0, 9.0]] * 2)
    np.testing.assert_allclose(ase_mod._forces_hartree_per_bohr(result), -np.asarray(gradient))


def test_forces_helper_unscales_the_native_ev_forces(ase_mod):
    forces_au = np.array([[0.01, -0.02, 0.03], [-0.01, 0.02, -0.03]])
    result = _FakeResult(energy_hartree=-5.0, forces_ev=(forces_au * ase_mod._NATIVE_FORCE_FACTOR).tolist())
    np.testing.assert_allclose(ase_mod._forces_hartree_per_bohr(result), forces_au, rtol=1e-13)


def test_forces_helper_returns_none_without_a_gradient(ase_mod):
    assert ase_mod._forces_hartree_per_bohr(_FakeResult(energy_hartree=-5.0)) is None


def test_origin_helper_converts_angstrom_to_bohr(ase_mod):
    Bohr, _, _ = units()
    assert ase_mod._origin_bohr(None) is None
    np.testing.assert_allclose(ase_mod._origin_bohr((1.0, -2.0, 3.5)), np.array([1.0, -2.0, 3.5]) / Bohr,
                               rtol=1e-14)


def test_gradient_dict_helper_applies_the_suffix_rule(ase_mod):
    Bohr, Hartree, _ = units()
    raw = {"energy_hartree": -3.5, "gradient": [[1.0, 2.0, 3.0]], "forces": [[-1.0, -2.0, -3.0]]}
    out = ase_mod._to_ase_gradient_dict(dict(raw), energy_keys=(("energy_hartree", "energy"),))
    assert out["energy"] == pytest.approx(-3.5 * Hartree, rel=1e-14)
    assert out["energy_hartree"] == -3.5
    np.testing.assert_allclose(out["gradient"], np.asarray(raw["gradient"]) * (Hartree / Bohr))
    np.testing.assert_allclose(out["gradient_hartree_per_bohr"], np.asarray(raw["gradient"]))
    np.testing.assert_allclose(out["forces"], np.asarray(raw["forces"]) * (Hartree / Bohr))
    np.testing.assert_allclose(out["forces_hartree_per_bohr"], np.asarray(raw["forces"]))


def test_gradient_dict_helper_is_a_noop_without_those_keys(ase_mod):
    assert ase_mod._to_ase_gradient_dict({"wavenumbers": [1.0, 2.0]}) == {"wavenumbers": [1.0, 2.0]}


def test_properties_are_advertised(ase_mod):
    for name in ("energy", "forces", "charges", "dipole"):
        assert name in ase_mod.GFN1RSCalculator.implemented_properties


class _Atoms:
    def __init__(self, numbers, positions):
        self.numbers = np.asarray(numbers)
        self.positions = np.asarray(positions, dtype=float)

    def get_atomic_numbers(self):
        return self.numbers

    def get_positions(self):
        return self.positions.copy()

    def get_pbc(self):
        return np.array([False] * 3)


class _FakeNative:
    """Records the call and returns a fixed atomic-unit result."""

    def __init__(self, result):
        self.result = result
        self.calls = []

    def calculate(self, numbers, positions, unit="angstrom", compute_gradient=False):
        self.calls.append((list(numbers), np.asarray(positions), unit, compute_gradient))
        return self.result

    def calculate_periodic(self, numbers, positions, cell, unit="angstrom", compute_gradient=False):
        self.calls.append((list(numbers), np.asarray(positions), unit, compute_gradient, np.asarray(cell)))
        return self.result


def test_calculate_converts_every_property_with_ase_units(ase_mod):
    """calculate() (ase.py:358-434) against the mocked native layer: energy,
    forces, charges and dipole in ASE units, the atomic-unit passthroughs
    unchanged, positions handed over in Angstrom."""
    Bohr, Hartree, _ = units()
    gradient = np.array([[0.01, 0.0, -0.02], [-0.01, 0.0, 0.02]])
    fake = _FakeNative(_FakeResult(energy_hartree=-5.25, gradient=gradient))
    calc = ase_mod.GFN1RSCalculator()
    calc._native = fake
    atoms = _Atoms([1, 1], [[0.0, 0.0, 0.0], [0.74, 0.0, 0.0]])
    calc.calculate(atoms, properties=("energy", "forces"))
    numbers, pos, unit, grad = fake.calls[-1]
    assert numbers == [1, 1] and unit == "angstrom" and grad is True
    np.testing.assert_array_equal(pos, atoms.positions)
    r = calc.results
    assert r["energy"] == -5.25 * Hartree
    np.testing.assert_allclose(r["forces"], -gradient * Hartree / Bohr, rtol=1e-15)
    np.testing.assert_allclose(r["native_forces_hartree_per_bohr"], -gradient)
    np.testing.assert_allclose(r["dipole"], np.array([0.1, -0.2, 0.35]) * Bohr, rtol=1e-15)
    np.testing.assert_allclose(r["native_dipole_au"], [0.1, -0.2, 0.35])
    np.testing.assert_allclose(r["charges"], [0.25, -0.25])
    assert r["native_energy_terms_hartree"]["repulsion"] == 0.125
    assert r["native_energy_terms_ev"]["repulsion"] == 0.125 * Hartree
    assert r["native_iterations"] == 7 and r["native_converged"] is True
    # energy only: no gradient requested, no forces reported
    calc.results = {}
    fake.result = _FakeResult(energy_hartree=-5.25)
    calc.calculate(atoms, properties=("energy",))
    assert fake.calls[-1][3] is False and "forces" not in calc.results


class _PeriodicAtoms(_Atoms):
    def __init__(self, numbers, positions, cell, pbc):
        super().__init__(numbers, positions)
        self.cell = np.asarray(cell, dtype=float)
        self.pbc = np.asarray(pbc)

    def get_pbc(self):
        return self.pbc

    def get_cell(self):
        return self.cell


def test_periodic_atoms_go_to_the_periodic_single_point(ase_mod):
    """A cell periodic in every direction is handed to calculate_periodic with
    its cell in Angstrom, and converted like a molecule's result; a partly
    periodic one is refused."""
    Bohr, Hartree, _ = units()
    gradient = np.array([[0.01, 0.0, -0.02], [-0.01, 0.0, 0.02]])
    fake = _FakeNative(_FakeResult(energy_hartree=-5.25, gradient=gradient))
    calc = ase_mod.GFN1RSCalculator()
    calc._native = fake
    cell = np.diag([5.0, 6.0, 7.0])
    atoms = _PeriodicAtoms([1, 1], [[0.0, 0.0, 0.0], [0.74, 0.0, 0.0]], cell, [True, True, True])
    calc.calculate(atoms, properties=("energy", "forces"))
    numbers, pos, unit, grad, c = fake.calls[-1]
    assert numbers == [1, 1] and unit == "angstrom" and grad is True
    np.testing.assert_array_equal(c, cell)
    assert calc.results["energy"] == -5.25 * Hartree
    np.testing.assert_allclose(calc.results["forces"], -gradient * Hartree / Bohr, rtol=1e-15)
    with pytest.raises(NotImplementedError):
        calc.calculate(_PeriodicAtoms([1, 1], [[0, 0, 0], [0.74, 0, 0]], cell, [True, False, True]))
