"""Cartesian Hessians on the device (through the C ABI) against the oracle's
finite-difference Hessian (oracle/hessian_oracle.py), same step and SCC
tolerances.  Bar: max |dH| <= 1e-6 Eh/bohr^2, the bar at which the reference
accepts its analytic Hessian against this construction
(tests/hessian.rs:691-760); 1e-9 for the repulsion, dispersion and halogen
parts of the breakdown, differences of gradients that do not depend on the SCC.
"""
import sys

import numpy as np
import pytest

import hessian_oracle
import optimizer_oracle
from gfn1_amd import (ElectronicOptions, GeometryOptimizationOptions, Gfn1NativeCalculator, HessianOptions,
                      NativeError, geometry, hessian, hessian_batch, optimize_geometry, vibrational_analysis)
from gfn1_amd import AnalyticGradientOptions, system_from_xyz

import reference_molecules

pytestmark = pytest.mark.gpu

H_TOL = 1e-6
A2B = geometry.ANGSTROM_TO_BOHR


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _systems():
    out = {"h2": ([1, 1], np.array([[0.0, 0.0, 0.0], [1.45, 0.1, 0.0]]))}
    for k, name in enumerate(("water", "methane", "ammonia")):
        z, x = geometry.MONOMERS[name]
        out[name] = geometry.perturbed(z, x * A2B, 21 + k, 0.03)
    return out


@pytest.mark.parametrize("name", ["h2", "water", "methane", "ammonia"])
def test_hessian_matches_oracle(name):
    z, x = _systems()[name]
    note("case hessian %s" % name)
    r = hessian(z, x)
    want, asym = hessian_oracle.fd_hessian(z, x)
    d = float(np.abs(r.hessian - want).max())
    note("%s: max|dH| %.2e, asymmetry device %.2e oracle %.2e" % (name, d, r.max_asymmetry, asym))
    assert d <= H_TOL
    assert np.array_equal(r.hessian, r.hessian.T)
    assert r.max_asymmetry < 1e-7


@pytest.mark.parametrize("name", sorted(reference_molecules.HESSIAN_MOLECULES))
def test_reference_hessian_molecules_match_oracle(name):
    # the other molecules of the reference's Hessian tests (tests/hessian.rs):
    # HF, chloromethanol and bromoethanol (halogens), Ni(CO)4 (a 3d metal)
    note("case hessian reference molecule %s" % name)
    z, x, _ = system_from_xyz(reference_molecules.HESSIAN_MOLECULES[name])
    r = hessian(z, x)
    want, asym = hessian_oracle.fd_hessian(z, x)
    d = float(np.abs(r.hessian - want).max())
    note("%s: max|dH| %.2e, asymmetry device %.2e oracle %.2e" % (name, d, r.max_asymmetry, asym))
    assert d <= H_TOL
    nat = len(z)
    assert np.abs(r.hessian.reshape(3 * nat, nat, 3).sum(axis=1)).max() < 1e-5


def test_batch_matches_single_and_reports_failures():
    systems = _systems()
    names = ["water", "methane", "ammonia"]
    batch = [systems[n] for n in names] + [([], np.zeros((0, 3))), ([1, 1], np.array([[0, 0, 0], [1.4, 0, 0]]), 10.0)]
    note("case hessian batch (max_batch 10)")
    hs, status = hessian_batch(batch, options=HessianOptions(max_batch=10))
    assert status[:3] == [0, 0, 0]
    assert status[3] == 3 and status[4] == 3
    assert hs[3] is None and hs[4] is None
    for n, h in zip(names, hs):
        single = hessian(*systems[n]).hessian
        assert np.abs(h - single).max() <= 1e-9, n
    with pytest.raises(NativeError) as e:
        hessian([], np.zeros((0, 3)))
    assert e.value.kind == "InvalidInput"


def test_term_selection_matches_oracle():
    z, x = _systems()["water"]
    note("case hessian terms")
    for terms in (dict(include_electronic=False, include_dispersion=False, include_halogen=False),
                  dict(include_repulsion=False, include_electronic=False),
                  dict(include_repulsion=False, include_dispersion=False)):
        r = hessian(z, x, options=HessianOptions(**terms))
        want, _ = hessian_oracle.fd_hessian(z, x, **terms)
        assert np.abs(r.hessian - want).max() <= H_TOL, terms


def test_hessian_breakdown_matches_oracle_terms():
    # the repulsion / dispersion / halogen parts of a halogen-bonded complex
    # (bromomethane and ammonia, all three classical terms nonzero) against the
    # oracle's Hessians of each term alone; the electronic rest against the
    # electronic term alone; one displaced set of evaluations on the device
    z = np.array([6, 35, 1, 1, 1, 7, 1, 1, 1])
    x = np.array([[0, 0, 0], [0, 0, 1.94], [1.03, 0, -0.36], [-0.51, 0.89, -0.36], [-0.51, -0.89, -0.36],
                  [0, 0, 4.9], [0.94, 0, 5.25], [-0.47, 0.81, 5.25], [-0.47, -0.81, 5.25]]) * A2B
    note("case hessian breakdown (CH3Br...NH3)")
    r = hessian(z, x, terms=True)
    full, _ = hessian_oracle.fd_hessian(z, x)
    assert np.abs(r.hessian - full).max() <= H_TOL
    only = dict(include_repulsion=False, include_electronic=False, include_dispersion=False, include_halogen=False)
    for part, flag in (("repulsion", "include_repulsion"), ("dispersion", "include_dispersion"),
                       ("halogen", "include_halogen"), ("electronic", "include_electronic")):
        want, _ = hessian_oracle.fd_hessian(z, x, **dict(only, **{flag: True}))
        got = getattr(r, part)
        d = float(np.abs(got - want).max())
        note("%s: max|dH| %.2e, largest element %.2e" % (part, d, float(np.abs(want).max())))
        # the classical parts are differences of SCC-independent gradients
        assert d <= (H_TOL if part == "electronic" else 1e-9), part
        assert np.abs(want).max() > 1e-5, part  # each term contributes here
    # a term switched off contributes nothing to the breakdown
    r2 = hessian(z, x, options=HessianOptions(include_halogen=False), terms=True)
    assert np.abs(r2.halogen).max() == 0.0
    assert np.abs(r2.repulsion - r.repulsion).max() <= 1e-9


def test_warm_and_cold_starts_agree():
    z, x = _systems()["methane"]
    note("case hessian warm / cold start")
    warm = hessian(z, x).hessian
    cold = hessian(z, x, options=HessianOptions(warm_start=False, scc_charge_tolerance=1e-9)).hessian
    assert np.abs(cold - warm).max() <= 1e-7
    want, _ = hessian_oracle.fd_hessian(z, x, warm_start=False, charge_tolerance=1e-9)
    assert np.abs(cold - want).max() <= H_TOL


def test_relaxed_water_frequencies_match_oracle():
    z, x = _systems()["water"]
    note("case relaxed water frequencies")
    tight = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-11, charge_tolerance=1e-9))
    r = optimize_geometry(z, x, options=GeometryOptimizationOptions(gradient_tolerance=1e-6, max_iterations=200),
                          gradient_options=tight)
    want = optimizer_oracle.optimize_gfn1(z, x, scc_options=dict(energy_tolerance=1e-11, charge_tolerance=1e-9),
                                          gradient_tolerance=1e-6, max_iterations=200)
    assert r.converged and want["converged"]
    assert np.abs(r.positions - want["positions"]).max() <= 1e-6
    w = vibrational_analysis(hessian(z, r.positions).hessian, z).wavenumbers
    w_ref = vibrational_analysis(hessian_oracle.fd_hessian(z, want["positions"])[0], z).wavenumbers
    note("water wavenumbers device %s oracle %s" % (np.round(w[6:], 3), np.round(w_ref[6:], 3)))
    assert np.abs(w[6:] - w_ref[6:]).max() < 0.01


def _hydronium():
    import math
    pts = [[0.0, 0.0, 0.0]]
    for k in range(3):
        phi = 2.0 * math.pi * k / 3.0
        pts.append([0.95 * math.cos(phi), 0.95 * math.sin(phi), 0.30])
    return [8, 1, 1, 1], np.array(pts) * A2B


@pytest.mark.parametrize("name", ["water", "methane", "ammonia", "hydronium"])
def test_dipole_derivatives_match_oracle(name):
    if name == "hydronium":
        (z, x), charge = _hydronium(), 1.0
    else:
        (z, x), charge = _systems()[name], 0.0
    note("case dipole derivatives %s" % name)
    r = hessian(z, x, charge=charge)
    _, _, ddip, mu = hessian_oracle.fd_hessian_dipoles(z, x, charge=charge)
    d = float(np.abs(r.dipole_derivatives - ddip).max())
    note("%s: max|d ddip| %.2e, |d mu| %.2e" % (name, d, float(np.abs(r.dipole - mu).max())))
    assert d <= 1e-6
    assert np.abs(r.dipole - mu).max() <= 1e-8
    born = r.dipole_derivatives.reshape(len(z), 3, 3)
    assert np.abs(born.sum(axis=0) - charge * np.eye(3)).max() < 1e-6


def test_calculator_ir_spectrum():
    from gfn1_amd.api import ir_spectrum_from
    z, x = _systems()["water"]
    note("case calculator ir spectrum")
    calc = Gfn1NativeCalculator()
    ir = calc.ir_spectrum(z, x, unit="bohr")
    assert set(ir) == {"wavenumbers", "intensities_km_per_mol", "intensities_au", "dipole_gradients"}
    r = hessian(z, x)
    want = ir_spectrum_from(r.hessian, r.dipole_derivatives, z)
    assert np.allclose(ir["intensities_km_per_mol"], want["intensities_km_per_mol"], rtol=1e-12, atol=1e-12)
    dd = calc.dipole_derivatives(z, x, unit="bohr")
    assert np.array_equal(dd["ddipole_dr"], r.dipole_derivatives)
    # the three vibrations of water are IR active
    assert np.all(ir["intensities_km_per_mol"][6:] > 1.0)

