"""Cartesian Hessians on the device (through the C ABI) against the oracle's
finite-difference Hessian (oracle/hessian_oracle.py), same step and SCC
tolerances.  Bar: max |dH| <= 1e-6 Eh/bohr^2, the bar at which the reference
accepts its analytic Hessian against this construction
(tests/hessian.rs:691-760).
"""
import sys

import numpy as np
import pytest

import hessian_oracle
import optimizer_oracle
from gfn1_amd import (ElectronicOptions, GeometryOptimizationOptions, HessianOptions, NativeError,
                      geometry, hessian, hessian_batch, optimize_geometry, vibrational_analysis)
from gfn1_amd import AnalyticGradientOptions

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
