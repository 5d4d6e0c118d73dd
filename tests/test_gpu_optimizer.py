"""Geometry optimisation on the device (through the C ABI) against the oracle's
restatement of the reference's L-BFGS (oracle/optimizer_oracle.py).

Both sides use tight SCC tolerances, so every energy and gradient agrees to
|dE| <= 1e-8 and the line searches take the same decisions: equal iteration
counts and convergence flags, final energies within 1e-8 Eh, positions within
1e-6 bohr.
"""
import os
import sys

import numpy as np
import pytest

import optimizer_oracle as opt_oracle
import oracle
from gfn1_amd import (AnalyticGradientOptions, BatchEngine, ElectronicOptions, GeometryOptimizationOptions,
                      Gfn1NativeCalculator, GuessModel, NativeError, optimize_geometry)
from gfn1_amd import geometry

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)
E_TOL = 1e-8
X_TOL = 1e-6
MAX_IT = 30  # the monomers converge within it; the water trimer does not (a non-converged case)
A2B = geometry.ANGSTROM_TO_BOHR


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tight_options():
    return AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))


def _systems():
    out = {"h2_stretched": ([1, 1], np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0]]) * A2B)}
    for k, name in enumerate(("water", "methane", "ammonia")):
        z, x = geometry.MONOMERS[name]
        out[name] = geometry.perturbed(z, x * A2B, 11 + k, 0.08)
    out["water3"] = geometry.water_cluster(3, 6.0, 5)
    return out


def _compare(name, got, want):
    note("%s: device it=%d E=%.12f maxg=%.3e | oracle it=%d E=%.12f maxg=%.3e" % (
        name, got.iterations, got.energy, got.max_gradient, want["iterations"], want["energy"],
        want["max_gradient"]))
    assert got.status == 0, got.message
    assert got.converged == want["converged"]
    assert got.iterations == want["iterations"]
    assert abs(got.energy - want["energy"]) <= E_TOL
    assert np.max(np.abs(got.positions - want["positions"])) <= X_TOL
    assert len(got.trajectory) == len(want["trajectory"])
    for a, b in zip(got.trajectory, want["trajectory"]):
        assert a.iteration == b[0]
        assert abs(a.energy - b[1]) <= E_TOL


def test_reference_h2_stretched():
    # tests/optimizer.rs:5-22
    z, x = _systems()["h2_stretched"]
    note("case h2_stretched single")
    r = optimize_geometry(z, x, options=GeometryOptimizationOptions(max_iterations=40, gradient_tolerance=2.0e-4))
    assert r.converged
    assert r.max_gradient < 2.0e-4
    want = opt_oracle.optimize_gfn1(z, x, max_iterations=40, gradient_tolerance=2.0e-4)
    _compare("h2_stretched", r, want)


def test_batch_matches_oracle_per_system():
    systems = _systems()
    names = sorted(systems)
    note("case batch of %s" % names)
    eng = BatchEngine([systems[n] for n in names], options=_tight_options())
    res = eng.optimize(GeometryOptimizationOptions(max_iterations=MAX_IT))
    note("stats %s" % eng.optimization_stats)
    evals, longest = 0, 0
    for n, r in zip(names, res):
        z, x = systems[n]
        want = opt_oracle.optimize_gfn1(z, x, scc_options=TIGHT, max_iterations=MAX_IT)
        _compare(n, r, want)
        evals += want["evaluations"]
        longest = max(longest, want["evaluations"])
    # each round evaluates just the systems still pending: the device made the
    # oracle's evaluations (start points and line-search trials), in as many
    # rounds as the system that needed the most
    assert eng.optimization_stats["evaluations"] == evals
    assert eng.optimization_stats["rounds"] == longest


def test_warm_start_reaches_the_same_minimum():
    systems = _systems()
    names = sorted(systems)
    note("case warm start batch")
    names.remove("water3")
    eng = BatchEngine([systems[n] for n in names], options=_tight_options())
    cold = eng.optimize(GeometryOptimizationOptions(max_iterations=MAX_IT))
    eng.set_positions([systems[n][1] for n in names])
    warm = eng.optimize(GeometryOptimizationOptions(max_iterations=MAX_IT, warm_start=True))
    for n, a, b in zip(names, cold, warm):
        note("%s cold it=%d E=%.10f warm it=%d E=%.10f" % (n, a.iterations, a.energy, b.iterations, b.energy))
        assert b.status == 0 and b.converged
        assert abs(a.energy - b.energy) <= 1e-6


def test_failures_stay_per_system():
    systems = _systems()
    zw, xw = systems["water"]
    batch = [(zw, xw), ([1, 1], np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), 10.0), ([], np.zeros((0, 3)))]
    note("case failures")
    eng = BatchEngine(batch, options=_tight_options())
    res = eng.optimize(GeometryOptimizationOptions(max_iterations=MAX_IT))
    assert res[0].status == 0 and res[0].converged
    assert res[1].status == 3  # electron count out of range: the start geometry does not evaluate
    assert res[2].status == 3 and "empty" in res[2].message
    want = opt_oracle.optimize_gfn1(zw, xw, scc_options=TIGHT, max_iterations=MAX_IT)
    _compare("water_with_failures", res[0], want)
    with pytest.raises(NativeError):
        optimize_geometry([], np.zeros((0, 3)))


def test_zero_iterations_and_single_step():
    z, x = _systems()["methane"]
    note("case max_iterations 0 / 1")
    r0 = optimize_geometry(z, x, options=GeometryOptimizationOptions(max_iterations=0),
                           gradient_options=_tight_options())
    assert r0.iterations == 0 and not r0.converged and len(r0.trajectory) == 1
    assert np.array_equal(r0.positions, np.asarray(x))
    ref = oracle.run(z, x, **TIGHT)
    assert abs(r0.energy - ref["energy"]) <= E_TOL
    assert np.max(np.abs(r0.gradient - ref["gradient"])) <= 1e-7
    r1 = optimize_geometry(z, x, options=GeometryOptimizationOptions(max_iterations=1),
                           gradient_options=_tight_options())
    want = opt_oracle.optimize_gfn1(z, x, scc_options=TIGHT, max_iterations=1)
    _compare("methane_one_step", r1, want)


def test_set_positions_matches_a_fresh_batch():
    systems = _systems()
    names = sorted(systems)
    moved = [geometry.perturbed(systems[n][0], np.asarray(systems[n][1]), 99, 0.03)[1] for n in names]
    note("case set_positions")
    eng = BatchEngine([systems[n] for n in names], options=_tight_options()).run()
    eng.set_positions(moved)
    eng.run()
    fresh = BatchEngine([(systems[n][0], m) for n, m in zip(names, moved)], options=_tight_options()).run()
    for i in range(len(names)):
        assert eng.status(i) == 0
        assert abs(eng.energy(i) - fresh.energy(i)) <= 1e-12
        assert np.max(np.abs(eng.gradient(i) - fresh.gradient(i))) <= 1e-12


def test_calculator_optimize_writes_trajectory(tmp_path):
    calc = Gfn1NativeCalculator()
    path = tmp_path / "traj.xyz"
    note("case calculator optimize")
    r = calc.optimize([1, 1], [[0.0, 0.0, 0.0], [1.1, 0.0, 0.0]], max_iterations=40, gradient_tolerance=2.0e-4,
                      trajectory_path=str(path))
    assert r.converged and r.max_gradient < 2.0e-4
    text = path.read_text().splitlines()
    assert len(text) == 4 * len(r.trajectory_energies_hartree)
    assert text[0] == "2" and text[1].startswith("iter 0 energy")
    assert r.to_xyz().splitlines()[2].split()[0] == "H"
    assert np.allclose(r.forces_ev_per_angstrom, -np.asarray(r.gradient_hartree_per_bohr) * 27.211386245988
                       * A2B)


def test_learned_mixed_warm_path_reaches_the_fp64_minima():
    """The throughput path (learned start, mixed precision, warm start): its line
    searches may decide differently at the 1e-6 level, so only the minima are
    compared, within the mixed path's energy bar."""
    systems = _systems()
    names = sorted(n for n in systems if n != "water3")
    note("case learned + mixed + warm optimisation")
    model = GuessModel.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                         "gfn1-rs_amd", "data", "guess_mlp.bin"))
    fast = BatchEngine([systems[n] for n in names], options=_tight_options(), guess=model, precision="mixed")
    got = fast.optimize(GeometryOptimizationOptions(max_iterations=MAX_IT, warm_start=True))
    ref = BatchEngine([systems[n] for n in names], options=_tight_options()).optimize(
        GeometryOptimizationOptions(max_iterations=MAX_IT))
    for n, a, b in zip(names, got, ref):
        note("%s mixed it=%d E=%.10f fp64 it=%d E=%.10f" % (n, a.iterations, a.energy, b.iterations, b.energy))
        assert a.status == 0 and a.converged and b.converged
        assert abs(a.energy - b.energy) <= 1e-5
        assert a.max_gradient <= 1e-4


def test_large_system_in_a_mixed_batch_matches_it_alone():
    """A 70-water cluster (560 AO, above the engine's own Cholesky limit, so its
    setup goes through rocSOLVER) optimised next to a small molecule: the rounds
    select different systems, and the large system must follow exactly the
    trajectory it follows alone; the small one matches the oracle."""
    big = geometry.water_cluster(70, 12.9, 3)
    zw, xw = _systems()["water"]
    note("case large system in a mixed batch (%d atoms)" % len(big[0]))
    opts = GeometryOptimizationOptions(max_iterations=3)
    both = BatchEngine([big, (zw, xw)], options=_tight_options()).optimize(opts)
    alone = BatchEngine([big], options=_tight_options()).optimize(opts)[0]
    a = both[0]
    note("large: batch it=%d E=%.12f alone it=%d E=%.12f" % (a.iterations, a.energy, alone.iterations, alone.energy))
    assert a.status == 0 and alone.status == 0
    assert a.iterations == alone.iterations == 3
    assert abs(a.energy - alone.energy) <= 1e-10
    assert np.abs(a.positions - alone.positions).max() <= 1e-9
    want = opt_oracle.optimize_gfn1(zw, xw, scc_options=TIGHT, max_iterations=3)
    _compare("water_next_to_large", both[1], want)


def test_batch_of_small_clusters_matches_oracle():
    """Eight small mixed H2O/CH4/NH3 clusters (2-3 molecules, seeds 1000-1007),
    15 L-BFGS iterations each, in one batch: every system's trajectory equals the
    oracle's."""
    systems = geometry.mixed_batch(8, first_seed=1000, kmin=2, kmax=3)
    note("case 8 small clusters")
    res = BatchEngine(systems, options=_tight_options()).optimize(GeometryOptimizationOptions(max_iterations=15))
    for k, ((z, x), r) in enumerate(zip(systems, res)):
        want = opt_oracle.optimize_gfn1(z, x, scc_options=TIGHT, max_iterations=15)
        _compare("cluster%d(%d atoms)" % (k, len(z)), r, want)

