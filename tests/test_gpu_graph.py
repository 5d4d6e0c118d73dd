"""SCC iterations by replays of a captured HIP graph against the host-launched
iteration loop.

The replay only decides which positions of a size bucket iterate (the gate
kernel at its head); each instance's arithmetic is the host loop's, so every
result must agree bit for bit.  Cases: a batch whose iterations are all plain
ones, the fallback ladder (level shift and smearing rungs hand the bucket back
to the host loop mid-run), reuse of the captured graphs by a second run, and
the mixed-precision path with the learned start.
"""
import sys

import numpy as np
import pytest

import oracle
from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, GuessModel, basis_counts, geometry

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def results(eng, systems):
    out = []
    for i, (z, _) in enumerate(systems):
        out.append((eng.status(i), eng.energy(i), eng.gradient(i), eng.shell_charges(i, basis_counts(z)[0]),
                    eng.iterations(i)))
    return out


def assert_identical(a, b):
    for i, (ra, rb) in enumerate(zip(a, b)):
        assert ra[0] == rb[0], i
        assert ra[4] == rb[4], i
        assert ra[1] == rb[1], (i, ra[1], rb[1])
        np.testing.assert_array_equal(ra[2], rb[2])
        np.testing.assert_array_equal(ra[3], rb[3])


def run_both(systems, **kw):
    g = BatchEngine(systems, scc_graphs=True, **kw).run()
    rg, sg = results(g, systems), g.launch_stats()
    h = BatchEngine(systems, scc_graphs=False, **kw).run()
    rh, sh = results(h, systems), h.launch_stats()
    note("graph mode: %s" % sg)
    note("host loop:  %s" % sh)
    return g, rg, sg, rh, sh


def test_replays_match_the_host_loop_bit_for_bit():
    systems = geometry.mixed_batch(96, first_seed=7000) + [geometry.water_cluster(48, 11.3, seed=0)]
    note("case plain batch of %d" % len(systems))
    _, rg, sg, rh, sh = run_both(systems)
    assert_identical(rg, rh)
    assert all(r[0] == 0 for r in rg)
    assert sg["graph_replays"] > 0 and sg["host_iterations"] == 0
    assert sh["graph_replays"] == 0 and sh["host_iterations"] > 0
    # one host call per replay against one per kernel
    assert sg["enqueues"] == sg["graph_replays"]
    assert sh["enqueues"] == sh["kernels"]
    assert sg["state_reads"] == sg["graph_replays"]


def test_fallback_ladder_hands_the_bucket_back_to_the_host_loop():
    """max_scc = 1 pushes every instance down the ladder (Broyden, damped
    mixing, level shift, smearing): the buckets leave the replays for the host
    loop at the first rung that is not plain, with unchanged results."""
    systems = [geometry.water_cluster(2, 6.0, seed=s) for s in range(3)] + geometry.mixed_batch(5, first_seed=7100)
    z, x = systems[0]
    note("case ladder batch of %d" % len(systems))
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(max_scc=1, **TIGHT))
    _, rg, sg, rh, _ = run_both(systems, options=opts)
    assert_identical(rg, rh)
    assert sg["graph_replays"] > 0 and sg["host_iterations"] > 0
    ref = oracle.run(z, x, max_scc=1, **TIGHT)
    assert rg[0][4] == ref["iterations"]
    assert abs(rg[0][1] - ref["energy"]) <= 1e-8


def test_captured_graphs_are_reused_by_the_next_run():
    systems = geometry.mixed_batch(40, first_seed=7200)
    note("case rerun of %d" % len(systems))
    eng = BatchEngine(systems).run()
    first, s1 = results(eng, systems), eng.launch_stats()
    eng.run()
    second, s2 = results(eng, systems), eng.launch_stats()
    note("first run %s, second run %s" % (s1, s2))
    assert s1["graphs_captured"] > 0 and s2["graphs_captured"] == 0
    assert s2["graph_replays"] == s1["graph_replays"]
    assert_identical(first, second)


def test_mixed_precision_with_the_learned_start():
    systems = geometry.mixed_batch(64, first_seed=7300)
    note("case mixed + learned start, %d clusters" % len(systems))
    model = GuessModel(seed=3)
    _, rg, sg, rh, _ = run_both(systems, precision="mixed", guess=model)
    assert_identical(rg, rh)
    assert sg["graph_replays"] > 0


def test_backtransform_widths_agree_bit_for_bit_and_autotune_keeps_results():
    """Every back-transformation width gives the same bits (each eigenvector's
    arithmetic does not depend on how many share a work-group), so the
    autotuner's choice, made from timings, cannot change a result."""
    systems = geometry.mixed_batch(48, first_seed=7400) + [geometry.water_cluster(48, 11.3, seed=0)]
    note("case back-transformation widths, %d clusters" % len(systems))
    eng = BatchEngine(systems)
    ref = None
    for cols in (32, 16, 64):
        eng.set_backtransform_width(cols).run()
        r = results(eng, systems)
        if ref is None:
            ref = r
        else:
            assert_identical(ref, r)
    table = eng.autotune()
    note("tuning: %s" % table)
    assert table and all(t["cols"] in t["ms"] for t in table)
    assert all(set(t["ms"]) == ({16, 32, 64} if t["ld"] <= 256 else {16, 32}) for t in table)
    eng.run()
    assert_identical(ref, results(eng, systems))
