"""Batched velocity-Verlet dynamics on the device (gfn1_amd.md over the C ABI)
against the oracle's restatement (oracle/md_oracle.py), and energy
conservation."""
import sys

import numpy as np
import pytest

import md_oracle
from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, geometry
from gfn1_amd.api import ATOMIC_MASSES
from gfn1_amd.md import maxwell_boltzmann, run_md

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)
A2B = geometry.ANGSTROM_TO_BOHR


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _systems():
    out = []
    for k, name in enumerate(("water", "methane", "ammonia")):
        z, x = geometry.MONOMERS[name]
        out.append((z, x * A2B))
    return out


def test_batched_dynamics_match_oracle():
    systems = _systems()
    v0 = [maxwell_boltzmann(z, 300.0, 11 + i) for i, (z, _) in enumerate(systems)]
    note("case md 3 monomers, 12 steps of 0.5 fs")
    eng = BatchEngine(systems, options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT)))
    res = run_md(eng, [x for _, x in systems], 12, timestep_fs=0.5, velocities=v0)
    assert res.status == [0, 0, 0]
    for i, (z, x) in enumerate(systems):
        xo, vo, pot, kin = md_oracle.md_gfn1(z, x, v0[i], [ATOMIC_MASSES[a - 1] for a in z], 12, 0.5, **TIGHT)
        dx = float(np.abs(res.positions[i] - xo).max())
        de = float(np.abs(res.potential[:, i] - pot).max())
        note("system %d: max|dx| %.2e max|dE| %.2e" % (i, dx, de))
        assert dx <= 1e-8 and float(np.abs(res.velocities[i] - vo).max()) <= 1e-10
        assert de <= 1e-8 and float(np.abs(res.kinetic[:, i] - kin).max()) <= 1e-10


def test_energy_conservation_and_reversibility():
    z, x = _systems()[0]
    note("case md water conservation")
    eng = BatchEngine([(z, x)], options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT)))
    res = run_md(eng, [x], 200, timestep_fs=0.2, temperature=300.0, seed=3)
    tot = res.total[:, 0]
    drift = float(np.abs(tot - tot[0]).max())
    note("water 200 x 0.2 fs: max |E_tot - E_tot(0)| %.2e, mean kinetic %.2e" % (drift, float(res.kinetic.mean())))
    assert drift < 3e-5
    back = run_md(eng, res.positions, 200, timestep_fs=0.2, velocities=[-res.velocities[0]])
    assert np.abs(back.positions[0] - x).max() < 1e-6


def test_warm_start_with_a_learned_model_and_restore():
    """A batch with a learned start-charge model: with warm_start the model
    starts step 0 only and every later SCC starts from the previous step's
    charges, so the later steps need fewer iterations than the same dynamics
    without warm start; afterwards the batch runs exactly as before."""
    import os

    from gfn1_amd import GuessModel
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    model = GuessModel.load(os.path.join(root, "gfn1-rs_amd", "data", "guess_mlp.bin"))
    systems = geometry.mixed_batch(4, first_seed=1000, kmax=3)
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    note("case md learned model + warm start, 4 clusters, 6 steps")
    eng = BatchEngine(systems, options=opts, guess=model)
    eng.run()
    before = [eng.iterations(i) for i in range(eng.n)]
    e_before = eng.energies()
    x0 = [x for _, x in systems]
    v0 = [maxwell_boltzmann(z, 300.0, 5 + i) for i, (z, _) in enumerate(systems)]
    warm = run_md(eng, x0, 6, timestep_fs=0.5, velocities=v0, warm_start=True)
    cold = run_md(eng, x0, 6, timestep_fs=0.5, velocities=v0, warm_start=False)
    assert warm.status == [0] * 4 and cold.status == [0] * 4
    note("warm iterations %s / cold %s" % (warm.scc_iterations.sum(axis=1).tolist(),
                                          cold.scc_iterations.sum(axis=1).tolist()))
    # step 0 is the same learned start on both
    assert (warm.scc_iterations[0] == cold.scc_iterations[0]).all()
    assert warm.scc_iterations[1:].sum() < cold.scc_iterations[1:].sum()
    assert np.abs(warm.potential - cold.potential).max() <= 1e-8
    # the caller's model and start charges are back in place
    eng.set_positions(x0)
    eng.run()
    assert [eng.iterations(i) for i in range(eng.n)] == before
    assert np.abs(eng.energies() - e_before).max() <= 1e-12
