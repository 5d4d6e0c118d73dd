"""How the SCC iteration count depends on the start charges (the oracle's
restatement of electronic.rs:1231-1239, :1843-1947).

The convergence test compares two successive energies, and the first iteration
always continues (E_{-1} = inf, electronic.rs:1854-1856), so even the exact
converged charges as the start take two iterations; from there the count
grows with the error of the start vector.  These are the numbers behind the
learned start's ceiling: at the default tolerances (1e-7 Eh, charge rms 2e-5)
an rms start error of 1e-2 e takes about six iterations, 1e-3 e five, 1e-4 e
four, the standard zero start seven or eight.
"""
import numpy as np

import oracle
from gfn1_amd import geometry


def test_iterations_grow_with_the_start_error():
    systems = geometry.mixed_batch(6, first_seed=1000, kmax=10)
    counts = {}
    for z, x in systems:
        q = oracle.run(z, x, energy_tolerance=1e-11, charge_tolerance=1e-9)["shell_charges"]
        rng = np.random.default_rng(0)
        for sig in (0.0, 1e-4, 1e-3, 1e-2):
            q0 = q + (rng.normal(0.0, sig, q.shape) if sig else 0.0)
            q0 -= q0.mean() - q.mean()  # the same total charge
            counts.setdefault(sig, []).append(oracle.run(z, x, initial_shell_charges=q0)["iterations"])
        counts.setdefault("zero", []).append(oracle.run(z, x)["iterations"])
    assert counts[0.0] == [2] * len(systems)
    means = [np.mean(counts[k]) for k in (0.0, 1e-4, 1e-3, 1e-2, "zero")]
    assert all(a <= b for a, b in zip(means, means[1:])), means
    assert means[1] >= 3 and means[3] < means[4]
