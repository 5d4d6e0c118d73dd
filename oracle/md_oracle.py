"""CPU restatement of the batched velocity-Verlet driver (gfn1_amd/md.py)
over the oracle's energy and analytic gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
MD driver.  The product never imports it.

Pinned by tests/test_md_oracle.py: the discrete harmonic oscillator (velocity
Verlet's exact recurrence, cos(w' dt) = 1 - (w dt)^2 / 2), time reversibility,
and momentum conservation under translation-invariant forces.
"""
import numpy as np

import oracle as _oracle

AMU_TO_ELECTRON_MASS = 1822.888486209
FS_PER_ATOMIC_TIME = 0.024188843265857


def verlet(evaluate, x0, v0, masses, steps, dt):
    """Velocity Verlet: returns positions, velocities and the potential and
    kinetic energies after every step (index 0 = start)."""
    x = np.array(x0, dtype=np.float64)
    v = np.array(v0, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64).reshape(-1, 1)
    e, g = evaluate(x)
    pot, kin = [e], [0.5 * float((m * v * v).sum())]
    for _ in range(steps):
        v = v - 0.5 * dt * g / m
        x = x + dt * v
        e, g = evaluate(x)
        v = v - 0.5 * dt * g / m
        pot.append(e)
        kin.append(0.5 * float((m * v * v).sum()))
    return x, v, np.array(pot), np.array(kin)


def md_gfn1(numbers, positions_bohr, velocities, masses_amu, steps, timestep_fs, warm_start=True, **scc):
    """The same dynamics over oracle.run, each SCC started from the previous
    step's charges when warm_start."""
    state = {"q": None}

    def evaluate(x):
        r = _oracle.run(numbers, x, initial_shell_charges=state["q"], **scc)
        if r["status"] != 0:
            raise RuntimeError(r["message"])
        if warm_start:
            state["q"] = r["shell_charges"]
        return r["energy"], r["gradient"]

    m = np.asarray(masses_amu) * AMU_TO_ELECTRON_MASS
    return verlet(evaluate, np.asarray(positions_bohr, dtype=np.float64).reshape(-1, 3), velocities, m, steps,
                  timestep_fs / FS_PER_ATOMIC_TIME)
