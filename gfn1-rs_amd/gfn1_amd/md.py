"""Batched molecular dynamics (NVE, velocity Verlet) over the batch engine.

The reference drives molecular dynamics through ASE (its calculator supplies
energies and forces to ``ase.md``; scripts/run_mgo_kpoint_md.py).  Here every
system of a batch advances together: each step is one batched energy +
gradient pass on the device, and the integration is host work on 3 nat
vectors.  Each SCC can start from the previous step's converged charges
through the reference's ``scc_initial_shell_charges`` hook.

Units: positions bohr, velocities bohr per atomic time unit, energies Eh,
time step in femtoseconds.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .api import ATOMIC_MASSES, BatchEngine, basis_counts

# CODATA 2018
AMU_TO_ELECTRON_MASS = 1822.888486209
FS_PER_ATOMIC_TIME = 0.024188843265857
BOLTZMANN_HARTREE_PER_K = 3.166811563e-6


def masses_au(numbers):
    """Atomic masses (electron masses) per atom, from the standard atomic weights."""
    return np.array([ATOMIC_MASSES[int(z) - 1] for z in numbers]) * AMU_TO_ELECTRON_MASS


def maxwell_boltzmann(numbers, temperature, seed=1):
    """Velocities drawn from the Maxwell-Boltzmann distribution at the given
    temperature (K), centre-of-mass motion removed."""
    m = masses_au(numbers)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(BOLTZMANN_HARTREE_PER_K * temperature / m)[:, None]
    v -= (m[:, None] * v).sum(axis=0) / m.sum()
    return v


def kinetic_energy(numbers, velocities):
    return 0.5 * float((masses_au(numbers)[:, None] * np.asarray(velocities) ** 2).sum())


@dataclass
class MdResult:
    """Per system: final positions and velocities; per recorded step (rows)
    and system (columns): potential (total free energy), kinetic and total
    energy; status per system (0, or the error that stopped it)."""
    positions: List[np.ndarray]
    velocities: List[np.ndarray]
    times_fs: np.ndarray
    potential: np.ndarray
    kinetic: np.ndarray
    status: List[int]
    trajectory: List[List[np.ndarray]] = field(default_factory=list)

    @property
    def total(self):
        return self.potential + self.kinetic


def run_md(engine: BatchEngine, positions, steps, timestep_fs=0.5, velocities=None, temperature=None, seed=1,
           warm_start=True, record_every=1, keep_trajectory=False) -> MdResult:
    """Velocity-Verlet NVE dynamics of every system of the batch.

    positions: one (nat, 3) array per system (bohr), the start geometry;
    velocities: one (nat, 3) array per system, or None for Maxwell-Boltzmann
    velocities at temperature (K; zero velocities when None too).
    """
    n = engine.n
    x = [np.asarray(p, dtype=np.float64).reshape(-1, 3).copy() for p in positions]
    z = engine.numbers
    m = [masses_au(zi)[:, None] for zi in z]
    if velocities is not None:
        v = [np.asarray(vi, dtype=np.float64).reshape(-1, 3).copy() for vi in velocities]
    elif temperature is not None:
        v = [maxwell_boltzmann(zi, temperature, seed + i) for i, zi in enumerate(z)]
    else:
        v = [np.zeros_like(xi) for xi in x]
    nsh = [basis_counts(zi)[0] if len(zi) else 0 for zi in z]
    dt = timestep_fs / FS_PER_ATOMIC_TIME
    status = [0] * n
    times, pot, kin, traj = [], [], [], [[] for _ in range(n)]

    q_start = None

    def evaluate():
        nonlocal q_start
        engine.set_positions(x)
        engine.run()
        st, e, g, q = engine.results()
        for i in range(n):
            if status[i] == 0:
                status[i] = int(st[i])
            if status[i] != 0:
                e[i] = np.nan
        if warm_start:
            # a failed system keeps the charges it last converged to
            q_start = [q[i] if status[i] == 0 else (q_start[i] if q_start is not None else np.zeros(nsh[i]))
                       for i in range(n)]
            engine.set_all_start_charges(q_start)
        return e, g

    def record(step, e):
        times.append(step * timestep_fs)
        pot.append(e.copy())
        kin.append([kinetic_energy(z[i], v[i]) if status[i] == 0 else np.nan for i in range(n)])
        if keep_trajectory:
            for i in range(n):
                traj[i].append(x[i].copy())

    e, g = evaluate()
    record(0, e)
    for step in range(1, steps + 1):
        for i in range(n):
            if status[i] == 0:
                v[i] -= 0.5 * dt * g[i] / m[i]
                x[i] += dt * v[i]
        e, g = evaluate()
        for i in range(n):
            if status[i] == 0:
                v[i] -= 0.5 * dt * g[i] / m[i]
        if step % max(record_every, 1) == 0 or step == steps:
            record(step, e)
    return MdResult(positions=x, velocities=v, times_fs=np.array(times), potential=np.array(pot),
                    kinetic=np.array(kin), status=status, trajectory=traj if keep_trajectory else [])
