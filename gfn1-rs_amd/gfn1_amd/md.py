"""Batched molecular dynamics (NVE, velocity Verlet) over the batch engine.

The reference drives molecular dynamics through ASE (its calculator supplies
energies and forces to ``ase.md``; scripts/run_mgo_kpoint_md.py).  Here every
system of a batch advances together: each step is one batched energy +
gradient pass on the device, and the integration is host work on 3 nat
vectors (the C++ driver behind gfn1_batch_md).  Each SCC can start from the
previous step's converged charges through the reference's
``scc_initial_shell_charges`` hook.

Units: positions bohr, velocities bohr per atomic time unit, energies Eh,
time step in femtoseconds.
"""
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .api import ATOMIC_MASSES, BatchEngine

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
    """Per system: final positions and velocities; per step (rows, step 0 =
    start) and system (columns): potential (total free energy) and kinetic
    energy; SCC iterations per step and system (-1 once it stopped); status per
    system (0, or the error that stopped it)."""
    positions: List[np.ndarray]
    velocities: List[np.ndarray]
    times_fs: np.ndarray
    potential: np.ndarray
    kinetic: np.ndarray
    status: List[int]
    scc_iterations: np.ndarray

    @property
    def total(self):
        return self.potential + self.kinetic


def run_md(engine: BatchEngine, positions, steps, timestep_fs=0.5, velocities=None, temperature=None, seed=1,
           warm_start=True) -> MdResult:
    """Velocity-Verlet NVE dynamics of every system of the batch, on the
    device through gfn1_batch_md.

    positions: one (nat, 3) array per system (bohr), the start geometry;
    velocities: one (nat, 3) array per system, or None for Maxwell-Boltzmann
    velocities at temperature (K; zero velocities when None too).
    """
    from . import _lib
    import ctypes
    n = engine.n
    z = engine.numbers
    if velocities is not None:
        v = [np.asarray(vi, dtype=np.float64).reshape(-1, 3) for vi in velocities]
    elif temperature is not None:
        v = [maxwell_boltzmann(zi, temperature, seed + i) for i, zi in enumerate(z)]
    else:
        v = [np.zeros((len(zi), 3)) for zi in z]
    engine.set_positions(positions)
    vf = np.ascontiguousarray(np.concatenate(v) if n else np.zeros((0, 3)))
    pot = np.zeros((steps + 1, n))
    kin = np.zeros((steps + 1, n))
    st = np.zeros(n, dtype=np.int32)
    its = np.zeros((steps + 1, n), dtype=np.int32)
    o = _lib.MdOptions()
    o.timestep_fs = float(timestep_fs)
    o.steps = int(steps)
    o.warm_start = int(bool(warm_start))
    err = _lib.Gfn1Error()
    _lib.check(engine._lib.gfn1_batch_md(engine._h, ctypes.byref(o), vf.ctypes.data_as(_lib._DP),
                                         pot.ctypes.data_as(_lib._DP), kin.ctypes.data_as(_lib._DP),
                                         st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         its.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(err)), err)
    xf = np.zeros((sum(engine.natoms), 3))
    if n:
        engine._lib.gfn1_batch_positions(engine._h, xf.ctypes.data_as(_lib._DP))
    off = np.cumsum([0] + engine.natoms)
    return MdResult(positions=[xf[off[i]:off[i + 1]].copy() for i in range(n)],
                    velocities=[vf[off[i]:off[i + 1]].copy() for i in range(n)],
                    times_fs=np.arange(steps + 1) * float(timestep_fs), potential=pot, kinetic=kin,
                    status=[int(s) for s in st], scc_iterations=its)
