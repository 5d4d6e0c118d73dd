"""Periodic Gamma-point path on the GPU (through the C ABI) against the oracle's
periodic restatement (tests/test_oracle_periodic.py pins it).

Bar: |dE| <= 1e-8 Eh per energy term, max |dg| <= 1e-7 Eh/bohr, S(Gamma),
H0(Gamma) to 1e-12, the same SCC iteration count, at tight SCC tolerances.
Cases: water boxes (8 and 32 molecules, images through every face), a
halogen-bonded pair in a skewed cell (the periodic halogen term), hydrogen
sulfide with water (d shells), periodic and molecular instances in one batch,
mixed precision, the include flags and the level shift the periodic path
ignores, a singular cell, and the pair gradient from both sides (the path of
cells above the ket-share limit) against the symmetry-unique one.
"""
import sys

import numpy as np
import pytest

import hessian_oracle
import optimizer_oracle
import oracle
import periodic_cells
from gfn1_amd import (AnalyticGradientOptions, BatchEngine, ElectronicOptions, GeometryOptimizationOptions,
                      NativeError, analytic_gradient, geometry, hessian, optimize_geometry, run_electronic,
                      system_from_xyz)

pytestmark = pytest.mark.gpu

A2B = 1.8897261246257702
TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)
KT300 = 300.0 * 3.1668115634556e-6
OPTS = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def halogen_cell():
    z = np.array([6, 35, 1, 1, 1, 7, 1, 1, 1])
    x = np.array([[0, 0, 0], [0, 0, 1.94], [1.03, 0, -0.36], [-0.51, 0.89, -0.36], [-0.51, -0.89, -0.36],
                  [0, 0, 4.9], [0.94, 0, 5.25], [-0.47, 0.81, 5.25], [-0.47, -0.81, 5.25]]) * A2B
    cell = np.array([[9.0, 0, 0], [0.5, 8.5, 0], [0.3, 0.2, 8.8]]) * A2B
    return z, x, cell


def sulfide_cell():
    z = np.array([16, 1, 1, 8, 1, 1])
    x = np.array([[0.0, 0.0, 0.0], [1.336, 0.0, 0.0], [-0.04, 1.335, 0.0],
                  [3.2, 2.4, 0.3], [3.9, 2.9, 0.6], [2.5, 3.0, 0.1]]) * A2B
    cell = np.array([[6.5, 0.0, 0.0], [0.0, 6.8, 0.0], [0.4, 0.0, 7.2]]) * A2B
    return z, x, cell


CASES = {
    "water8": lambda: geometry.water_box(8, 6.2, seed=3),
    "halogen": halogen_cell,
    "sulfide": sulfide_cell,
}


def check_against_oracle(label, z, x, cell, res, ref, e_tol=1e-8, g_tol=1e-7):
    el = res.electronic_result
    de = abs(res.total_energy - ref["energy"])
    dg = float(np.abs(res.gradient - ref["gradient"]).max())
    note("%s: |dE| %.2e  max|dg| %.2e  iterations %d / %d" % (label, de, dg, el.iterations, ref["iterations"]))
    assert el.converged
    assert de <= e_tol
    for k, v in ref["energies"].items():
        assert abs(getattr(el.terms, k) - v) <= e_tol, k
    assert dg <= g_tol
    assert el.iterations == ref["iterations"]
    for k in ("repulsion_gradient", "dispersion_gradient", "halogen_gradient"):
        assert np.abs(getattr(res, k) - ref[k]).max() <= 1e-11, k
    assert np.abs(res.gradient.sum(axis=0)).max() <= 1e-10


@pytest.mark.parametrize("name", sorted(CASES))
def test_periodic_matches_oracle(name):
    note("case periodic %s" % name)
    z, x, cell = CASES[name]()
    ref = oracle.run(z, x, cell=cell, want_matrices=True, **TIGHT)
    assert ref["status"] == 0
    res = analytic_gradient(z, x, options=OPTS, lattice=cell)
    check_against_oracle(name, z, x, cell, res, ref)
    el = res.electronic_result
    assert np.abs(el.overlap - ref["overlap"]).max() <= 1e-12
    assert np.abs(el.h0 - ref["h0"]).max() <= 1e-12
    assert np.abs(el.coordination_numbers - ref["coordination_numbers"]).max() <= 1e-12
    assert np.abs(el.density - ref["density"]).max() <= 1e-7
    assert np.abs(el.energy_weighted_density - ref["energy_weighted_density"]).max() <= 1e-7
    assert np.abs(el.fock - ref["fock"]).max() <= 1e-8
    assert np.abs(el.shell_charges - ref["shell_charges"]).max() <= 1e-8
    assert np.abs(el.shell_scc_potential - ref["shell_scc_potential"]).max() <= 1e-8
    assert np.abs(el.dipole - ref["dipole"]).max() <= 1e-8
    # the Fermi level sits anywhere on the flat region of the electron count
    # inside the gap (see test_gpu_parity.assert_fermi_level_matches)
    assert abs(el.fermi_level - ref["fermi_level"]) <= 3.0 * KT300
    nocc = int(round(ref["nelec"] / 2))
    assert ref["orbital_energies"][nocc - 1] < el.fermi_level < ref["orbital_energies"][nocc]
    assert abs(sum(el.shell_charges)) < 1e-10


@pytest.mark.parametrize("name", sorted(periodic_cells.CELLS))
def test_reference_cells_match_oracle(name):
    # the periodic cells of the reference's own tests: diamond, rocksalt, LiH and
    # H2 in large cells, metallic Ni2 and an H2 chain at 3000 K, skewed two-atom
    # diamond and BN cells (many images per atom pair)
    note("case periodic reference cell %s" % name)
    z, x, q, cell = system_from_xyz(periodic_cells.CELLS[name], with_lattice=True)
    extra = periodic_cells.SMEARED.get(name, {})
    ref = oracle.run(z, x, cell=cell, **dict(TIGHT, **extra))
    assert ref["status"] == 0
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**dict(TIGHT, **extra)))
    res = analytic_gradient(z, x, options=opts, lattice=cell)
    check_against_oracle(name, z, x, cell, res, ref)
    if name == "diamond":
        assert np.abs(res.gradient).max() < 1e-4  # tests/pbc_integration.rs:62-80


def test_periodic_32_water_box_matches_oracle():
    note("case periodic water32")
    z, x, cell = geometry.water_box(32, 9.86, seed=3)
    ref = oracle.run(z, x, cell=cell, **TIGHT)
    assert ref["status"] == 0
    res = analytic_gradient(z, x, options=OPTS, lattice=cell)
    check_against_oracle("water32", z, x, cell, res, ref)


def test_periodic_and_molecular_instances_in_one_batch():
    note("case periodic batch")
    systems = []
    for name in sorted(CASES):
        z, x, cell = CASES[name]()
        systems.append((z, x, 0.0, cell))
        systems.append((z, x))  # the same atoms as a molecule
    systems += [(z, x) for z, x in geometry.mixed_batch(3, first_seed=1000, kmax=6)]
    eng = BatchEngine(systems, options=OPTS).run()
    for i, s in enumerate(systems):
        cell = s[3] if len(s) > 3 else None
        ref = oracle.run(s[0], s[1], cell=cell, **TIGHT)
        de = abs(eng.energy(i) - ref["energy"])
        dg = float(np.abs(eng.gradient(i) - ref["gradient"]).max())
        note("batch %d (%s): |dE| %.2e max|dg| %.2e" % (i, "periodic" if cell is not None else "molecule", de, dg))
        assert eng.status(i) == 0 and de <= 1e-8 and dg <= 1e-7
        assert eng.iterations(i) == ref["iterations"]


def test_periodic_pair_gradient_both_sides_matches_unique_pairs(monkeypatch):
    # cells above the ket-share limit evaluate every (atom, image) pair from both
    # sides; the limit set to 0 takes that path for a small cell
    note("case periodic both-sides pair gradient")
    z, x, cell = sulfide_cell()
    unique = BatchEngine([(z, x, 0.0, cell)], options=OPTS).run()
    monkeypatch.setenv("GFN1_KET_SHARE_MAX_ATOMS", "0")
    both = BatchEngine([(z, x, 0.0, cell)], options=OPTS).run()
    ref = oracle.run(z, x, cell=cell, **TIGHT)
    dg = float(np.abs(both.gradient(0) - unique.gradient(0)).max())
    note("both sides vs unique pairs: max|dg| %.2e" % dg)
    assert abs(both.energy(0) - unique.energy(0)) <= 1e-12
    assert dg <= 1e-11
    assert np.abs(both.gradient(0) - ref["gradient"]).max() <= 1e-7


def test_periodic_mixed_precision():
    note("case periodic mixed")
    z, x, cell = geometry.water_box(8, 6.2, seed=3)
    a = BatchEngine([(z, x, 0.0, cell)], options=OPTS).run()
    b = BatchEngine([(z, x, 0.0, cell)], options=OPTS, precision="mixed").run()
    assert abs(a.energy(0) - b.energy(0)) <= 1e-6
    assert np.abs(a.gradient(0) - b.gradient(0)).max() <= 1e-5


def test_periodic_gradient_ignores_include_flags_and_level_shift():
    note("case periodic flags")
    z, x, cell = halogen_cell()
    full = analytic_gradient(z, x, options=OPTS, lattice=cell)
    off = AnalyticGradientOptions(electronic=ElectronicOptions(level_shift=0.3, **TIGHT), include_repulsion=False,
                                  include_dispersion=False, include_hamiltonian=False, include_scc=False,
                                  include_halogen=False)
    res = analytic_gradient(z, x, options=off, lattice=cell)
    assert res.electronic_result.iterations == full.electronic_result.iterations
    assert abs(res.total_energy - full.total_energy) <= 1e-12
    assert np.abs(res.gradient - full.gradient).max() <= 1e-12
    assert np.abs(res.repulsion_gradient - full.repulsion_gradient).max() <= 1e-14


def test_periodic_errors_and_xyz_input():
    note("case periodic errors")
    z, x, _ = halogen_cell()
    with pytest.raises(NativeError) as e:
        run_electronic(z, x, lattice=np.array([[5.0, 0, 0], [0, 5.0, 0], [10.0, 0, 0]]))
    assert e.value.kind == "SingularCell"
    text = '3\nLattice="6.0 0 0 0 6.0 0 0 0 6.0" pbc="T T T"\nO 0 0 0\nH 0.757 0.586 0\nH -0.757 0.586 0\n'
    zz, xx, q, cell = system_from_xyz(text, with_lattice=True)
    res = run_electronic(zz, xx, lattice=cell, options=ElectronicOptions(**TIGHT))
    ref = oracle.run(zz, xx, cell=cell, **TIGHT)
    assert abs(res.total_free - ref["energy"]) <= 1e-8


def test_periodic_fixed_cell_optimisation_matches_oracle():
    # optimize_geometry over the periodic energy and gradient, cell fixed
    # (optimizer.rs:78-185 through the periodic dispatch of gradient.rs:61-66)
    note("case periodic optimize")
    z, x, cell = geometry.water_box(2, 4.6, seed=7)
    r = optimize_geometry(z, x, options=GeometryOptimizationOptions(max_iterations=12), gradient_options=OPTS,
                          lattice=cell)
    want = optimizer_oracle.optimize_gfn1(z, x, scc_options=dict(cell=cell, **TIGHT), max_iterations=12)
    note("periodic optimize: device it=%d E=%.12f | oracle it=%d E=%.12f" % (r.iterations, r.energy,
                                                                              want["iterations"], want["energy"]))
    assert r.status == 0 and r.iterations == want["iterations"] and r.converged == want["converged"]
    assert abs(r.energy - want["energy"]) <= 1e-8
    assert np.abs(r.positions - want["positions"]).max() <= 1e-6
    for a, b in zip(r.trajectory, want["trajectory"]):
        assert abs(a.energy - b[1]) <= 1e-8


def test_periodic_diamond_relaxation_matches_oracle():
    # the reference's periodic optimisation test (tests/optimizer.rs:27-80): a
    # diamond cell with one displaced atom relaxes at fixed cell, the energy and
    # the largest gradient drop; here also iterate for iterate against the
    # oracle's L-BFGS over the oracle's periodic energy
    note("case periodic diamond relaxation")
    z, x, _, cell = system_from_xyz(periodic_cells.DIAMOND_PERTURBED, with_lattice=True)
    r = optimize_geometry(z, x, options=GeometryOptimizationOptions(max_iterations=20, gradient_tolerance=1e-3),
                          lattice=cell)
    want = optimizer_oracle.optimize_gfn1(z, x, scc_options=dict(cell=cell), max_iterations=20,
                                          gradient_tolerance=1e-3)
    note("diamond relaxation: device it=%d E=%.12f | oracle it=%d E=%.12f" % (r.iterations, r.energy,
                                                                              want["iterations"], want["energy"]))
    assert r.status == 0
    assert r.energy < r.trajectory[0].energy - 1e-6
    assert r.max_gradient < r.trajectory[0].max_gradient
    assert r.iterations == want["iterations"] and r.converged == want["converged"]
    assert abs(r.energy - want["energy"]) <= 1e-8
    assert np.abs(r.positions - want["positions"]).max() <= 1e-6


def test_periodic_hessian_matches_oracle():
    note("case periodic hessian")
    z = np.array([8, 1, 1])
    x = np.array([[0.0, 0.0, 0.0], [0.78, 0.58, 0.02], [-0.74, 0.6, -0.03]]) * A2B
    cell = np.array([[5.2, 0.0, 0.0], [0.3, 5.5, 0.0], [0.0, 0.2, 5.8]]) * A2B
    r = hessian(z, x, lattice=cell)
    want, asym = hessian_oracle.fd_hessian(z, x, cell=cell)
    d = float(np.abs(r.hessian - want).max())
    note("periodic hessian: max|dH| %.2e" % d)
    assert d <= 1e-6
    # translational sum rule of the periodic Hessian
    assert np.abs(r.hessian.reshape(9, 3, 3).sum(axis=1)).max() < 1e-5


def test_periodic_box_above_512_ao_matches_oracle():
    # 70 waters, 560 AO: the large-matrix branch (whole-device reduction,
    # rocSOLVER potrf/trtri, dstedc, dormtr) with a periodic instance
    note("case periodic water70 (560 AO)")
    z, x, cell = geometry.water_box(70, 12.8, seed=9)
    ref = oracle.run(z, x, cell=cell, **TIGHT)
    assert ref["status"] == 0
    res = analytic_gradient(z, x, options=OPTS, lattice=cell)
    check_against_oracle("water70", z, x, cell, res, ref)
