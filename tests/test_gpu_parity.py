"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Bar (FP64 path): |dE| <= 1e-8 Eh, max |dg| <= 1e-7 Eh/bohr, at tight SCC
tolerances so that both sides sit on the same fixed point.
"""
import sys

import numpy as np
import pytest

import oracle
from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, NativeError, analytic_gradient
from gfn1_amd import geometry, system_from_xyz

import reference_molecules
from test_oracle import MOLECULES, mol

pytestmark = pytest.mark.gpu

TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)
E_TOL = 1e-8
G_TOL = 1e-7


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(MOLECULES))
def test_single_molecule_matches_oracle(name):
    note("case single %s" % name)
    z, x = mol(name)
    check_single_against_oracle(z, x)


@pytest.mark.parametrize("name", sorted(reference_molecules.MOLECULES))
def test_reference_molecule_matches_oracle(name):
    # the reference's own test and example molecules: ferrocene and Ni(CO)4
    # (3d metals), caffeine, and two bromine compounds (halogen bond probe,
    # bromoethanol)
    note("case reference molecule %s" % name)
    z, x, _ = system_from_xyz(reference_molecules.MOLECULES[name])
    check_single_against_oracle(z, x)


@pytest.mark.parametrize("name", sorted(reference_molecules.SMEARED_MOLECULES))
def test_smeared_reference_molecule_matches_oracle(name):
    # the Fermi-smeared systems of the reference's smearing tests (distorted
    # Ni(CO)4 at 3000 K, formaldehyde at 10000 K, NiO at 3000 K): every
    # eigenvector kept, many fractional occupations
    note("case smeared reference molecule %s" % name)
    xyz, temp = reference_molecules.SMEARED_MOLECULES[name]
    z, x, _ = system_from_xyz(xyz)
    opts = dict(TIGHT, enable_dispersion=False, electronic_temperature=temp, max_scc=500)
    ref = oracle.run(z, x, **dict(opts, enable_dispersion=0))
    assert ref["status"] == 0
    res = analytic_gradient(z, x, options=AnalyticGradientOptions(electronic=ElectronicOptions(**opts)))
    el = res.electronic_result
    de = abs(res.total_energy - ref["energy"])
    dg = float(np.abs(res.gradient - ref["gradient"]).max())
    note("%s: |dE| %.2e max|dg| %.2e iterations %d / %d" % (name, de, dg, el.iterations, ref["iterations"]))
    assert el.converged and de <= E_TOL and dg <= G_TOL
    assert el.iterations == ref["iterations"]
    assert np.abs(el.occupations - ref["occupations"]).max() <= 1e-8
    assert abs(el.fermi_level - ref["fermi_level"]) <= 1e-8
    # the reference's guard that the fixture is meaningfully smeared (smearing.rs:161-190)
    f = el.occupations
    frac = f[(f > 1e-8) & (np.abs(f - 2.0) > 1e-8)]
    assert len(frac) >= 2 and np.minimum(np.abs(frac - 2.0), frac).max() > 1e-4


def check_single_against_oracle(z, x):
    ref = oracle.run(z, x, want_matrices=True, **TIGHT)
    assert ref["status"] == 0
    res = analytic_gradient(z, x, options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT)))
    el = res.electronic_result
    assert el.converged
    assert abs(res.total_energy - ref["energy"]) <= E_TOL
    for k, v in ref["energies"].items():
        assert abs(getattr(el.terms, k) - v) <= E_TOL, k
    assert np.abs(res.gradient - ref["gradient"]).max() <= G_TOL
    assert np.abs(res.dispersion_gradient - ref["dispersion_gradient"]).max() <= 1e-12
    assert np.abs(res.repulsion_gradient - ref["repulsion_gradient"]).max() <= 1e-12
    assert np.abs(res.halogen_gradient - ref["halogen_gradient"]).max() <= 1e-12
    assert np.abs(el.overlap - ref["overlap"]).max() <= 1e-13
    assert np.abs(el.h0 - ref["h0"]).max() <= 1e-13
    assert np.abs(el.density - ref["density"]).max() <= 1e-7
    assert np.abs(el.shell_charges - ref["shell_charges"]).max() <= 1e-8
    assert np.abs(el.coordination_numbers - ref["coordination_numbers"]).max() <= 1e-13
    assert np.abs(el.orbital_energies - ref["orbital_energies"]).max() <= 1e-8
    assert abs(el.nelec - ref["nelec"]) < 1e-12
    # the rest of the result contract (electronic.rs:600-654)
    assert np.abs(el.fock - ref["fock"]).max() <= 1e-8
    assert np.abs(el.energy_weighted_density - ref["energy_weighted_density"]).max() <= 1e-7
    assert np.abs(el.shell_scc_potential - ref["shell_scc_potential"]).max() <= 1e-8
    assert np.abs(el.occupations - ref["occupations"]).max() <= 1e-8
    assert np.abs(el.atomic_charges - ref["atomic_charges"]).max() <= 1e-8
    assert np.abs(el.dipole - ref["dipole"]).max() <= 1e-8
    assert el.electronic_temperature == ref["electronic_temperature"] == 300.0
    assert el.charge_order == 3
    assert_fermi_level_matches(el, ref)


KT300 = 300.0 * 3.1668115634556e-6


def assert_fermi_level_matches(el, ref):
    """The Fermi level is the f64 fixed point of a bisection on sum f(mu) = nelec
    (electronic.rs:3853-3921).  Across a HOMO-LUMO gap of many kT that sum is
    exactly nelec in floating point over a wide range of mu; the bisection
    stops at the edge of that range, about 35 kT above the HOMO, where the
    occupation deficit 2 exp(-(mu - e_HOMO)/kT) per HOMO orbital drops below
    the last bit of the sum.  How the partial sums round places that edge
    anywhere within a factor of about 8 in the deficit, i.e. ln 8 = 2.1 kT in
    mu (summing the same eigenvalues in reverse order on the CPU already moves
    it by 0.5 kT).  Both sides must therefore lie in the gap and agree to
    3 kT; the occupations, which do not depend on where mu sits inside the
    flat range, are compared to 1e-8 above.  With an odd electron count the
    singly occupied orbital pins mu (the sum changes by about 1/kT per unit of
    mu there), so both sides must agree closely instead."""
    ne = int(round(ref["nelec"]))
    nocc = ne // 2
    eps = ref["orbital_energies"]
    assert abs(el.fermi_level - ref["fermi_level"]) <= 3.0 * KT300
    if ne % 2 == 1:
        assert abs(el.fermi_level - ref["fermi_level"]) <= 1e-9
    elif 0 < nocc < len(eps) and eps[nocc] - eps[nocc - 1] > 100 * KT300:
        assert eps[nocc - 1] < el.fermi_level < eps[nocc]


def test_fock_and_potential_semantics_at_default_tolerances():
    """At the default SCC tolerances the input and output charges of the last
    iteration differ visibly.  The Fock matrix is built from the input charges
    of the last step and shell_scc_potential from the output charges
    (electronic.rs:1957-1991); both sides must agree with the oracle on each,
    far closer than the input/output difference."""
    note("case default-tolerance result semantics (water)")
    z, x = mol("water")
    ref = oracle.run(z, x, want_matrices=True)
    el = analytic_gradient(z, x).electronic_result
    assert el.iterations == ref["iterations"]
    # potential implied by the Fock diagonal (S_mumu = 1): v_in = H0 - F
    v_in = np.diag(ref["h0"]) - np.diag(ref["fock"])
    # AO -> shell map of water: O 2s | O 2p (3 AO) | H 1s | H 2s | H 1s | H 2s
    ao_shell = [0, 1, 1, 1, 2, 3, 4, 5]
    v_out = ref["shell_scc_potential"][ao_shell]
    gap = np.abs(v_in - v_out).max()
    assert gap > 1e-7
    assert np.abs(el.fock - ref["fock"]).max() <= 1e-3 * gap
    assert np.abs(el.shell_scc_potential - ref["shell_scc_potential"]).max() <= 1e-3 * gap
    assert np.abs(el.density - ref["density"]).max() <= 1e-9
    assert_fermi_level_matches(el, ref)


def test_default_tolerances_match_iteration_count():
    note("case default-tolerance water")
    z, x = mol("water")
    ref = oracle.run(z, x)
    res = analytic_gradient(z, x)
    assert res.electronic_result.iterations == ref["iterations"]
    assert abs(res.total_energy - ref["energy"]) <= 1e-9


def test_batch_of_mixed_clusters_matches_oracle():
    note("case batch mixed clusters")
    systems = geometry.mixed_batch(12, first_seed=1000, kmax=12)
    eng = BatchEngine(systems, options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))).run()
    for i, (z, x) in enumerate(systems):
        ref = oracle.run(z, x, **TIGHT)
        assert eng.status(i) == 0
        assert abs(eng.energy(i) - ref["energy"]) <= E_TOL, i
        assert np.abs(eng.gradient(i) - ref["gradient"]).max() <= G_TOL, i


def test_water48_matches_oracle():
    note("case water48")
    z, x = geometry.water_cluster(48, 11.3, seed=0)
    ref = oracle.run(z, x, **TIGHT)
    res = analytic_gradient(z, x, options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT)))
    assert abs(res.total_energy - ref["energy"]) <= E_TOL
    assert np.abs(res.gradient - ref["gradient"]).max() <= G_TOL


def test_water70_matches_oracle():
    """560 AO: above the batched kernels' limit (512), so the setup goes
    through rocSOLVER and every SCC eigensolve through the large-matrix branch
    (own Householder reduction, rocSOLVER tridiagonal solver)."""
    note("case water70 (large-matrix branch)")
    z, x = geometry.water_cluster(70, 12.9, 3)
    ref = oracle.run(z, x, **TIGHT)
    res = analytic_gradient(z, x, options=AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT)))
    note("water70: dE %.2e max|dg| %.2e" % (abs(res.total_energy - ref["energy"]),
                                            float(np.abs(res.gradient - ref["gradient"]).max())))
    assert abs(res.total_energy - ref["energy"]) <= E_TOL
    assert np.abs(res.gradient - ref["gradient"]).max() <= G_TOL


def test_gfn1_si_parameters_match_oracle():
    """The bundled GFN1(Si)-xTB set ("builtin:si", params.rs:192-210) on
    silanol, whose Si element block and O-Si pair scaling differ from GFN1-xTB."""
    import os

    from gfn1_amd import Gfn1Parameters
    note("case GFN1(Si)-xTB silanol")
    z = [14, 8, 1, 1, 1, 1]
    x = np.array([[0.0, 0.0, 0.0], [1.65, 0.0, 0.0], [2.100693, 0.84763, 0.0], [-0.494034, 1.395109, 0.0],
                  [-0.494034, -0.697555, 1.2082], [-0.494034, -0.697555, -1.2082]]) * geometry.ANGSTROM_TO_BOHR
    si_file = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gfn1-rs_amd", "data",
                           "param_gfn1-si-xtb.txt")
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    res = analytic_gradient(z, x, params=Gfn1Parameters.builtin_si(), options=opts)
    ref = oracle.run(z, x, param_file=si_file, **TIGHT)
    note("silanol GFN1(Si): dE %.2e max|dg| %.2e" % (abs(res.total_energy - ref["energy"]),
                                                    float(np.abs(res.gradient - ref["gradient"]).max())))
    assert abs(res.total_energy - ref["energy"]) <= E_TOL
    assert np.abs(res.gradient - ref["gradient"]).max() <= G_TOL
    base = analytic_gradient(z, x, options=opts)
    assert abs(base.total_energy - res.total_energy) > 1e-3


def test_fallback_ladder_matches_oracle():
    note("case fallback ladder")
    z, x = mol("water")
    ref = oracle.run(z, x, max_scc=1, **TIGHT)
    res = analytic_gradient(z, x, options=AnalyticGradientOptions(
        electronic=ElectronicOptions(max_scc=1, **TIGHT)))
    assert res.electronic_result.converged
    assert res.electronic_result.iterations == ref["iterations"]
    assert abs(res.total_energy - ref["energy"]) <= E_TOL


def test_error_paths():
    note("case error paths")
    with pytest.raises(NativeError) as e:
        analytic_gradient([1, 1, 8], np.array([[0, 0, 0], [0, 0, 1e-7], [1.8, 0, 0]]))
    assert e.value.kind == "InvalidInput"
    with pytest.raises(NativeError) as e:
        analytic_gradient([8, 1, 1], np.array([[0, 0, 0], [1.8, 0, 0], [float("nan"), 0, 0]]))
    assert e.value.kind == "InvalidInput"
    with pytest.raises(NativeError) as e:
        analytic_gradient([1, 1], np.array([[0, 0, 0], [1.4, 0, 0]]), charge=40.0)
    assert e.value.kind == "InvalidInput"
    with pytest.raises(NativeError) as e:
        analytic_gradient([1, 1], np.array([[0, 0, 0], [1.4, 0, 0]]), options=AnalyticGradientOptions(
            electronic=ElectronicOptions(max_scc=2, energy_tolerance=-1.0)))
    assert e.value.kind == "SccNotConverged"
    # an empty system is invalid input, and the calls after it are unaffected
    with pytest.raises(NativeError) as e:
        analytic_gradient([], np.zeros((0, 3)))
    assert e.value.kind == "InvalidInput"
    assert np.isfinite(analytic_gradient([1, 1], np.array([[0, 0, 0], [1.4, 0, 0]])).total_energy)


MIXED_E_TOL = 1e-6
MIXED_G_TOL = 1e-5


def test_mixed_precision_batch_matches_oracle():
    """FP32 density/energy-weighted-density products and gradient contraction:
    |dE| <= 1e-6 Eh, max |dg| <= 1e-5 Eh/bohr against the FP64 oracle."""
    systems = geometry.mixed_batch(8, first_seed=4200) + [geometry.water_cluster(48, 11.3, seed=0)]
    note("case mixed batch of %d" % len(systems))
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    eng = BatchEngine(systems, options=opts, precision="mixed").run()
    for i, (z, x) in enumerate(systems):
        ref = oracle.run(z, x, **TIGHT)
        assert eng.status(i) == 0
        de = abs(eng.energy(i) - ref["energy"])
        dg = float(np.abs(eng.gradient(i) - ref["gradient"]).max())
        note("instance %d: n_atoms %d  dE %.2e  dg %.2e" % (i, len(z), de, dg))
        assert de <= MIXED_E_TOL
        assert dg <= MIXED_G_TOL


def test_batch_matches_committed_fixtures():
    """All fixture systems in one batch against tests/golden/oracle_fixtures.json."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures.json")) as f:
        fx = json.load(f)
    names = sorted(fx["systems"])
    systems = [(np.array(fx["systems"][k]["numbers"]), np.array(fx["systems"][k]["positions"])) for k in names]
    note("case fixtures batch of %d" % len(systems))
    eng = BatchEngine(systems, options=AnalyticGradientOptions(electronic=ElectronicOptions(**fx["tolerances"]))).run()
    for i, k in enumerate(names):
        s = fx["systems"][k]
        assert eng.status(i) == 0, k
        assert abs(eng.energy(i) - s["energy"]) <= E_TOL, k
        assert np.abs(eng.gradient(i) - np.array(s["gradient"])).max() <= G_TOL, k
        q = eng.shell_charges(i, len(s["shell_charges"]))
        assert np.abs(q - np.array(s["shell_charges"])).max() <= 1e-7, k


def test_full_batch_reruns_agree_and_largest_match_oracle():
    """The 4096-cluster batched workload run twice on one engine at tight SCC
    tolerances: every energy agrees between the runs within 1e-8 Eh (a race
    between the size buckets' streams once changed ~1% of the large instances by
    up to 1e-2 Eh from run to run), and the four largest clusters match the
    oracle.  Tight tolerances make the comparison independent of the iteration
    at which a run stops."""
    systems = geometry.mixed_batch(4096, first_seed=1000)
    note("case full batch rerun, %d clusters" % len(systems))
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    eng = BatchEngine(systems, options=opts).run()
    e1 = eng.energies()
    eng.run()
    e2 = eng.energies()
    assert all(eng.status(i) == 0 for i in range(len(systems)))
    d = np.abs(e1 - e2)
    note("max |dE| between runs %.2e at instance %d" % (d.max(), int(d.argmax())))
    assert d.max() <= E_TOL
    nao = np.array([geometry.basis_size(z)[1] for z, _ in systems])
    for i in np.argsort(-nao)[:4]:
        z, x = systems[i]
        ref = oracle.run(z, x, **TIGHT)
        de = abs(eng.energy(i) - ref["energy"])
        dg = float(np.abs(eng.gradient(i) - ref["gradient"]).max())
        note("instance %d (%d AO): dE %.2e  dg %.2e" % (i, nao[i], de, dg))
        assert de <= E_TOL
        assert dg <= G_TOL


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_rocsolver_eigensolvers_agree_with_the_own_solver(solver):
    """eigensolver 1/2/3 (rocSOLVER syevd, syevj, syevdj) in place of the own
    kernels for a small mixed batch: same energies and gradients within the
    FP64 bar, same iteration counts."""
    systems = geometry.mixed_batch(6, first_seed=4300) + [mol("water")]
    note("case rocSOLVER eigensolver %d, %d systems" % (solver, len(systems)))
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    own = BatchEngine(systems, options=opts).run()
    other = BatchEngine(systems, options=opts, eigensolver=solver).run()
    for i in range(len(systems)):
        assert other.status(i) == 0
        assert abs(other.energy(i) - own.energy(i)) <= E_TOL
        assert np.abs(other.gradient(i) - own.gradient(i)).max() <= G_TOL


@pytest.mark.parametrize("nmol,side,seed", [(70, 13.0, 2), (150, 16.6, 3)])
def test_own_reduction_above_512_ao_matches_rocsolver_syevd(nmol, side, seed):
    """Above 512 AO the automatic eigensolver is the engine's own Householder
    reduction with rocSOLVER's tridiagonal solver and back-transformation;
    rocSOLVER dsyevd (eigensolver 1) solves the same problem by another route.
    560 and 1200 AO: same iterations, energies within 1e-8 Eh, gradients
    within 1e-7 Eh/bohr."""
    z, x = geometry.water_cluster(nmol, side, seed)
    nao = geometry.basis_size(z)[1]
    note("case own reduction vs syevd, %d waters, %d AO" % (nmol, nao))
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(**TIGHT))
    own = BatchEngine([(z, x)], options=opts).run()
    ref = BatchEngine([(z, x)], options=opts, eigensolver=1).run()
    assert own.status(0) == 0 and ref.status(0) == 0
    de = abs(own.energy(0) - ref.energy(0))
    dg = float(np.abs(own.gradient(0) - ref.gradient(0)).max())
    note("dE %.2e  dg %.2e  iterations %d / %d" % (de, dg, own.iterations(0), ref.iterations(0)))
    assert own.iterations(0) == ref.iterations(0)
    assert de <= E_TOL and dg <= G_TOL
