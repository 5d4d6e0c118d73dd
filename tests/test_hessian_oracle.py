"""Pins of the oracle's finite-difference Hessian (oracle/hessian_oracle.py)
and of the host vibrational analysis (gfn1_amd.vibrational_analysis).

The Hessian is checked against second differences of the energy (the energy
and the analytic gradient are separate code paths, so a transposed index, a
wrong sign or a dropped term in the assembly fails), the translational sum
rule and symmetry.  The vibrational analysis is checked with the reference's
own unit tests (vibrational.rs:88-129) and the closed-form diatomic.
"""
import numpy as np
import pytest

import hessian_oracle
import oracle
from gfn1_amd import geometry, vibrational_analysis
from gfn1_amd.api import WAVENUMBER_PER_SQRT_AU

A2B = geometry.ANGSTROM_TO_BOHR


def _water():
    z, x = geometry.MONOMERS["water"]
    return z, geometry.perturbed(z, x * A2B, 5, 0.03)[1]


def test_fd_hessian_matches_energy_second_differences():
    z, x = _water()
    h, asym = hessian_oracle.fd_hessian(z, x)
    x0 = np.asarray(x).reshape(-1)
    tight = dict(energy_tolerance=1e-12, charge_tolerance=1e-10)
    d = 2.0e-3

    def e(dx):
        return oracle.run(z, (x0 + dx).reshape(-1, 3), **tight)["energy"]

    n = x0.size
    e0 = e(np.zeros(n))
    worst = 0.0
    for i in range(n):
        ei = np.zeros(n)
        ei[i] = d
        hii = (e(ei) - 2.0 * e0 + e(-ei)) / (d * d)
        worst = max(worst, abs(hii - h[i, i]))
        for j in range(i + 1, n):
            ej = np.zeros(n)
            ej[j] = d
            hij = (e(ei + ej) - e(ei - ej) - e(-ei + ej) + e(-ei - ej)) / (4.0 * d * d)
            worst = max(worst, abs(hij - h[i, j]))
    assert worst < 5e-5, worst


def test_fd_hessian_sum_rule_and_symmetry():
    z, x = _water()
    h, asym = hessian_oracle.fd_hessian(z, x)
    assert asym < 1e-7
    nat = len(z)
    # translating every atom leaves the gradient unchanged: sum over B of H[(A,a),(B,b)] = 0
    rows = h.reshape(3 * nat, nat, 3).sum(axis=1)
    assert np.abs(rows).max() < 1e-6


def test_fd_hessian_term_selection_adds_up():
    z, x = _water()
    full, _ = hessian_oracle.fd_hessian(z, x)
    parts = [hessian_oracle.fd_hessian(z, x, include_repulsion=r, include_electronic=el, include_dispersion=di,
                                       include_halogen=False)[0]
             for r, el, di in ((True, False, False), (False, True, False), (False, False, True))]
    assert np.abs(sum(parts) - full).max() < 1e-8


def test_vibrational_single_oscillator():
    # vibrational.rs:88-106: one H atom, k = 1 Eh/bohr^2 on x
    h = np.zeros((3, 3))
    h[0, 0] = 1.0
    v = vibrational_analysis(h, [1])
    expected = np.sqrt(1.0 / 1.008) * WAVENUMBER_PER_SQRT_AU
    assert abs(v.wavenumbers[-1] - expected) < 1e-6
    assert abs(v.wavenumbers[0]) < 1e-6 and abs(v.wavenumbers[1]) < 1e-6


def test_vibrational_translations_and_diatomic_closed_form():
    # vibrational.rs:108-129 plus the two-body closed form omega^2 = k (1/m1 + 1/m2)
    k = 0.5
    h = np.zeros((6, 6))
    h[0, 0] = h[3, 3] = k
    h[0, 3] = h[3, 0] = -k
    v = vibrational_analysis(h, [1, 1])
    assert int(np.sum(np.abs(v.wavenumbers) < 1e-6)) == 5 and v.wavenumbers[-1] > 0.0
    v = vibrational_analysis(h, [1, 8])
    # zero modes up to the rounding of the mass-weighted eigenvalues (sqrt(1e-16) ~ 1e-8)
    assert np.all(np.abs(v.wavenumbers[:5]) < 1e-3)
    want = np.sqrt(k * (1.0 / 1.008 + 1.0 / 15.999)) * WAVENUMBER_PER_SQRT_AU
    assert abs(v.wavenumbers[-1] - want) < 1e-6
    # the mode moves the two atoms against each other, the light one more
    mode = v.modes[-1]
    assert mode[0] * mode[3] < 0 and abs(mode[0]) > abs(mode[3])
    # negative curvature gives a negative (imaginary) wavenumber
    assert vibrational_analysis(-h, [1, 8]).wavenumbers[0] < 0


def test_relaxed_water_has_six_zero_modes():
    import optimizer_oracle
    z, x = _water()
    r = optimizer_oracle.optimize_gfn1(z, x, scc_options=dict(energy_tolerance=1e-11, charge_tolerance=1e-9),
                                       gradient_tolerance=1e-6, max_iterations=200)
    assert r["converged"]
    h, _ = hessian_oracle.fd_hessian(z, r["positions"])
    w = vibrational_analysis(h, z).wavenumbers
    assert np.all(np.abs(w[:6]) < 30.0), w[:6]
    assert 1300.0 < w[6] < 1800.0 and 3000.0 < w[7] < w[8] < 4000.0, w[6:]


def _hydronium():
    # H3O+: O at the origin, three H at 0.98 Angstrom in a flattened pyramid
    import math
    pts = [[0.0, 0.0, 0.0]]
    for k in range(3):
        phi = 2.0 * math.pi * k / 3.0
        pts.append([0.95 * math.cos(phi), 0.95 * math.sin(phi), 0.30])
    return [8, 1, 1, 1], np.array(pts) * A2B


@pytest.mark.parametrize("case", ["water", "hydronium"])
def test_dipole_derivative_sum_rules(case):
    z, x = _water() if case == "water" else _hydronium()
    charge = 0.0 if case == "water" else 1.0
    _, _, ddip, mu = hessian_oracle.fd_hessian_dipoles(z, x, charge=charge)
    nat = len(z)
    born = ddip.reshape(nat, 3, 3)  # [atom][beta][alpha]
    # translation by t moves mu by Q t: the Born charges sum to Q * identity
    assert np.abs(born.sum(axis=0) - charge * np.eye(3)).max() < 1e-6
    # a small rotation about n moves every atom by n x R and mu by n x mu
    xr = np.asarray(x).reshape(-1, 3)
    for n in np.eye(3):
        dmu = np.einsum("bja,bj->a", born, np.cross(n, xr))
        assert np.abs(dmu - np.cross(n, mu)).max() < 1e-6


def test_ir_intensity_constant_and_point_charge_closed_form():
    from gfn1_amd.api import IR_INTENSITY_AU_TO_KM_PER_MOL, ir_spectrum_from
    # the reference tabulates the conversion (properties.rs:35)
    assert abs(IR_INTENSITY_AU_TO_KM_PER_MOL / 974.8801118 - 1.0) < 1e-9
    # rigid point charges +q, -q on a spring along x: d mu / d Q = q (L_1 - L_2)
    q, k = 0.4, 0.5
    h = np.zeros((6, 6))
    h[0, 0] = h[3, 3] = k
    h[0, 3] = h[3, 0] = -k
    ddip = np.zeros((6, 3))
    for b, qb in ((0, q), (1, -q)):
        ddip[3 * b:3 * b + 3] = qb * np.eye(3)
    ir = ir_spectrum_from(h, ddip, [1, 8])
    m1, m2 = 1.008, 15.999
    mu_red = m1 * m2 / (m1 + m2)
    # the stretch, normalised in mass-weighted coordinates: |dmu/dQ|^2 = q^2 / mu_red
    assert abs(ir["intensities_au"][-1] - q * q / mu_red) < 1e-12
    # the modes are an orthonormal basis of the mass-weighted coordinates, so the
    # intensities of all modes add up to sum_c |d mu / d R_c|^2 / m_c
    masses = np.repeat([m1, m2], 3)
    assert abs(ir["intensities_au"].sum() - float((ddip ** 2).sum(axis=1) @ (1.0 / masses))) < 1e-12

