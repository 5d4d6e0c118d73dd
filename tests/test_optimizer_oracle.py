"""Pins of the oracle's L-BFGS restatement (oracle/optimizer_oracle.py).

Each check is fixed by mathematics or by the reference's own test, not by
restating the formula: the two-loop recursion against the explicit BFGS
inverse-Hessian recursion, the secant condition, the step cap's closed form,
exact minimisers of quadratics, and the stretched-H2 test of
tests/optimizer.rs:5-22.
"""
import math

import numpy as np
import pytest

import optimizer_oracle as opt
from gfn1_amd import geometry


def _pairs(rng, n, m):
    a = rng.normal(size=(n, n))
    hess = a @ a.T + n * np.eye(n)  # SPD: every pair has y^T s > 0
    s = [rng.normal(size=n) for _ in range(m)]
    y = [hess @ v for v in s]
    return s, y


def _explicit_direction(g, s, y):
    """-H g with H from H0 = gamma I and the BFGS inverse update
    H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T (Nocedal & Wright, eq. 7.19)."""
    n = len(g)
    gamma = float(y[-1] @ s[-1]) / float(y[-1] @ y[-1])
    h = gamma * np.eye(n)
    for sk, yk in zip(s, y):
        rho = 1.0 / float(yk @ sk)
        v = np.eye(n) - rho * np.outer(yk, sk)
        h = v.T @ h @ v + rho * np.outer(sk, sk)
    return -(h @ g), h


@pytest.mark.parametrize("m", [1, 2, 5])
def test_two_loop_matches_explicit_bfgs_inverse(m):
    rng = np.random.default_rng(7 + m)
    n = 9
    s, y = _pairs(rng, n, m)
    g = rng.normal(size=n)
    rho = [1.0 / float(yk @ sk) for sk, yk in zip(s, y)]
    d = np.array(opt.lbfgs_direction(g.tolist(), [v.tolist() for v in s], [v.tolist() for v in y], rho))
    want, h = _explicit_direction(g, s, y)
    assert np.allclose(d, want, rtol=1e-10, atol=1e-12)
    # secant condition of the newest pair: H y = s
    sec = -np.array(opt.lbfgs_direction(y[-1].tolist(), [v.tolist() for v in s], [v.tolist() for v in y], rho))
    assert np.allclose(sec, s[-1], rtol=1e-9, atol=1e-12)


def test_two_loop_without_history_is_steepest_descent():
    assert opt.lbfgs_direction([1.0, -2.0, 0.5], [], [], []) == [-1.0, 2.0, -0.5]


def test_two_loop_isotropic_curvature_is_exact_newton():
    # y = c s for every pair: H = I / c, the exact inverse Hessian
    c = 3.0
    s = [[1.0, 0.0, 2.0], [0.5, -1.0, 0.0]]
    y = [[c * v for v in sk] for sk in s]
    rho = [1.0 / opt.dot(yk, sk) for sk, yk in zip(s, y)]
    g = [0.3, -0.6, 0.9]
    d = opt.lbfgs_direction(g, s, y, rho)
    assert np.allclose(d, [-v / c for v in g], rtol=1e-14)


def test_limit_atom_step_closed_form():
    d = [0.3, 0.4, 0.0, 1.2, 0.0, 0.5]  # atom norms 0.5 and 1.3
    capped = opt.limit_atom_step(d, 0.26)
    assert math.isclose(math.hypot(*capped[3:6]), 0.26, rel_tol=1e-14)
    assert np.allclose(np.array(capped) * (1.3 / 0.26), d, rtol=1e-14)
    assert opt.limit_atom_step(d, 2.0) == d
    assert opt.limit_atom_step(d, 0.0) == d


def test_unit_quadratic_is_solved_in_one_step():
    # f = |x - x*|^2 / 2: the first steepest-descent step with alpha = 1 lands on x*
    xs = np.array([0.1, -0.2, 0.05, 0.3, 0.0, -0.1])
    x0 = xs + np.array([0.1, 0.05, -0.08, 0.0, 0.12, 0.02])

    def f(x):
        r = np.array(x) - xs
        return 0.5 * float(r @ r), r.tolist()

    res = opt.optimize(f, x0.tolist())
    assert res["iterations"] == 1 and res["converged"]
    assert np.allclose(res["positions"].reshape(-1), xs, atol=1e-15)


def test_convex_quadratic_converges_with_armijo_steps():
    rng = np.random.default_rng(3)
    n = 12
    a = rng.normal(size=(n, n))
    hess = a @ a.T / n + 0.5 * np.eye(n)
    xs = rng.normal(size=n)

    def f(x):
        r = np.array(x) - xs
        return 0.5 * float(r @ hess @ r), (hess @ r).tolist()

    res = opt.optimize(f, (xs + rng.normal(size=n)).tolist(), gradient_tolerance=1e-9, max_iterations=200)
    assert res["converged"]
    assert np.allclose(res["positions"].reshape(-1), xs, atol=1e-7)
    traj = res["trajectory"]
    for prev, cur in zip(traj, traj[1:]):
        assert cur[1] <= prev[1]  # Armijo steps never raise the energy
    for (_, _, _, sn, x1), (_, _, _, _, x0) in zip(traj[1:], traj):
        assert sn <= 0.30 * math.sqrt(n / 3) + 1e-12  # no atom moves more than the cap


def test_reference_h2_stretched_converges():
    # tests/optimizer.rs:5-22: H at 0 and 1.1 Angstrom, 40 iterations, gtol 2e-4
    x = np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0]]) * geometry.ANGSTROM_TO_BOHR
    res = opt.optimize_gfn1([1, 1], x, max_iterations=40, gradient_tolerance=2.0e-4)
    assert res["converged"]
    assert res["max_gradient"] < 2.0e-4
    # the bond shortens towards the GFN1 equilibrium (about 0.74 Angstrom)
    r = np.linalg.norm(res["positions"][1] - res["positions"][0]) / geometry.ANGSTROM_TO_BOHR
    assert 0.70 < r < 0.80
    assert res["energy"] < res["trajectory"][0][1]
