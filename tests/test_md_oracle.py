"""Pins of the oracle's velocity-Verlet restatement (oracle/md_oracle.py)."""
import math

import numpy as np

import md_oracle


def test_discrete_harmonic_oscillator():
    # x'' = -w^2 x: velocity Verlet advances (x, v) by a fixed 2 x 2 map whose
    # rotation angle theta obeys cos(theta) = 1 - (w dt)^2 / 2
    k, m, dt, steps = 0.7, 2.0, 0.3, 40
    w = math.sqrt(k / m)

    def f(x):
        return 0.5 * k * float((x * x).sum()), k * x

    x, v, pot, kin = md_oracle.verlet(f, np.array([[1.0, 0.0, 0.0]]), np.zeros((1, 3)), [m], steps, dt)
    theta = math.acos(1.0 - 0.5 * (w * dt) ** 2)
    assert abs(x[0, 0] - math.cos(steps * theta)) < 1e-12
    # the shadow energy 1/2 m v^2 (1 - (w dt/2)^2)^-1 + 1/2 k x^2 is conserved exactly
    shadow = kin / (1.0 - (0.5 * w * dt) ** 2) + pot
    assert np.abs(shadow - shadow[0]).max() < 1e-12


def test_time_reversibility_and_momentum():
    rng = np.random.default_rng(5)
    a = rng.normal(size=(6, 6))
    hess = a @ a.T + 6 * np.eye(6)

    def pair(x):
        # a translation-invariant potential of two atoms: depends on x1 - x0 only
        d = (x[1] - x[0]).reshape(-1)
        e = 0.5 * float(d @ hess[:3, :3] @ d)
        g1 = hess[:3, :3] @ d
        return e, np.stack([-g1, g1])

    m = [1.0, 3.0]
    x0 = rng.normal(size=(2, 3))
    v0 = rng.normal(size=(2, 3)) * 0.1
    x1, v1, _, _ = md_oracle.verlet(pair, x0, v0, m, 25, 0.05)
    x2, v2, _, _ = md_oracle.verlet(pair, x1, -v1, m, 25, 0.05)
    assert np.abs(x2 - x0).max() < 1e-12 and np.abs(-v2 - v0).max() < 1e-12
    p0 = (np.array(m)[:, None] * v0).sum(axis=0)
    p1 = (np.array(m)[:, None] * v1).sum(axis=0)
    assert np.abs(p1 - p0).max() < 1e-12
