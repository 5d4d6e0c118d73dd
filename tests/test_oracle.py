"""Pins of the CPU oracle against the reference's golden vectors, finite
differences, closed forms and invariants (CPU only)."""
import math
import os

import numpy as np
import pytest
from scipy import integrate

import oracle
import reference_molecules
from gfn1_amd import system_from_xyz

AA = 1.8897261246257702
TIGHT = dict(energy_tolerance=1e-11, charge_tolerance=1e-9)

MOLECULES = {
    # tests/nonpbc_smoke.rs:14-60 fixtures (Angstrom)
    "hydrogen": ([1, 1], [[0, 0, 0], [0.74, 0, 0]]),
    "hydrogen_fluoride": ([1, 9], [[0, 0, 0], [0.917, 0, 0]]),
    "water": ([8, 1, 1], [[0, 0, 0], [0.757, 0.586, 0], [-0.757, 0.586, 0]]),
    "phosphine": ([15, 1, 1, 1], [[0, 0, 0], [1.193, 0, 0.768], [-0.5965, 1.0333, 0.768],
                                  [-0.5965, -1.0333, 0.768]]),
    "borane": ([5, 1, 1, 1], [[0, 0, 0], [1.19, 0, 0], [-0.595, 1.03057, 0], [-0.595, -1.03057, 0]]),
    # examples/methane.xyz, examples/ammonia.xyz
    "methane": ([6, 1, 1, 1, 1], [[0, 0, 0], [0.6276, 0.6276, 0.6276], [0.6276, -0.6276, -0.6276],
                                  [-0.6276, 0.6276, -0.6276], [-0.6276, -0.6276, 0.6276]]),
    "ammonia": ([7, 1, 1, 1], [[0, 0, 0.1], [0, 0.942, -0.267], [0.816, -0.471, -0.267],
                               [-0.816, -0.471, -0.267]]),
    # bromomethane ... ammonia: exercises the halogen-bond term (Br carries CXB)
    "bromomethane_ammonia": ([35, 6, 1, 1, 1, 7, 1, 1, 1],
                             [[0, 0, 0], [-1.94, 0, 0], [-2.30, 1.03, 0], [-2.30, -0.51, 0.89],
                              [-2.30, -0.51, -0.89], [3.0, 0.1, 0.0], [3.4, 1.0, 0.2], [3.4, -0.5, 0.8],
                              [3.35, -0.45, -0.8]]),
}


def mol(name):
    z, x = MOLECULES[name]
    return np.array(z), np.array(x, dtype=float) * AA


def test_d3_h2_golden():
    # dispersion.rs:3389-3396: D3 energy of H2 at 0.74 A
    e, _ = oracle.dispersion([1, 1], np.array([[0, 0, 0], [0.74, 0, 0]]) * AA)
    assert abs(e - (-3.5462980e-5)) < 5e-10


def test_h2_engrad_golden():
    # examples/h2.engrad: E = -1.036116470529 Eh, g_x = +-0.025116267882 Eh/bohr at R = 1.3983973 bohr
    x = np.array([[0.0, 0.0, 0.0], [1.3983973, 0.0, 0.0]])
    r = oracle.run([1, 1], x)
    assert r["status"] == 0, r["message"]
    assert abs(r["energy"] - (-1.036116470529)) < 1e-9
    assert abs(r["gradient"][0, 0] - 0.025116267882) < 1e-7
    assert abs(r["gradient"][1, 0] + 0.025116267882) < 1e-7
    assert np.abs(r["gradient"][:, 1:]).max() < 1e-12


@pytest.mark.parametrize("name", sorted(MOLECULES))
def test_gradient_matches_finite_difference(name):
    # tests/nonpbc_smoke.rs:109-150 (central FD of total_free, h = 1e-4 bohr)
    z, x = mol(name)
    r = oracle.run(z, x, **TIGHT)
    assert r["status"] == 0, r["message"]
    h = 1e-4
    fd = np.zeros_like(x)
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            ep = oracle.run(z, xp, compute_gradient=0, **TIGHT)["energy"]
            em = oracle.run(z, xm, compute_gradient=0, **TIGHT)["energy"]
            fd[a, k] = (ep - em) / (2 * h)
    assert np.abs(fd - r["gradient"]).max() < 2e-7, np.abs(fd - r["gradient"]).max()


@pytest.mark.parametrize("name", ["ferrocene", "ni_co4"])
def test_gradient_of_3d_metal_complexes_matches_finite_difference(name):
    # the 3d-metal molecules of the reference's own tests and examples: central
    # differences of the energy for the metal and two ligand atoms
    z, x, _ = system_from_xyz(reference_molecules.MOLECULES[name])
    r = oracle.run(z, x, **TIGHT)
    assert r["status"] == 0, r["message"]
    h = 1e-4
    worst = 0.0
    for a in (0, 1, len(z) - 1):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            ep = oracle.run(z, xp, compute_gradient=0, **TIGHT)["energy"]
            em = oracle.run(z, xm, compute_gradient=0, **TIGHT)["energy"]
            worst = max(worst, abs((ep - em) / (2 * h) - r["gradient"][a, k]))
    assert worst < 2e-7, worst
    assert np.abs(r["gradient"].sum(axis=0)).max() < 1e-10


def test_halogen_term_is_active_and_fd_consistent():
    z, x = mol("bromomethane_ammonia")
    e, g = oracle.halogen(z, x)
    assert e != 0.0
    h = 1e-5
    fd = np.zeros_like(x)
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            fd[a, k] = (oracle.halogen(z, xp)[0] - oracle.halogen(z, xm)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-9


@pytest.mark.parametrize("term", ["dispersion", "repulsion"])
def test_pair_term_gradients(term):
    z, x = mol("ammonia")
    f = getattr(oracle, term)
    e, g = f(z, x)
    h = 1e-5
    fd = np.zeros_like(x)
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            fd[a, k] = (f(z, xp)[0] - f(z, xm)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-9


def test_invariants():
    # physical_consistency.rs:286 (sum of gradient), :481 (rotational invariance);
    # electronic.rs:4058 (term sum); Mulliken charge conservation.
    z, x = mol("ammonia")
    r = oracle.run(z, x, want_matrices=True, **TIGHT)
    assert np.abs(r["gradient"].sum(axis=0)).max() < 1e-10
    e = r["energies"]
    comp = sum(v for k, v in e.items() if k not in ("total_free", "total_internal", "electronic_entropy"))
    assert abs(comp - e["total_internal"]) < 1e-12
    assert abs(e["total_free"] - e["total_internal"] - e["electronic_entropy"]) < 1e-12
    assert abs(r["atomic_charges"].sum()) < 1e-10
    assert abs(np.sum(r["density"] * r["overlap"]) - r["nelec"]) < 1e-10
    # Tr(W S) = sum f eps
    assert abs(np.sum(r["energy_weighted_density"] * r["overlap"])
               - np.sum(r["occupations"] * r["orbital_energies"])) < 1e-10
    th = 0.7
    rot = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    rot = rot @ np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    r2 = oracle.run(z, x @ rot.T + 1.3, **TIGHT)
    assert abs(r2["energy"] - r["energy"]) < 1e-10
    # covariance of the gradient (physical_consistency.rs:514)
    assert np.abs(r2["gradient"] - r["gradient"] @ rot.T).max() < 1e-9


def test_fixed_point_is_mixer_independent():
    # tests/scf_accelerators.rs:42-66: linear mixing reaches the Broyden energy
    z, x = mol("water")
    a = oracle.run(z, x, **TIGHT)
    b = oracle.run(z, x, scc_broyden=0, max_scc=500, **TIGHT)
    assert a["status"] == 0 and b["status"] == 0
    assert abs(a["energy"] - b["energy"]) < 1e-9
    assert b["iterations"] > a["iterations"]


def test_fallback_ladder_recovers():
    # electronic.rs:1390-1447: rung 0 with one iteration fails, rung 1 (linear
    # mixing + level shift) converges to the same fixed point
    z, x = mol("water")
    ref = oracle.run(z, x, **TIGHT)
    r = oracle.run(z, x, max_scc=1, **TIGHT)
    assert r["status"] == 0 and r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-8
    assert r["electronic_temperature"] == 300.0


def test_error_paths():
    z, x = mol("water")
    # two hydrogens 1e-7 bohr apart carry identical AOs: the overlap has an
    # eigenvalue below eigen_tolerance (linalg.rs:204-208)
    r = oracle.run([1, 1, 8], np.array([[0, 0, 0], [0, 0, 1e-7], [1.8, 0, 0]]))
    assert r["status"] == 4 and "positive definite" in r["message"]
    x3 = x.copy()
    x3[0, 0] = float("nan")
    assert oracle.run(z, x3)["status"] != 0
    assert oracle.run(z, x, charge=40.0)["status"] != 0
    # SCC that cannot converge: every rung exhausted
    r = oracle.run([1, 1], np.array([[0, 0, 0], [1.4, 0, 0]]), max_scc=2, energy_tolerance=-1.0)
    assert r["status"] == 6 and not r["converged"]


def test_eigensolver_matches_lapack():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 60, 131):
        a = rng.standard_normal((n, n))
        a = a + a.T
        w, v = oracle.eigh(a)
        wl = np.linalg.eigh(a)[0]
        assert np.abs(w - wl).max() < 1e-11 * max(1.0, np.abs(wl).max())
        assert np.abs(a @ v - v * w).max() < 1e-11 * max(1.0, np.abs(wl).max())
        assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12
    # degenerate spectrum
    q = np.linalg.qr(rng.standard_normal((20, 20)))[0]
    d = np.repeat([1.0, 2.0, 3.0, 4.0], 5)
    w, v = oracle.eigh(q @ np.diag(d) @ q.T)
    assert np.abs(w - d).max() < 1e-12


# ---- overlap integrals against numerical quadrature of the STO-nG functions
def _sto(tag_type, ng, zeta, l):
    """Normalised primitives from the Stewart table (sto_ng.txt) written from the
    Gaussian normalisation N = (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!)."""
    path = os.path.join(oracle.DATA_DIR, "sto_ng.txt")
    for line in open(path):
        if line.startswith("#"):
            continue
        f = line.split()
        if f[0] == str(tag_type) and int(f[1]) == ng:
            v = [float(t) for t in f[2:]]
            a = np.array(v[:ng]) * zeta ** 2
            c = np.array(v[ng:])
            c = c * (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2.0) / math.sqrt([1, 1, 3][l])
            return a, c
    raise KeyError(tag_type)


def _radial(a, c, r2):
    return sum(ci * np.exp(-ai * r2) for ai, ci in zip(a, c))


def test_sto_normalisation_and_overlap_quadrature():
    # H 1s: STO-4G, type 1, zeta 1.207940 ; O 2p: STO-6G, type 6 (2p), zeta 2.153060
    ah, ch = _sto(1, 4, 1.207940, 0)
    ao, co = _sto(6, 6, 2.153060, 1)
    # <1s|1s> = 1 in spherical coordinates
    n1 = integrate.quad(lambda r: 4 * math.pi * r * r * _radial(ah, ch, r * r) ** 2, 0, 30, limit=200)[0]
    assert abs(n1 - 1.0) < 1e-9
    # <pz|pz> = 1 : radial part r * R(r), angular factor 4 pi / 3
    n2 = integrate.quad(lambda r: 4 * math.pi / 3 * r ** 4 * _radial(ao, co, r * r) ** 2, 0, 30, limit=200)[0]
    assert abs(n2 - 1.0) < 1e-9
    # <H 1s at origin | O pz at z = R> by cylindrical quadrature vs the oracle
    R = 1.8
    def integrand(rho, zc):
        r2h = rho * rho + zc * zc
        r2o = rho * rho + (zc - R) ** 2
        return 2 * math.pi * rho * _radial(ah, ch, r2h) * (zc - R) * _radial(ao, co, r2o)
    val = integrate.dblquad(lambda rho, zc: integrand(rho, zc), -12, 14, 0, 14, epsabs=1e-12, epsrel=1e-10)[0]
    S = oracle.overlap([1, 8], np.array([[0, 0, 0], [0, 0, R]]))
    # AO order: H 1s, H 2s, O 2s, O px, py, pz
    assert abs(S[0, 5] - val) < 1e-9, (S[0, 5], val)
    # Stewart's 10-digit expansions normalise to 1 within 1e-9; like the
    # reference, contracted shells are not renormalised (sto.rs:867-873)
    assert abs(S[0, 0] - 1.0) < 1e-9 and abs(S[5, 5] - 1.0) < 1e-9
    assert abs(S[0, 0] - n1) < 1e-12
    # H 2s is orthogonalised against H 1s on the same centre (basis.rs:318-344);
    # the residue is <1s|2s'> = ov (1 - <1s|1s>) / |2s - ov 1s|, i.e. the 1s
    # normalisation defect above amplified by the near-linear-dependence
    assert abs(S[0, 1]) < 2e-9 and abs(S[1, 1] - 1.0) < 1e-12


def test_overlap_derivative_matches_fd():
    # integrals.rs:1307 (GFN1-basis overlap derivative vs FD, 1e-7)
    z, x = mol("phosphine")
    S = oracle.overlap(z, x)
    rng = np.random.default_rng(0)
    W = rng.standard_normal(S.shape)
    _, dS = oracle.overlap(z, x, weights=W)
    h = 1e-5
    for a in range(len(z)):
        for k in range(3):
            xp = x.copy(); xp[a, k] += h
            xm = x.copy(); xm[a, k] -= h
            fd = (np.sum(W * oracle.overlap(z, xp)) - np.sum(W * oracle.overlap(z, xm))) / (2 * h)
            assert abs(fd - dS[a, k]) < 1e-8


def _fixtures():
    import json
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_fixtures.json")) as f:
        return json.load(f)


def test_oracle_reproduces_committed_fixtures():
    """Regression pin: the oracle still produces the committed fixtures
    (tests/golden/make_golden.py)."""
    fx = _fixtures()
    for name, s in fx["systems"].items():
        r = oracle.run(np.array(s["numbers"]), np.array(s["positions"]), **fx["tolerances"])
        assert r["status"] == 0, name
        assert abs(r["energy"] - s["energy"]) <= 1e-10, name
        assert np.abs(r["gradient"] - np.array(s["gradient"])).max() <= 1e-9, name
        assert np.abs(r["shell_charges"] - np.array(s["shell_charges"])).max() <= 1e-8, name
