"""Pins of the oracle's heteroatomic terms, which the reference's two H-H golden
vectors (D3 of H2, examples/h2.engrad) leave open (CPU only).

Each check is chosen so that a plausible restatement mistake fails it:
- the parameter parser against an independent field-name reading of the
  parameter files (a wrong column or a misassigned shell fails);
- H0 off-diagonal elements of O-H and C-N diatomics against hand values from
  the file and the published Pauling electronegativities and xTB atomic radii,
  and ratios under single-parameter edits of the file (enscale sign and the
  Delta EN^2 dependence, pair scaling lookup, ksp / kdiff assignment, the shell
  polynomial), hamiltonian.rs:91-166, basis.rs:281-293;
- gamma: the same-atom harmonic mean of the LPAR-scaled hardnesses and the 1/R
  limit (coulomb.rs:82-107, :239-248, params.rs:153-160);
- CN: exactly 1/2 at R = R_cov,A + R_cov,B with slope -4/R_cov there
  (coordination.rs:96-106, data_tables.rs:208-219);
- repulsion of O-H and C-N by hand (repulsion.rs:139-159);
- D3(BJ): finite at R -> 0 and C6/R^6 at R -> infinity (dispersion.rs:1271-1296);
- halogen bond: zero without donor or acceptor, the angular factor
  ((1 - cos theta)/2)^6 and the value at R = R0 (halogen.rs:64-160, :716-784).
"""
import math
import os
import re

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "gfn1-rs_amd", "data")
PARAM = os.path.join(DATA, "param_gfn1-xtb.txt")
PARAM_SI = os.path.join(DATA, "param_gfn1-si-xtb.txt")
AA = 1.8897261246257702          # system.rs:8
EV = 1.0 / 27.21138505           # basis.rs:18
# Pauling electronegativities (published values; data_tables.rs:165-173)
EN = {1: 2.20, 6: 2.55, 7: 3.04, 8: 3.44}
# xTB atomic radii in Angstrom (xtb's atomicRad table; data_tables.rs:14-21)
RAD = {1: 0.32, 6: 0.75, 7: 0.71, 8: 0.64, 35: 1.17}
# Pyykko-Atsumi 2009 single-bond covalent radii, Angstrom (data_tables.rs:29-37)
RCOV = {1: 0.32, 6: 0.75, 7: 0.71, 8: 0.63}
LSYM = "spdf"


def read_param_file(path):
    """Independent reader of the xtb parameter format: globals and pair
    factors by name, element records by field name."""
    text = open(path).read()
    glob, pairs, elements = {}, {}, {}
    for block in re.findall(r"\$globpar\n(.*?)\$end", text, re.S):
        for line in block.splitlines():
            parts = line.split()
            if len(parts) == 2:
                glob[parts[0].lower()] = float(parts[1])
    for block in re.findall(r"\$pairpar\n(.*?)\$end", text, re.S):
        for line in block.splitlines():
            parts = line.split()
            if len(parts) == 3:
                a, b = sorted((int(parts[0]), int(parts[1])))
                pairs[(a, b)] = float(parts[2])
    for z, block in re.findall(r"\$Z=\s*(\d+)[^\n]*\n(.*?)\$end", text, re.S):
        rec = {}
        for line in block.splitlines():
            if "=" not in line:
                continue
            key, val = line.split("=", 1)
            key = key.strip().upper()
            if key == "AO":
                rec["AO"] = [(int(n), LSYM.index(l.lower())) for n, l in re.findall(r"(\d+)([spdfSPDF])", val)]
            else:
                rec[key] = [float(v) for v in val.split()]
        elements[int(z)] = rec
    return glob, pairs, elements


GLOB, PAIRS, ELEMENTS = read_param_file(PARAM)


def shell_poly(rec, l):
    return rec.get("POLY" + "SPDF"[l], [0.0])[0]


def shell_lpar(rec, l):
    return rec.get("LPAR" + "SPDF"[l], [0.0])[0]


@pytest.mark.parametrize("path", [PARAM, PARAM_SI], ids=["gfn1", "gfn1-si"])
def test_parser_matches_an_independent_reading(path):
    glob, pairs, elements = read_param_file(path)
    pf = None if path == PARAM else path
    for k, v in glob.items():
        assert oracle.global_value(k, param_file=pf) == v, k
    for (a, b), v in pairs.items():
        assert oracle.pair_scaling(a, b, param_file=pf) == v
        assert oracle.pair_scaling(b, a, param_file=pf) == v
    # a pair without an entry (both main group) scales by 1
    assert oracle.pair_scaling(6, 8, param_file=pf) == 1.0
    assert len(elements) >= 86
    for z, rec in elements.items():
        ev = lambda key, sh=0: oracle.element_value(z, key, sh, param_file=pf)  # noqa: E731
        assert ev("present") == 1.0
        assert ev("nshell") == len(rec["AO"]), z
        for key in ("GAM", "GAM3", "REPA", "REPB", "CXB"):
            assert ev(key.lower()) == rec.get(key, [0.0])[0], (z, key)
        for i, (n, l) in enumerate(rec["AO"]):
            assert ev("n", i) == n and ev("l", i) == l, (z, i)
            assert ev("lev", i) == rec["LEV"][i], (z, i)
            assert ev("exp", i) == rec["EXP"][i], (z, i)
            assert ev("poly", i) == shell_poly(rec, l), (z, i)
            assert ev("lpar", i) == shell_lpar(rec, l), (z, i)


def diatomic(za, zb, r):
    return np.array([za, zb]), np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])


def h0_over_s(za, zb, r, param_file=None, cn=0):
    """H0/S of the off-atom AO pairs (first AO of each shell of A, first AO
    of each shell of B, with B on the x axis so that p_x carries the overlap),
    keyed by (shell of A, shell of B)."""
    z, x = diatomic(za, zb, r)
    res = oracle.run(z, x, want_matrices=True, enable_cn_hamiltonian=cn, max_scc=3, param_file=param_file)
    sh = oracle.shells(z, param_file=param_file)
    first = np.concatenate([[0], np.cumsum(sh["nao"])[:-1]])
    out = {}
    for i in np.where(sh["atom"] == 0)[0]:
        for j in np.where(sh["atom"] == 1)[0]:
            mu, nu = first[i], first[j]  # s, or p_x of a p shell
            s = res["overlap"][mu, nu]
            assert abs(s) > 1e-3, (i, j)
            out[(int(i), int(j))] = res["h0"][mu, nu] / s
    return out, sh, res


def hand_h0_over_s(za, zb, r, glob=GLOB, pairs=PAIRS):
    """hamiltonian.rs:91-166 from the file values: 1/2 (h_i + h_j) K poly_i poly_j."""
    ra, rb = ELEMENTS[za], ELEMENTS[zb]
    kl = {0: glob["ks"], 1: glob["kp"], 2: glob["kd"]}

    def kpair(li, lj):
        return glob["ksp"] if {li, lj} == {0, 1} else 0.5 * (kl[li] + kl[lj])

    rr = math.sqrt(r / ((RAD[za] + RAD[zb]) * AA))
    out = {}
    shells_a = [(i, n, l, l not in [ll for _, ll in ra["AO"][:i]]) for i, (n, l) in enumerate(ra["AO"])]
    shells_b = [(j, n, l, l not in [ll for _, ll in rb["AO"][:j]]) for j, (n, l) in enumerate(rb["AO"])]
    for i, _, li, vi in shells_a:
        for j, _, lj, vj in shells_b:
            if vi and vj:
                k = pairs.get(tuple(sorted((za, zb))), 1.0) * kpair(li, lj) * \
                    (1.0 + glob["enscale"] * 0.01 * (EN[za] - EN[zb]) ** 2)
            elif vi:
                k = 0.5 * (kpair(li, li) + glob["kdiff"])
            elif vj:
                k = 0.5 * (kpair(lj, lj) + glob["kdiff"])
            else:
                k = glob["kdiff"]
            h = 0.5 * (ra["LEV"][i] + rb["LEV"][j]) * EV
            poly = (1.0 + 0.01 * shell_poly(ra, li) * rr) * (1.0 + 0.01 * shell_poly(rb, lj) * rr)
            out[(i, len(ra["AO"]) + j)] = h * k * poly
    return out


@pytest.mark.parametrize("pair", [(8, 1), (6, 7)], ids=["O-H", "C-N"])
def test_h0_off_atom_elements_by_hand(pair):
    za, zb = pair
    r = 2.1
    got, _, _ = h0_over_s(za, zb, r)
    want = hand_h0_over_s(za, zb, r)
    assert set(got) == set(want)
    for k in want:
        assert abs(got[k] - want[k]) <= 1e-12 * abs(want[k]), (k, got[k], want[k])


def edited(tmp_path, name, edit):
    text = open(PARAM).read()
    new = edit(text)
    assert new != text
    p = tmp_path / name
    p.write_text(new)
    return str(p)


def test_h0_en_dependence_and_sign(tmp_path):
    """With enscale = 0 the valence-valence O-H elements grow by exactly
    1 / (1 - 0.007 (EN_O - EN_H)^2); elements with H's second (non-valence)
    s shell do not change.  A flipped sign or a wrong EN entry fails."""
    pf = edited(tmp_path, "enscale0.txt", lambda t: re.sub(r"enscale\s+-0\.70000", "enscale     0.00000", t))
    base, sh, _ = h0_over_s(8, 1, 2.1)
    mod, _, _ = h0_over_s(8, 1, 2.1, param_file=pf)
    factor = 1.0 - 0.007 * (3.44 - 2.20) ** 2
    for (i, j), v in base.items():
        ratio = v / mod[(i, j)]
        if sh["valence"][i] and sh["valence"][j]:
            assert abs(ratio - factor) < 1e-13
        else:
            assert abs(ratio - 1.0) < 1e-13
    assert factor < 1.0  # enscale < 0 shrinks the coupling of unlike atoms


def test_h0_pair_scaling_lookup(tmp_path):
    """N-H carries the pairpar entry 1.04 in either atom order; removing it
    scales the valence-valence elements by exactly 1/1.04."""
    pf = edited(tmp_path, "nopair.txt", lambda t: t.replace("  1  7   1.0400\n", ""))
    for za, zb in ((7, 1), (1, 7)):
        base, sh, _ = h0_over_s(za, zb, 1.9)
        mod, _, _ = h0_over_s(za, zb, 1.9, param_file=pf)
        for (i, j), v in base.items():
            want = 1.04 if (sh["valence"][i] and sh["valence"][j]) else 1.0
            assert abs(v / mod[(i, j)] - want) < 1e-13


def test_h0_ksp_kdiff_and_polynomial(tmp_path):
    """Single edits of ksp, kdiff and O's POLYS change exactly the elements
    the reference's hscale and shell polynomial assign them to."""
    base, sh, _ = h0_over_s(8, 1, 2.1)
    # shells: O 2s (0), O 2p (1), H 1s (2, valence), H 2s (3, not valence)
    assert list(sh["l"]) == [0, 1, 0, 0] and list(sh["valence"]) == [True, True, True, False]
    pf = edited(tmp_path, "ksp.txt", lambda t: t.replace("ksp         2.08000", "ksp         3.00000"))
    mod, _, _ = h0_over_s(8, 1, 2.1, param_file=pf)
    assert abs(mod[(1, 2)] / base[(1, 2)] - 3.0 / 2.08) < 1e-13       # O 2p - H 1s uses ksp
    for k in ((0, 2), (0, 3), (1, 3)):
        assert abs(mod[k] / base[k] - 1.0) < 1e-13
    pf = edited(tmp_path, "kdiff.txt", lambda t: t.replace("kdiff       2.85000", "kdiff       1.00000"))
    mod, _, _ = h0_over_s(8, 1, 2.1, param_file=pf)
    assert abs(mod[(0, 3)] / base[(0, 3)] - (1.85 + 1.0) / (1.85 + 2.85)) < 1e-13  # s - nonvalence s
    assert abs(mod[(1, 3)] / base[(1, 3)] - (2.25 + 1.0) / (2.25 + 2.85)) < 1e-13  # p - nonvalence s
    for k in ((0, 2), (1, 2)):
        assert abs(mod[k] / base[k] - 1.0) < 1e-13
    pf = edited(tmp_path, "polys.txt", lambda t: t.replace("POLYS=  -13.729047", "POLYS=    0.000000"))
    mod, _, _ = h0_over_s(8, 1, 2.1, param_file=pf)
    rr = math.sqrt(2.1 / ((0.64 + 0.32) * AA))
    for k, v in base.items():
        want = (1.0 - 0.13729047 * rr) if sh["l"][k[0]] == 0 else 1.0
        assert abs(v / mod[k] - want) < 1e-12, k


def test_h0_diagonal_self_energies_and_cn_shift():
    """Diagonal H0 = level (eV -> Eh) without the CN term and
    level (1 + 0.01 k_l CN) with it (k_s = cns, k_p = cnp for H and O)."""
    z, x = diatomic(8, 1, 1.85)
    sh = oracle.shells(z)
    first = np.concatenate([[0], np.cumsum(sh["nao"])[:-1]])
    cn = oracle.coordination(z, x)
    for cn_on in (0, 1):
        res = oracle.run(z, x, want_matrices=True, enable_cn_hamiltonian=cn_on, max_scc=3)
        for s in range(len(sh["l"])):
            zz = int(z[sh["atom"][s]])
            idx = sum(1 for t in range(s) if sh["atom"][t] == sh["atom"][s])
            lev = ELEMENTS[zz]["LEV"][idx] * EV
            fac = {0: GLOB["cns"], 1: GLOB["cnp"]}[int(sh["l"][s])]
            want = lev * (1.0 + 0.01 * fac * cn[sh["atom"][s]] * cn_on)
            mu = first[s]
            assert abs(res["h0"][mu, mu] / res["overlap"][mu, mu] - want) < 1e-13


@pytest.mark.parametrize("pair", [(8, 1), (6, 7)], ids=["O-H", "C-N"])
def test_gamma_limits_and_hardness(pair):
    za, zb = pair
    eta = {}
    for z in pair:
        rec = ELEMENTS[z]
        eta[z] = [rec["GAM"][0] * (1.0 + 0.1 * shell_lpar(rec, l)) for _, l in rec["AO"]]
    sh = oracle.shells(np.array(pair))
    etas = eta[za] + eta[zb]
    assert np.allclose(sh["hardness"], etas, rtol=0, atol=1e-15)
    for r in (2.3, 1e7):
        z, x = diatomic(za, zb, r)
        g = oracle.gamma(z, x)
        for i in range(len(etas)):
            for j in range(len(etas)):
                h = 2.0 / (1.0 / etas[i] + 1.0 / etas[j])
                if sh["atom"][i] == sh["atom"][j]:
                    assert abs(g[i, j] - h) < 1e-15        # same atom: harmonic mean
                else:
                    assert abs(g[i, j] - 1.0 / math.sqrt(r * r + 1.0 / h ** 2)) < 1e-15
                    if r > 1e6:
                        assert abs(g[i, j] * r - 1.0) < 1e-12   # bare Coulomb at long range
    # unlike hardnesses: the mean differs from both (a swapped operand would not)
    assert len(set(np.round(etas, 12))) > 2


@pytest.mark.parametrize("pair", [(8, 1), (6, 7), (1, 1)], ids=["O-H", "C-N", "H-H"])
def test_cn_half_at_covalent_distance(pair):
    rc = 4.0 / 3.0 * (RCOV[pair[0]] + RCOV[pair[1]]) * AA
    z, x = diatomic(*pair, rc)
    cn = oracle.coordination(z, x)
    assert np.abs(cn - 0.5).max() < 1e-15
    h = 1e-5
    d = (oracle.coordination(*diatomic(*pair, rc + h))[0] - oracle.coordination(*diatomic(*pair, rc - h))[0]) / (2 * h)
    assert abs(d - (-4.0 / rc)) < 1e-8


@pytest.mark.parametrize("pair", [(8, 1), (6, 7)], ids=["O-H", "C-N"])
def test_repulsion_by_hand(pair):
    za, zb = pair
    for r in (1.8, 2.6, 4.0):
        e, g = oracle.repulsion(*diatomic(za, zb, r))
        a, b = ELEMENTS[za], ELEMENTS[zb]
        want = a["REPB"][0] * b["REPB"][0] * math.exp(-math.sqrt(a["REPA"][0] * b["REPA"][0]) * r ** GLOB["kexp"]) / r
        assert abs(e - want) < 1e-14 * max(1.0, abs(want))
    # beyond the 25 bohr cutoff (model.rs:27) the pair is dropped
    assert oracle.repulsion(*diatomic(za, zb, 25.5))[0] == 0.0


def test_d3_bj_limits():
    """E_AB = -C6 (s6/(R^6 + R0^6) + s8 r/(R^8 + R0^8)), R0 = a1 sqrt(r) + a2:
    finite as R -> 0, -C6/R^6 as R -> infinity; the damping form (with the
    file's a1, a2, s8) predicts intermediate distances from a fit of C6 and
    r = 3 r4r2_A r4r2_B at two of them."""
    e1 = oracle.dispersion(*diatomic(8, 1, 1e-3))[0]
    e2 = oracle.dispersion(*diatomic(8, 1, 2e-3))[0]
    assert e1 < 0 and np.isfinite(e1) and abs(e1 / e2 - 1.0) < 1e-12
    a1, a2, s8 = GLOB["a1"], GLOB["a2"], GLOB["s8"]

    def form(r, c6, rr):
        r0 = a1 * math.sqrt(rr) + a2
        return -c6 * (1.0 / (r ** 6 + r0 ** 6) + s8 * rr / (r ** 8 + r0 ** 8))

    ra, rb = 14.0, 22.0
    ea, eb = (oracle.dispersion(*diatomic(8, 1, r))[0] for r in (ra, rb))
    from scipy.optimize import brentq

    def resid(rr):
        return form(rb, ea / form(ra, 1.0, rr), rr) - eb

    rr = brentq(resid, 1.0, 60.0, xtol=1e-14)
    c6 = ea / form(ra, 1.0, rr)
    assert c6 > 0 and rr > 0
    for r in (16.0, 18.0, 30.0, 45.0):
        e = oracle.dispersion(*diatomic(8, 1, r))[0]
        assert abs(e / form(r, c6, rr) - 1.0) < 1e-6, r
    # long range: -E R^6 / C6 -> 1 + s8 r / R^2 (the damping terms are O((R0/R)^6))
    e = oracle.dispersion(*diatomic(8, 1, 49.0))[0]
    assert abs(-e * 49.0 ** 6 / (c6 * (1.0 + s8 * rr / 49.0 ** 2)) - 1.0) < 5e-5
    # beyond the 50 bohr pair cutoff nothing is left
    assert oracle.dispersion(*diatomic(8, 1, 50.5))[0] == 0.0


def halogen_geometry(theta_deg, r_jx, acceptor=7, donor=35):
    th = math.radians(theta_deg)
    z = np.array([1, donor, acceptor])
    x = np.array([[-2.7, 0.0, 0.0], [0.0, 0.0, 0.0], [-r_jx * math.cos(th), r_jx * math.sin(th), 0.0]])
    return z, x


def test_halogen_angular_factor_and_value_at_r0():
    """Donor Br with its nearest neighbour H and acceptor N: the energy is
    0.1 CXB ((1 - cos theta)/2)^6 (t^12 - xbdamp t^6)/(1 + t^12), t = R0/R,
    R0 = xbrad (r_Br + r_N), theta the H-Br...N angle."""
    r0 = GLOB["xbrad"] * (RAD[35] + RAD[7]) * AA
    cc = 0.1 * ELEMENTS[35]["CXB"][0]
    e_lin = oracle.halogen(*halogen_geometry(180.0, r0))[0]
    assert abs(e_lin - cc * (1.0 - GLOB["xbdamp"]) / 2.0) < 1e-15
    e_90 = oracle.halogen(*halogen_geometry(90.0, r0))[0]
    assert abs(e_90 / e_lin - 1.0 / 64.0) < 1e-13
    for theta in (120.0, 150.0):
        e = oracle.halogen(*halogen_geometry(theta, 1.3 * r0))[0]
        t = 1.0 / 1.3
        want = cc * ((1.0 - math.cos(math.radians(theta))) / 2.0) ** 6 * (t ** 12 - GLOB["xbdamp"] * t ** 6) / (1 + t ** 12)
        assert abs(e - want) < 1e-15


def test_halogen_needs_donor_and_acceptor():
    r = 5.0
    # no acceptor (C is not one of N, O, P, S)
    assert oracle.halogen(*halogen_geometry(180.0, r, acceptor=6))[0] == 0.0
    # no halogen donor: Cl and F have no CXB entry; H is not a halogen
    assert "CXB" not in ELEMENTS[17] and "CXB" not in ELEMENTS[9]
    assert oracle.halogen(*halogen_geometry(180.0, r, donor=17))[0] == 0.0
    assert oracle.halogen(*halogen_geometry(180.0, r, donor=6))[0] == 0.0
    # acceptor O gives a term
    assert oracle.halogen(*halogen_geometry(180.0, r, acceptor=8))[0] != 0.0
