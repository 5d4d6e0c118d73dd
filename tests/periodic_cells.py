"""Periodic cells of the reference's own tests, as extended-XYZ text (Angstrom).

diamond: tests/pbc_integration.rs:18-27 (ideal diamond, an equilibrium by
symmetry: the reference asserts convergence, |sum q| < 1e-6 and a largest
gradient component below 1e-4, :30-80); nacl: src/pbc/ewald.rs:341; lih40:
src/pbc/ewald.rs:399; h2_30: src/pbc/ewald.rs:242; ni2 and h2chain:
tests/pbc_finite_t.rs:14-20 (run there at finite electronic temperature
without dispersion); diamond_skew and bn_skew: src/pbc/gamma_response.rs:655-668
(two-atom cells with skewed lattice vectors).
"""

CELLS = {
    "diamond": ("8\nLattice=\"3.567 0 0 0 3.567 0 0 0 3.567\" pbc=\"T T T\"\n"
                "C 0.000000 0.000000 0.000000\nC 0.891750 0.891750 0.891750\n"
                "C 0.000000 1.783500 1.783500\nC 0.891750 2.675250 2.675250\n"
                "C 1.783500 0.000000 1.783500\nC 2.675250 0.891750 2.675250\n"
                "C 1.783500 1.783500 0.000000\nC 2.675250 2.675250 0.891750\n"),
    "nacl": "2\nLattice=\"5.6 0 0 0 5.6 0 0 0 5.6\" pbc=\"T T T\"\nNa 0 0 0\nCl 2.8 2.8 2.8\n",
    "lih40": "2\nLattice=\"40 0 0 0 40 0 0 0 40\" pbc=\"T T T\"\nLi 0 0 0\nH 0 0 1.6\n",
    "h2_30": "2\nLattice=\"30 0 0 0 30 0 0 0 30\" pbc=\"T T T\"\nH 0 0 0\nH 0 0 0.74\n",
    "ni2": ("2\nLattice=\"3.52 0 0 0 3.52 0 0 0 3.52\" pbc=\"T T T\"\n"
            "Ni 0.120000 0.040000 0.000000\nNi 1.760000 1.760000 1.760000\n"),
    "h2chain": ("2\nLattice=\"4.2 0 0 0 12 0 0 0 12\" pbc=\"T T T\"\n"
                "H 0.000000 0.300000 0.000000\nH 0.950000 -0.200000 0.000000\n"),
    "diamond_skew": ("2\nLattice=\"0.06 1.83 1.75 1.75 0.04 1.81 1.82 1.76 0.03\" pbc=\"T T T\"\n"
                     "C 0.000000 0.000000 0.000000\nC 0.930000 0.880000 0.905000\n"),
    "bn_skew": ("2\nLattice=\"0.06 1.86 1.78 1.78 0.04 1.84 1.85 1.79 0.03\" pbc=\"T T T\"\n"
                "B 0.000000 0.000000 0.000000\nN 0.940000 0.890000 0.920000\n"),
}

# tests/optimizer.rs:27-80: the diamond cell with atom 0 displaced about 0.15 A,
# relaxed at fixed cell (max_iterations 20, gradient_tolerance 1e-3)
DIAMOND_PERTURBED = ("8\nLattice=\"3.567 0 0 0 3.567 0 0 0 3.567\" pbc=\"T T T\"\n"
                     "C 0.150000 0.080000 0.030000\nC 0.891750 0.891750 0.891750\n"
                     "C 0.000000 1.783500 1.783500\nC 0.891750 2.675250 2.675250\n"
                     "C 1.783500 0.000000 1.783500\nC 2.675250 0.891750 2.675250\n"
                     "C 1.783500 1.783500 0.000000\nC 2.675250 2.675250 0.891750\n")

# the reference's finite-temperature options for the metallic cells (tests/pbc_finite_t.rs:26-35)
SMEARED = {"ni2": dict(enable_dispersion=0, electronic_temperature=3000.0, max_scc=500),
           "h2chain": dict(enable_dispersion=0, electronic_temperature=3000.0, max_scc=500)}
