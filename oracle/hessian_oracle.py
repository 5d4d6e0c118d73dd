"""CPU Hessian (and dipole derivatives) by central differences of the oracle's
analytic gradient (and Mulliken dipole).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
Hessian.  The product never imports it.

H[:, c] = (g(x + h e_c) - g(x - h e_c)) / 2h at tight SCC tolerances, the
construction the reference validates its analytic Hessian against
(tests/hessian.rs:691-760), then symmetrised.  Pinned by
tests/test_hessian_oracle.py against second differences of the energy (a
separate code path), the translational sum rule, symmetry, and the
reference's own vibrational-analysis unit tests (vibrational.rs:88-129); the
dipole derivatives by the translational (sum of Born charges = total charge)
and rotational (sum_B dmu/dR_B . (n x R_B) = n x mu) sum rules.
"""
import numpy as np

import oracle as _oracle


def gradient_terms(include_repulsion=True, include_electronic=True, include_dispersion=True, include_halogen=True):
    return dict(include_repulsion=int(include_repulsion), include_hamiltonian=int(include_electronic),
                include_scc=int(include_electronic), include_dispersion=int(include_dispersion),
                include_halogen=int(include_halogen))


def fd_hessian(numbers, positions_bohr, **kw):
    """(symmetrised Hessian, max asymmetry before symmetrisation)."""
    h, asym, _, _ = fd_hessian_dipoles(numbers, positions_bohr, **kw)
    return h, asym


def fd_hessian_dipoles(numbers, positions_bohr, step=1.0e-4, energy_tolerance=1.0e-11, charge_tolerance=1.0e-11,
                       charge=0.0, warm_start=True, cell=None, **terms):
    """(Hessian, asymmetry, dipole derivatives [3 nat, 3] = d mu_alpha / d R_c by
    the same central differences, Mulliken dipole of the undisplaced geometry).
    cell: the cell vectors (rows, bohr) of a periodic system."""
    x0 = np.asarray(positions_bohr, dtype=np.float64).reshape(-1)
    ndof = x0.size
    opts = dict(energy_tolerance=energy_tolerance, charge_tolerance=charge_tolerance, charge=charge, cell=cell)
    opts.update(gradient_terms(**terms))
    base = _oracle.run(numbers, x0.reshape(-1, 3), **opts)
    if base["status"] != 0:
        raise RuntimeError(base["message"])
    q0 = base["shell_charges"] if warm_start else None
    h = np.zeros((ndof, ndof))
    ddip = np.zeros((ndof, 3))
    for c in range(ndof):
        xp = x0.copy()
        xp[c] += step
        xm = x0.copy()
        xm[c] -= step
        gp = _oracle.run(numbers, xp.reshape(-1, 3), initial_shell_charges=q0, **opts)
        gm = _oracle.run(numbers, xm.reshape(-1, 3), initial_shell_charges=q0, **opts)
        if gp["status"] or gm["status"]:
            raise RuntimeError("displaced evaluation failed")
        h[:, c] = (gp["gradient"].reshape(-1) - gm["gradient"].reshape(-1)) / (2.0 * step)
        ddip[c] = (gp["dipole"] - gm["dipole"]) / (2.0 * step)
    asym = float(np.abs(h - h.T).max()) if ndof else 0.0
    return 0.5 * (h + h.T), asym, ddip, base["dipole"]
