"""Reference-shaped Python API over the C ABI.

Mirrors the reference's run_electronic / analytic_gradient entry points
(src/electronic.rs:800, src/gradient.rs:56), their option and result structs
(electronic.rs:157-452, :600-712; gradient.rs:21-54) and the Python calculator
(src/python.rs:115-381, :2845-2935).
"""
import ctypes
import os
from dataclasses import dataclass, field, fields, replace
from typing import Dict, Optional

import numpy as np

from . import _lib

ANGSTROM_TO_BOHR = 1.8897261246257702        # system.rs:8
HARTREE_TO_EV = 27.211386245988              # constants.rs:17
FORCE_AU_TO_EV_PER_ANGSTROM = HARTREE_TO_EV * ANGSTROM_TO_BOHR


class Gfn1Parameters:
    """Loaded parameter set (params.rs:165-230); one device copy per GPU."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def resolve(cls, spec: Optional[str] = None):
        lib = _lib.load()
        h = ctypes.c_void_p()
        err = _lib.Gfn1Error()
        _lib.check(lib.gfn1_params_load(spec.encode() if spec else None, ctypes.byref(h), ctypes.byref(err)), err)
        return cls(h)

    @classmethod
    def builtin(cls):
        return cls.resolve("builtin")

    @classmethod
    def from_str(cls, text: str):
        """Gfn1Parameters::from_str (params.rs:272): the text of a parameter file."""
        lib = _lib.load()
        h = ctypes.c_void_p()
        err = _lib.Gfn1Error()
        _lib.check(lib.gfn1_params_from_text(text.encode(), ctypes.byref(h), ctypes.byref(err)), err)
        return cls(h)

    @classmethod
    def builtin_si(cls):
        """The bundled GFN1(Si)-xTB reparametrization (params.rs:192-196)."""
        return cls.resolve("builtin:si")

    def source_description(self) -> str:
        """Name, source and references (params.rs:236-270)."""
        lib = _lib.load()
        n = lib.gfn1_params_source(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib.gfn1_params_source(self._h, buf, n + 1)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.gfn1_params_free(self._h)
            self._h = None


@dataclass
class ElectronicOptions:
    """In-scope ElectronicOptions with the reference defaults (electronic.rs:395-452)."""
    charge: Optional[float] = None
    max_scc: int = 250
    energy_tolerance: float = 1.0e-6
    charge_tolerance: float = 2.0e-5
    mixing: float = 0.4
    scc_broyden: bool = True
    scc_broyden_size: int = 250
    level_shift: float = 0.0
    nprim: int = 0
    eigen_tolerance: float = 1.0e-12
    electronic_temperature: float = 300.0
    enable_dispersion: bool = True
    integral_cutoff: float = 30.0
    coordination_cutoff: float = 30.0
    enable_cn_hamiltonian: bool = True
    scc_initial_shell_charges: Optional[np.ndarray] = None

    def to_c(self):
        o = _lib.ElectronicOptions()
        _lib.load().gfn1_electronic_options_default(ctypes.byref(o))
        o.has_charge = 0 if self.charge is None else 1
        o.charge = 0.0 if self.charge is None else float(self.charge)
        for name in ("max_scc", "scc_broyden_size", "nprim"):
            setattr(o, name, int(getattr(self, name)))
        for name in ("energy_tolerance", "charge_tolerance", "mixing", "level_shift", "eigen_tolerance",
                     "electronic_temperature", "integral_cutoff", "coordination_cutoff"):
            setattr(o, name, float(getattr(self, name)))
        o.scc_broyden = int(bool(self.scc_broyden))
        o.enable_dispersion = int(bool(self.enable_dispersion))
        o.enable_cn_hamiltonian = int(bool(self.enable_cn_hamiltonian))
        keep = None
        if self.scc_initial_shell_charges is not None:
            keep = np.ascontiguousarray(self.scc_initial_shell_charges, dtype=np.float64)
            o.scc_initial_shell_charges = keep.ctypes.data_as(_lib._DP)
            o.scc_initial_shell_charges_len = len(keep)
        return o, keep


@dataclass
class AnalyticGradientOptions:
    """AnalyticGradientOptions (gradient.rs:21-41)."""
    electronic: ElectronicOptions = field(default_factory=ElectronicOptions)
    include_repulsion: bool = True
    include_dispersion: bool = True
    include_hamiltonian: bool = True
    include_scc: bool = True
    include_halogen: bool = True

    def to_c(self):
        e, keep = self.electronic.to_c()
        g = _lib.GradientOptions()
        g.electronic = e
        for name in ("include_repulsion", "include_dispersion", "include_hamiltonian", "include_scc",
                     "include_halogen"):
            setattr(g, name, int(bool(getattr(self, name))))
        return g, keep


@dataclass
class EnergyTerms:
    total_free: float = 0.0
    total_internal: float = 0.0
    repulsion: float = 0.0
    electronic: float = 0.0
    isotropic_scc: float = 0.0
    third_order: float = 0.0
    higher_order: float = 0.0
    dispersion: float = 0.0
    halogen: float = 0.0
    multipole: float = 0.0
    exchange: float = 0.0
    spin_polarization: float = 0.0
    plus_u: float = 0.0
    external_field: float = 0.0
    electronic_entropy: float = 0.0

    def named_values(self):
        return [(f.name, getattr(self, f.name)) for f in fields(self)]


@dataclass
class ElectronicResult:
    """ElectronicResult (electronic.rs:600-654); moment integrals are not produced."""
    overlap: np.ndarray
    h0: np.ndarray
    fock: np.ndarray
    density: np.ndarray
    energy_weighted_density: np.ndarray
    orbital_energies: np.ndarray
    occupations: np.ndarray
    electronic_temperature: float
    fermi_level: float
    charge_order: int
    shell_charges: np.ndarray
    atomic_charges: np.ndarray
    shell_scc_potential: np.ndarray
    coordination_numbers: np.ndarray
    terms: EnergyTerms
    dipole: np.ndarray
    nelec: float
    iterations: int
    converged: bool

    @property
    def total_free(self):
        return self.terms.total_free

    @property
    def total_internal(self):
        return self.terms.total_internal

    def energy_terms(self):
        return self.terms


@dataclass
class AnalyticGradientResult:
    """AnalyticGradientResult (gradient.rs:44-54)."""
    electronic_result: ElectronicResult
    total_energy: float
    gradient: np.ndarray
    forces: np.ndarray
    electronic_gradient: np.ndarray
    repulsion_gradient: np.ndarray
    dispersion_gradient: np.ndarray
    halogen_gradient: np.ndarray
    max_gradient: float


def system_from_xyz(text: str, charge: float = 0.0, coordinates_are_bohr: bool = False, with_lattice=False):
    """PeriodicSystem::from_xyz_str (system.rs:45-99): returns (numbers,
    positions in bohr as an (nat, 3) array, charge), and with with_lattice the
    cell vectors as a fourth item (a (3, 3) array in bohr from an extended-XYZ
    Lattice="..." comment, None without one).  Raises NativeError
    (InvalidInput, Parse for a bad coordinate, SingularCell) like the
    reference's Err."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    err = _lib.Gfn1Error()
    _lib.check(lib.gfn1_system_from_xyz(text.encode(), float(charge), int(bool(coordinates_are_bohr)),
                                        ctypes.byref(h), ctypes.byref(err)), err)
    try:
        s = lib.gfn1_xyz_system(h).contents
        nat = s.nat
        numbers = np.ctypeslib.as_array(s.numbers, shape=(nat,)).copy() if nat else np.zeros(0, dtype=np.int32)
        pos = np.ctypeslib.as_array(s.positions, shape=(3 * nat,)).reshape(nat, 3).copy() if nat else np.zeros((0, 3))
        if not with_lattice:
            return numbers, pos, s.charge
        cell = np.ctypeslib.as_array(s.lattice, shape=(9,)).reshape(3, 3).copy() if s.lattice else None
        return numbers, pos, s.charge, cell
    finally:
        lib.gfn1_xyz_free(h)


def system_from_xyz_file(path, charge: float = 0.0, coordinates_are_bohr: bool = False, with_lattice=False):
    """PeriodicSystem::from_xyz_file (system.rs:35-43)."""
    with open(path) as f:
        return system_from_xyz(f.read(), charge, coordinates_are_bohr, with_lattice)


def _params(params):
    return params if params is not None else Gfn1Parameters.resolve()


def _read_electronic(lib, h):
    nat, nsh, nao = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib.gfn1_electronic_result_sizes(h, ctypes.byref(nat), ctypes.byref(nsh), ctypes.byref(nao))
    n, ns, na = nao.value, nsh.value, nat.value

    def mat(k):
        p = lib.gfn1_electronic_result_matrix(h, k)
        return np.ctypeslib.as_array(p, shape=(n * n,)).reshape(n, n).copy() if n else np.zeros((0, 0))

    def vec(k, length):
        p = lib.gfn1_electronic_result_vector(h, k)
        return np.ctypeslib.as_array(p, shape=(length,)).copy() if length else np.zeros(0)

    e = (ctypes.c_double * 15)()
    lib.gfn1_electronic_result_energy_terms(h, e)
    terms = EnergyTerms(*list(e))
    sc = lambda k: lib.gfn1_electronic_result_scalar(h, k)  # noqa: E731
    return ElectronicResult(
        overlap=mat(0), h0=mat(1), fock=mat(2), density=mat(3), energy_weighted_density=mat(4),
        orbital_energies=vec(0, n), occupations=vec(1, n), electronic_temperature=sc(0), fermi_level=sc(1),
        charge_order=int(sc(5)), shell_charges=vec(2, ns), atomic_charges=vec(3, na),
        shell_scc_potential=vec(4, ns), coordination_numbers=vec(5, na), terms=terms, dipole=vec(6, 3),
        nelec=sc(2), iterations=int(sc(3)), converged=bool(sc(4)))


def run_electronic(numbers, positions_bohr, params: Optional[Gfn1Parameters] = None,
                   options: Optional[ElectronicOptions] = None, charge: float = 0.0,
                   lattice=None) -> ElectronicResult:
    """run_electronic (electronic.rs:800): positions in bohr.  lattice: the
    three cell vectors (rows, bohr) of a fully periodic cell, which takes the
    Gamma-point path (electronic.rs:829-832)."""
    lib = _lib.load()
    p = _params(params)
    opts = options or ElectronicOptions()
    sysc, keep_sys = _lib.make_system(numbers, positions_bohr, charge, lattice)
    oc, keep_q = opts.to_c()
    h = ctypes.c_void_p()
    err = _lib.Gfn1Error()
    rc = lib.gfn1_run_electronic(ctypes.byref(sysc), p._h, ctypes.byref(oc), ctypes.byref(h), ctypes.byref(err))
    _lib.check(rc, err)
    try:
        return _read_electronic(lib, h)
    finally:
        lib.gfn1_electronic_result_free(h)


def analytic_gradient(numbers, positions_bohr, params: Optional[Gfn1Parameters] = None,
                      options: Optional[AnalyticGradientOptions] = None,
                      charge: float = 0.0, lattice=None) -> AnalyticGradientResult:
    """analytic_gradient (gradient.rs:56): gradient in Eh/bohr.  With a lattice
    (rows, bohr) the periodic Gamma-point gradient (gradient.rs:61-66), which
    has every term whatever the include flags."""
    lib = _lib.load()
    p = _params(params)
    opts = options or AnalyticGradientOptions()
    sysc, keep_sys = _lib.make_system(numbers, positions_bohr, charge, lattice)
    oc, keep_q = opts.to_c()
    h = ctypes.c_void_p()
    err = _lib.Gfn1Error()
    rc = lib.gfn1_analytic_gradient(ctypes.byref(sysc), p._h, ctypes.byref(oc), ctypes.byref(h), ctypes.byref(err))
    _lib.check(rc, err)
    try:
        el = _read_electronic(lib, lib.gfn1_gradient_result_electronic(h))
        nat = len(np.asarray(numbers))

        def g(k):
            ptr = lib.gfn1_gradient_result_field(h, k)
            return np.ctypeslib.as_array(ptr, shape=(3 * nat,)).reshape(nat, 3).copy() if nat else np.zeros((0, 3))

        return AnalyticGradientResult(
            electronic_result=el, total_energy=lib.gfn1_gradient_result_total_energy(h), gradient=g(0),
            forces=g(1), electronic_gradient=g(2), repulsion_gradient=g(3), dispersion_gradient=g(4),
            halogen_gradient=g(5), max_gradient=lib.gfn1_gradient_result_max_gradient(h))
    finally:
        lib.gfn1_gradient_result_free(h)


@dataclass
class GeometryOptimizationOptions:
    """GeometryOptimizationOptions (optimizer.rs:18-45) with its defaults
    (optimizer.rs:10-15).  warm_start and keep_trajectory have no reference
    counterpart: warm_start=True starts each evaluation's SCC from the charges
    of the last accepted geometry (the reference starts every evaluation
    afresh); the trajectory is always kept by the reference."""
    max_iterations: int = 250
    gradient_tolerance: float = 1.0e-4
    step_tolerance: float = 1.0e-7
    history: int = 12
    initial_step: float = 1.0
    max_atom_step: float = 0.30
    warm_start: bool = False
    keep_trajectory: bool = True

    def to_c(self):
        o = _lib.OptimizationOptions()
        o.max_iterations = int(self.max_iterations)
        o.gradient_tolerance = float(self.gradient_tolerance)
        o.step_tolerance = float(self.step_tolerance)
        o.history = int(self.history)
        o.initial_step = float(self.initial_step)
        o.max_atom_step = float(self.max_atom_step)
        o.warm_start = int(bool(self.warm_start))
        o.keep_trajectory = int(bool(self.keep_trajectory))
        return o


@dataclass
class GeometryOptimizationStep:
    """GeometryOptimizationStep (optimizer.rs:47-55); positions in bohr."""
    iteration: int
    energy: float
    max_gradient: float
    step_norm: float
    positions: np.ndarray


@dataclass
class GeometryOptimizationResult:
    """GeometryOptimizationResult (optimizer.rs:57-67); positions in bohr,
    gradient in Eh/bohr.  status/message: 0 and "" for a finished
    optimisation, else the error optimize_geometry returned for the system."""
    energy: float
    positions: np.ndarray
    gradient: np.ndarray
    forces: np.ndarray
    iterations: int
    converged: bool
    max_gradient: float
    trajectory: list
    status: int = 0
    message: str = ""


def _read_optimization(lib, h, i, nat):
    status = lib.gfn1_optimization_status(h, i)
    msg = (lib.gfn1_optimization_message(h, i) or b"").decode(errors="replace")
    pos = np.zeros((nat, 3))
    grad = np.zeros((nat, 3))
    traj = []
    if status == _lib.GFN1_OK:
        lib.gfn1_optimization_positions(h, i, pos.ctypes.data_as(_lib._DP))
        lib.gfn1_optimization_gradient(h, i, grad.ctypes.data_as(_lib._DP))
        for k in range(lib.gfn1_optimization_trajectory_length(h, i)):
            it, e, mg, sn = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
            p = np.zeros((nat, 3))
            lib.gfn1_optimization_trajectory_step(h, i, k, ctypes.byref(it), ctypes.byref(e), ctypes.byref(mg),
                                                  ctypes.byref(sn), p.ctypes.data_as(_lib._DP))
            traj.append(GeometryOptimizationStep(it.value, e.value, mg.value, sn.value, p))
    return GeometryOptimizationResult(
        energy=lib.gfn1_optimization_energy(h, i), positions=pos, gradient=grad, forces=-grad,
        iterations=lib.gfn1_optimization_iterations(h, i), converged=bool(lib.gfn1_optimization_converged(h, i)),
        max_gradient=lib.gfn1_optimization_max_gradient(h, i), trajectory=traj, status=status, message=msg)


def optimize_geometry(numbers, positions_bohr, params: Optional[Gfn1Parameters] = None,
                      options: Optional[GeometryOptimizationOptions] = None,
                      gradient_options: Optional[AnalyticGradientOptions] = None,
                      charge: float = 0.0, lattice=None) -> GeometryOptimizationResult:
    """optimize_geometry (optimizer.rs:78-185): L-BFGS relaxation of one
    system; positions in bohr.  With a lattice (rows, bohr) the atoms relax in
    the fixed cell over the periodic Gamma-point energy and gradient.  Raises
    NativeError where the reference returns Err."""
    lib = _lib.load()
    p = _params(params)
    oc = (options or GeometryOptimizationOptions()).to_c()
    gc, keep_q = (gradient_options or AnalyticGradientOptions()).to_c()
    sysc, keep_sys = _lib.make_system(numbers, positions_bohr, charge, lattice)
    h = ctypes.c_void_p()
    err = _lib.Gfn1Error()
    rc = lib.gfn1_optimize_geometry(ctypes.byref(sysc), p._h, ctypes.byref(gc), ctypes.byref(oc), ctypes.byref(h),
                                    ctypes.byref(err))
    _lib.check(rc, err)
    try:
        return _read_optimization(lib, h, 0, len(np.asarray(numbers)))
    finally:
        lib.gfn1_optimization_free(h)


@dataclass
class HessianOptions:
    """Options of the Cartesian Hessian (central differences of the analytic
    gradient, the construction the reference validates analytic_hessian
    against, tests/hessian.rs:691-760).  include_electronic stands for the
    reference's four electronic flags (hessian/mod.rs:28-37) together; the
    SCC tolerances replace those of the electronic options."""
    step: float = 1.0e-4
    scc_energy_tolerance: float = 1.0e-11
    scc_charge_tolerance: float = 1.0e-11
    include_repulsion: bool = True
    include_electronic: bool = True
    include_dispersion: bool = True
    include_halogen: bool = True
    warm_start: bool = True
    max_batch: int = 4096

    def to_c(self):
        o = _lib.HessianOptions()
        o.step = float(self.step)
        o.scc_energy_tolerance = float(self.scc_energy_tolerance)
        o.scc_charge_tolerance = float(self.scc_charge_tolerance)
        for name in ("include_repulsion", "include_electronic", "include_dispersion", "include_halogen",
                     "warm_start"):
            setattr(o, name, int(bool(getattr(self, name))))
        o.max_batch = int(self.max_batch)
        return o


@dataclass
class HessianResult:
    """hessian: (3 nat, 3 nat) in Eh/bohr^2, index 3*atom + axis, symmetrised
    (AnalyticHessianResult.hessian, hessian/mod.rs:55-65); from the same
    displaced evaluations, dipole_derivatives (3 nat, 3): d mu_alpha / d R_c in
    e (DipoleDerivatives.ddipole_dr, properties.rs:80-88) and the Mulliken
    dipole (e bohr, origin 0)."""
    hessian: np.ndarray
    max_asymmetry: float
    dipole_derivatives: np.ndarray = None
    dipole: np.ndarray = None
    # with terms=True: the repulsion / dispersion / halogen sub-results'
    # .hessian (hessian/mod.rs:55-65) and the electronic rest
    repulsion: np.ndarray = None
    dispersion: np.ndarray = None
    halogen: np.ndarray = None
    electronic: np.ndarray = None


def hessian(numbers, positions_bohr, params: Optional[Gfn1Parameters] = None,
            electronic: Optional[ElectronicOptions] = None, options: Optional[HessianOptions] = None,
            charge: float = 0.0, lattice=None, terms: bool = False) -> HessianResult:
    """Cartesian Hessian of one system (replaces analytic_hessian(...).hessian,
    hessian/mod.rs:112-134, at the reference's own 1e-6 Eh/bohr^2 acceptance).
    With a lattice (rows, bohr): the fixed-cell Gamma-point Hessian of a
    periodic system (pbc_gamma_hessian, pbc/hessian.rs), by the same central
    differences of the periodic gradient.  terms=True also returns the
    repulsion, dispersion and halogen parts (differences of those terms'
    gradients) and the electronic rest (no dipole derivatives then)."""
    lib = _lib.load()
    p = _params(params)
    eo, keep_q = (electronic or ElectronicOptions()).to_c()
    ho = (options or HessianOptions()).to_c()
    sysc, keep_sys = _lib.make_system(numbers, positions_bohr, charge, lattice)
    ndof = 3 * len(np.asarray(numbers))
    if terms:
        out = np.zeros((ndof, ndof))
        parts = [np.zeros((ndof, ndof)) for _ in range(3)]
        asym = ctypes.c_double()
        err = _lib.Gfn1Error()
        _lib.check(lib.gfn1_hessian_terms(ctypes.byref(sysc), p._h, ctypes.byref(eo), ctypes.byref(ho),
                                          out.ctypes.data_as(_lib._DP), ctypes.byref(asym),
                                          *[h.ctypes.data_as(_lib._DP) for h in parts], ctypes.byref(err)), err)
        rep, disp, hal = parts
        return HessianResult(out, asym.value, repulsion=rep, dispersion=disp, halogen=hal,
                             electronic=out - rep - disp - hal)
    out = np.zeros((ndof, ndof))
    ddip = np.zeros((ndof, 3))
    dip = np.zeros(3)
    asym = ctypes.c_double()
    err = _lib.Gfn1Error()
    _lib.check(lib.gfn1_hessian(ctypes.byref(sysc), p._h, ctypes.byref(eo), ctypes.byref(ho),
                                out.ctypes.data_as(_lib._DP), ctypes.byref(asym), ddip.ctypes.data_as(_lib._DP),
                                dip.ctypes.data_as(_lib._DP), ctypes.byref(err)), err)
    return HessianResult(out, asym.value, ddip, dip)


def hessian_batch(systems, params: Optional[Gfn1Parameters] = None, electronic: Optional[ElectronicOptions] = None,
                  options: Optional[HessianOptions] = None):
    """Hessians of many systems ((numbers, positions_bohr[, charge]) tuples) in
    shared device passes: a list of (3 nat, 3 nat) arrays, None where the
    system failed, and the list of per-system statuses."""
    lib = _lib.load()
    p = _params(params)
    eo, keep_q = (electronic or ElectronicOptions()).to_c()
    ho = (options or HessianOptions()).to_c()
    arr, keep = _system_array(systems)
    sizes = [3 * len(np.asarray(s[0])) for s in systems]
    flat = np.zeros(sum(k * k for k in sizes))
    status = np.zeros(len(systems), dtype=np.int32)
    err = _lib.Gfn1Error()
    _lib.check(lib.gfn1_hessian_batch(len(systems), arr, p._h, ctypes.byref(eo), ctypes.byref(ho),
                                      flat.ctypes.data_as(_lib._DP), None,
                                      status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(err)), err)
    out, off = [], 0
    for k, st in zip(sizes, status):
        out.append(flat[off:off + k * k].reshape(k, k).copy() if st == 0 else None)
        off += k * k
    return out, [int(v) for v in status]


# Standard atomic weights (IUPAC conventional values; mass number of the
# longest-lived isotope for Tc, Pm, Po, At, Rn), Z = 1..86, as the reference
# tabulates them (data_tables.rs:40-128).
ATOMIC_MASSES = (
    1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999, 18.998403163, 20.1797,
    22.98976928, 24.305, 26.9815385, 28.085, 30.973761998, 32.06, 35.45, 39.948, 39.0983, 40.078,
    44.955908, 47.867, 50.9415, 51.9961, 54.938044, 55.845, 58.933194, 58.6934, 63.546, 65.38,
    69.723, 72.630, 74.921595, 78.971, 79.904, 83.798, 85.4678, 87.62, 88.90584, 91.224,
    92.90637, 95.95, 98.0, 101.07, 102.90550, 106.42, 107.8682, 112.414, 114.818, 118.710,
    121.760, 127.60, 126.90447, 131.293, 132.90545196, 137.327, 138.90547, 140.116, 140.90766, 144.242,
    145.0, 150.36, 151.964, 157.25, 158.92535, 162.500, 164.93033, 167.259, 168.93422, 173.045,
    174.9668, 178.49, 180.94788, 183.84, 186.207, 190.23, 192.217, 195.084, 196.966569, 200.592,
    204.38, 207.2, 208.98040, 209.0, 210.0, 222.0)
WAVENUMBER_PER_SQRT_AU = 5140.4843  # cm^-1 per sqrt(Eh / (bohr^2 amu)), vibrational.rs:20


# Integrated molar absorption (km/mol) per atomic-unit intensity |d mu / d Q|^2
# (e^2 / u): N_A e^2 / (12 eps0 c^2 u), CODATA 2018 constants; the reference
# tabulates 974.8801118 (properties.rs:32-35).
IR_INTENSITY_AU_TO_KM_PER_MOL = (6.02214076e23 * 1.602176634e-19 ** 2
                                 / (12.0 * 8.8541878128e-12 * 299792458.0 ** 2 * 1.66053906660e-27)) / 1000.0


def ir_spectrum_from(hessian_matrix, dipole_derivatives, numbers):
    """Harmonic IR spectrum (properties.rs:381-420): per mode the wavenumber,
    d mu / d Q_k = sum_c d mu / d R_c L_ck over the Cartesian mode L, and its
    squared norm in e^2/u and km/mol."""
    v = vibrational_analysis(hessian_matrix, numbers)
    dmu = v.modes @ np.asarray(dipole_derivatives, dtype=np.float64)
    inten = np.einsum("ka,ka->k", dmu, dmu)
    return {"wavenumbers": v.wavenumbers, "intensities_km_per_mol": inten * IR_INTENSITY_AU_TO_KM_PER_MOL,
            "intensities_au": inten, "dipole_gradients": dmu}


@dataclass
class VibrationalModes:
    """VibrationalModes (vibrational.rs:22-32)."""
    wavenumbers: np.ndarray  # cm^-1, ascending; negative = imaginary
    eigenvalues: np.ndarray  # Eh / (bohr^2 amu)
    modes: np.ndarray        # (3 nat, 3 nat), one Cartesian displacement per row


def vibrational_analysis(hessian_matrix, numbers) -> VibrationalModes:
    """Harmonic analysis of a Cartesian Hessian (vibrational.rs:37-84): mass-
    weighted, symmetrised, diagonalised; wavenumber = sign(l) sqrt(|l|) *
    5140.4843 cm^-1.  Host linear algebra (numpy): a 3 nat square problem."""
    h = np.asarray(hessian_matrix, dtype=np.float64)
    z = [int(v) for v in numbers]
    mass = np.array([max(ATOMIC_MASSES[v - 1] if 1 <= v <= len(ATOMIC_MASSES) else 0.0, 1.0e-12)
                     for v in z for _ in range(3)])
    mw = h / np.sqrt(np.outer(mass, mass))
    mw = 0.5 * (mw + mw.T)
    lam, vec = np.linalg.eigh(mw)
    w = np.sqrt(np.abs(lam)) * WAVENUMBER_PER_SQRT_AU
    return VibrationalModes(np.where(lam < 0.0, -w, w), lam, (vec / np.sqrt(mass)[:, None]).T)


# element symbols Z = 1..86 (the elements GFN1-xTB parametrises)
_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn").split()


def _xyz_frame(numbers, positions_angstrom, comment):
    """One XYZ frame (python.rs:3035-3045 layout)."""
    lines = ["%d" % len(numbers), comment]
    for z, p in zip(numbers, positions_angstrom):
        sym = _SYMBOLS[int(z) - 1] if 1 <= int(z) <= len(_SYMBOLS) else "X"
        lines.append("%-2s %18.10f %18.10f %18.10f" % (sym, p[0], p[1], p[2]))
    return "\n".join(lines) + "\n"


@dataclass
class OptimizationResult:
    """python.rs PyOptimizationResult (:2954-3078)."""
    energy_hartree: float
    energy_ev: float
    positions_angstrom: np.ndarray
    gradient_hartree_per_bohr: np.ndarray
    forces_ev_per_angstrom: np.ndarray
    iterations: int
    converged: bool
    max_gradient: float
    numbers: np.ndarray
    trajectory_positions_angstrom: list
    trajectory_energies_hartree: list

    @classmethod
    def from_result(cls, numbers, r: GeometryOptimizationResult):
        b2a = 1.0 / ANGSTROM_TO_BOHR
        return cls(
            energy_hartree=r.energy, energy_ev=r.energy * HARTREE_TO_EV, positions_angstrom=r.positions * b2a,
            gradient_hartree_per_bohr=r.gradient, forces_ev_per_angstrom=r.forces * FORCE_AU_TO_EV_PER_ANGSTROM,
            iterations=r.iterations, converged=r.converged, max_gradient=r.max_gradient,
            numbers=np.asarray(numbers, dtype=np.int32),
            trajectory_positions_angstrom=[s.positions * b2a for s in r.trajectory],
            trajectory_energies_hartree=[s.energy for s in r.trajectory])

    def to_xyz(self, comment=None):
        """Final geometry as an XYZ string (Angstrom)."""
        return _xyz_frame(self.numbers, self.positions_angstrom,
                          comment if comment is not None else "gfn1 optimized geometry")

    def trajectory_xyz(self):
        """Every accepted step as a multi-frame XYZ string (Angstrom)."""
        return "".join(_xyz_frame(self.numbers, f, "iter %d energy %.10f Ha" % (i, e))
                       for i, (f, e) in enumerate(zip(self.trajectory_positions_angstrom,
                                                      self.trajectory_energies_hartree)))


@dataclass
class CalculationResult:
    """python.rs CalculationResult (:2845-2935)."""
    energy_hartree: float
    gradient_hartree_per_bohr: Optional[np.ndarray]
    charges: np.ndarray
    dipole: np.ndarray
    iterations: int
    converged: bool
    _terms: EnergyTerms

    @property
    def energy_ev(self):
        return self.energy_hartree * HARTREE_TO_EV

    @property
    def forces_ev_per_angstrom(self):
        if self.gradient_hartree_per_bohr is None:
            return None
        return -self.gradient_hartree_per_bohr * FORCE_AU_TO_EV_PER_ANGSTROM

    def energy_terms_hartree(self) -> Dict[str, float]:
        return dict(self._terms.named_values())


class Gfn1NativeCalculator:
    """Mirror of the PyO3 Gfn1NativeCalculator (python.rs:115-381)."""

    def __init__(self, param_path=None, charge=0.0, multiplicity=None, max_scc=250, energy_tolerance=1e-6,
                 charge_tolerance=2e-5, mixing=0.4, scc_broyden=True, scc_broyden_size=250,
                 electronic_temperature=300.0, nprim=0, eigen_tolerance=1e-12, enable_dispersion=True,
                 enable_cn_hamiltonian=True, level_shift=0.0):
        if multiplicity not in (None, 1):
            raise ValueError("only closed-shell (multiplicity 1) calculations are supported")
        self.params = Gfn1Parameters.resolve(param_path)
        self.charge = charge
        self.options = ElectronicOptions(
            max_scc=max_scc, energy_tolerance=energy_tolerance, charge_tolerance=charge_tolerance, mixing=mixing,
            scc_broyden=scc_broyden, scc_broyden_size=scc_broyden_size,
            electronic_temperature=electronic_temperature, nprim=nprim, eigen_tolerance=eigen_tolerance,
            enable_dispersion=enable_dispersion, enable_cn_hamiltonian=enable_cn_hamiltonian,
            level_shift=level_shift)
        self._pbc_guess = None  # shell charges of the last periodic result

    def param_source(self) -> str:
        """python.rs:328-330."""
        return self.params.source_description()

    def __repr__(self):
        return "Gfn1NativeCalculator(parameters: %s)" % self.params.source_description()

    def calculate(self, numbers, positions, unit="angstrom", compute_gradient=False) -> CalculationResult:
        unit_l = unit.lower()
        if unit_l in ("angstrom", "ang", "a"):
            scale = ANGSTROM_TO_BOHR
        elif unit_l in ("bohr", "au", "a0"):
            scale = 1.0
        else:
            raise ValueError("unit must be 'angstrom' or 'bohr'")
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * scale
        if compute_gradient:
            r = analytic_gradient(numbers, x, self.params, AnalyticGradientOptions(electronic=self.options),
                                  charge=self.charge)
            el = r.electronic_result
            return CalculationResult(r.total_energy, r.gradient, el.atomic_charges, el.dipole, el.iterations,
                                     el.converged, el.terms)
        el = run_electronic(numbers, x, self.params, self.options, charge=self.charge)
        return CalculationResult(el.total_free, None, el.atomic_charges, el.dipole, el.iterations, el.converged,
                                 el.terms)

    def calculate_periodic(self, numbers, positions, cell, pbc=(True, True, True), kgrid=None, unit="angstrom",
                           compute_gradient=False, compute_stress=False) -> CalculationResult:
        """Periodic single point (python.rs:381-470) at the Gamma point: cell is
        the three lattice vectors (rows) in the chosen unit.  As in the
        reference, each periodic calculation starts its SCC from the shell
        charges of the previous one when the shell count matches (its
        PbcSccGuess).  Only fully periodic cells and the Gamma point are
        built; a k-point mesh or the stress tensor raises NotImplementedError."""
        if tuple(bool(p) for p in pbc) != (True, True, True):
            raise ValueError("only fully periodic cells (pbc = (True, True, True)) are supported")
        if kgrid is not None and tuple(kgrid) != (1, 1, 1):
            raise NotImplementedError("k-point meshes are not built; only the Gamma point (kgrid None or (1, 1, 1))")
        if compute_stress:
            raise NotImplementedError("the periodic stress tensor is not built")
        scale = _unit_scale(unit)
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * scale
        lattice = np.asarray(cell, dtype=np.float64).reshape(3, 3) * scale
        opts = replace(self.options, scc_initial_shell_charges=self._pbc_guess)
        if compute_gradient:
            r = analytic_gradient(numbers, x, self.params, AnalyticGradientOptions(electronic=opts),
                                  charge=self.charge, lattice=lattice)
            el = r.electronic_result
            out = CalculationResult(r.total_energy, r.gradient, el.atomic_charges, el.dipole, el.iterations,
                                    el.converged, el.terms)
        else:
            el = run_electronic(numbers, x, self.params, opts, charge=self.charge, lattice=lattice)
            out = CalculationResult(el.total_free, None, el.atomic_charges, el.dipole, el.iterations, el.converged,
                                    el.terms)
        self._pbc_guess = np.array(el.shell_charges)
        return out


    def optimize(self, numbers, positions, unit="angstrom", max_iterations=250, gradient_tolerance=1.0e-4,
                 step_tolerance=1.0e-7, history=12, max_atom_step=0.30, trajectory_path=None) -> OptimizationResult:
        """Calculator optimize (python.rs:486-542), non-periodic systems.  The
        trajectory file, when asked for, is written once the optimisation has
        finished (one frame per accepted step, optimizer.rs:187-213 layout)."""
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * _unit_scale(unit)
        opts = GeometryOptimizationOptions(max_iterations=max_iterations, gradient_tolerance=gradient_tolerance,
                                           step_tolerance=step_tolerance, history=history,
                                           max_atom_step=max_atom_step)
        r = optimize_geometry(numbers, x, self.params, opts, AnalyticGradientOptions(electronic=self.options),
                              charge=self.charge)
        if trajectory_path is not None:
            b2a = 1.0 / ANGSTROM_TO_BOHR
            with open(trajectory_path, "w") as f:
                for step in r.trajectory:
                    f.write(_xyz_frame(numbers, step.positions * b2a, "iter %d energy %.10f Ha max_grad %.3e"
                                       % (step.iteration, step.energy, step.max_gradient)))
        return OptimizationResult.from_result(numbers, r)


    def hessian(self, numbers, positions, unit="angstrom"):
        """Calculator hessian (python.rs:1753-1774): the (3N, 3N) Cartesian
        Hessian in Eh/bohr^2 as a numpy array (central differences of the
        analytic gradient on the device)."""
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * _unit_scale(unit)
        return hessian(numbers, x, self.params, self.options, charge=self.charge).hessian

    def vibrational_frequencies(self, numbers, positions, unit="angstrom"):
        """Calculator vibrational_frequencies (python.rs:1825-1846): dict with
        wavenumbers (cm^-1, ascending, imaginary as negative) and modes (one
        Cartesian displacement per row)."""
        v = vibrational_analysis(self.hessian(numbers, positions, unit), numbers)
        return {"wavenumbers": v.wavenumbers, "modes": v.modes}

    def dipole_derivatives(self, numbers, positions, unit="angstrom", origin=None):
        """Calculator dipole_derivatives (python.rs:575-600): dipole (e bohr,
        about origin, bohr) and ddipole_dr ([coord][alpha], e), by central
        differences on the device."""
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * _unit_scale(unit)
        r = hessian(numbers, x, self.params, self.options, charge=self.charge)
        o = np.zeros(3) if origin is None else np.asarray(origin, dtype=np.float64)
        return {"dipole": r.dipole - self.charge * o, "ddipole_dr": r.dipole_derivatives}

    def ir_spectrum(self, numbers, positions, unit="angstrom", origin=None):
        """Calculator ir_spectrum (python.rs:600-647): wavenumbers (cm^-1),
        intensities_km_per_mol, intensities_au, dipole_gradients.  The
        derivatives do not depend on the origin (the total charge is fixed)."""
        x = np.asarray(positions, dtype=np.float64).reshape(-1, 3) * _unit_scale(unit)
        r = hessian(numbers, x, self.params, self.options, charge=self.charge)
        return ir_spectrum_from(r.hessian, r.dipole_derivatives, numbers)


def _unit_scale(unit):
    unit_l = unit.lower()
    if unit_l in ("angstrom", "ang", "a"):
        return ANGSTROM_TO_BOHR
    if unit_l in ("bohr", "au", "a0"):
        return 1.0
    raise ValueError("unit must be 'angstrom' or 'bohr'")


def basis_counts(numbers, params: Optional[Gfn1Parameters] = None):
    """(shells,) of a system in the loaded parameters."""
    p = _params(params)
    lib = _lib.load()
    nsh = 0
    for z in numbers:
        k = lib.gfn1_params_shell_count(p._h, int(z))
        if k < 0:
            raise _lib.NativeError(4, "missing element Z=%d" % int(z))
        nsh += k
    return (nsh,)


def _system_array(systems):
    """(numbers, positions_bohr[, charge[, lattice]]) tuples -> ctypes System array + keep-alive list."""
    arr = (_lib.System * len(systems))()
    keep = []
    for i, s in enumerate(systems):
        charge = s[2] if len(s) > 2 else 0.0
        sc, k = _lib.make_system(s[0], s[1], charge, s[3] if len(s) > 3 else None)
        arr[i] = sc
        keep.append(k)
    return arr, keep


class GuessModel:
    """Learned SCC start charges: a per-atom MLP on radial-environment descriptors
    (68 -> 64 SiLU -> 64 SiLU -> 2 shell charges), trained and evaluated on the GPU
    by the library's own kernels.  Its output is the SCC start vector (the
    reference's scc_initial_shell_charges, electronic.rs:386-392); the converged
    result does not depend on it."""

    PARAMS = _lib.GUESS_PARAMS
    FEATURES = _lib.GUESS_FEATURES

    def __init__(self, seed=1, _handle=None):
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            err = _lib.Gfn1Error()
            _lib.check(self._lib.gfn1_guess_model_create(int(seed), ctypes.byref(self._h), ctypes.byref(err)), err)

    @classmethod
    def load(cls, path):
        lib = _lib.load()
        h = ctypes.c_void_p()
        err = _lib.Gfn1Error()
        _lib.check(lib.gfn1_guess_model_load(os.fsencode(path), ctypes.byref(h), ctypes.byref(err)), err)
        return cls(_handle=h)

    def save(self, path):
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_guess_model_save(self._h, os.fsencode(path), ctypes.byref(err)), err)

    def params(self):
        theta = np.zeros(self.PARAMS, dtype=np.float32)
        norm = np.zeros(2 * self.FEATURES, dtype=np.float32)
        rc = self._lib.gfn1_guess_model_get_params(self._h, theta.ctypes.data_as(_lib._FP),
                                                   norm.ctypes.data_as(_lib._FP))
        if rc != 0:
            raise ValueError("model has no parameters")
        return theta, norm

    def set_params(self, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.size != self.PARAMS:
            raise ValueError("expected %d parameters" % self.PARAMS)
        if self._lib.gfn1_guess_model_set_params(self._h, theta.ctypes.data_as(_lib._FP)) != 0:
            raise ValueError("invalid parameters")

    def train(self, systems, shell_charges, params: Optional[Gfn1Parameters] = None, epochs=40, batch_atoms=1024,
              learning_rate=1e-3, seed=1):
        """Fit to converged shell charges (one array per system); returns the
        per-epoch mean squared error."""
        p = _params(params)
        arr, keep = _system_array(systems)
        q = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in shell_charges]))
        o = _lib.GuessTrainOptions()
        self._lib.gfn1_guess_train_options_default(ctypes.byref(o))
        o.epochs, o.batch_atoms, o.learning_rate, o.seed = int(epochs), int(batch_atoms), float(learning_rate), int(seed)
        loss = np.zeros(max(int(epochs), 1), dtype=np.float32)
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_guess_train(self._h, len(systems), arr, p._h, q.ctypes.data_as(_lib._DP),
                                              ctypes.byref(o), loss.ctypes.data_as(_lib._FP), ctypes.byref(err)), err)
        del keep
        return loss[:int(epochs)]

    def predict(self, systems, params: Optional[Gfn1Parameters] = None, noalloc=False):
        """Start vectors, one array of shell charges per system.  noalloc: the
        variant that allocates no device memory (gfn1_guess_predict_noalloc;
        at most NOALLOC_MAX_ATOMS atoms)."""
        p = _params(params)
        arr, keep = _system_array(systems)
        nsh = [basis_counts(s[0], p)[0] for s in systems]
        out = np.zeros(sum(nsh))
        err = _lib.Gfn1Error()
        fn = self._lib.gfn1_guess_predict_noalloc if noalloc else self._lib.gfn1_guess_predict
        _lib.check(fn(self._h, len(systems), arr, p._h, out.ctypes.data_as(_lib._DP), ctypes.byref(err)), err)
        del keep
        return np.split(out, np.cumsum(nsh)[:-1])

    NOALLOC_MAX_ATOMS = 16384
    NOALLOC_MAX_SYSTEMS = 16384

    def loss_gradient(self, systems, shell_charges, params: Optional[Gfn1Parameters] = None):
        p = _params(params)
        arr, keep = _system_array(systems)
        q = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in shell_charges]))
        loss = ctypes.c_double()
        grad = np.zeros(self.PARAMS, dtype=np.float32)
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_guess_loss_gradient(self._h, len(systems), arr, p._h, q.ctypes.data_as(_lib._DP),
                                                      ctypes.byref(loss), grad.ctypes.data_as(_lib._FP),
                                                      ctypes.byref(err)), err)
        del keep
        return loss.value, grad

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.gfn1_guess_model_free(self._h)
            self._h = None


def guess_features(systems):
    """Descriptors of every atom (natoms x 68, FP32), computed on the GPU."""
    lib = _lib.load()
    arr, keep = _system_array(systems)
    nat = sum(len(s[0]) for s in systems)
    out = np.zeros((nat, _lib.GUESS_FEATURES), dtype=np.float32)
    err = _lib.Gfn1Error()
    _lib.check(lib.gfn1_guess_features(len(systems), arr, out.ctypes.data_as(_lib._FP), ctypes.byref(err)), err)
    del keep
    return out


class BatchEngine:
    """Energy + gradient of many independent systems in one device pass."""

    def __init__(self, systems, params: Optional[Gfn1Parameters] = None,
                 options: Optional[AnalyticGradientOptions] = None, eigensolver=0, keep_matrices=False,
                 guess: Optional[GuessModel] = None, precision="fp64", scc_graphs=True):
        lib = _lib.load()
        self._lib = lib
        self.params = _params(params)
        opts = options or AnalyticGradientOptions()
        bo = _lib.BatchOptions()
        lib.gfn1_batch_options_default(ctypes.byref(bo))
        g, self._keep_q = opts.to_c()
        bo.gradient = g
        bo.eigensolver = int(eigensolver)
        bo.keep_matrices = int(bool(keep_matrices))
        # buckets of the own eigensolver iterate by replays of a captured HIP
        # graph of one SCC iteration (False: host-launched kernel sequences)
        bo.scc_graphs = int(bool(scc_graphs))
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        # mixed: FP32 MFMA density/energy-weighted-density products and FP32
        # gradient contraction; eigensolve, mixing and energy sums stay FP64
        bo.precision = 1 if precision == "mixed" else 0
        arr = (_lib.System * len(systems))()
        self._keep = []
        self.natoms = []
        for i, s in enumerate(systems):
            # (numbers, positions_bohr[, charge[, lattice]]): a lattice (rows,
            # bohr) makes the instance periodic (Gamma point)
            numbers, pos = s[0], s[1]
            charge = s[2] if len(s) > 2 else 0.0
            sc, keep = _lib.make_system(numbers, pos, charge, s[3] if len(s) > 3 else None)
            arr[i] = sc
            self._keep.append(keep)
            self.natoms.append(len(keep[0]))
        self.numbers = [np.asarray(k[0], dtype=np.int32) for k in self._keep]
        self._h = ctypes.c_void_p()
        err = _lib.Gfn1Error()
        _lib.check(lib.gfn1_batch_create(len(systems), arr, self.params._h, ctypes.byref(bo), ctypes.byref(self._h),
                                         ctypes.byref(err)), err)
        self.n = len(systems)
        self.set_guess(guess)

    def set_guess(self, model: Optional["GuessModel"]):
        """Start the SCC from the model's charges (None: the standard start)."""
        self._guess = model
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_set_guess(self._h, model._h if model is not None else None,
                                                  ctypes.byref(err)), err)

    def run(self):
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_run(self._h, ctypes.byref(err)), err)
        return self

    def set_backtransform_width(self, cols):
        """Fix the back-transformation width (16, 32, or 64 for buckets of at
        most 256 rows) in place of the tuned one."""
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_set_backtransform_width(self._h, int(cols), ctypes.byref(err)), err)
        return self

    def autotune(self):
        """Time every back-transformation width on each size bucket's data from
        the last run and keep the fastest for later runs (results unchanged).
        Returns one dict per tuned bucket: ld, instances, cols (chosen) and
        ms (width -> mean launch milliseconds)."""
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_autotune(self._h, ctypes.byref(err)), err)
        out = []
        for k in range(self._lib.gfn1_batch_tuning_count(self._h)):
            iv = (ctypes.c_int32 * 3)()
            ms = (ctypes.c_double * 3)()
            if self._lib.gfn1_batch_tuning_entry(self._h, k, iv, ms) != 0:
                raise RuntimeError("tuning entry %d" % k)
            out.append({"ld": iv[0], "instances": iv[1], "cols": iv[2],
                        "ms": {w: m for w, m in zip((16, 32, 64), ms) if m >= 0.0}})
        return out

    def status(self, i):
        return self._lib.gfn1_batch_status(self._h, i)

    def energy(self, i):
        return self._lib.gfn1_batch_energy(self._h, i)

    def energies(self):
        return np.array([self.energy(i) for i in range(self.n)])

    def energy_terms(self, i):
        e = (ctypes.c_double * 15)()
        self._lib.gfn1_batch_energy_terms(self._h, i, e)
        return dict(zip(_lib.ENERGY_NAMES, list(e)))

    def iterations(self, i):
        return self._lib.gfn1_batch_iterations(self._h, i)

    def gradient(self, i):
        out = np.zeros((self.natoms[i], 3))
        self._lib.gfn1_batch_gradient(self._h, i, out.ctypes.data_as(_lib._DP))
        return out

    def shell_charges(self, i, nsh):
        out = np.zeros(nsh)
        self._lib.gfn1_batch_shell_charges(self._h, i, out.ctypes.data_as(_lib._DP))
        return out

    def set_positions(self, positions_bohr):
        """New positions (bohr) for the next runs: one (nat, 3) array per system."""
        x = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1, 3)
                                                 for p in positions_bohr]) if self.n else np.zeros((0, 3)))
        if x.shape[0] != sum(self.natoms):
            raise ValueError("positions do not match the batch's atom counts")
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_set_positions(self._h, x.ctypes.data_as(_lib._DP), ctypes.byref(err)), err)

    def results(self):
        """All systems at once: statuses (n), energies (n, NaN where failed),
        gradients (list of (nat, 3)) and shell charges (list of (nsh,))."""
        lib = self._lib
        if not hasattr(self, "_nsh"):
            self._nsh = [basis_counts(z, self.params)[0] if len(z) else 0 for z in self.numbers]
        st = np.zeros(self.n, dtype=np.int32)
        e = np.zeros(self.n)
        g = np.zeros((sum(self.natoms), 3))
        q = np.zeros(sum(self._nsh))
        if self.n:
            for rc in (lib.gfn1_batch_statuses(self._h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                       lib.gfn1_batch_energies(self._h, e.ctypes.data_as(_lib._DP)),
                       lib.gfn1_batch_gradients(self._h, g.ctypes.data_as(_lib._DP)),
                       lib.gfn1_batch_all_shell_charges(self._h, q.ctypes.data_as(_lib._DP))):
                if rc != 0:
                    raise RuntimeError("batch results are not available (run the batch first)")
        ga = np.cumsum([0] + self.natoms)
        qa = np.cumsum([0] + self._nsh)
        return (st, e, [g[ga[i]:ga[i + 1]] for i in range(self.n)], [q[qa[i]:qa[i + 1]] for i in range(self.n)])

    def set_all_start_charges(self, shell_charges):
        """Start charges of every system (one (nsh,) array each) for the next runs."""
        q = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64) for v in shell_charges])
                                 if self.n else np.zeros(0))
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_set_all_start_charges(self._h, q.ctypes.data_as(_lib._DP),
                                                              ctypes.byref(err)), err)

    def set_start_charges(self, i, shell_charges):
        """Start the next SCCs of system i from these shell charges (the
        reference's scc_initial_shell_charges hook, electronic.rs:386-392)."""
        q = np.ascontiguousarray(shell_charges, dtype=np.float64)
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_set_start_charges(self._h, int(i), q.ctypes.data_as(_lib._DP),
                                                          ctypes.byref(err)), err)

    def optimize(self, options: Optional["GeometryOptimizationOptions"] = None):
        """optimize_geometry (optimizer.rs:78) of every system at once; one
        GeometryOptimizationResult per system (status != 0: the error that
        optimize_geometry would have returned for it)."""
        oc = (options or GeometryOptimizationOptions()).to_c()
        h = ctypes.c_void_p()
        err = _lib.Gfn1Error()
        _lib.check(self._lib.gfn1_batch_optimize(self._h, ctypes.byref(oc), ctypes.byref(h), ctypes.byref(err)), err)
        try:
            res = [_read_optimization(self._lib, h, i, self.natoms[i]) for i in range(self.n)]
            rounds, evals, ms = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_double()
            self._lib.gfn1_optimization_stats(h, ctypes.byref(rounds), ctypes.byref(evals), ctypes.byref(ms))
            self.optimization_stats = {"rounds": rounds.value, "evaluations": evals.value, "total_ms": ms.value}
            return res
        finally:
            self._lib.gfn1_optimization_free(h)

    def timings(self):
        t = (ctypes.c_double * 4)()
        self._lib.gfn1_batch_timings(self._h, t)
        return dict(zip(("setup_ms", "scc_ms", "gradient_ms", "total_ms"), list(t)))

    KERNEL_TRIDIAG = 0

    def kernel_stats(self, kernel=KERNEL_TRIDIAG):
        """(summed launch milliseconds, launches, algorithmic bytes) of one kernel
        over the last run, timed with HIP events on the launch streams."""
        ms, nl, nb = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        rc = self._lib.gfn1_batch_kernel_stats(self._h, kernel, ctypes.byref(ms), ctypes.byref(nl),
                                               ctypes.byref(nb))
        if rc != 0:
            raise ValueError("unknown kernel id %r" % (kernel,))
        return ms.value, nl.value, nb.value

    LAUNCH_STATS = ("kernels", "enqueues", "graph_replays", "host_iterations", "state_reads", "graphs_captured",
                    "capture_us")

    def launch_stats(self):
        """How the last run's SCC phase drove the device: kernels run, host
        enqueue calls, graph replays, host-launched iterations, host reads of
        gathered states, graphs captured and the capture time (microseconds)."""
        out = (ctypes.c_int64 * len(self.LAUNCH_STATS))()
        if self._lib.gfn1_batch_launch_stats(self._h, out) != 0:
            raise RuntimeError("no batch")
        return dict(zip(self.LAUNCH_STATS, (int(v) for v in out)))

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.gfn1_batch_free(self._h)
            self._h = None
