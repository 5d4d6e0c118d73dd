"""ctypes binding of the CPU oracle (oracle/gfn1_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgfn1_oracle.so")
DATA_DIR = os.path.join(os.path.dirname(HERE), "gfn1-rs_amd", "data")

ENERGY_NAMES = (
    "total_free", "total_internal", "repulsion", "electronic", "isotropic_scc",
    "third_order", "higher_order", "dispersion", "halogen", "multipole", "exchange",
    "spin_polarization", "plus_u", "external_field", "electronic_entropy",
)


class Options(ctypes.Structure):
    _fields_ = [
        ("charge", ctypes.c_double),
        ("max_scc", ctypes.c_int),
        ("energy_tolerance", ctypes.c_double),
        ("charge_tolerance", ctypes.c_double),
        ("mixing", ctypes.c_double),
        ("scc_broyden", ctypes.c_int),
        ("scc_broyden_size", ctypes.c_int),
        ("level_shift", ctypes.c_double),
        ("nprim", ctypes.c_int),
        ("eigen_tolerance", ctypes.c_double),
        ("electronic_temperature", ctypes.c_double),
        ("enable_dispersion", ctypes.c_int),
        ("integral_cutoff", ctypes.c_double),
        ("coordination_cutoff", ctypes.c_double),
        ("enable_cn_hamiltonian", ctypes.c_int),
        ("compute_gradient", ctypes.c_int),
        ("include_repulsion", ctypes.c_int),
        ("include_dispersion", ctypes.c_int),
        ("include_hamiltonian", ctypes.c_int),
        ("include_scc", ctypes.c_int),
        ("include_halogen", ctypes.c_int),
    ]


_DP = ctypes.POINTER(ctypes.c_double)


class Result(ctypes.Structure):
    _fields_ = [
        ("energies", ctypes.c_double * 15),
        ("fermi_level", ctypes.c_double),
        ("electronic_temperature", ctypes.c_double),
        ("nelec", ctypes.c_double),
        ("dipole", ctypes.c_double * 3),
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("gradient", _DP),
        ("electronic_gradient", _DP),
        ("repulsion_gradient", _DP),
        ("dispersion_gradient", _DP),
        ("halogen_gradient", _DP),
        ("shell_charges", _DP),
        ("atomic_charges", _DP),
        ("shell_scc_potential", _DP),
        ("coordination_numbers", _DP),
        ("orbital_energies", _DP),
        ("occupations", _DP),
        ("overlap", _DP),
        ("h0", _DP),
        ("fock", _DP),
        ("density", _DP),
        ("energy_weighted_density", _DP),
        ("message", ctypes.c_char * 256),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ctx = None


def lib():
    global _lib, _ctx
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_create.restype = ctypes.c_void_p
        _lib.oracle_create.argtypes = [ctypes.c_char_p]
        for name in ("oracle_dispersion", "oracle_repulsion", "oracle_halogen"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_double
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(Options), ctypes.c_void_p, ctypes.POINTER(Result)]
        _lib.oracle_run_cell.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_void_p,
                                         ctypes.POINTER(Result)]
        _lib.oracle_basis_size.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib.oracle_overlap.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_eigh.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_coordination.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_void_p]
        _lib.oracle_run_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _lib.oracle_default_options.argtypes = [ctypes.POINTER(Options)]
        _ctx = _lib.oracle_create(DATA_DIR.encode())
        if not _ctx:
            raise RuntimeError("oracle_create failed")
    return _lib


_param_ctx = {}


def ctx(param_file=None):
    """The oracle context of the bundled GFN1-xTB parameters, or of another
    parameter file (one context per file, kept for the process)."""
    lib()
    if param_file is None:
        return _ctx
    if param_file not in _param_ctx:
        _lib.oracle_create_params.restype = ctypes.c_void_p
        _lib.oracle_create_params.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        c = _lib.oracle_create_params(DATA_DIR.encode(), param_file.encode())
        if not c:
            raise RuntimeError("oracle_create_params failed for %s" % param_file)
        _param_ctx[param_file] = c
    return _param_ctx[param_file]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def default_options(**kw):
    o = Options()
    lib().oracle_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def basis_size(numbers, param_file=None):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    nsh, nao = ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_basis_size(ctx(param_file), len(z), _ptr(z), ctypes.byref(nsh), ctypes.byref(nao))
    if rc:
        raise ValueError("basis build failed (%d)" % rc)
    return nsh.value, nao.value


def run(numbers, positions_bohr, initial_shell_charges=None, want_matrices=False, param_file=None, cell=None,
        **opts):
    """Energy (+ gradient) of one system; returns a dict of numpy arrays.
    param_file: another GFN1 parameter file (default: the bundled one).
    cell: 3 x 3 cell vectors in bohr (rows) for the periodic Gamma-point path."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64).reshape(-1, 3)
    nat = len(z)
    nsh, nao = basis_size(z, param_file)
    o = default_options(**opts)
    out = {
        "gradient": np.zeros((nat, 3)), "electronic_gradient": np.zeros((nat, 3)),
        "repulsion_gradient": np.zeros((nat, 3)), "dispersion_gradient": np.zeros((nat, 3)),
        "halogen_gradient": np.zeros((nat, 3)), "shell_charges": np.zeros(nsh),
        "atomic_charges": np.zeros(nat), "shell_scc_potential": np.zeros(nsh),
        "coordination_numbers": np.zeros(nat), "orbital_energies": np.zeros(nao),
        "occupations": np.zeros(nao),
    }
    if want_matrices:
        for k in ("overlap", "h0", "fock", "density", "energy_weighted_density"):
            out[k] = np.zeros((nao, nao))
    r = Result()
    for k, v in out.items():
        setattr(r, k, v.ctypes.data_as(_DP))
    q0 = None
    if initial_shell_charges is not None:
        q0 = np.ascontiguousarray(initial_shell_charges, dtype=np.float64)
        assert q0.shape == (nsh,)
    c = None if cell is None else np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
    rc = lib().oracle_run_cell(ctx(param_file), nat, _ptr(z), _ptr(x), _ptr(c) if c is not None else None,
                               ctypes.byref(o), _ptr(q0) if q0 is not None else None, ctypes.byref(r))
    out["status"] = rc
    out["message"] = r.message.decode()
    out["energies"] = dict(zip(ENERGY_NAMES, list(r.energies)))
    out["energy"] = r.energies[0]
    out["fermi_level"] = r.fermi_level
    out["electronic_temperature"] = r.electronic_temperature
    out["nelec"] = r.nelec
    out["dipole"] = np.array(r.dipole[:])
    out["iterations"] = r.iterations
    out["converged"] = bool(r.converged)
    return out


def dispersion(numbers, positions_bohr):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    g = np.zeros((len(z), 3))
    e = lib().oracle_dispersion(ctx(), len(z), _ptr(z), _ptr(x), _ptr(g))
    return e, g


def repulsion(numbers, positions_bohr):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    g = np.zeros((len(z), 3))
    e = lib().oracle_repulsion(ctx(), len(z), _ptr(z), _ptr(x), _ptr(g))
    return e, g


def halogen(numbers, positions_bohr):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    g = np.zeros((len(z), 3))
    e = lib().oracle_halogen(ctx(), len(z), _ptr(z), _ptr(x), _ptr(g))
    return e, g


def coordination(numbers, positions_bohr, cutoff=30.0):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    cn = np.zeros(len(z))
    lib().oracle_coordination(len(z), _ptr(z), _ptr(x), cutoff, _ptr(cn))
    return cn


def overlap(numbers, positions_bohr, cutoff=30.0, weights=None):
    """Overlap matrix; with ``weights`` also sum_{mu,nu} W dS/dR per atom."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    nsh, nao = basis_size(z)
    S = np.zeros((nao, nao))
    dS = np.zeros((len(z), 3))
    W = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    rc = lib().oracle_overlap(ctx(), len(z), _ptr(z), _ptr(x), cutoff, _ptr(S),
                              _ptr(W) if W is not None else None, _ptr(dS) if W is not None else None)
    if rc:
        raise ValueError("overlap failed (%d)" % rc)
    return (S, dS) if W is not None else S


def eigh(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    if lib().oracle_eigh(n, _ptr(a), _ptr(w), _ptr(v)):
        raise RuntimeError("oracle eigh failed")
    return w, v


def run_batch(systems, nthreads=0, **opts):
    """Energies/gradients of many systems on the host cores (OpenMP)."""
    offs = np.zeros(len(systems) + 1, dtype=np.int32)
    for i, (zz, _) in enumerate(systems):
        offs[i + 1] = offs[i] + len(zz)
    z = np.concatenate([np.asarray(s[0], dtype=np.int32) for s in systems])
    x = np.concatenate([np.asarray(s[1], dtype=np.float64).reshape(-1, 3) for s in systems])
    e = np.zeros(len(systems))
    g = np.zeros((offs[-1], 3))
    it = np.zeros(len(systems), dtype=np.int32)
    st = np.zeros(len(systems), dtype=np.int32)
    o = default_options(**opts)
    lib().oracle_run_batch(ctx(), len(systems), _ptr(offs), _ptr(z), _ptr(x), ctypes.byref(o),
                           _ptr(e), _ptr(g), _ptr(it), _ptr(st), nthreads)
    return e, g, it, st


def element_value(z, key, shell=0, param_file=None):
    """Parsed value of one parameter of element z (see oracle_element_value);
    None when the element or shell has no such entry."""
    f = lib().oracle_element_value
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    out = ctypes.c_double()
    rc = f(ctx(param_file), int(z), key.encode(), int(shell), ctypes.byref(out))
    return None if rc else out.value


def global_value(key, param_file=None):
    f = lib().oracle_global
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    return f(ctx(param_file), key.encode())


def pair_scaling(a, b, param_file=None):
    f = lib().oracle_pair_scaling
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return f(ctx(param_file), int(a), int(b))


def shells(numbers, param_file=None):
    """Per shell (basis order): atom, l, valence, nao, hardness, CN-free self energy."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    nsh, _ = basis_size(z, param_file)
    ints = [np.zeros(nsh, dtype=np.int32) for _ in range(4)]
    dbl = [np.zeros(nsh) for _ in range(2)]
    lib().oracle_shells.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 7
    rc = lib().oracle_shells(ctx(param_file), len(z), _ptr(z), *[_ptr(a) for a in ints + dbl])
    if rc:
        raise ValueError("basis build failed (%d)" % rc)
    return dict(atom=ints[0], l=ints[1], valence=ints[2].astype(bool), nao=ints[3], hardness=dbl[0], hdiag=dbl[1])


def gamma(numbers, positions_bohr, param_file=None):
    """The shell Coulomb matrix of the SCC (nsh x nsh)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    nsh, _ = basis_size(z, param_file)
    out = np.zeros((nsh, nsh))
    lib().oracle_gamma.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    rc = lib().oracle_gamma(ctx(param_file), len(z), _ptr(z), _ptr(x), _ptr(out))
    if rc:
        raise ValueError("basis build failed (%d)" % rc)
    return out


# ---- pieces of the periodic Gamma-point path

def _cellarr(cell):
    return np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)


def ewald_atoms(positions_bohr, cell, alpha=0.0):
    """Ewald (1/R) potential matrix of unit point charges and their images,
    with the neutralising background and the self term (nat x nat)."""
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64).reshape(-1, 3)
    c = _cellarr(cell)
    out = np.zeros((len(x), len(x)))
    f = lib().oracle_ewald_atoms
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    if f(len(x), _ptr(x), _ptr(c), alpha, _ptr(out)):
        raise ValueError("singular cell")
    return out


def gamma_cell(numbers, positions_bohr, cell, alpha=0.0):
    """Periodic shell Coulomb matrix of the Gamma-point SCC (nsh x nsh)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    c = _cellarr(cell)
    nsh, _ = basis_size(z)
    out = np.zeros((nsh, nsh))
    f = lib().oracle_gamma_cell
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_double, ctypes.c_void_p]
    if f(ctx(), len(z), _ptr(z), _ptr(x), _ptr(c), alpha, _ptr(out)):
        raise ValueError("gamma_cell failed")
    return out


def gamma_cell_gradient(numbers, positions_bohr, cell, q):
    """Gradient of 1/2 q^T Gamma q at fixed shell charges (nat x 3)."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    c = _cellarr(cell)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    g = np.zeros((len(z), 3))
    f = lib().oracle_gamma_cell_gradient
    f.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5
    if f(ctx(), len(z), _ptr(z), _ptr(x), _ptr(c), _ptr(qq), _ptr(g)):
        raise ValueError("gamma_cell_gradient failed")
    return g


def classical_cell(numbers, positions_bohr, cell, which):
    """Periodic repulsion / dispersion / halogen energy and gradient."""
    idx = {"repulsion": 0, "dispersion": 1, "halogen": 2}[which]
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    c = _cellarr(cell)
    g = np.zeros((len(z), 3))
    f = lib().oracle_classical_cell
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p]
    e = f(ctx(), len(z), _ptr(z), _ptr(x), _ptr(c), idx, _ptr(g))
    return e, g


def coordination_cell(numbers, positions_bohr, cell, cutoff=30.0):
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    c = _cellarr(cell)
    cn = np.zeros(len(z))
    f = lib().oracle_coordination_cell
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    if f(len(z), _ptr(z), _ptr(x), _ptr(c), cutoff, _ptr(cn)):
        raise ValueError("singular cell")
    return cn


def hs(numbers, positions_bohr, cell=None, cn_on=True):
    """S and H0 of a molecule (pairs within 30 bohr), or S(Gamma) and H0(Gamma)
    of a cell, with or without the CN shift of the self energies."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64)
    _, nao = basis_size(z)
    S = np.zeros((nao, nao))
    H = np.zeros((nao, nao))
    c = None if cell is None else _cellarr(cell)
    f = lib().oracle_hs
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_void_p]
    if f(ctx(), len(z), _ptr(z), _ptr(x), _ptr(c) if c is not None else None, int(bool(cn_on)), _ptr(S), _ptr(H)):
        raise ValueError("hs failed")
    return S, H
