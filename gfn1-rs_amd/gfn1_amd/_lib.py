"""ctypes binding of libgfn1_amd.so (include/gfn1_amd.h).

The library is the only compute path: importing this module on a machine
without the built library, or calling into it without a HIP device, raises.
"""
import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libgfn1_amd.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "gfn1_amd.h")

GFN1_OK = 0
STATUS_NAMES = {
    0: "Ok", 1: "Io", 2: "Parse", 3: "InvalidInput", 4: "MissingElement", 5: "MissingGlobal",
    6: "LinearAlgebra", 7: "SingularCell", 8: "SccNotConverged", 100: "Device",
}
ENERGY_NAMES = (
    "total_free", "total_internal", "repulsion", "electronic", "isotropic_scc",
    "third_order", "higher_order", "dispersion", "halogen", "multipole", "exchange",
    "spin_polarization", "plus_u", "external_field", "electronic_entropy",
)

_DP = ctypes.POINTER(ctypes.c_double)


class Gfn1Error(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("iterations", ctypes.c_int), ("rms", ctypes.c_double),
                ("message", ctypes.c_char * 256)]


class System(ctypes.Structure):
    _fields_ = [("nat", ctypes.c_int32), ("numbers", ctypes.POINTER(ctypes.c_int32)),
                ("positions", _DP), ("charge", ctypes.c_double), ("lattice", _DP)]


class ElectronicOptions(ctypes.Structure):
    _fields_ = [
        ("has_charge", ctypes.c_int32), ("charge", ctypes.c_double), ("max_scc", ctypes.c_int32),
        ("energy_tolerance", ctypes.c_double), ("charge_tolerance", ctypes.c_double),
        ("mixing", ctypes.c_double), ("scc_broyden", ctypes.c_int32), ("scc_broyden_size", ctypes.c_int32),
        ("level_shift", ctypes.c_double), ("nprim", ctypes.c_int32), ("eigen_tolerance", ctypes.c_double),
        ("electronic_temperature", ctypes.c_double), ("enable_dispersion", ctypes.c_int32),
        ("integral_cutoff", ctypes.c_double), ("coordination_cutoff", ctypes.c_double),
        ("enable_cn_hamiltonian", ctypes.c_int32), ("scc_initial_shell_charges", _DP),
        ("scc_initial_shell_charges_len", ctypes.c_int32),
    ]


class GradientOptions(ctypes.Structure):
    _fields_ = [("electronic", ElectronicOptions), ("include_repulsion", ctypes.c_int32),
                ("include_dispersion", ctypes.c_int32), ("include_hamiltonian", ctypes.c_int32),
                ("include_scc", ctypes.c_int32), ("include_halogen", ctypes.c_int32)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("gradient", GradientOptions), ("precision", ctypes.c_int32), ("guess", ctypes.c_int32),
                ("eigensolver", ctypes.c_int32), ("keep_matrices", ctypes.c_int32),
                ("scc_graphs", ctypes.c_int32)]


class OptimizationOptions(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int32), ("gradient_tolerance", ctypes.c_double),
                ("step_tolerance", ctypes.c_double), ("history", ctypes.c_int32),
                ("initial_step", ctypes.c_double), ("max_atom_step", ctypes.c_double),
                ("warm_start", ctypes.c_int32), ("keep_trajectory", ctypes.c_int32)]


class HessianOptions(ctypes.Structure):
    _fields_ = [("step", ctypes.c_double), ("scc_energy_tolerance", ctypes.c_double),
                ("scc_charge_tolerance", ctypes.c_double), ("include_repulsion", ctypes.c_int32),
                ("include_electronic", ctypes.c_int32), ("include_dispersion", ctypes.c_int32),
                ("include_halogen", ctypes.c_int32), ("warm_start", ctypes.c_int32), ("max_batch", ctypes.c_int32)]


class MdOptions(ctypes.Structure):
    _fields_ = [("timestep_fs", ctypes.c_double), ("steps", ctypes.c_int32), ("warm_start", ctypes.c_int32)]


# (name, restype, argtypes) of every exported entry point
_VP = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)


class GuessTrainOptions(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_int32), ("batch_atoms", ctypes.c_int32), ("learning_rate", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


GUESS_PARAMS = 8706
GUESS_FEATURES = 68
SIGNATURES = [
    ("gfn1_params_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_params_free", None, [_VP]),
    ("gfn1_electronic_options_default", None, [ctypes.POINTER(ElectronicOptions)]),
    ("gfn1_gradient_options_default", None, [ctypes.POINTER(GradientOptions)]),
    ("gfn1_run_electronic", ctypes.c_int, [ctypes.POINTER(System), _VP, ctypes.POINTER(ElectronicOptions),
                                           ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_analytic_gradient", ctypes.c_int, [ctypes.POINTER(System), _VP, ctypes.POINTER(GradientOptions),
                                              ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_electronic_result_free", None, [_VP]),
    ("gfn1_gradient_result_free", None, [_VP]),
    ("gfn1_electronic_result_sizes", None, [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                            ctypes.POINTER(ctypes.c_int32)]),
    ("gfn1_electronic_result_energy_terms", None, [_VP, _DP]),
    ("gfn1_electronic_result_matrix", _DP, [_VP, ctypes.c_int]),
    ("gfn1_electronic_result_vector", _DP, [_VP, ctypes.c_int]),
    ("gfn1_electronic_result_scalar", ctypes.c_double, [_VP, ctypes.c_int]),
    ("gfn1_gradient_result_electronic", _VP, [_VP]),
    ("gfn1_gradient_result_total_energy", ctypes.c_double, [_VP]),
    ("gfn1_gradient_result_max_gradient", ctypes.c_double, [_VP]),
    ("gfn1_gradient_result_field", _DP, [_VP, ctypes.c_int]),
    ("gfn1_batch_options_default", None, [ctypes.POINTER(BatchOptions)]),
    ("gfn1_batch_create", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(System), _VP, ctypes.POINTER(BatchOptions),
                                         ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_run", ctypes.c_int, [_VP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_status", ctypes.c_int, [_VP, ctypes.c_int32]),
    ("gfn1_batch_energy", ctypes.c_double, [_VP, ctypes.c_int32]),
    ("gfn1_batch_energy_terms", None, [_VP, ctypes.c_int32, _DP]),
    ("gfn1_batch_iterations", ctypes.c_int32, [_VP, ctypes.c_int32]),
    ("gfn1_batch_gradient", ctypes.c_int, [_VP, ctypes.c_int32, _DP]),
    ("gfn1_batch_shell_charges", ctypes.c_int, [_VP, ctypes.c_int32, _DP]),
    ("gfn1_batch_timings", None, [_VP, _DP]),
    ("gfn1_batch_kernel_stats", ctypes.c_int,
     [_VP, ctypes.c_int, _DP, ctypes.POINTER(ctypes.c_int64), _DP]),
    ("gfn1_batch_launch_stats", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int64)]),
    ("gfn1_batch_autotune", ctypes.c_int, [_VP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_set_backtransform_width", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_tuning_count", ctypes.c_int32, [_VP]),
    ("gfn1_batch_tuning_entry", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _DP]),
    ("gfn1_batch_free", None, [_VP]),
    ("gfn1_params_shell_count", ctypes.c_int, [_VP, ctypes.c_int]),
    ("gfn1_guess_model_create", ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_model_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_model_save", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_model_get_params", ctypes.c_int, [_VP, _FP, _FP]),
    ("gfn1_guess_model_set_params", ctypes.c_int, [_VP, _FP]),
    ("gfn1_guess_model_free", None, [_VP]),
    ("gfn1_guess_train_options_default", None, [ctypes.POINTER(GuessTrainOptions)]),
    ("gfn1_guess_train", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(System), _VP, _DP,
                                        ctypes.POINTER(GuessTrainOptions), _FP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_predict", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(System), _VP, _DP,
                                          ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_predict_noalloc", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(System), _VP, _DP,
                                          ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_noalloc_scratch_bytes", ctypes.c_size_t, []),
    ("gfn1_guess_features", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(System), _FP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_guess_loss_gradient", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.POINTER(System), _VP, _DP, _DP, _FP,
                                                ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_set_guess", ctypes.c_int, [_VP, _VP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_set_positions", ctypes.c_int, [_VP, _DP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_set_start_charges", ctypes.c_int, [_VP, ctypes.c_int32, _DP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_statuses", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int32)]),
    ("gfn1_batch_positions", ctypes.c_int, [_VP, _DP]),
    ("gfn1_md_options_default", None, [ctypes.POINTER(MdOptions)]),
    ("gfn1_batch_md", ctypes.c_int, [_VP, ctypes.POINTER(MdOptions), _DP, _DP, _DP, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_batch_energies", ctypes.c_int, [_VP, _DP]),
    ("gfn1_batch_gradients", ctypes.c_int, [_VP, _DP]),
    ("gfn1_batch_all_shell_charges", ctypes.c_int, [_VP, _DP]),
    ("gfn1_batch_set_all_start_charges", ctypes.c_int, [_VP, _DP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_optimization_options_default", None, [ctypes.POINTER(OptimizationOptions)]),
    ("gfn1_batch_optimize", ctypes.c_int, [_VP, ctypes.POINTER(OptimizationOptions), ctypes.POINTER(_VP),
                                           ctypes.POINTER(Gfn1Error)]),
    ("gfn1_optimize_geometry", ctypes.c_int, [ctypes.POINTER(System), _VP, ctypes.POINTER(GradientOptions),
                                              ctypes.POINTER(OptimizationOptions), ctypes.POINTER(_VP),
                                              ctypes.POINTER(Gfn1Error)]),
    ("gfn1_optimization_count", ctypes.c_int32, [_VP]),
    ("gfn1_optimization_status", ctypes.c_int, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_message", ctypes.c_char_p, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_energy", ctypes.c_double, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_max_gradient", ctypes.c_double, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_iterations", ctypes.c_int32, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_converged", ctypes.c_int32, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_positions", ctypes.c_int, [_VP, ctypes.c_int32, _DP]),
    ("gfn1_optimization_gradient", ctypes.c_int, [_VP, ctypes.c_int32, _DP]),
    ("gfn1_optimization_trajectory_length", ctypes.c_int32, [_VP, ctypes.c_int32]),
    ("gfn1_optimization_trajectory_step", ctypes.c_int,
     [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _DP, _DP, _DP, _DP]),
    ("gfn1_optimization_stats", None, [_VP, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64), _DP]),
    ("gfn1_optimization_free", None, [_VP]),
    ("gfn1_hessian_options_default", None, [ctypes.POINTER(HessianOptions)]),
    ("gfn1_hessian", ctypes.c_int, [ctypes.POINTER(System), _VP, ctypes.POINTER(ElectronicOptions),
                                    ctypes.POINTER(HessianOptions), _DP, _DP, _DP, _DP, ctypes.POINTER(Gfn1Error)]),
    ("gfn1_hessian_terms", ctypes.c_int, [ctypes.POINTER(System), _VP, ctypes.POINTER(ElectronicOptions),
                                          ctypes.POINTER(HessianOptions), _DP, _DP, _DP, _DP, _DP,
                                          ctypes.POINTER(Gfn1Error)]),
    ("gfn1_hessian_batch", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(System), _VP,
                                          ctypes.POINTER(ElectronicOptions), ctypes.POINTER(HessianOptions), _DP,
                                          _DP, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_system_from_xyz", ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(_VP),
                                            ctypes.POINTER(Gfn1Error)]),
    ("gfn1_xyz_system", ctypes.POINTER(System), [_VP]),
    ("gfn1_xyz_free", None, [_VP]),
    ("gfn1_params_source", ctypes.c_int32, [_VP, ctypes.c_char_p, ctypes.c_int32]),
    ("gfn1_params_from_text", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.POINTER(Gfn1Error)]),
    ("gfn1_version", ctypes.c_char_p, []),
]

_lib = None


def request_hw_queues(count=16):
    """Ask the HIP runtime for `count` hardware queues (GPU_MAX_HW_QUEUES).

    The batch engine runs one HIP stream per size bucket (up to 14 for
    matrices of 32..448 AO); the runtime maps streams onto GPU_MAX_HW_QUEUES
    hardware queues (4 by default), and streams that share a queue run one
    after another.  The variable is read once, when the HIP runtime starts, so
    this must be called before anything in the process uses the GPU; a value
    the caller already exported is left alone.  Returns the value in effect.
    Opt-in: loading the library changes nothing in the process environment."""
    os.environ.setdefault("GPU_MAX_HW_QUEUES", str(int(count)))
    return int(os.environ["GPU_MAX_HW_QUEUES"])


def load():
    """Load the built library; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgfn1_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


class NativeError(RuntimeError):
    def __init__(self, code, message, iterations=0, rms=0.0):
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))
        self.iterations = iterations
        self.rms = rms
        super().__init__("%s: %s" % (self.kind, message))


def check(code, err):
    if code != GFN1_OK:
        raise NativeError(code, err.message.decode(errors="replace"), err.iterations, err.rms)


def make_system(numbers, positions_bohr, charge=0.0, lattice=None):
    """lattice: None for a molecule, or the three cell vectors (rows, bohr) of a
    fully periodic cell."""
    z = np.ascontiguousarray(numbers, dtype=np.int32)
    x = np.ascontiguousarray(positions_bohr, dtype=np.float64).reshape(-1, 3)
    if len(z) != len(x):
        raise ValueError("numbers and positions differ in length")
    cell = None
    if lattice is not None:
        cell = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    s = System(len(z), z.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), x.ctypes.data_as(_DP), float(charge),
               cell.ctypes.data_as(_DP) if cell is not None else None)
    return s, (z, x, cell)  # keep the arrays alive with the struct
