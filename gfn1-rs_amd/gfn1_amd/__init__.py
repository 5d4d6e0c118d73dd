"""MI355X GFN1-xTB SCC energy + analytic gradient engine (Python host mirror).

Python surface of the C ABI in include/gfn1_amd.h:
  Gfn1Parameters.resolve()      params.rs:204-230
  system_from_xyz(...)          system.rs:45-99 (PeriodicSystem::from_xyz_str, non-periodic)
  run_electronic(...)           electronic.rs:800
  analytic_gradient(...)        gradient.rs:56
  Gfn1NativeCalculator          python.rs:115-381
  BatchEngine                   batched energy + gradient of many systems
  GuessModel                    learned SCC start charges (GPU-trained MLP)
  optimize_geometry(...)        optimizer.rs:78 (L-BFGS; BatchEngine.optimize for many systems)
  hessian(...), hessian_batch   analytic_hessian(...).hessian, hessian/mod.rs:112 (central differences
                                of the analytic gradient on the device)
  vibrational_analysis(...)     vibrational.rs:37
  md.run_md(...)                batched velocity-Verlet dynamics over BatchEngine
  ir_spectrum_from(...)         properties.rs:381 (with the dipole derivatives of hessian(...))
"""
from .api import (  # noqa: F401
    ANGSTROM_TO_BOHR,
    AnalyticGradientOptions,
    AnalyticGradientResult,
    BatchEngine,
    CalculationResult,
    ElectronicOptions,
    ElectronicResult,
    Gfn1NativeCalculator,
    Gfn1Parameters,
    GeometryOptimizationOptions,
    GeometryOptimizationResult,
    GeometryOptimizationStep,
    GuessModel,
    HessianOptions,
    HessianResult,
    VibrationalModes,
    OptimizationResult,
    HARTREE_TO_EV,
    basis_counts,
    guess_features,
    analytic_gradient,
    hessian,
    hessian_batch,
    ir_spectrum_from,
    optimize_geometry,
    vibrational_analysis,
    run_electronic,
    system_from_xyz,
    system_from_xyz_file,
)
from ._lib import NativeError, request_hw_queues  # noqa: F401
