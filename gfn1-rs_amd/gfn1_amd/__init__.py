"""MI355X GFN1-xTB SCC energy + analytic gradient engine (Python host mirror).

Python surface of the C ABI in include/gfn1_amd.h:
  Gfn1Parameters.resolve()      params.rs:204-230
  run_electronic(...)           electronic.rs:800
  analytic_gradient(...)        gradient.rs:56
  Gfn1NativeCalculator          python.rs:115-381
  BatchEngine                   batched energy + gradient of many systems
  GuessModel                    learned SCC start charges (GPU-trained MLP)
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
    GuessModel,
    HARTREE_TO_EV,
    basis_counts,
    guess_features,
    analytic_gradient,
    run_electronic,
)
from ._lib import NativeError  # noqa: F401
