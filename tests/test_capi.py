"""The C-ABI library: it loads, exports every symbol include/gfn1_amd.h declares,
and its host-side input handling behaves (no device compute on the CPU)."""
import ctypes
import os
import re

import pytest

from gfn1_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gfn1_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gfn1_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the header exactly
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_version_string():
    assert b"gfn1" in _lib.load().gfn1_version()


def test_params_load_and_errors():
    lib = _lib.load()
    h = ctypes.c_void_p()
    err = _lib.Gfn1Error()
    assert lib.gfn1_params_load(None, ctypes.byref(h), ctypes.byref(err)) == 0
    lib.gfn1_params_free(h)
    rc = lib.gfn1_params_load(b"/nonexistent/param.txt", ctypes.byref(h), ctypes.byref(err))
    assert rc == 1 and b"cannot read" in err.message


def test_option_defaults_match_reference():
    # electronic.rs:395-452 and gradient.rs:31-41
    o = _lib.GradientOptions()
    _lib.load().gfn1_gradient_options_default(ctypes.byref(o))
    e = o.electronic
    assert (e.max_scc, e.energy_tolerance, e.charge_tolerance, e.mixing) == (250, 1e-6, 2e-5, 0.4)
    assert (e.scc_broyden, e.scc_broyden_size, e.level_shift, e.nprim) == (1, 250, 0.0, 0)
    assert (e.eigen_tolerance, e.electronic_temperature, e.enable_dispersion) == (1e-12, 300.0, 1)
    assert (e.integral_cutoff, e.coordination_cutoff, e.enable_cn_hamiltonian) == (30.0, 30.0, 1)
    assert (o.include_repulsion, o.include_dispersion, o.include_hamiltonian, o.include_scc,
            o.include_halogen) == (1, 1, 1, 1, 1)


def test_calculator_rejects_bad_units():
    from gfn1_amd import Gfn1NativeCalculator
    calc = Gfn1NativeCalculator()
    with pytest.raises(ValueError):
        calc.calculate([1, 1], [[0, 0, 0], [0.74, 0, 0]], unit="furlong")
