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


def test_parameter_sources_and_descriptions(tmp_path, monkeypatch):
    """Gfn1Parameters::resolve specs and source_description (params.rs:204-270,
    tests params.rs:1104-1145): bundled GFN1-xTB and GFN1(Si)-xTB, an explicit
    path, and the GFN1_XTB_PARAM variable."""
    import shutil

    from gfn1_amd import Gfn1Parameters, Gfn1NativeCalculator
    base = Gfn1Parameters.resolve("builtin")
    d = base.source_description()
    assert d.startswith("GFN1-xTB (builtin, grimme-lab/xtb@2b5cd48, LGPL-3.0-or-later; refs: ")
    assert "10.1021/acs.inorgchem.7b01950" in d
    for spec in ("builtin:si", "BUILTIN:GFN1-SI", " builtin-si "):
        si = Gfn1Parameters.resolve(spec).source_description()
        assert si.startswith("GFN1(Si)-xTB (builtin") and "10.1021/acs.jcim.1c01170" in si
    src = os.path.join(ROOT, "gfn1-rs_amd", "data", "param_gfn1-xtb.txt")
    dst = tmp_path / "param_copy.txt"
    shutil.copy(src, dst)
    assert Gfn1Parameters.resolve(str(dst)).source_description() == "GFN1-xTB (%s)" % dst
    monkeypatch.setenv("GFN1_XTB_PARAM", str(dst))
    assert Gfn1Parameters.resolve().source_description() == "GFN1-xTB (%s (via GFN1_XTB_PARAM))" % dst
    assert Gfn1Parameters.resolve("builtin").source_description() == d
    monkeypatch.delenv("GFN1_XTB_PARAM")
    assert "GFN1-xTB (builtin" in repr(Gfn1NativeCalculator())
    # Gfn1Parameters::from_str: the text of the bundled file, source "inline text"
    with open(src) as f:
        inline = Gfn1Parameters.from_str(f.read())
    assert inline.source_description() == "GFN1-xTB (inline text)"
    from gfn1_amd import NativeError
    with pytest.raises(NativeError) as e:
        Gfn1Parameters.from_str("$globpar\nks not-a-number\n")
    assert e.value.kind in ("Parse", "MissingGlobal")

