"""A minimal stand-in for the two ASE modules gfn1_amd.ase imports, for
machines without ASE: the Calculator base (parameters, results, reset) and
ase.units with ASE's own CODATA 2014 values.  With ASE installed the real
package is used and nothing is replaced.  Test helper only."""
import sys
import types

ASE_BOHR = 0.52917721067           # ase.units (CODATA 2014)
ASE_HARTREE = 27.211386024367243
ASE_INVCM = 1.0 / 8065.544005


def install_stand_in():
    """Put the stand-in in sys.modules unless ASE imports; True when it did."""
    try:
        import ase  # noqa: F401
        return False
    except ImportError:
        pass
    ase = types.ModuleType("ase")
    calc_pkg = types.ModuleType("ase.calculators")
    calc_mod = types.ModuleType("ase.calculators.calculator")
    units = types.ModuleType("ase.units")
    units.Bohr, units.Hartree, units.invcm = ASE_BOHR, ASE_HARTREE, ASE_INVCM

    class Parameters(dict):
        def __getattr__(self, k):
            return self[k]

    class Calculator:
        default_parameters = {}

        def __init__(self, restart=None, label=None, atoms=None, **kwargs):
            self.parameters = Parameters(self.default_parameters)
            self.parameters.update(kwargs)
            self.atoms = atoms
            self.results = {}

        def set(self, **kwargs):
            changed = {k: v for k, v in kwargs.items() if self.parameters.get(k) != v}
            self.parameters.update(changed)
            if changed:
                self.results = {}
            return changed

        def calculate(self, atoms=None, properties=("energy",), system_changes=None):
            if atoms is not None:
                self.atoms = atoms

        def reset(self):
            self.results = {}

    calc_mod.Calculator = Calculator
    calc_mod.all_changes = ["positions", "numbers", "cell", "pbc"]
    sys.modules.update({"ase": ase, "ase.calculators": calc_pkg, "ase.calculators.calculator": calc_mod,
                        "ase.units": units})
    return True


