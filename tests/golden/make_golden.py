"""Write tests/golden/oracle_fixtures.json: energies, energy terms, gradients and
converged shell charges of the FP64 oracle (tight SCC tolerances) for the test
molecules, a water dimer and one mixed cluster of the batched workload.

These are self-generated fixtures (the reference holds none for these systems):
they pin the oracle against regressions and let GPU tests compare with fixed
numbers.  Run from the repository root after `make -C oracle`:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "gfn1-rs_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from gfn1_amd import geometry  # noqa: E402
from test_oracle import MOLECULES, TIGHT, mol  # noqa: E402


def systems():
    out = {name: mol(name) for name in sorted(MOLECULES)}
    out["water_dimer"] = geometry.water_cluster(2, 4.0, seed=7)
    out["mixed_cluster_1003"] = geometry.mixed_cluster(1003)
    return out


def main():
    fixtures = {"tolerances": TIGHT, "units": "bohr, hartree", "systems": {}}
    for name, (z, x) in systems().items():
        r = oracle.run(z, x, **TIGHT)
        assert r["status"] == 0, name
        fixtures["systems"][name] = {
            "numbers": [int(v) for v in z], "positions": np.asarray(x).tolist(),
            "energy": r["energy"], "energies": {k: float(v) for k, v in r["energies"].items()},
            "gradient": np.asarray(r["gradient"]).tolist(),
            "shell_charges": np.asarray(r["shell_charges"]).tolist(),
        }
        print("%-24s %3d atoms  E = %.12f" % (name, len(z), r["energy"]))
    with open(os.path.join(HERE, "oracle_fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1)


if __name__ == "__main__":
    main()
