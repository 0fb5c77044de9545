"""Writes tests/golden/set_weights.json: the barotropic time-averaging weights
(weight1, weight2, nfast) that the reference's own set_weights
(nonlinear/set_weights.py:45-231) computes for several splitting ratios.

Run in a container that holds the reference sources (default /root/reference, or
the path given as the first argument):  python tests/golden/make_set_weights_fixture.py

The reference module imports `misc` (utility/misc.py, plain Python) and `cupy`,
which its set_weights body never uses; an empty module stands in for `cupy` so that
the file imports without a GPU.  The weights are plain Python floats computed by the
reference's own loops; the fixture stores them as JSON numbers (repr round-trips FP64
exactly)."""
from __future__ import annotations

import json
import sys
import types
from pathlib import Path

REF = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
OUT = Path(__file__).resolve().parent / "set_weights.json"
NDTFAST = (10, 15, 20, 30, 40, 60)


class _Times:
    """The CompTimes attributes set_weights reads and writes (mod_comptimes.py:7-108)."""

    def __init__(self, ndtfast):
        import numpy as np
        self.ndtfast = ndtfast
        self.nfast = 0
        self.weight1 = np.zeros(257)
        self.weight2 = np.zeros(257)


def main():
    sys.modules.setdefault("cupy", types.ModuleType("cupy"))
    sys.path.insert(0, str(REF / "utility"))
    sys.path.insert(0, str(REF / "nonlinear"))
    import set_weights as ref   # the reference module
    out = {"source": "nonlinear/set_weights.py:45-231 (reference), POWER_LAW filter, Falpha=2, Fbeta=4, Fgamma=0.284",
           "cases": []}
    for nd in NDTFAST:
        t = _Times(nd)
        ref.set_weights(t)
        n = int(t.nfast)
        out["cases"].append({"ndtfast": nd, "nfast": n,
                             "weight1": [float(x) for x in t.weight1[: n + 2]],
                             "weight2": [float(x) for x in t.weight2[: n + 2]]})
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT} ({len(out['cases'])} cases)")


if __name__ == "__main__":
    main()
