"""Writes generators.json: edge counts and FNV-1a 64 hashes of synthetic CSRs.

The reference has no generator, so these are self-fixtures of this project's
RMAT/grid specification, computed by the serial oracle (oracle/oracle.c).
The device generators are checked against the same pins by the GPU tests.
Run: python tests/golden/make_generator_pins.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle_py as O  # noqa: E402

SPECS = {
    "rmat_s10": {"kind": "rmat", "scale": 10},
    "rmat_s12_w255": {"kind": "rmat", "scale": 12, "weights": [1, 255], "wseed": 3},
    "rmat_s16": {"kind": "rmat", "scale": 16},
    "grid_8x8": {"kind": "grid", "rows": 8, "cols": 8},
    "grid_64x96_w1000": {"kind": "grid", "rows": 64, "cols": 96, "weights": [1, 1000],
                         "wseed": 2},
}

if __name__ == "__main__":
    out = {}
    for k, s in SPECS.items():
        if s["kind"] == "rmat":
            g = O.gen_rmat(s["scale"], weights=s.get("weights"), wseed=s.get("wseed", 0))
        else:
            g = O.gen_grid(s["rows"], s["cols"], weights=s.get("weights"), wseed=s.get("wseed", 0))
        out[k] = dict(s, E=g.E, fnv="%016x" % g.fnv())
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))
