"""Write the parameters of the HD+H example's neural-network surface as data.

The reference's example (EXAMPLES/HD+H_collision/pes/FFmodules.f90, module
H3_pot) fixes its H3 surface by Fortran DATA statements: the input scaling
xstandard (mean, std of the three invariants), the output scaling ylinear,
and the weights and biases of a 3-20-20-1 network (w, b, w1, b1, w2, b2).
This script reads those statements where the reference tree exists and
stores every number as the decimal string written there (they are
default-real, i.e. single precision, literals, so the string is kept
exactly) in h3_nn_params.json, indexed like the Fortran arrays (first index
= input unit).  Run only where the reference tree exists; the output is
committed.
"""
import json
import os
import re

REF = os.environ.get("MW3_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "EXAMPLES", "HD+H_collision", "pes", "FFmodules.f90")
HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = {"xstandard": (3, 2), "ylinear": (2,), "w": (3, 20), "b": (20,), "w1": (20, 20),
          "b1": (20,), "w2": (20, 1), "b2": (1,)}


def _range(t):
    if ":" in t:
        a, b = t.split(":")
        return list(range(int(a), int(b) + 1))
    return [int(t)]


def main():
    arrays = {k: [[None] * s[1] for _ in range(s[0])] if len(s) == 2 else [None] * s[0]
              for k, s in SHAPES.items()}
    pat = re.compile(r"^\s*data\s+(\w+)\(([^)]*)\)\s*/([^/]*)/", re.I)
    for ln in open(SRC):
        m = pat.match(ln)
        if not m:
            continue
        name, idx, vals = m.group(1).lower(), m.group(2), m.group(3)
        vals = [v.strip() for v in vals.split(",") if v.strip()]
        parts = [p.strip() for p in idx.split(",")]
        shape = SHAPES[name]
        if len(shape) == 1:
            for i, v in zip(_range(parts[0]) if parts[0] != ":" else range(1, shape[0] + 1), vals):
                arrays[name][i - 1] = v
        else:
            rows = _range(parts[0]) if parts[0] != ":" else list(range(1, shape[0] + 1))
            cols = _range(parts[1]) if parts[1] != ":" else list(range(1, shape[1] + 1))
            k = 0
            # a DATA list fills the named section in array-element order
            # (first index fastest)
            for cidx in cols:
                for ridx in rows:
                    arrays[name][ridx - 1][cidx - 1] = vals[k]
                    k += 1
    for k, a in arrays.items():
        flat = a if isinstance(a[0], str) or a[0] is None else [x for r in a for x in r]
        if any(v is None for v in flat):
            raise ValueError("%s not fully set" % k)
    with open(os.path.join(HERE, "h3_nn_params.json"), "w") as f:
        json.dump(arrays, f, indent=0)
    print("written", {k: SHAPES[k] for k in arrays})


if __name__ == "__main__":
    main()
