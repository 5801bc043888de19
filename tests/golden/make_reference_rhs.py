"""Write tests/golden/reference_rhs.json: the reference's own network
definitions (pre_odevec/example_systems.py) evaluated at seeded points.

Runs only in the build container, where the reference is mounted at
/root/reference. load_orego/load_primordial use `symbols` and `Matrix`
without importing them (pre_odevec/example_systems.py:27,33,65-69) and
load_orego returns a `k` it never defines (:37); the
reference's preprocessor would have them from its own `from sympy import *`
(pre_odevec.py:34), so they are put into the module's namespace here.
ROBER's rates are the reference's literals (0.04, 3e7, 1e4).
"""
import json
import os
import sys

import numpy as np
import sympy

REF = os.environ.get("ODEVEC_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_rhs.json")


def main():
    sys.path.insert(0, REF)
    import pre_odevec.example_systems as ex
    ex.symbols = sympy.symbols
    ex.Matrix = sympy.Matrix
    ex.k = 0   # load_orego returns an undefined k (pre_odevec/example_systems.py:37)
    rng = np.random.default_rng(11)
    out = {"source": "pre_odevec/example_systems.py", "sympy": sympy.__version__, "cases": {}}
    for name, loader in (("ROBER", ex.load_rober), ("OREGO", ex.load_orego), ("PRIMORDIAL", ex.load_primordial)):
        neq, k, nrea, y, rhs = loader()
        ks = list(k) if isinstance(k, (list, tuple)) else []
        f = sympy.lambdify(list(y) + ks, list(rhs), modules="math")
        pts = []
        for _ in range(6):
            yv = [float(v) for v in np.exp(rng.uniform(np.log(1e-3), np.log(10.0), neq))]
            kv = [float(v) for v in np.exp(rng.uniform(np.log(1e-3), np.log(10.0), len(ks)))]
            fv = [float(v) for v in f(*(yv + kv))]
            pts.append({"y": yv, "k": kv, "rhs": fv})
        out["cases"][name] = {"neq": int(neq), "nrate": len(ks), "points": pts}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
