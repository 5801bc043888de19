"""Write tests/golden/primordial_fail_cells.json: the cells of the seeded
1,048,576-cell PRIMORDIAL workload (seed 3, rtol 1e-6, atol 1e-20) on which
the oracle -- the reference's algorithm -- ends with a status other than OK
(the reference would STOP there, src/odevec.F90:380-389). About two minutes
on 8 cores."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from odevec_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    w = W.primordial(1048576, seed=3)
    net = O.Network.load(w["network"])
    y, st, dto, cnt = net.solve_cells(w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], rtol=w["rtol"], atol=w["atol"],
                                      dt_min=w["dt_min"], threads=os.cpu_count() or 1)
    bad = np.where(st != 0)[0]
    ok = st == 0
    cons, quant = {}, {}
    for name, v in (("H", [0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0]),
                    ("He", [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0]),
                    ("D", [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0])):
        v = np.array(v, dtype=float)
        a, b = v @ w["y"][:, ok], v @ y[:, ok]
        r = np.abs(a - b) / a
        cons[name] = float(r.max())
        quant[name] = {"q999": float(np.quantile(r, 0.999)), "q99": float(np.quantile(r, 0.99)),
                       "n_gt_1e-6": int((r > 1e-6).sum())}
    out = {"workload": "primordial(1048576, seed=3)", "cells": [int(i) for i in bad],
           "status": [int(st[i]) for i in bad], "ncells": 1048576,
           "mean_attempts": float(cnt[0].mean()), "mean_steps": float(cnt[1].mean()),
           # largest relative drift of the nuclei counts over the finished cells:
           # the method conserves them only up to rounding and tolerances
           "oracle_conservation_max": cons, "oracle_conservation_quantiles": quant}
    with open(os.path.join(HERE, "primordial_fail_cells.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote primordial_fail_cells.json:", len(bad), "cells")


if __name__ == "__main__":
    main()
