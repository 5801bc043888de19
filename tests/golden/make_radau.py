"""Write tests/golden/radau_known_answers.json: tight-tolerance solutions of
seeded cells by SciPy's Radau IIA integrator (an implicit Runge-Kutta method,
independent of the BDF code under test), rtol 1e-12.

The right-hand sides and analytic Jacobians are evaluated from the network
term tables (pinned to the reference's network definitions by
reference_rhs.json); PRIMORDIAL's rate coefficients come from the oracle
(pinned to the reference's rate text by reference_rates.json). SYNTH64 has
no reference definition: its right-hand side and Jacobian are the mass-action
law of the generator's reaction list (A + B -> C + D, k = A exp(-Ea/T)),
evaluated here directly from that list, without the preprocessor's term
table or the oracle. KROME_ISM's come from the KROME setup file's own SymPy
expressions (lambdified), again without the term table or the oracle.
"""
import json
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from odevec_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def make_fun(ir, k):
    n = ir["neq"]
    terms = ir["terms"]

    def f(t, y):
        out = np.zeros(n)
        for row, coef, rate, fac in terms:
            v = coef * (k[rate] if rate >= 0 else 1.0)
            for s in fac:
                v *= y[s]
            out[row] += v
        return out

    def jac(t, y):
        J = np.zeros((n, n))
        for row, coef, rate, fac in terms:
            base = coef * (k[rate] if rate >= 0 else 1.0)
            for j in set(fac):
                red = list(fac)
                red.remove(j)
                v = base * fac.count(j)
                for s in red:
                    v *= y[s]
                J[row, j] += v
        return J

    return f, jac


def solve(net_id, w, idx, t_stop=None, rtol=1e-12, atol=1e-30):
    net = O.Network.load(net_id)
    out = []
    for i in idx:
        y0 = w["y"][:, i].copy()
        prm = None if w["params"] is None else w["params"][:, i]
        k = net.rates(y0, prm)
        f, jac = make_fun(net.ir, k)
        ts = float(w["t_stop"][i] if t_stop is None else t_stop)
        r = solve_ivp(f, (0.0, ts), y0, method="Radau", rtol=rtol, atol=atol, jac=jac)
        assert r.success, r.message
        out.append({"cell": int(i), "y0": y0.tolist(), "t_stop": ts,
                    "params": None if prm is None else prm.tolist(), "y": r.y[:, -1].tolist(),
                    "nfev": int(r.nfev)})
    return out


def synth_fun(reactions, A, Ea, T):
    """Mass-action f and J of two-body reactions (a, b) -> (c, d) with
    k_r = A_r exp(-Ea_r / T), straight from the reaction list."""
    ra = np.array([r[0][0] for r in reactions])
    rb = np.array([r[0][1] for r in reactions])
    pc = np.array([r[1][0] for r in reactions])
    pd = np.array([r[1][1] for r in reactions])
    k = np.asarray(A) * np.exp(-np.asarray(Ea) / T)

    def f(t, y):
        v = k * y[ra] * y[rb]
        out = np.zeros(y.size)
        for idx, sign in ((ra, -1.0), (rb, -1.0), (pc, 1.0), (pd, 1.0)):
            np.add.at(out, idx, sign * v)
        return out

    def jac(t, y):
        J = np.zeros((y.size, y.size))
        da, db = k * y[rb], k * y[ra]     # d(rate)/dy_a, d(rate)/dy_b
        for idx, sign in ((ra, -1.0), (rb, -1.0), (pc, 1.0), (pd, 1.0)):
            np.add.at(J, (idx, ra), sign * da)
            np.add.at(J, (idx, rb), sign * db)
        return J

    return f, jac


def solve_synth(w, idx, rtol=1e-12, atol=1e-30):
    from odevec_amd.pre_odevec.networks import synth_reactions
    reactions, A, Ea = synth_reactions(64, 300)
    out = []
    for i in idx:
        y0 = w["y"][:, i].copy()
        T = float(w["params"][0, i])
        f, jac = synth_fun(reactions, A, Ea, T)
        ts = float(w["t_stop"][i])
        r = solve_ivp(f, (0.0, ts), y0, method="Radau", rtol=rtol, atol=atol, jac=jac)
        assert r.success, r.message
        out.append({"cell": int(i), "y0": y0.tolist(), "t_stop": ts, "params": [T], "y": r.y[:, -1].tolist(),
                    "nfev": int(r.nfev)})
    return out


def krome_fun(path):
    """f and J of a KROME-style setup file straight from its SymPy ``rhs``
    (executed as pre_odevec.py:97-105 executes it), without the
    preprocessor's term table or the oracle; k(:,i) are the file's ``k``."""
    import sympy as sp
    scope = {}
    exec("from sympy import *", scope)
    with open(path) as fh:
        exec(fh.read(), scope)
    y, k, rhs = list(scope["y"]), list(scope["k"]), sp.Matrix(scope["rhs"])
    fl = sp.lambdify([y, k], list(rhs), "numpy")
    jl = sp.lambdify([y, k], rhs.jacobian(y).tolist(), "numpy")
    return (lambda kv: (lambda t, yv: np.array(fl(yv, kv), dtype=float)),
            lambda kv: (lambda t, yv: np.array(jl(yv, kv), dtype=float)))


def solve_krome(w, idx, rtol=1e-12, atol=1e-30):
    from odevec_amd.build import KROME_SETUPFILE
    mk_f, mk_j = krome_fun(KROME_SETUPFILE)
    out = []
    for i in idx:
        y0 = w["y"][:, i].copy()
        kv = w["params"][:, i].copy()
        ts = float(w["t_stop"][i])
        r = solve_ivp(mk_f(kv), (0.0, ts), y0, method="Radau", rtol=rtol, atol=atol, jac=mk_j(kv))
        assert r.success, r.message
        out.append({"cell": int(i), "y0": y0.tolist(), "t_stop": ts, "params": kv.tolist(),
                    "y": r.y[:, -1].tolist(), "nfev": int(r.nfev)})
    return out


def main():
    res = {"method": "scipy.integrate.solve_ivp Radau", "rtol": 1e-12, "cases": {}}
    res["cases"]["ROBER"] = {"network": "ROBER/NONE/DENSE", "atol": 1e-30,
                             "cells": solve("ROBER/NONE/DENSE", W.rober(32), range(4))}
    res["cases"]["OREGO"] = {"network": "OREGO/NONE/DENSE", "atol": 1e-30,
                             "cells": solve("OREGO/NONE/DENSE", W.orego(32), range(4), t_stop=30.0)}
    res["cases"]["PRIMORDIAL"] = {"network": "PRIMORDIAL/CMK/CSC", "atol": 1e-40,
                                  "cells": solve("PRIMORDIAL/CMK/CSC", W.primordial(64), range(6), atol=1e-40)}
    res["cases"]["SYNTH64"] = {"network": "SYNTH64/FF/CSC", "atol": 1e-30,
                               "cells": solve_synth(W.synth64(64, seed=7), range(4))}
    res["cases"]["KROME_ISM"] = {"network": "KROME/FF/CSC", "atol": 1e-30,
                                 "cells": solve_krome(W.krome_ism(64, seed=5), range(6))}
    with open(os.path.join(HERE, "radau_known_answers.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print("wrote radau_known_answers.json")


if __name__ == "__main__":
    main()
