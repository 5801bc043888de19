"""Accuracy of the benchmarked mixed form against the reference's algorithm
and against a tight-tolerance solution, with no magnitude filter: every
integrated species of every cell, errors in units of the reference's
error-weight scale rtol |y_ref| + atol (src/odevec.F90:706-716).

Why the gate is distributional. The reference's algorithm at rtol 1e-6 is
itself far from the rtol-1e-12 solution on a quarter of these cells (global
error accumulates over up to 1e14 s of decaying trace species), and two FP64
builds that only round differently (with and without FMA contraction)
disagree beyond 10 tolerances on about 5% of the cells: the trajectory of a
BDF solve on the hardest cells depends on the last bit of the step-size
decisions. So the mixed form is held to what those two facts allow:

* against the reference (the +REPRO FP64 build, bit-identical to the oracle):
  nearly as many cells within 10 tolerances as the FP64 default build;
* against the tight solution: no larger errors than the reference's own
  (99th percentile and maximum), and nearly as many cells within 10;
* the H, He and D nuclei conserved as well as the oracle conserves them on
  the full 1M-cell workload (its recorded maxima and 99.9th percentiles,
  tests/golden/primordial_fail_cells.json), within a factor 10, or within
  10 rtol for the largest drift (the corrector's invariant test admits one
  tolerance per step; the oracle's exact corrections keep the abundant H
  nuclei to 1e-7, the mixed form to a few 1e-6).
"""
import json
import os
import sys

import numpy as np
import pytest

from odevec_amd import workloads as W
from odevec_amd import _lib

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
NUCLEI = {"H": [0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0],
          "He": [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0],
          "D": [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0]}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solve(w, nid, precision, mode, fallback=0, schedule=0, rtol=None, atol=None, max_attempts=0):
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    say("GPU solve", nid, "cells", n, "precision", precision, "mode", mode, "fallback", fallback, "rtol", rtol)
    ode = Odevec(nid, n, precision=precision, mode=mode, schedule=schedule, max_attempts=max_attempts)
    ode.RelativeTolerance = w["rtol"] if rtol is None else rtol
    ode.AbsoluteTolerance = w["atol"] if atol is None else atol
    ode.MinimalTimestep = w["dt_min"]
    if fallback:
        ode.set_fp64_fallback(fallback)
    y = np.array(w["y"], order="C", copy=True)
    st, dto, _ = ode.solve_cells(y, w["t_stop"], t0=w["t0"], dt0=w["dt0"], raise_on_fail=False)
    ode.CloseOdevec()
    return y, st


def cell_err(y, yref, rtol=1e-6, atol=1e-20):
    return (np.abs(y[:12] - yref[:12]) / (rtol * np.abs(yref[:12]) + atol)).max(axis=0)


@pytest.fixture(scope="module")
def runs():
    full = W.primordial(1048576, seed=3)
    idx = np.sort(np.random.default_rng(21).choice(1048576, 65536, replace=False))
    w = W.subset(full, idx)
    net = w["network"]
    out = {"w": w}
    # the bench configuration: persistent, cost-ordered queue, FP64 fallback
    # at 128 convergence failures, the default invariant test
    out["mixed"] = solve(w, net, _lib.MIXED, _lib.PERSISTENT, fallback=128, schedule=1)
    out["ref"] = solve(w, net + "+REPRO", _lib.FP64, _lib.PERSISTENT)
    out["fp64"] = solve(w, net, _lib.FP64, _lib.PERCELL)
    tidx = np.arange(0, 65536, 32)
    out["tidx"] = tidx
    out["tight"] = solve(W.subset(w, tidx), net + "+REPRO", _lib.FP64, _lib.PERSISTENT, rtol=1e-12, atol=1e-26,
                         max_attempts=200000)
    return out


def test_mixed_cells_finish(runs):
    ym, sm = runs["mixed"]
    yr, sr = runs["ref"]
    say("status mixed", np.bincount(sm), "reference", np.bincount(sr))
    assert (sm != 0).sum() <= (sr != 0).sum()


def test_mixed_against_the_reference(runs):
    ym, sm = runs["mixed"]
    yr, sr = runs["ref"]
    yd, sd = runs["fp64"]
    ok = (sm == 0) & (sr == 0) & (sd == 0)
    em, ed = cell_err(ym[:, ok], yr[:, ok]), cell_err(yd[:, ok], yr[:, ok])
    fm, fd = (em <= 10).mean(), (ed <= 10).mean()
    say("cells within 10 tolerances of the reference: mixed %.4f, FP64 default build %.4f; median %.3g / %.3g"
        % (fm, fd, np.median(em), np.median(ed)))
    assert fm >= 0.88 and fm >= fd - 0.07
    assert np.median(em) <= 1.0


def test_mixed_against_tight_solution(runs):
    tidx = runs["tidx"]
    yt, st = runs["tight"]
    ym, sm = runs["mixed"]
    yr, sr = runs["ref"]
    ok = (st == 0) & (sm[tidx] == 0) & (sr[tidx] == 0)
    em, er = cell_err(ym[:, tidx][:, ok], yt[:, ok]), cell_err(yr[:, tidx][:, ok], yt[:, ok])
    say("vs rtol 1e-12 (%d cells): mixed q99 %.3g max %.3g within10 %.4f | reference q99 %.3g max %.3g within10 %.4f"
        % (ok.sum(), np.quantile(em, .99), em.max(), (em <= 10).mean(), np.quantile(er, .99), er.max(),
           (er <= 10).mean()))
    assert ok.sum() >= 0.9 * len(tidx)
    assert np.quantile(em, 0.99) <= 1.5 * np.quantile(er, 0.99)
    assert em.max() <= 1.5 * er.max()
    assert (em <= 10).mean() >= (er <= 10).mean() - 0.05


def test_mixed_conserves_nuclei_like_the_oracle(runs):
    fail = json.load(open(os.path.join(GOLD, "primordial_fail_cells.json")))
    w = runs["w"]
    ym, sm = runs["mixed"]
    ok = sm == 0
    for name, v in NUCLEI.items():
        v = np.array(v, dtype=float)
        r = np.abs(v @ w["y"][:, ok] - v @ ym[:, ok]) / (v @ w["y"][:, ok])
        omax = fail["oracle_conservation_max"][name]
        oq = fail["oracle_conservation_quantiles"][name]["q999"]
        say("%s drift max %.3g q99.9 %.3g (oracle, 1M cells: max %.3g q99.9 %.3g)"
            % (name, r.max(), np.quantile(r, 0.999), omax, oq))
        # the corrector's invariant test admits a residual of one tolerance
        # per step: the abundant H nuclei may drift by a few rtol
        assert r.max() <= max(10 * omax, 10 * w["rtol"]) + 1e-14
        assert np.quantile(r, 0.999) <= 10 * oq + 1e-14


def test_tight_run_is_the_oracle(oracle):
    """The tight run above is the +REPRO build at rtol 1e-12; it equals the
    oracle at the same tolerances bit for bit (so the tight errors are the
    oracle's)."""
    w = W.subset(W.primordial(1048576, seed=3), np.arange(0, 1048576, 131072))
    y, st = solve(w, w["network"] + "+REPRO", _lib.FP64, _lib.PERSISTENT, rtol=1e-12, atol=1e-26)
    yo, so, _, _ = oracle.Network.load(w["network"]).solve_cells(
        w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], rtol=1e-12, atol=1e-26, dt_min=w["dt_min"])
    assert np.array_equal(st, so) and np.array_equal(y, yo)
