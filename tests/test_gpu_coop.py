"""Cooperative kernels on the GPU (ODEVEC_COOP: one cell per group of lanes,
one species per lane).

* the +REPRO build equals the oracle bit for bit (states, status, last step,
  counters) for the CSC networks: ROBER, OREGO (groups of 4 lanes),
  PRIMORDIAL CMK and FF (16), SYNTH64 (64 species, one wavefront per cell),
  with a group per cell and with one workgroup working through the queue;
* the default build and the mixed form: every cell finishes and the states
  agree with the oracle's FP64 solution to a bound set by the tolerance.
"""
import numpy as np
import pytest

from odevec_amd import _lib
from odevec_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def coop_solve(nid, w, precision=_lib.FP64, grid=0):
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    say("GPU coop solve", nid, "cells", n, "precision", precision, "grid", grid)
    ode = Odevec(nid, max(n, 1), precision=precision, mode=_lib.COOP, grid=grid)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    y = np.array(w["y"], order="C", copy=True)
    st, dto, stats = ode.solve_cells(y, w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], stats=True,
                                     raise_on_fail=False)
    ode.CloseOdevec()
    return y, st, dto, stats


def oracle_cells(oracle, nid, w, counters=False):
    return oracle.Network.load(nid.replace("+REPRO", "")).solve_cells(
        w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"],
        dt_min=w["dt_min"])


def rober_csc(n):
    w = W.rober(n, seed=31)
    w["network"] = "ROBER/NONE/CSC"
    return w


def orego_csc(n):
    w = W.orego(n, seed=32)
    w["network"] = "OREGO/NONE/CSC"
    return w


CASES = {
    "ROBER": lambda: rober_csc(100),
    "OREGO": lambda: orego_csc(70),
    "PRIMORDIAL": lambda: W.primordial(300, seed=33),
    "PRIMORDIAL_FF": lambda: W.primordial(130, seed=34, network="PRIMORDIAL/FF/CSC"),
    "SYNTH64": lambda: W.synth64(40, seed=35),
}


@pytest.mark.parametrize("grid", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_coop_repro_bitwise(oracle, case, grid):
    """grid 0: a group per cell where the device holds them all; grid 1: one
    workgroup whose groups take every cell from the queue."""
    w = CASES[case]()
    w["t_stop"] = np.array(np.broadcast_to(w["t_stop"], (w["y"].shape[1],)))
    w["t_stop"][::9] = 0.0                      # skipped cells
    y, st, dto, stats = coop_solve(w["network"] + "+REPRO", w, grid=grid)
    yr, sr, dr, cr = oracle_cells(oracle, w["network"], w)
    assert np.array_equal(st, sr), case
    assert np.array_equal(y, yr), case
    assert np.array_equal(dto, dr), case
    assert stats["attempts"] == int(cr[0].sum()) and stats["jac"] == int(cr[3].sum())
    assert (st[::9] == 5).all()


@pytest.mark.parametrize("case", ["PRIMORDIAL", "SYNTH64"])
@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
def test_coop_default_build_accuracy(oracle, case, precision):
    w = CASES[case]()
    y, st, dto, stats = coop_solve(w["network"], w, precision)
    yr, sr, dr, cr = oracle_cells(oracle, w["network"], w)
    ok = (sr == 0)
    assert (st[ok] == 0).mean() > 0.99
    ok &= st == 0
    # every species, in units of the error weights' scale rtol |y| + atol
    e = (np.abs(y - yr)[:, ok] / (w["rtol"] * np.abs(yr[:, ok]) + w["atol"])).max(axis=0)
    print("coop %s precision %d: median %.3g, within 10 tolerances %.4f" % (case, precision, np.median(e),
                                                                         (e <= 10).mean()), flush=True)
    assert np.median(e) < 1.0 and (e <= 10).mean() > 0.8


@pytest.mark.parametrize("case", ["ROBER", "SYNTH64"])
def test_repro_coop_nan_state_bitwise(oracle, case):
    """A NaN or an infinity in a cell's state: the group's NaN vote fails
    the corrector exactly as the oracle's per-cell solve does, and the
    other cells are unaffected."""
    w = W.subset(CASES[case](), np.arange(8))
    y = w["y"].copy()
    y[0, 0] = np.nan
    y[y.shape[0] - 1, 1] = np.nan
    y[1, 3] = np.inf
    w = dict(w, y=y)
    nid = w["network"] + "+REPRO"
    yg, st, dto, stats = coop_solve(nid, w)
    yr, sr, dr, cr = oracle_cells(oracle, nid, w)
    assert np.array_equal(st, sr) and (st[[0, 1, 3]] != 0).all()
    assert np.array_equal(yg, yr, equal_nan=True)
    assert np.array_equal(dto, dr, equal_nan=True)
    assert stats["attempts"] == int(cr[0].sum())
