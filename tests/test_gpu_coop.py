"""Cooperative kernels on the GPU (ODEVEC_COOP: one cell per group of lanes,
one species per lane).

* the +REPRO build equals the oracle bit for bit (states, status, last step,
  counters) for the CSC networks: ROBER, OREGO (groups of 4 lanes),
  PRIMORDIAL CMK and FF (16), SYNTH64 (64 species, one wavefront per cell),
  SYNTH40 (40 species in a 64-lane group), KROME (16),
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


def coop_solve(nid, w, precision=_lib.FP64, grid=0, mixed_test=None):
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    say("GPU coop solve", nid, "cells", n, "precision", precision, "grid", grid, "mixed test", mixed_test)
    ode = Odevec(nid, max(n, 1), precision=precision, mode=_lib.COOP, grid=grid)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    if mixed_test:
        ode.set_mixed_convergence_test(mixed_test)
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
    "KROME": lambda: W.krome_ism(120, seed=38),     # 12 species, 21 rate parameters (groups of 16)
    # 40 species in a 64-lane group: 24 padding lanes beside rows held in LDS
    "SYNTH40": lambda: W.synth64(60, seed=39, network="SYNTH40/FF/CSC"),
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


def test_coop_mixed_rms_test(oracle):
    """SYNTH64's mixed form under the reference's RMS corrector test: every
    cell finishes within the accuracy bound of the max-norm form, with fewer
    corrector iterations; on an FP64 handle the setting changes nothing."""
    w = W.synth64(256, seed=36)
    yr, sr, dr, cr = oracle_cells(oracle, w["network"], w)
    res = {}
    for t in ("max", "rms"):
        y, st, dto, stats = coop_solve(w["network"], w, _lib.MIXED, mixed_test=t)
        ok = (sr == 0) & (st == 0)
        assert (st[sr == 0] == 0).all(), t
        e = (np.abs(y - yr)[:, ok] / (w["rtol"] * np.abs(yr[:, ok]) + w["atol"])).max(axis=0)
        say("mixed test %s: newton %d, median %.3g, within 10 tolerances %.4f" % (t, stats["newton"], np.median(e),
                                                                                 (e <= 10).mean()))
        assert np.median(e) < 1.0 and (e <= 10).mean() > 0.8, t
        res[t] = stats["newton"]
    assert res["rms"] < res["max"]
    w8 = W.subset(w, np.arange(32))
    a = coop_solve(w["network"] + "+REPRO", w8)
    b = coop_solve(w["network"] + "+REPRO", w8, mixed_test="rms")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


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


@pytest.mark.parametrize("net", ["SYNTH64/FF/CSC", "SYNTH40/FF/CSC"])
@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("grid", [0, 1])
def test_coop_continue_equals_one_solve(oracle, precision, grid, net):
    """Warm restart of the cooperative kernel (SYNTH64 and SYNTH40,
    networks no per-lane kernel runs; SYNTH40's groups carry padding lanes,
    whose history planes are kept too): over [-S, 0] (the first step then does not depend
    on t_stop), solve(-S -> t1) continued to 0 equals one solve bit for bit
    (states, status, last steps; attempts, steps, Newton iterations and
    Jacobians add up), and in FP64 equals the oracle's solve; a cell whose
    state changed between the calls starts cold and equals a fresh solve."""
    from odevec_amd.api import Odevec
    w = W.synth64(48, seed=36, network=net)
    nid = w["network"] + "+REPRO"
    n = w["y"].shape[1]
    span = np.broadcast_to(np.asarray(w["t_stop"], dtype=np.float64), (n,)).copy()
    t0, t1, t2 = -span, -span * 0.63, np.zeros(n)
    say("GPU coop history, precision", precision, "grid", grid)

    def run(hist, pieces, y, prm):
        ode = Odevec(nid, n, precision=precision, mode=_lib.COOP, grid=grid)
        ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
        if hist:
            ode.enable_history(n)
        out = []
        for a, b, yy in pieces:
            st, dto, stats = ode.solve_cells(yy, b, t0=a, dt0=w["dt0"], params=prm, stats=True, raise_on_fail=False)
            out.append((st, dto, stats))
        ode.CloseOdevec()
        return out

    y_one = np.array(w["y"], order="C", copy=True)
    (st1, dt1, s1), = run(False, [(t0, t2, y_one)], None, w["params"])
    y = np.array(w["y"], order="C", copy=True)
    (sta, _, sa), (stb, dtb, sb) = run(True, [(t0, t1, y), (t1, t2, y)], None, w["params"])
    assert (st1 == 0).all() and (sta == 0).all() and (stb == 0).all()
    assert np.array_equal(y, y_one) and np.array_equal(dtb, dt1)
    for k in ("attempts", "steps", "newton", "jac"):
        assert sa[k] + sb[k] == s1[k], k
    if precision == _lib.FP64:
        yo, so, do, _ = oracle.Network.load(w["network"]).solve_cells(
            w["y"], t2, t0=t0, dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"], dt_min=w["dt_min"])
        assert np.array_equal(so, stb) and np.array_equal(yo, y) and np.array_equal(do, dtb)
    # changed cells (every third) start cold at t1
    ode = Odevec(nid, n, precision=precision, mode=_lib.COOP, grid=grid)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    ode.enable_history(n)
    y = np.array(w["y"], order="C", copy=True)
    ode.solve_cells(y, t1, t0=t0, dt0=w["dt0"], params=w["params"], raise_on_fail=False)
    y[5, 0::3] *= 1.0 + 1e-12
    y_mid = y.copy()
    st2, dt2, _ = ode.solve_cells(y, t2, t0=t1, dt0=w["dt0"], params=w["params"], raise_on_fail=False)
    ode.CloseOdevec()
    y_cold = y_mid.copy()
    (stc, dtc, _), = run(False, [(t1, t2, y_cold)], None, w["params"])
    ch = np.zeros(n, dtype=bool)
    ch[0::3] = True
    assert np.array_equal(y[:, ch], y_cold[:, ch]) and np.array_equal(dt2[ch], dtc[ch])
    assert np.array_equal(y[:, ~ch], y_one[:, ~ch])


def test_coop_history_only_for_networks_without_per_lane_kernels():
    """A network the per-lane kernels run keeps their history layout: the
    cooperative mode refuses a history for it."""
    from odevec_amd.api import Odevec
    ode = Odevec("PRIMORDIAL/CMK/CSC", 64, mode=_lib.COOP)
    with pytest.raises(_lib.OdevecError):
        ode.enable_history(64)
    ode.CloseOdevec()
