"""Warm restart through the C ABI (odevec_set_history): no reference
counterpart, so the bar is the property that defines it. Steps are never
clamped to t_stop, so a solve to t1 continued to t2 takes exactly the steps
of one solve to t2 -- provided the first step is the same, and the initial
step size depends on t_stop through W0 = max(|t0|, |t_stop|)
(src/odevec.F90:228-237). The networks are autonomous, so the tests run the
span [-S, 0] instead of [0, S]: W0 = S for every t_stop in it. Then states,
last step sizes and summed event counters of the continued solve equal the
single solve's bit for bit, per-cell and persistent, FP64 and mixed, and
splitting the continuation further changes nothing. A cell whose state,
parameters or start time changed between calls starts cold, and then
equals a fresh solve from that state."""
import numpy as np
import pytest

from odevec_amd import _lib
from odevec_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make(nid, n, precision, mode, hist=True):
    from odevec_amd.api import Odevec
    ode = Odevec(nid, n, precision=precision, mode=mode, grid=2 if mode == _lib.PERSISTENT else 0)
    if hist:
        ode.enable_history(n)
    return ode


def solve(ode, w, y, t0, ts, params=None):
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    st, dto, stats = ode.solve_cells(y, ts, t0=t0, dt0=w["dt0"], params=params, stats=True, raise_on_fail=False)
    return st, dto, stats


def split_times(w, frac=0.37):
    """[t0, t_stop] shifted to [-S, 0] (S = t_stop - t0), and a point inside."""
    ts = np.broadcast_to(w["t_stop"], (w["y"].shape[1],)).astype(np.float64)
    t0 = np.broadcast_to(w["t0"], ts.shape).astype(np.float64)
    span = ts - t0
    return -span, -span * (1.0 - frac), np.zeros_like(span)


CASES = {
    "PRIMORDIAL": lambda: W.primordial(700, seed=41),
    "ROBER": lambda: W.rober(300, seed=42),          # per-cell rate parameters
    "OREGO": lambda: W.orego(300, seed=43),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_continue_equals_one_solve(oracle, case, precision, mode):
    """FP64: the single solve, and so the continued one, is also the
    oracle's solve over the same span bit for bit."""
    w = CASES[case]()
    nid = w["network"] + "+REPRO"
    n = w["y"].shape[1]
    t0, t1, t2 = split_times(w)
    say("history", nid, "precision", precision, "mode", mode, "cells", n)
    one = make(nid, n, precision, mode, hist=False)
    y_one = np.array(w["y"], order="C", copy=True)
    st_one, dt_one, s_one = solve(one, w, y_one, t0, t2, w["params"])
    one.CloseOdevec()
    ode = make(nid, n, precision, mode)
    y = np.array(w["y"], order="C", copy=True)
    st_a, _, s_a = solve(ode, w, y, t0, t1, w["params"])
    st_b, dt_b, s_b = solve(ode, w, y, t1, t2, w["params"])
    ode.CloseOdevec()
    assert (st_one == 0).all() and (st_a == 0).all() and (st_b == 0).all()
    assert np.array_equal(y, y_one)
    assert np.array_equal(dt_b, dt_one)
    for k in ("attempts", "steps", "newton", "jac"):
        assert s_a[k] + s_b[k] == s_one[k], k
    if precision == _lib.FP64:
        yo, so, do, _ = oracle.Network.load(w["network"]).solve_cells(
            w["y"], t2, t0=t0, dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"], dt_min=w["dt_min"])
        assert np.array_equal(so, st_b) and np.array_equal(yo, y) and np.array_equal(do, dt_b)


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_split_continuation(mode):
    """From [0, S] (t_stop-dependent first step): continuing through one or
    through two intermediate end times gives the same result."""
    w = W.primordial(500, seed=46)
    nid = w["network"] + "+REPRO"
    n = w["y"].shape[1]
    ts = np.asarray(w["t_stop"], dtype=np.float64)
    t1, t15 = 0.3 * ts, 0.55 * ts
    runs = []
    for cuts in ([t1], [t1, t15]):
        ode = make(nid, n, _lib.MIXED, mode)
        y = np.array(w["y"], order="C", copy=True)
        t = np.zeros(n)
        for c in cuts + [ts]:
            st, dto, _ = solve(ode, w, y, t, c)
            assert (st == 0).all()
            t = c
        ode.CloseOdevec()
        runs.append((y, dto))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_changed_cells_start_cold():
    w = W.rober(256, seed=44)
    nid = w["network"] + "+REPRO"
    n = w["y"].shape[1]
    t0, t1, t2 = split_times(w, 0.2)
    t1 = np.asarray(t1)
    ode = make(nid, n, _lib.FP64, _lib.PERSISTENT)
    y = np.array(w["y"], order="C", copy=True)
    st, _, _ = solve(ode, w, y, t0, t1, w["params"])
    assert (st == 0).all()
    # cells 0::4: the host changes the state; 1::4: a parameter; 2::4: the
    # start time; 3::4 continue
    y[0, 0::4] *= 1.0 + 1e-12
    prm = np.array(w["params"], copy=True)
    prm[1, 1::4] *= 1.0 + 1e-12
    t1b = t1.copy()
    t1b[2::4] = t1[2::4] * (1.0 - 1e-9)
    y_mid = y.copy()
    st2, dt2, _ = solve(ode, w, y, t1b, t2, prm)
    ode.CloseOdevec()
    assert (st2 == 0).all()
    cold = make(nid, n, _lib.FP64, _lib.PERSISTENT, hist=False)
    y_cold = y_mid.copy()
    stc, dtc, _ = solve(cold, w, y_cold, t1b, t2, prm)
    cold.CloseOdevec()
    changed = np.zeros(n, dtype=bool)
    changed[0::4] = changed[1::4] = changed[2::4] = True
    assert np.array_equal(y[:, changed], y_cold[:, changed]) and np.array_equal(dt2[changed], dtc[changed])
    # the continued cells differ from a cold restart at t1 (else the test proves nothing)
    assert not np.array_equal(y[:, ~changed], y_cold[:, ~changed])


def test_history_rejects_other_sizes_and_modes():
    from odevec_amd.api import Odevec
    w = W.rober(64, seed=45)
    ode = make(w["network"], 64, _lib.FP64, _lib.PERCELL)
    y = np.array(w["y"], order="C", copy=True)
    with pytest.raises(_lib.OdevecError):
        ode.solve_cells(y[:, :32].copy(), np.asarray(w["t_stop"])[:32].copy(), params=w["params"][:, :32].copy(),
                        raise_on_fail=False)
    ode.CloseOdevec()
    lock = Odevec(w["network"], 64, mode=_lib.LOCKSTEP, group=32)
    with pytest.raises(_lib.OdevecError):
        lock.enable_history(64)
    lock.CloseOdevec()


def _dev_solve(ode, w, y_dev, t0, ts):
    """One device-resident solve on `ode` (torch tensors): status, last
    steps and the 8 per-cell counters."""
    import torch
    n = y_dev.shape[1]
    dev = y_dev.device
    stream = torch.cuda.Stream()
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    t0d = torch.tensor(np.ascontiguousarray(t0, dtype=np.float64), device=dev)
    tsd = torch.tensor(np.ascontiguousarray(ts, dtype=np.float64), device=dev)
    dt0 = torch.tensor(np.broadcast_to(w["dt0"], (n,)).astype(np.float64), device=dev)
    st = torch.zeros(n, dtype=torch.int32, device=dev)
    dto = torch.zeros(n, dtype=torch.float64, device=dev)
    cnt = torch.zeros(8 * n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ode.solve_cells_async(n, tsd.data_ptr(), y_dev.data_ptr(), t0_ptr=t0d.data_ptr(), dt0_ptr=dt0.data_ptr(),
                          status_ptr=st.data_ptr(), dt_out_ptr=dto.data_ptr(), counters_ptr=cnt.data_ptr(),
                          stream=stream.cuda_stream)
    stream.synchronize()
    return st.cpu().numpy(), dto.cpu().numpy(), cnt.cpu().numpy().reshape(8, n)


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_continue_with_fp64_fallback_equals_one_solve(mode):
    """Mixed form with the FP64 fallback (limit 2 convergence failures, so
    that many cells go on in FP64) and a history: a solve split at t1 equals
    the single solve with the fallback bit for bit (states, status, last
    steps; attempts, steps, Newton iterations and Jacobians add up), both
    for cells handed to FP64 lanes before t1 -- the second call hands them
    straight to FP64 lanes, which continue from the history -- and for cells
    handed over only after t1, whose convergence-failure count carried over."""
    import torch
    w = W.primordial(3000, seed=47)
    nid = w["network"] + "+REPRO"
    n = w["y"].shape[1]
    t0, t1, t2 = split_times(w, 0.45)
    limit = 2
    say("history + fallback", nid, "mode", mode, "cells", n)
    dev = torch.device("cuda:0")
    one = make(nid, n, _lib.MIXED, mode, hist=False)
    one.set_fp64_fallback(limit)
    y_one = torch.tensor(np.ascontiguousarray(w["y"]), device=dev)
    st1, dt1, c1 = _dev_solve(one, w, y_one, t0, t2)
    one.CloseOdevec()
    ode = make(nid, n, _lib.MIXED, mode)
    ode.set_fp64_fallback(limit)
    y = torch.tensor(np.ascontiguousarray(w["y"]), device=dev)
    sta, _, ca = _dev_solve(ode, w, y, t0, t1)
    ya = y.cpu().numpy()
    stb, dtb, cb = _dev_solve(ode, w, y, t1, t2)
    ode.CloseOdevec()
    say("status counts: single", np.bincount(st1, minlength=7).tolist(), "to t1",
        np.bincount(sta, minlength=7).tolist(), "t1 to t2", np.bincount(stb, minlength=7).tolist())
    # a cell the algorithm fails on before t1 starts cold in the second call
    ok = (sta == 0) & (st1 == 0)
    assert ok.mean() > 0.99
    assert np.array_equal(stb[sta == 0], st1[sta == 0])
    before = ca[5] >= limit                                      # convergence failures by t1
    handed = c1[5] >= limit                                      # the single solve handed them over
    say("handed over: before t1", int((before & handed & ok).sum()), "after t1", int((~before & handed & ok).sum()))
    assert (before & handed & ok).sum() >= 5 and (~before & handed & ok).sum() >= 5
    assert np.array_equal(y.cpu().numpy()[:, ok], y_one.cpu().numpy()[:, ok])
    assert np.array_equal(dtb[ok], dt1[ok])
    for k, name in enumerate(("attempts", "steps", "newton", "jac")):
        assert np.array_equal(ca[k][ok] + cb[k][ok], c1[k][ok]), name
    # the first call's states are those of a single solve to t1
    one = make(nid, n, _lib.MIXED, mode, hist=False)
    one.set_fp64_fallback(limit)
    y_t1 = torch.tensor(np.ascontiguousarray(w["y"]), device=dev)
    st_t1, _, _ = _dev_solve(one, w, y_t1, t0, t1)
    one.CloseOdevec()
    assert np.array_equal(st_t1, sta) and np.array_equal(ya[:, ok], y_t1.cpu().numpy()[:, ok])
