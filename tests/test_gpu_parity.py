"""GPU parity through the C ABI.

* The +REPRO build (no FMA contraction) must equal the oracle bit for bit:
  states, last step sizes and event counters, in every mode.
* The default build differs only by contraction; there the bar is accuracy:
  10*rtol*|y| + atol against the oracle where the problem is well conditioned
  (ROBER), and against tight Radau solutions otherwise.
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_solve(nid, w, n=None, precision=0, mode=0, group=0, fallback=0, **kw):
    from odevec_amd.api import Odevec
    n = w["y"].shape[1] if n is None else n
    say("GPU solve", nid, "cells", n, "precision", precision, "mode", mode, "group", group, "fallback", fallback)
    ode = Odevec(nid, max(n, 1), precision=precision, mode=mode, group=group, **kw)
    if fallback:
        ode.set_fp64_fallback(fallback)
    ode.RelativeTolerance = w["rtol"]
    ode.AbsoluteTolerance = w["atol"]
    ode.MinimalTimestep = w["dt_min"]
    y = np.array(w["y"], order="C", copy=True)
    st, dto, stats = ode.solve_cells(y, w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], stats=True,
                                     raise_on_fail=False)
    ode.CloseOdevec()
    return y, st, dto, stats


def oracle_cells(oracle, nid, w):
    return oracle.Network.load(nid.replace("+REPRO", "")).solve_cells(
        w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"],
        dt_min=w["dt_min"])


CASES = {
    "ROBER": lambda: W.rober(32, seed=1),          # the CPU configuration: 32 cells
    "OREGO": lambda: W.orego(512, seed=2),
    "PRIMORDIAL": lambda: W.primordial(512, seed=3),
    "PRIMORDIAL_FF": lambda: W.primordial(256, seed=3, network="PRIMORDIAL/FF/CSC"),
    "PRIMORDIAL_DENSE": lambda: W.primordial(256, seed=3, network="PRIMORDIAL/NONE/DENSE"),
    # LUmethod 3: the Jacobian by forward differences of the RHS
    "ROBER_LU3": lambda: dict(W.rober(32, seed=1), network="ROBER/NONE/DENSE/LU3"),
    "PRIMORDIAL_LU3": lambda: W.primordial(256, seed=3, network="PRIMORDIAL/NONE/DENSE/LU3"),
    # LUmethod 2: the L and U factors in closed form
    "ROBER_LU2": lambda: dict(W.rober(32, seed=1), network="ROBER/NONE/DENSE/LU2"),
    "OREGO_LU2": lambda: dict(W.orego(512, seed=2), network="OREGO/NONE/DENSE/LU2"),
    # a network built from a KROME setup file; its 21 rate coefficients per cell
    "KROME": lambda: W.krome_ism(512, seed=5),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_repro_bitwise_percell(oracle, case, mode):
    w = CASES[case]()
    nid = w["network"] + "+REPRO"
    y, st, dto, stats = gpu_solve(nid, w, mode=mode)
    yr, sr, dr, cr = oracle_cells(oracle, nid, w)
    assert np.array_equal(st, sr) and (st == 0).all()
    assert np.array_equal(y, yr)
    assert np.array_equal(dto, dr)
    assert stats["attempts"] == int(cr[0].sum()) and stats["newton"] == int(cr[2].sum())


@pytest.mark.parametrize("case", ["ROBER", "PRIMORDIAL"])
@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_repro_nan_state_bitwise(oracle, case, mode):
    """Cells whose state holds a NaN or an infinity fail through the
    corrector's NaN check exactly as in the oracle (status, returned state,
    last step, counters); the other cells are unaffected. PRIMORDIAL's FP64
    kernels keep the error weights in LDS, ROBER's in registers."""
    w = CASES[case]()
    w = W.subset(w, np.arange(16))
    y = w["y"].copy()
    y[0, 0] = np.nan                 # an integrated species
    y[y.shape[0] - 1, 1] = np.nan    # PRIMORDIAL: a frozen species
    y[1, 3] = np.inf
    w = dict(w, y=y)
    nid = w["network"] + "+REPRO"
    yg, st, dto, stats = gpu_solve(nid, w, mode=mode)
    yr, sr, dr, cr = oracle_cells(oracle, nid, w)
    assert np.array_equal(st, sr) and (st[[0, 1, 3]] != 0).all() and (np.delete(st, [0, 1, 3]) == 0).all()
    assert np.array_equal(yg, yr, equal_nan=True)
    assert np.array_equal(dto, dr, equal_nan=True)
    assert stats["attempts"] == int(cr[0].sum())


@pytest.mark.parametrize("case,G", [("ROBER", 32), ("OREGO", 64), ("PRIMORDIAL", 64), ("ROBER_LU3", 32),
                                    ("PRIMORDIAL_LU3", 64), ("ROBER_LU2", 32), ("OREGO_LU2", 64), ("KROME", 64)])
def test_repro_bitwise_lockstep(oracle, case, G):
    """Lockstep groups reproduce the reference at nvector = G (shared t_stop)."""
    w = CASES[case]()
    n = min(w["y"].shape[1], 4 * G) - 7      # a ragged last group
    w = W.subset(w, np.arange(n))
    w["t_stop"] = np.full(n, w["t_stop"][0])
    nid = w["network"] + "+REPRO"
    y, st, dto, stats = gpu_solve(nid, w, mode=_lib.LOCKSTEP, group=G)
    net = oracle.Network.load(w["network"])
    yr = np.zeros_like(y)
    for g0 in range(0, n, G):
        sl = slice(g0, min(g0 + G, n))
        prm = None if w["params"] is None else w["params"][:, sl]
        yv, s, t, d, cnt = net.solve_vector(w["y"][:, sl], 0.0, w["t_stop"][0], w["dt0"][0], w["rtol"], w["atol"],
                                            w["dt_min"], params=prm)
        assert s == 0
        yr[:, sl] = yv
    assert (st == 0).all() and np.array_equal(y, yr)


def test_repro_lockstep_nan_group(oracle):
    """Lockstep (nvector = 32): a NaN in one cell fails its whole group
    through the vector's NaN check, as in the oracle; the next group is
    unaffected."""
    G, n = 32, 64
    w = W.rober(n, seed=1)
    w["t_stop"] = np.full(n, w["t_stop"][0])
    w["y"] = w["y"].copy()
    w["y"][1, 5] = np.nan
    y, st, dto, stats = gpu_solve("ROBER/NONE/DENSE+REPRO", w, mode=_lib.LOCKSTEP, group=G)
    net = oracle.Network.load("ROBER/NONE/DENSE")
    for g0 in (0, G):
        sl = slice(g0, g0 + G)
        prm = None if w["params"] is None else w["params"][:, sl]
        yv, s, t, d, cnt = net.solve_vector(w["y"][:, sl], 0.0, w["t_stop"][0], w["dt0"][0], w["rtol"], w["atol"],
                                            w["dt_min"], params=prm)
        assert (st[sl] == s).all() and np.array_equal(y[:, sl], yv, equal_nan=True)
    assert (st[:G] != 0).all() and (st[G:] == 0).all()


def test_repro_lockstep_mask(oracle):
    w = W.rober(32)
    mask = np.zeros(32, np.int8)
    mask[::3] = 1
    from odevec_amd.api import Odevec
    ode = Odevec("ROBER/NONE/DENSE+REPRO", 32, mode=_lib.LOCKSTEP, group=32)
    ode.y[:] = w["y"].T
    say("GPU SolveODE masked lockstep")
    t, dt = ode.SolveODE(0.0, 1e-10, 1e3, Mask=mask)
    yv, s, tt, dd, cnt = oracle.Network.load("ROBER/NONE/DENSE").solve_vector(
        w["y"], 0.0, 1e3, 1e-10, 1e-6, 1e-10, params=np.array([0.04, 3e7, 1e4])[:, None] * np.ones((1, 32)),
        mask=mask)
    assert t == 1e3 and np.array_equal(ode.y.T, yv) and dt == dd
    assert np.array_equal(ode.y[mask == 0], w["y"].T[mask == 0])
    ode.CloseOdevec()


def test_reference_rober_driver_pattern(oracle):
    """tests/rober.f90:39-62: one handle, nvector 32, repeated SolveODE calls
    over 21 output decades with dt = 1e-10 before every call."""
    from odevec_amd.api import Odevec
    ode = Odevec("ROBER/NONE/DENSE+REPRO", 32, mode=_lib.LOCKSTEP, group=32)
    rtol = 10.0 ** (-2.0 - 24 * 0.25)
    ode.RelativeTolerance = rtol
    ode.AbsoluteTolerance = 1e-6 * rtol
    ode.MinimalTimestep = 1e-20
    ode.y[:] = np.array([1.0, 0.0, 0.0])
    net = oracle.Network.load("ROBER/NONE/DENSE")
    yref = np.zeros((3, 32))
    yref[0] = 1.0
    prm = np.array([0.04, 3e7, 1e4])[:, None] * np.ones((1, 32))
    t_start, t_stop = 0.0, 1e-10
    say("GPU reference driver pattern")
    while t_stop < 1e11:
        t_prev = t_start
        t_start, dt = ode.SolveODE(t_start, 1e-10, t_stop)
        yref, s, tr, dr, cnt = net.solve_vector(yref, t_prev, t_stop, 1e-10, rtol, 1e-6 * rtol, 1e-20, params=prm)
        assert t_start == t_stop == tr
        assert s == 0 and np.array_equal(ode.y.T, yref), t_stop
        t_stop = t_stop * 10.0
    assert abs(ode.y[0].sum() - 1.0) < 1e-6     # mass conservation of ROBER
    ode.CloseOdevec()


def test_fortran_module_driver_equals_oracle(oracle):
    """The Fortran module (odevec_amd/fortran) driving ROBER in the reference
    driver's pattern (tests/rober.f90:39-62; nvector 32, lockstep over the
    vector like the reference, reproducible build): every printed state
    equals the oracle's at nvector = 32 bit for bit."""
    import subprocess
    from odevec_amd import build as B
    if not os.path.exists(B.FORTRAN_DRIVER):
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    say("GPU Fortran driver", B.FORTRAN_DRIVER)
    r = subprocess.run([B.FORTRAN_DRIVER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = [line for line in r.stdout.splitlines() if line.strip()]
    rows = np.array([[float(v) for v in line.split()] for line in lines if not line.startswith(("RHS", "JAC"))])
    rhs = np.array([[float(v) for v in line[3:].split()] for line in lines if line.startswith("RHS")])
    jac = np.array([[float(v) for v in line[3:].split()] for line in lines if line.startswith("JAC")])
    net = oracle.Network.load("ROBER/NONE/DENSE")
    yref = np.zeros((3, 32))
    yref[0] = 1.0
    prm = np.array([0.04, 3e7, 1e4])[:, None] * np.ones((1, 32))
    t_start, t_stop, k = 0.0, 1e-10, 0
    while t_stop < 1e11:
        yref, s, tr, dr, cnt = net.solve_vector(yref, t_start, t_stop, 1e-10, 1e-6, 1e-10, 1e-20, params=prm)
        t_start = tr
        blk = rows[k * 32:(k + 1) * 32]
        assert np.array_equal(blk[:, 0], np.full(32, t_stop))
        assert np.array_equal(blk[:, 1], np.arange(1, 33))
        assert np.array_equal(blk[:, 2:5], yref.T), t_stop
        assert (blk[:, 5] == 0).all()
        t_stop, k = t_stop * 10.0, k + 1
    assert k * 32 == len(rows)
    # the plugin's GetRHS / GetJac at the final state (default rate constants)
    kd = np.array([0.04, 3e7, 1e4])
    assert np.array_equal(rhs[:, 0], np.arange(1, 33)) and np.array_equal(jac[:, 0], np.arange(1, 33))
    for i in range(32):
        assert np.array_equal(rhs[i, 1:], net.rhs(yref[:, i], kd))
        assert np.array_equal(jac[i, 1:].reshape(3, 3).T, net.p_dense(yref[:, i], kd, 0.5 * 1e-3))


@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_workgroup_sizes_bitwise(oracle, block, precision, mode):
    """Each wavefront owns an LDS slice and a workspace slot: any workgroup
    size gives the same cells' results (FP64: history, weights and previous
    correction in LDS, rates in the workspace; mixed: rates and history in
    LDS, previous correction in the workspace)."""
    w = W.primordial(1000, seed=11)
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    y, st, dto, stats = gpu_solve(nid, w, precision=precision, mode=mode, block=block, grid=3 if mode else 0)
    y0, st0, dto0, _ = gpu_solve(nid, w, precision=precision, mode=_lib.PERCELL, block=64)
    assert np.array_equal(st, st0) and np.array_equal(y, y0) and np.array_equal(dto, dto0)
    if precision == 0:
        yr, sr, dr, cr = oracle_cells(oracle, nid, w)
        assert np.array_equal(st, sr) and np.array_equal(y, yr) and np.array_equal(dto, dr)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("fallback", [0, 2])
def test_cost_ordered_schedule_changes_nothing(oracle, precision, fallback):
    """The persistent queue in cost order (costliest cells first, spread one
    per wavefront; the solve runs on copies of the inputs in queue order and
    scatters the results back) and in index order give the same cells'
    results and counters, bit for bit, with and without the FP64 fallback
    (limit 2: many hand-backs); few lanes (grid 2) so the queue does the
    work."""
    w = W.primordial(3000, seed=12)
    w["t_stop"][::7] = 0.0                        # skipped cells rank last
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    a = gpu_solve(nid, w, precision=precision, mode=_lib.PERSISTENT, grid=2, schedule=1, fallback=fallback)
    b = gpu_solve(nid, w, precision=precision, mode=_lib.PERSISTENT, grid=2, schedule=0, fallback=fallback)
    assert all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3]))
    assert a[3]["attempts"] == b[3]["attempts"] and a[3]["newton"] == b[3]["newton"]
    assert (a[1][::7] == 5).all()
    if precision == 0:
        yr, sr, dr, cr = oracle_cells(oracle, nid, w)
        assert np.array_equal(a[0], yr) and np.array_equal(a[1], sr)


@pytest.mark.parametrize("n", [1, 63, 257, 1000])
@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_ragged_sizes_and_skips(oracle, n, mode):
    w = W.primordial(n, seed=9)
    w["t_stop"][::5] = 0.0           # t_stop == t0: skipped, y unchanged
    y, st, dto, stats = gpu_solve("PRIMORDIAL/CMK/CSC+REPRO", w, mode=mode)
    yr, sr, dr, cr = oracle_cells(oracle, "PRIMORDIAL/CMK/CSC", w)
    assert np.array_equal(st, sr) and np.array_equal(y, yr)
    assert (st[::5] == 5).all() and np.array_equal(y[:, ::5], w["y"][:, ::5])


def test_rober_default_build_within_10rtol(oracle):
    w = W.rober(512)
    for mode in (_lib.PERCELL, _lib.PERSISTENT):
        y, st, dto, stats = gpu_solve("ROBER/NONE/DENSE", w, mode=mode)
        yr, sr, dr, cr = oracle_cells(oracle, "ROBER/NONE/DENSE", w)
        assert (st == 0).all()
        assert np.all(np.abs(y - yr) <= 10 * w["rtol"] * np.abs(yr) + w["atol"])


@pytest.mark.parametrize("precision", [0, 1])
def test_default_build_accuracy_vs_radau(oracle, precision):
    """Distance to the rtol-1e-12 Radau answers is of the size of the
    oracle's own (the problems amplify rounding differences, so two builds
    that round differently cannot agree to 10*rtol with each other)."""
    gold = json.load(open(os.path.join(GOLD, "radau_known_answers.json")))["cases"]
    dt0 = {"ROBER": 1e-10, "PRIMORDIAL": W.PRIMORDIAL_DT_HINT, "KROME_ISM": 1e-6}
    for name in ("ROBER", "PRIMORDIAL", "KROME_ISM"):
        case = gold[name]
        cells = case["cells"]
        n = len(cells)
        w = {"network": case["network"], "y": np.array([c["y0"] for c in cells]).T.copy(),
             "t_stop": np.array([c["t_stop"] for c in cells]), "t0": np.zeros(n),
             "dt0": np.full(n, dt0[name]),
             "params": None if cells[0]["params"] is None else np.array([c["params"] for c in cells]).T.copy(),
             "rtol": 1e-6, "atol": 1e-10 if name == "ROBER" else 1e-20, "dt_min": 1e-20}
        y, st, _, _ = gpu_solve(case["network"], w, precision=precision)
        yo, so, _, _ = oracle_cells(oracle, case["network"], w)
        ref = np.array([c["y"] for c in cells]).T
        big = np.abs(ref) > 1e-6 * np.abs(ref).max(axis=0)
        eg = np.max(np.abs(y - ref)[big] / np.abs(ref[big]))
        eo = np.max(np.abs(yo - ref)[big] / np.abs(ref[big]))
        say(name, "precision", precision, "GPU err vs Radau %.3g, oracle %.3g" % (eg, eo))
        assert (st == 0).all() and eg <= max(10 * eo, 1e-5)


def test_mixed_precision_primordial_all_cells_finish():
    w = W.primordial(8192, seed=3)
    y, st, dto, stats = gpu_solve("PRIMORDIAL/CMK/CSC", w, precision=1, mode=_lib.PERSISTENT)
    assert (st == 0).all() and stats["flops32"] > 0


def test_krome_mixed_and_fp64_against_tight_solve():
    """The KROME setup-file network in the default build: FP64 and mixed
    precision (with the FP64 fallback) both finish every cell, and the mixed
    form's distance to a tight FP64 solve (rtol 1e-11, atol 1e-26), in units
    of rtol|y| + atol, is of the size of the FP64 run's own (median and 90th
    percentile within twice FP64's, or 10)."""
    w = W.krome_ism(4096, seed=5)
    yt, stt, _, _ = gpu_solve("KROME/FF/CSC", dict(w, rtol=1e-11, atol=1e-26, dt_min=1e-30),
                              mode=_lib.PERSISTENT)
    yd, sd, _, _ = gpu_solve("KROME/FF/CSC", w, mode=_lib.PERSISTENT)
    ym, sm, _, stats = gpu_solve("KROME/FF/CSC", w, precision=1, mode=_lib.PERSISTENT, fallback=128)
    assert (stt == 0).all() and (sd == 0).all() and (sm == 0).all() and stats["flops32"] > 0
    sc = w["rtol"] * np.abs(yt) + w["atol"]
    ed = (np.abs(yd - yt) / sc).max(axis=0)
    em = (np.abs(ym - yt) / sc).max(axis=0)
    for q in (0.5, 0.9):
        say("KROME q%.2f: FP64 %.3g mixed %.3g" % (q, np.quantile(ed, q), np.quantile(em, q)))
        assert np.quantile(em, q) <= max(2 * np.quantile(ed, q), 10.0)


def test_async_device_pointers_equal_sync():
    import torch
    w = W.primordial(4096, seed=12)
    from odevec_amd.api import Odevec
    ode = Odevec("PRIMORDIAL/CMK/CSC", 4096, mode=_lib.PERSISTENT)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = 1e-6, 1e-20, 1e-20
    y_sync = w["y"].copy()
    st_sync, _, _ = ode.solve_cells(y_sync, w["t_stop"], dt0=w["dt0"])
    dev = torch.device("cuda:0")
    y = torch.tensor(w["y"], dtype=torch.float64, device=dev)
    ts = torch.tensor(w["t_stop"], dtype=torch.float64, device=dev)
    dt0 = torch.tensor(w["dt0"], dtype=torch.float64, device=dev)
    st = torch.zeros(4096, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    say("GPU async solve on torch tensors")
    ode.solve_cells_async(4096, ts.data_ptr(), y.data_ptr(), dt0_ptr=dt0.data_ptr(), status_ptr=st.data_ptr(),
                          stream=stream.cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), y_sync) and (st.cpu().numpy() == st_sync).all()
    ode.CloseOdevec()


def test_repro_reference_failures(oracle):
    """The cells of the 1M workload on which the reference algorithm fails
    (it would STOP; golden/primordial_fail_cells.json) fail on the GPU with
    the same status and the same returned state."""
    fail = json.load(open(os.path.join(GOLD, "primordial_fail_cells.json")))
    full = W.primordial(fail["ncells"], seed=3)
    idx = np.array(sorted(set(fail["cells"]) | set(range(0, fail["ncells"], fail["ncells"] // 25))))
    w = W.subset(full, idx)
    y, st, dto, stats = gpu_solve(w["network"] + "+REPRO", w, mode=_lib.PERSISTENT)
    yr, sr, dr, cr = oracle_cells(oracle, w["network"], w)
    assert np.array_equal(st, sr) and np.array_equal(y, yr)
    assert set(idx[st != 0]) == set(fail["cells"])


def test_primordial_full_size_conservation():
    """Configuration size (1,048,576 cells, per-cell end times 1e8..1e14 s):
    cells finish at the reference algorithm's rate (it fails 39 of them), and
    the finished ones conserve H, He and D nuclei as well as the reference
    algorithm does (it conserves them up to rounding and tolerances: the
    oracle's largest drifts and 99.9th percentiles are recorded in the
    fixture), in FP64 and in the mixed form's benchmarked configuration."""
    fail = json.load(open(os.path.join(GOLD, "primordial_fail_cells.json")))
    w = W.primordial(1048576, seed=3)
    V = {"H": [0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0],
         "He": [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0],
         "D": [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0]}
    for precision in (0, 1):
        # the mixed form in its benchmarked configuration (FP64 fallback at
        # 128 convergence failures or 32 failed invariant tests)
        y, st, dto, stats = gpu_solve("PRIMORDIAL/CMK/CSC", w, precision=precision, mode=_lib.PERSISTENT,
                                      fallback=128 if precision else 0, schedule=1)
        ok = st == 0
        say("precision", precision, "status counts", np.bincount(st))
        assert ok.sum() >= 1048576 - 200 and set(np.unique(st[~ok])) <= {2}
        for name, v in V.items():
            v = np.array(v, dtype=float)
            a, b = v @ w["y"][:, ok], v @ y[:, ok]
            r = np.abs(a - b) / a
            drift = r.max()
            oq = fail["oracle_conservation_quantiles"][name]
            say(name, "drift max %.3g q99.9 %.3g cells>1e-6 %d (oracle max %.3g q99.9 %.3g cells>1e-6 %d)" % (
                drift, np.quantile(r, 0.999), (r > 1e-6).sum(), fail["oracle_conservation_max"][name], oq["q999"],
                oq["n_gt_1e-6"]))
            # both forms: the mixed form's corrector test keeps the invariants
            # (without it, FP32 factors let deuterium drift by up to 0.1); it
            # admits a residual of one tolerance per step, so the abundant H
            # nuclei may drift by a few rtol (measured 3.4e-6) where the
            # oracle's exact corrections keep 1e-7
            bound = 10 * fail["oracle_conservation_max"][name]
            if precision == 1:
                bound = max(bound, 10 * w["rtol"])
            assert drift <= bound + 1e-14
            assert np.quantile(r, 0.999) <= 10 * oq["q999"] + 1e-14
        assert np.array_equal(y[12:], w["y"][12:])    # frozen species untouched


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT, _lib.COOP])
def test_timeline_records_every_cell(mode):
    """With a timeline attached each cell's start and end ticks are written
    (end >= start, skipped cells included) and the results do not change."""
    import torch
    from odevec_amd.api import Odevec
    w = W.primordial(700, seed=13)
    w["t_stop"][::11] = 0.0
    n = w["y"].shape[1]
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    say("GPU timeline, mode", mode)
    ref = gpu_solve(nid, w, mode=mode, grid=2 if mode == _lib.PERSISTENT else 0)
    ode = Odevec(nid, n, mode=mode, grid=2 if mode == _lib.PERSISTENT else 0)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    times, hz = ode.enable_timeline(n)
    y = np.array(w["y"], order="C", copy=True)
    st, dto, _ = ode.solve_cells(y, w["t_stop"], t0=w["t0"], dt0=w["dt0"], raise_on_fail=False)
    t = times.cpu().numpy().reshape(2, n)
    ode.CloseOdevec()
    assert hz > 0
    assert np.array_equal(y, ref[0]) and np.array_equal(st, ref[1]) and np.array_equal(dto, ref[2])
    assert (t[0] > 0).all() and (t[1] >= t[0]).all()
    assert (t[1] - t[0]).max() / hz < 60.0


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_check_negatives_per_cell(oracle, mode):
    """CheckNegatives in the per-cell forms (src/odevec.F90:265-291): cells
    whose predicted state goes negative are reset with a quarter step, as in
    the oracle, bit for bit; cells that start negative end at dt_min."""
    from odevec_amd.api import Odevec
    w = W.primordial(96, seed=14)
    w["y"][4, 48:] = -1e-3                      # negative H2 in the second half
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    n = w["y"].shape[1]
    say("GPU check negatives, mode", mode)
    ode = Odevec(nid, n, mode=mode)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], 1e-3
    ode.CheckNegatives = True
    y = np.array(w["y"], order="C", copy=True)
    st, dto, stats = ode.solve_cells(y, w["t_stop"], t0=w["t0"], dt0=w["dt0"], stats=True, raise_on_fail=False)
    ode.CloseOdevec()
    yr, sr, dr, cr = oracle.Network.load("PRIMORDIAL/CMK/CSC").solve_cells(
        w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], rtol=w["rtol"], atol=w["atol"], dt_min=1e-3,
        check_negatives=True)
    assert np.array_equal(st, sr) and np.array_equal(y, yr) and np.array_equal(dto, dr)
    assert (st[48:] == 1).all()
    assert stats["attempts"] == int(cr[0].sum())


def test_solveode_mask_per_cell(oracle):
    """SolveODE's Mask with per-cell step control: masked-out cells come back
    unchanged, the others equal the oracle's per-cell solves."""
    from odevec_amd.api import Odevec
    w = W.rober(48, seed=15)
    mask = np.zeros(48, np.int8)
    mask[1::4] = 1
    mask[2::4] = 1
    ode = Odevec("ROBER/NONE/DENSE+REPRO", 48, mode=_lib.PERCELL)
    ode.y[:] = w["y"].T
    say("GPU SolveODE masked per cell")
    # per-cell rate constants are the network defaults here (SolveODE has no parameter argument)
    t, dt = ode.SolveODE(0.0, 1e-10, 1e3, Mask=mask)
    y = ode.y.T.copy()
    ode.CloseOdevec()
    yr, sr, dr, cr = oracle.Network.load("ROBER/NONE/DENSE").solve_cells(w["y"], 1e3, dt0=1e-10, rtol=1e-6,
                                                                         atol=1e-10)
    on = mask == 1
    assert t == 1e3
    assert np.array_equal(y[:, ~on], w["y"][:, ~on])
    assert np.array_equal(y[:, on], yr[:, on])


@pytest.mark.parametrize("nid", ["ROBER/NONE/DENSE", "OREGO/NONE/CSC", "PRIMORDIAL/CMK/CSC", "PRIMORDIAL/NONE/DENSE",
                                 "ROBER/NONE/DENSE/LU2", "OREGO/NONE/DENSE/LU2", "KROME/FF/CSC"])
def test_plugin_getrhs_getjac(oracle, nid):
    """The generated plugin's GetRHS / GetJac through the C ABI (evaluated
    on the GPU, +REPRO build) equal the oracle's RHS and permuted P bit for
    bit, in the reference's layouts (DENSE: all neq x neq entries; CSC: the
    full LU pattern, fill entries 0; PRIMORDIAL's frozen species included).
    For a LUmethod 2 variant the same entry point returns GetLU's closed-form
    factors in the DENSE layout."""
    from odevec_amd.api import Odevec
    w = {"ROBER": lambda: W.rober(40, seed=16), "OREGO": lambda: W.orego(40, seed=17),
         "PRIMORDIAL": lambda: W.primordial(40, seed=18), "KROME": lambda: W.krome_ism(40, seed=19)}[
        nid.split("/")[0]]()
    n = w["y"].shape[1]
    # KROME: the rate coefficients are the caller's per-cell parameters
    prm = w["params"] if nid.startswith("KROME") else None
    say("GPU GetRHS/GetJac", nid)
    ode = Odevec(nid + "+REPRO", n)
    f = ode.GetRHS(w["y"], prm)
    beta, dt = 6.0 / 11.0, 3e-4
    jac = ode.GetJac(beta, w["y"], dt, prm)
    ode.CloseOdevec()
    net = oracle.Network.load(nid)
    neq = net.neq
    for i in range(n):
        k = net.rates(w["y"][:, i], None if prm is None else prm[:, i])
        assert np.array_equal(f[:, i], net.rhs(w["y"][:, i], k))
        P = net.p_dense(w["y"][:, i], k, beta * dt)
        if net.ir["lu_method"] == 2:
            # GetLU: the closed-form factors (L below the diagonal, U on and above)
            P = net.factor(w["y"][:, i], k, beta * dt).reshape(neq, neq)
        if net.csc:
            ref = np.array([P[r, c] for r, c in net.ir["lu_entries"]])
        else:
            ref = P.T.reshape(-1)          # entry c*neq + r
        assert len(ref) == jac.shape[0] and np.array_equal(jac[:, i], ref), i


def _solve_dev(nid, w, precision, mode, fallback=0, inv=None):
    """Device-resident solve with per-cell counters: (y, status, dt_out, counters[8, n])."""
    import torch
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    say("GPU solve (device arrays)", nid, "cells", n, "precision", precision, "mode", mode, "fallback", fallback)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    ode = Odevec(nid, n, precision=precision, mode=mode)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    if fallback:
        ode.set_fp64_fallback(fallback)
    if inv is not None:
        ode.set_invariant_test(*inv)
    y = torch.tensor(np.ascontiguousarray(w["y"]), device=dev)
    ts = torch.tensor(np.ascontiguousarray(w["t_stop"], dtype=np.float64), device=dev)
    dt0 = torch.tensor(np.broadcast_to(w["dt0"], (n,)).astype(np.float64), device=dev)
    st = torch.zeros(n, dtype=torch.int32, device=dev)
    dto = torch.zeros(n, dtype=torch.float64, device=dev)
    cnt = torch.zeros(8 * n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ode.solve_cells_async(n, ts.data_ptr(), y.data_ptr(), dt0_ptr=dt0.data_ptr(), status_ptr=st.data_ptr(),
                          dt_out_ptr=dto.data_ptr(), counters_ptr=cnt.data_ptr(), stream=stream.cuda_stream)
    stream.synchronize()
    out = (y.cpu().numpy(), st.cpu().numpy(), dto.cpu().numpy(), cnt.cpu().numpy().reshape(8, n))
    ode.CloseOdevec()
    return out


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
@pytest.mark.parametrize("nid,limit", [("PRIMORDIAL/CMK/CSC+REPRO", 2), ("OREGO/NONE/DENSE+REPRO", 1)])
def test_fallback_handoff_is_lossless(mode, nid, limit):
    """The fallback's hand-off on an FP64 handle: cells that reach the limit
    stop in the first launch, their integrator state goes through the
    hand-off record and the second launch continues them -- the results
    (states, status, last steps, all 8 counters) equal the solve without the
    fallback bit for bit, and so the oracle's for this +REPRO build. Covers
    the state kept in LDS and in the workspace (PRIMORDIAL) and in registers
    (OREGO, whose weights also carry the NaN flag)."""
    w = W.primordial(3000, seed=13) if nid.startswith("PRIMORDIAL") else W.orego(4096, seed=2)
    w["t_stop"][::11] = 0.0                      # skipped cells stay skipped
    yd, sd, dd, cd = _solve_dev(nid, w, 0, mode)
    yf, sf, df, cf = _solve_dev(nid, w, 0, mode, fallback=limit)
    handed = (cd[5] >= limit) & (sd == 0)
    say("cells handed off", int(handed.sum()), "of", len(sd))
    assert handed.sum() >= 5
    assert np.array_equal(yf, yd) and np.array_equal(sf, sd)
    assert np.array_equal(df, dd) and np.array_equal(cf, cd)


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_mixed_fp64_fallback(mode):
    """Mixed form with the FP64 fallback: cells whose mixed solve stays below
    the limit of convergence failures carry the plain mixed result bit for
    bit (state, status, last step, counters); the others continue in FP64
    from their hand-off, finish, keep the counts of both launches, and land
    near the FP64 form's result; per-cell and persistent modes agree bit for
    bit. The invariant test is off here, so that the handed-back cells are
    exactly those whose convergence failures reach the limit (the
    invariant-triggered hand-back is tested below)."""
    w = W.primordial(3000, seed=13)
    w["t_stop"][::11] = 0.0                      # skipped cells stay skipped
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    limit = 8
    off = (0.0, 0)
    ym, sm, dm, cm = _solve_dev(nid, w, 1, mode, inv=off)
    yd, sd, dd, cd = _solve_dev(nid, w, 0, _lib.PERCELL)
    yf, sf, df, cf = _solve_dev(nid, w, 1, mode, fallback=limit, inv=off)
    prom = (cm[5] >= limit) & (sm != 2)
    say("promoted cells", int(prom.sum()), "of", len(prom))
    assert 20 <= prom.sum() <= len(prom) // 2
    keep = cm[5] < limit
    assert np.array_equal(yf[:, keep], ym[:, keep]) and np.array_equal(sf[keep], sm[keep])
    assert np.array_equal(df[keep], dm[keep]) and np.array_equal(cf[:, keep], cm[:, keep])
    assert (sf[::11] == 5).all()
    # handed-back cells: finished, their failure count includes the limit
    # reached in the mixed launch, and as close to an rtol-1e-12 solution as
    # the FP64 form's results for the same cells (these are the hard cells,
    # whose trajectories any change of rounding moves by many tolerances)
    ok = prom & (sf == 0) & (sd == 0)
    assert ok.sum() >= 0.9 * prom.sum()
    assert (cf[5][prom] >= limit).all()
    idx = np.nonzero(ok)[0]
    yt, stt, _, _ = _solve_dev(nid, dict(W.subset(w, idx), rtol=1e-12, atol=1e-26), 0, _lib.PERSISTENT)
    good = stt == 0

    def err(y):
        yy, tt = y[:12, idx][:, good], yt[:12, good]
        return (np.abs(yy - tt) / (w["rtol"] * np.abs(tt) + w["atol"])).max(axis=0)
    ef, ed = err(yf), err(yd)
    say("handed-back cells vs rtol 1e-12 (%d): continued median %.3g q90 %.3g | FP64 form median %.3g q90 %.3g"
        % (good.sum(), np.median(ef), np.quantile(ef, 0.9), np.median(ed), np.quantile(ed, 0.9)))
    # (the rtol-1e-12 solve itself stops early on some of these cells)
    assert good.sum() >= 0.5 * len(idx)
    assert np.median(ef) <= max(2.0 * np.median(ed), 10.0)
    assert np.quantile(ef, 0.9) <= max(2.0 * np.quantile(ed, 0.9), 10.0)
    if mode == _lib.PERSISTENT:
        yp, sp, dp, cp = _solve_dev(nid, w, 1, _lib.PERCELL, fallback=limit, inv=off)
        assert np.array_equal(yp, yf) and np.array_equal(sp, sf) and np.array_equal(dp, df)
        assert np.array_equal(cp, cf)


@pytest.mark.parametrize("mode", [_lib.PERCELL, _lib.PERSISTENT])
def test_mixed_fp64_fallback_on_invariant_failures(mode):
    """With the fallback counting failed invariant tests (here: the first
    one hands the cell back, the convergence-failure limit out of reach),
    every cell carries either its plain mixed result bit for bit (no failed
    test) or, handed back, a continued result that finishes; both kinds
    occur, and per-cell and persistent modes agree."""
    w = W.primordial(3000, seed=13)
    nid = "PRIMORDIAL/CMK/CSC+REPRO"
    ym, sm, dm, cm = _solve_dev(nid, w, 1, mode, inv=(3.0, 1))
    yf, sf, df, cf = _solve_dev(nid, w, 1, mode, fallback=1 << 30, inv=(3.0, 1))
    as_mixed = np.all(yf == ym, axis=0) & (sf == sm) & (df == dm) & np.all(cf == cm, axis=0)
    handed = ~as_mixed
    say("handed back", int(handed.sum()), "kept", int(as_mixed.sum()))
    assert handed.sum() >= 1 and as_mixed.sum() >= len(sf) // 2
    assert ((sf[handed] == 0) | (sm[handed] != 0)).mean() >= 0.9
    yp, sp, dp, cp = _solve_dev(nid, w, 1, _lib.PERCELL if mode == _lib.PERSISTENT else _lib.PERSISTENT,
                                fallback=1 << 30, inv=(3.0, 1))
    assert np.array_equal(yp, yf) and np.array_equal(sp, sf) and np.array_equal(cp, cf)


def test_int_index_lists_refuse_2pow31_cells():
    """The FP64 fallback's list and the cost-ordered queue store cell indices
    as int: a solve of 2^31 cells that would use them is refused before
    anything is launched (the pointers are never read)."""
    from odevec_amd.api import Odevec
    fake = 1 << 12
    for prec, sched, fb in ((_lib.MIXED, 0, 128), (_lib.FP64, 1, 0)):
        say("refusal of 2^31 cells, precision", prec, "schedule", sched, "fallback", fb)
        ode = Odevec("PRIMORDIAL/CMK/CSC", 16, precision=prec, mode=_lib.PERSISTENT, schedule=sched)
        if fb:
            ode.set_fp64_fallback(fb)
        with pytest.raises(_lib.OdevecError) as ex:
            ode.solve_cells_async(1 << 31, fake, fake)
        assert ex.value.code == _lib.EINVAL and "2^31" in str(ex.value)
        ode.CloseOdevec()
