"""Batching form on the GPU: one launch over segments of different networks,
precisions and sizes (odevec_batch_*).

* +REPRO bundle: every segment's cells equal the oracle bit for bit (states,
  status, last step, counters), uneven and empty segments included.
* default bundle: a segment's results are the same alone and inside the mix
  (bitwise: per-cell step control, no data shared between cells).
* graph mode: replays of the captured launch, with eager work and
  synchronizations between them, each equal the eager launch.
* cost order: segments with more cells than the launch has lanes get their
  queue ranked by cost; on a one-workgroup grid the +REPRO bundle still
  equals the oracle, and with the FP64 fallback results equal a full grid's.
* the batching workload at full size (2,097,152 cells) through the
  reproducible build of the tuned bundle, fallback and graph replay: seeded
  samples equal the single-network kernels bit for bit.
"""
import numpy as np
import pytest

from odevec_amd import _lib
from odevec_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def to_dev(w, precision):
    import torch
    dev = torch.device("cuda:0")
    n = w["y"].shape[1]
    seg = {"network": w["network"], "precision": precision,
           "y": torch.tensor(w["y"], dtype=torch.float64, device=dev),
           "t0": torch.tensor(np.broadcast_to(w["t0"], (n,)).copy(), dtype=torch.float64, device=dev),
           "t_stop": torch.tensor(np.broadcast_to(w["t_stop"], (n,)).copy(), dtype=torch.float64, device=dev),
           "dt0": torch.tensor(np.broadcast_to(w["dt0"], (n,)).copy(), dtype=torch.float64, device=dev),
           "status": torch.full((n,), -9, dtype=torch.int32, device=dev),
           "dt_out": torch.zeros(n, dtype=torch.float64, device=dev),
           "counters": torch.zeros((8, n), dtype=torch.int32, device=dev),
           "rtol": w["rtol"], "atol": w["atol"], "dt_min": w["dt_min"]}
    if w.get("params") is not None:
        seg["params"] = torch.tensor(w["params"], dtype=torch.float64, device=dev)
    return seg


def host(seg):
    return (seg["y"].cpu().numpy(), seg["status"].cpu().numpy(), seg["dt_out"].cpu().numpy(),
            seg["counters"].cpu().numpy())


_STREAM = []


def stream():
    """A created stream (graph capture is not possible on the null stream)."""
    import torch
    if not _STREAM:
        _STREAM.append(torch.cuda.Stream())
    return _STREAM[0]


def run(bundle, segs, use_graph=False, batch=None):
    import torch
    from odevec_amd.api import Batch
    b = batch or Batch(bundle, use_graph=use_graph)
    torch.cuda.synchronize()            # the segments' inputs were written on the current stream
    b.solve_async(segs, stream().cuda_stream)
    torch.cuda.synchronize()
    if batch is None:
        b.close()
    return [host(s) for s in segs]


def mixed_workload():
    rob = W.rober(100, seed=11)
    rob["t_stop"] = W._logu(np.random.default_rng(1), 1e-1, 1e5, 100)
    ore = W.orego(63, seed=12)
    ore["t_stop"] = W._logu(np.random.default_rng(2), 3.6e-4, 360.0, 63)
    pri = W.primordial(300, seed=13)
    pri["t_stop"][::17] = 0.0            # skipped cells
    empty = W.rober(0, seed=14)
    one = W.primordial(1, seed=15)
    return [pri, rob, ore, empty, one]


def test_batch_repro_equals_oracle(oracle):
    ws = mixed_workload()
    say("GPU batch MIX3+REPRO, segments", [w["y"].shape[1] for w in ws])
    segs = [to_dev(w, _lib.FP64) for w in ws]
    out = run("MIX3+REPRO", segs)
    for w, (y, st, dto, cnt) in zip(ws, out):
        n = w["y"].shape[1]
        if n == 0:
            continue
        yr, sr, dr, cr = oracle.Network.load(w["network"]).solve_cells(
            w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"],
            dt_min=w["dt_min"])
        assert np.array_equal(st, sr), w["network"]
        assert np.array_equal(y, yr), w["network"]
        assert np.array_equal(dto, dr), w["network"]
        assert np.array_equal(cnt, cr), w["network"]


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
def test_batch_segment_same_alone_and_in_mix(precision):
    ws = mixed_workload()
    say("GPU batch mix invariance, precision", precision)
    together = run("MIX3", [to_dev(w, precision) for w in ws])
    for k, w in enumerate(ws):
        alone = run("MIX3", [to_dev(w, precision)])[0]
        for a, b in zip(alone, together[k]):
            assert np.array_equal(a, b), (w["network"], precision)
    pri = together[0]
    assert (pri[1][::17] == 5).all() and (np.delete(pri[1], np.s_[::17]) == 0).all()


def test_batch_graph_replays_equal_eager():
    import torch
    from odevec_amd.api import Batch
    ws = mixed_workload()
    say("GPU batch graph replays")
    eager = run("MIX3", [to_dev(w, _lib.MIXED) for w in ws])
    b = Batch("MIX3", use_graph=True)
    for rep in range(3):
        segs = [to_dev(w, _lib.MIXED) for w in ws]     # fresh state: eager host->device work between replays
        got = run("MIX3", segs, batch=b)
        torch.cuda.synchronize()
        for e, g in zip(eager, got):
            for a, c in zip(e, g):
                assert np.array_equal(a, c), rep
    # a different table through the same captured graph (the table is read
    # from device memory at run time; in graph mode the launch shape does not
    # depend on the batch)
    ws2 = [W.rober(50, seed=21), W.orego(40, seed=22)]
    got = run("MIX3", [to_dev(w, _lib.FP64) for w in ws2], batch=b)
    ref = run("MIX3", [to_dev(w, _lib.FP64) for w in ws2])
    for e, g in zip(ref, got):
        for a, c in zip(e, g):
            assert np.array_equal(a, c)
    b.close()


def test_batch_rejects_entries_not_in_bundle():
    import torch
    from odevec_amd.api import Batch
    w = W.primordial(4)
    seg = to_dev(dict(w, network="PRIMORDIAL/FF/CSC"), _lib.FP64)
    with Batch("MIX3") as b:
        with pytest.raises(_lib.OdevecError) as e:
            b.solve_async([seg], stream().cuda_stream)
        assert e.value.code == _lib.ENOTFOUND


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_fp64_fallback(use_graph):
    """Mixed segments with the FP64 fallback: every cell carries the result
    of the single-network mixed kernel with the same fallback bit for bit
    (handed-back cells continued in FP64 from their hand-off records); cells
    below the limit the batch's plain mixed result; an FP64 segment in the
    same batch is unaffected; graph replays agree."""
    from odevec_amd.api import Batch, Odevec
    import torch
    pri = W.primordial(1500, seed=21)
    pri["t_stop"][::13] = 0.0
    rob = W.rober(100, seed=22)
    limit = 8
    say("GPU batch fallback, graph", use_graph)
    # the invariant test off: the handed-back cells are then exactly those
    # whose convergence failures reach the limit
    pb = Batch("MIX3")
    pb.set_invariant_test(0.0, 0)
    plain = run("MIX3", [to_dev(pri, _lib.MIXED), to_dev(rob, _lib.FP64)], batch=pb)
    pb.close()
    b = Batch("MIX3", use_graph=use_graph, fp64_fallback=limit)
    b.set_invariant_test(0.0, 0)
    segs = [to_dev(pri, _lib.MIXED), to_dev(rob, _lib.FP64)]
    fb = run("MIX3", segs, batch=b)
    if use_graph:                        # a replay over fresh inputs gives the same
        segs2 = [to_dev(pri, _lib.MIXED), to_dev(rob, _lib.FP64)]
        fb2 = run("MIX3", segs2, batch=b)
        assert all(np.array_equal(x, y) for p, q in zip(fb, fb2) for x, y in zip(p, q))
    b.close()
    # the same cells through a handle: the single-network mixed per-cell
    # kernel with the same fallback
    n = pri["y"].shape[1]
    ode = Odevec(pri["network"], n, precision=1, mode=_lib.PERCELL)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = pri["rtol"], pri["atol"], pri["dt_min"]
    ode.set_fp64_fallback(limit)
    ode.set_invariant_test(0.0, 0)
    dev = to_dev(pri, _lib.MIXED)
    torch.cuda.synchronize()
    ode.solve_cells_async(n, dev["t_stop"].data_ptr(), dev["y"].data_ptr(), t0_ptr=dev["t0"].data_ptr(),
                          dt0_ptr=dev["dt0"].data_ptr(), status_ptr=dev["status"].data_ptr(),
                          dt_out_ptr=dev["dt_out"].data_ptr(), counters_ptr=dev["counters"].data_ptr(),
                          stream=stream().cuda_stream)
    torch.cuda.synchronize()
    ode.CloseOdevec()
    yd, sd, dd, cd = host(dev)
    ym, sm, dm, cm = plain[0]
    yf, sf, df, cf = fb[0]
    prom = cm[5] >= limit
    say("handed back", int(prom.sum()), "of", n)
    assert 10 <= prom.sum() <= n // 2
    assert np.array_equal(yf, yd) and np.array_equal(sf, sd)
    assert np.array_equal(df, dd) and np.array_equal(cf, cd)
    keep = ~prom
    assert np.array_equal(yf[:, keep], ym[:, keep]) and np.array_equal(sf[keep], sm[keep])
    assert np.array_equal(df[keep], dm[keep]) and np.array_equal(cf[:, keep], cm[:, keep])
    assert all(np.array_equal(x, y) for x, y in zip(fb[1], plain[1]))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_cost_order_small_grid_equals_oracle(oracle, use_graph):
    """One workgroup of 64 lanes: every per-lane segment has more cells than
    the launch has lanes, so each gets its queue in cost order (ranked before
    the launch); the +REPRO bundle still equals the oracle bit for bit."""
    from odevec_amd.api import Batch
    ws = mixed_workload()
    say("GPU batch cost order, grid 1 x 64, graph", use_graph)
    b = Batch("MIX3+REPRO", block=64, grid=1, use_graph=use_graph)
    out = run("MIX3+REPRO", [to_dev(w, _lib.FP64) for w in ws], batch=b)
    b.close()
    assert sum(w["y"].shape[1] > 64 for w in ws) >= 2       # PRIMORDIAL and ROBER get ranked queues
    for w, (y, st, dto, cnt) in zip(ws, out):
        if w["y"].shape[1] == 0:
            continue
        yr, sr, dr, cr = oracle.Network.load(w["network"]).solve_cells(
            w["y"], w["t_stop"], t0=w["t0"], dt0=w["dt0"], params=w["params"], rtol=w["rtol"], atol=w["atol"],
            dt_min=w["dt_min"])
        assert np.array_equal(st, sr) and np.array_equal(y, yr), w["network"]
        assert np.array_equal(dto, dr) and np.array_equal(cnt, cr), w["network"]


def test_batch_cost_order_with_fallback_equals_full_grid():
    """Cost-ordered segment queues on a small grid, together with the FP64
    fallback (which reuses the ranked list's machinery): results equal a
    full-grid launch (index order, no ranking) bit for bit."""
    from odevec_amd.api import Batch
    pri = W.primordial(2000, seed=31)
    rob = W.rober(500, seed=32)
    say("GPU batch cost order + fallback, grid 2 x 64")
    outs = []
    for grid in (2, 0):
        b = Batch("MIX3", block=64, grid=grid, fp64_fallback=8)
        outs.append(run("MIX3", [to_dev(pri, _lib.MIXED), to_dev(rob, _lib.FP64)], batch=b))
        b.close()
    for p, q in zip(*outs):
        for x, y in zip(p, q):
            assert np.array_equal(x, y)
    assert (outs[0][0][1] == 0).mean() > 0.99


def test_batching_configuration_full_size():
    """The batching workload at full size (2,097,152 cells, 40/30/30
    ROBER/OREGO/PRIMORDIAL, seed 4) through the reproducible build of the
    tuned bundle and precisions, with its FP64 fallback and graph replay:
    every cell the reference's algorithm finishes finishes, and seeded
    samples of each segment equal the single-network persistent kernels
    (same precision, same fallback) bit for bit."""
    import torch
    from odevec_amd import autotune as A
    from odevec_amd.api import Batch, Odevec
    segs = W.batch_mix(2097152, seed=4)
    nets = [w["network"] for w in segs]
    t = A.lookup(nets) or {"bundle": "MIX3_MIXED", "block": 64, "fp64_fallback": 128,
                           "precision": {"ROBER/NONE/DENSE": 0, "OREGO/NONE/DENSE": 0, "PRIMORDIAL/CMK/CSC": 1}}
    bundle = t["bundle"] + "+REPRO"
    fb = t.get("fp64_fallback", 0)
    say("GPU batching configuration:", bundle, "block", t["block"], "fallback", fb, "cells", [w["y"].shape[1] for w in segs])
    b = Batch(bundle, block=t["block"], use_graph=True, fp64_fallback=fb)
    dsegs = [to_dev(w, t["precision"][w["network"]]) for w in segs]
    out = run(bundle, dsegs, batch=b)
    b.close()
    del dsegs
    rng = np.random.default_rng(44)
    for w, (y, st, dto, cnt) in zip(segs, out):
        n = w["y"].shape[1]
        prec = t["precision"][w["network"]]
        say(w["network"], "precision", prec, "status", np.bincount(st))
        assert (st == 0).mean() > 0.9999
        idx = np.sort(rng.choice(n, 4096, replace=False))
        s = W.subset(w, idx)
        ode = Odevec(w["network"] + "+REPRO", idx.size, precision=prec, mode=_lib.PERSISTENT)
        ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = s["rtol"], s["atol"], s["dt_min"]
        if prec == _lib.MIXED and fb:
            ode.set_fp64_fallback(fb)
        ys = np.array(s["y"], order="C", copy=True)
        ss, ds, _ = ode.solve_cells(ys, s["t_stop"], t0=s["t0"], dt0=s["dt0"], params=s["params"], raise_on_fail=False)
        ode.CloseOdevec()
        assert np.array_equal(ss, st[idx]) and np.array_equal(ys, y[:, idx]), w["network"]
        assert np.array_equal(ds, dto[idx]), w["network"]
    torch.cuda.empty_cache()


def test_batch_history_with_fallback_equals_one_solve():
    """A mixed segment with a warm-restart history and the FP64 fallback
    (limit 2): the batch solve split at t1 -- cells handed to FP64 lanes
    before t1 are handed straight to them again in the second launch and
    continued from the history -- equals the single batch solve with the
    fallback bit for bit (states, status, last steps; attempts, steps,
    Newton iterations and Jacobians add up). Autonomous network over
    [-S, 0], so the first step does not depend on t_stop. The reproducible
    bundle: the history's kernel is another instantiation than the plain
    one, and without contraction both do the same arithmetic."""
    from odevec_amd.api import Batch
    import torch
    w = W.primordial(2000, seed=48)
    n = w["y"].shape[1]
    span = np.asarray(w["t_stop"], dtype=np.float64)
    limit = 2
    say("GPU batch history + fallback")
    single = dict(w, t0=-span, t_stop=np.zeros(n))
    b = Batch("MIX3+REPRO", fp64_fallback=limit)
    one = run("MIX3+REPRO", [to_dev(single, _lib.MIXED)], batch=b)[0]
    planes = _lib.history_planes(w["network"])
    hist = torch.zeros(planes * n, dtype=torch.float64, device="cuda")
    seg = to_dev(dict(w, t0=-span, t_stop=-span * 0.55), _lib.MIXED)
    seg["hist"] = hist
    ya, sa, _, ca = run("MIX3+REPRO", [seg], batch=b)[0]
    seg["t0"] = seg["t_stop"].clone()
    seg["t_stop"] = torch.zeros(n, dtype=torch.float64, device="cuda")
    seg["counters"].zero_()
    yb, sb, db, cb = run("MIX3+REPRO", [seg], batch=b)[0]
    b.close()
    y1, s1, d1, c1 = one
    ok = (sa == 0) & (s1 == 0)
    say("status: single", np.bincount(s1, minlength=7).tolist(), "to t1", np.bincount(sa, minlength=7).tolist())
    assert ok.mean() > 0.99 and np.array_equal(sb[sa == 0], s1[sa == 0])
    before, handed = ca[5] >= limit, c1[5] >= limit
    say("handed over: before t1", int((before & handed & ok).sum()), "after", int((~before & handed & ok).sum()))
    assert (before & handed & ok).sum() >= 5 and (~before & handed & ok).sum() >= 5
    assert np.array_equal(yb[:, ok], y1[:, ok]) and np.array_equal(db[ok], d1[ok])
    for k, name in enumerate(("attempts", "steps", "newton", "jac")):
        assert np.array_equal(ca[k][ok] + cb[k][ok], c1[k][ok]), name


def test_batch_quantum_launches_equal_one_launch(oracle):
    """Launch quantum (continuous batching): a batch whose cells may run at
    most 12 attempts per launch, relaunched until no cell is paused -- with
    a second segment joining at the third launch -- gives every cell the
    result of the solve in one launch bit for bit (states, status, last
    steps, counters), FP64 and mixed with its FP64 fallback (limit 2); the
    FP64 segment also equals the oracle."""
    from odevec_amd.api import Batch
    import torch
    pri = W.primordial(1200, seed=23)
    rob = W.rober(900, seed=24)
    pri["t_stop"][::17] = 0.0
    limit = 2
    say("GPU batch quantum")

    def seg(w, prec, quantum):
        s = to_dev(w, prec)
        s["hist"] = torch.zeros(_lib.history_planes(w["network"]) * w["y"].shape[1], dtype=torch.float64,
                                device="cuda")
        if quantum:
            s["quantum"] = quantum
            s["status"].fill_(_lib.FRESH)
        return s

    b = Batch("MIX3+REPRO", fp64_fallback=limit)
    one = run("MIX3+REPRO", [seg(pri, _lib.MIXED, 0), seg(rob, _lib.FP64, 0)], batch=b)
    sp, sr = seg(pri, _lib.MIXED, 12), seg(rob, _lib.FP64, 12)
    launches, active = 0, [sp]
    while True:
        if launches == 2:
            active.append(sr)          # joins the running work
        run("MIX3+REPRO", active, batch=b)
        launches += 1
        paused = sum(int((s["status"] == _lib.PAUSED).sum()) for s in active)
        if paused == 0 and len(active) == 2:
            break
        assert launches < 10000
    b.close()
    say("quantum launches", launches)
    assert launches > 20
    for got, ref in zip((host(sp), host(sr)), one):
        assert all(np.array_equal(x, y) for x, y in zip(got, ref))
    ys, ss, ds, _ = host(sr)
    yo, so, do, _ = oracle.Network.load(rob["network"]).solve_cells(
        rob["y"], rob["t_stop"], t0=rob["t0"], dt0=rob["dt0"], params=rob["params"], rtol=rob["rtol"],
        atol=rob["atol"], dt_min=rob["dt_min"])
    assert np.array_equal(ys, yo) and np.array_equal(ss, so) and np.array_equal(ds, do)
    assert (one[0][1][::17] == 5).all() and (one[0][1] == 0).mean() > 0.9
