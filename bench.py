"""Benchmark: PRIMORDIAL cell-solves/s, mixed precision vs the plain FP64
kernel, on one GPU.

A step integrates 1,048,576 PRIMORDIAL cells (16 species, CSC + CMK
ordering, seeded inputs: n_H logU[1,1e12], T logU[10,1e4], x(H2) logU[1e-6,1e-2],
x(e) logU[1e-8,1e-3], t_stop logU[1e8,1e14] s; rtol 1e-6, atol 1e-20) from
its initial state (restored by a device copy at the start of the step) to
each cell's end time. Inputs are resident in HBM before the timed region.
The mixed form (FP32 iteration matrix, LU factors and triangular solves;
FP64 residual, history, norms and control) with the persistent
work-compacting kernel is the headline value; the plain FP64 kernel (one
lane per cell, no compaction) runs the same steps for the ratio.

Also on the line (skipped with --no-batch): the batching form (one launch
over a 2,097,152-cell ROBER/OREGO/PRIMORDIAL batch), SYNTH64 (64 species)
through the cooperative kernel, and the serving form (continuous batching
under Poisson arrivals: p50/p99 latency and cell-solves/s, all four networks
at 1 request/s and the three small ones at 40/s), and (skipped with
--no-cpu) the CPU baseline; and serving with warm restart (sessions whose
blocks stay on the device and are advanced in closed loop) against the same
advances cold-started. `--only batching|synth64|sessions` runs that part
alone and prints its object, `--only headline` the headline solve alone (for
profiling one kernel).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-batch] [--no-cpu] [--only PART]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NCELLS = 1048576
# mixed form: cells reaching this many convergence failures (or 32 failed
# invariant tests) continue in FP64 from where they stopped, in the same launch
# (odevec_set_fp64_fallback); on the headline workload limits 128 / 64 / 32
# (invariant tests 32 / 16 / 8) gave 119-121 / 121-124 / 132-137 ms per solve
FALLBACK = 128
# persistent queue in cost order (odevec_opts_t.schedule = 1): starting the
# costliest cells first -- among them every cell the fallback takes -- keeps
# the launch bound by its bulk (results unchanged)
SCHEDULE = 1
# serving: Poisson rates swept (requests/s), and the rate whose p50/p99 the
# line reports (the device sustains it: about a fifth of its SYNTH64 capacity)
SERVE_RATES = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0]
SERVE_SUSTAINED = 4.0
PEAK_FP64 = 78.6e12     # MI355X vector/matrix FP64, FLOP/s
PEAK_FP32 = 157.3e12    # MI355X FP32, FLOP/s
PEAK_HBM = 8.0e12       # B/s
PROFILE = os.path.join(ROOT, "profiles", "primordial_1m_mixed_persistent.md")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--cells", type=int, default=NCELLS)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--only", choices=["headline", "batching", "synth64", "sessions", "tail", "serving"], default=None)
    p.add_argument("--fallback", type=int, default=FALLBACK,
                   help="mixed form: FP64 fallback after this many convergence failures (0: off)")
    return p.parse_args()


class Runner:
    """One configuration of the solver on device-resident inputs."""

    def __init__(self, w, precision, mode, fallback=0, schedule=0, max_attempts=0):
        import torch
        from odevec_amd.api import Odevec
        self.torch = torch
        n = w["y"].shape[1]
        self.n = n
        dev = torch.device("cuda:0")
        self.y0 = torch.tensor(w["y"], dtype=torch.float64, device=dev)
        self.y = torch.empty_like(self.y0)
        self.ts = torch.tensor(w["t_stop"], dtype=torch.float64, device=dev)
        self.dt0 = torch.tensor(w["dt0"], dtype=torch.float64, device=dev)
        self.prm = torch.tensor(w["params"], dtype=torch.float64, device=dev) if w.get("params") is not None else None
        self.st = torch.zeros(n, dtype=torch.int32, device=dev)
        self.flops = torch.zeros(2, dtype=torch.float64, device=dev)
        self.ode = Odevec(w["network"], n, precision=precision, mode=mode, schedule=schedule,
                          max_attempts=max_attempts)
        self.ode.RelativeTolerance = w["rtol"]
        self.ode.AbsoluteTolerance = w["atol"]
        self.ode.MinimalTimestep = w["dt_min"]
        if fallback:
            self.ode.set_fp64_fallback(fallback)

    def restore(self):
        self.y.copy_(self.y0)

    def solve(self, count_flops=False):
        stream = self.torch.cuda.current_stream().cuda_stream
        assert stream, "a created stream is required (the null handle selects the solver's own stream)"
        self.ode.solve_cells_async(self.n, self.ts.data_ptr(), self.y.data_ptr(), dt0_ptr=self.dt0.data_ptr(),
                                   params_ptr=self.prm.data_ptr() if self.prm is not None else None,
                                   status_ptr=self.st.data_ptr(),
                                   flops_ptr=self.flops.data_ptr() if count_flops else None, stream=stream)

    def timed(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.restore()
            self.solve()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            # a step: restore the initial state, then the solve (the events
            # bracket the solver's launches on this stream)
            self.restore()
            a.record()
            self.solve()
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernel_ms = [a.elapsed_time(b) for a, b in ev]
        return wall, kernel_ms

    def count_flops(self):
        self.flops.zero_()
        self.restore()
        self.solve(count_flops=True)
        self.torch.cuda.synchronize()
        st = self.st.cpu().numpy()
        return self.flops.cpu().numpy(), st


# linear invariants of PRIMORDIAL (H, He and D nuclei; species order of the state)
NUCLEI = {"H": [0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0],
          "He": [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0],
          "D": [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0]}
TIGHT_CELLS = 8192


def cell_errors(y, yref, rtol, atol, nsp=12):
    """Per cell: the largest over its integrated species (the first nsp:
    PRIMORDIAL's 12; its last 4 are constant) of |y - y_ref| /
    (rtol |y_ref| + atol), the reference's error-weight scale
    (src/odevec.F90:706-716); 1 = one tolerance."""
    e = np.abs(y[:nsp] - yref[:nsp]) / (rtol * np.abs(yref[:nsp]) + atol)
    return e.max(axis=0)


# integrated species of each network (the rest of the state is constant)
ACTIVE = {"ROBER/NONE/DENSE": 3, "OREGO/NONE/DENSE": 3, "PRIMORDIAL/CMK/CSC": 12, "SYNTH64/FF/CSC": 64}


def err_vs_tight(w, y, st, mode, sample=2048, seed=34):
    """Errors of a solve's results (y, st over w's cells) against the FP64
    reproducible build at rtol 1e-12 (atol 1e-6 of w's) on `sample` seeded
    cells, in units of w's rtol |y_ref| + atol."""
    from odevec_amd import workloads as W
    n = w["y"].shape[1]
    idx = np.sort(np.random.default_rng(seed).choice(n, min(sample, n), replace=False))
    s = W.subset(w, idx)
    tight = Runner(dict(s, network=w["network"] + "+REPRO", rtol=1e-12, atol=w["atol"] * 1e-6), precision=0,
                   mode=mode, max_attempts=200000 if mode != 3 else 0)
    tight.restore()
    tight.solve()
    tight.torch.cuda.synchronize()
    y_t, s_t = tight.y.cpu().numpy(), tight.st.cpu().numpy()
    tight.ode.CloseOdevec()
    ok = (s_t == 0) & (st[idx] == 0)
    out = error_summary(cell_errors(y[:, idx][:, ok], y_t[:, ok], w["rtol"], w["atol"], ACTIVE[w["network"]]))
    out["sampled"] = int(idx.size)
    return out


def error_summary(e):
    return {"max": float(e.max()), "q999": float(np.quantile(e, 0.999)), "q99": float(np.quantile(e, 0.99)),
            "q50": float(np.median(e)), "within_10": float((e <= 10.0).mean()), "cells": int(e.size)}


def drift(y0, y):
    out = {}
    for name, v in NUCLEI.items():
        v = np.array(v, dtype=float)
        r = np.abs(v @ y0 - v @ y) / (v @ y0)
        out[name] = {"max": float(r.max()), "q999": float(np.quantile(r, 0.999))}
    return out


def accuracy(w, results):
    """Errors of each solve in `results` ({name: (y, status)}, full workload)
    against (a) the FP64 reproducible build, which equals the oracle -- the
    reference's algorithm -- bit for bit, over every cell, and (b) an FP64
    run at rtol 1e-12, atol 1e-26 (the reproducible build again, attempt
    budget 200,000) on TIGHT_CELLS seeded cells; errors in units of the
    configuration's tolerance scale rtol |y_ref| + atol. Also the largest
    drift of the H, He and D nuclei over the solve."""
    import torch
    n = w["y"].shape[1]
    ref = Runner(dict(w, network=w["network"] + "+REPRO"), precision=0, mode=2, schedule=1)
    ref.restore()
    ref.solve()
    torch.cuda.synchronize()
    y_ref, s_ref = ref.y.cpu().numpy(), ref.st.cpu().numpy()
    ref.ode.CloseOdevec()
    idx = np.sort(np.random.default_rng(33).choice(n, min(TIGHT_CELLS, n), replace=False))
    from odevec_amd import workloads as W
    s = W.subset(w, idx)
    tight = Runner(dict(s, network=w["network"] + "+REPRO", rtol=1e-12, atol=1e-26), precision=0, mode=2,
                   schedule=1, max_attempts=200000)
    tight.restore()
    tight.solve()
    torch.cuda.synchronize()
    y_t, s_t = tight.y.cpu().numpy(), tight.st.cpu().numpy()
    tight.ode.CloseOdevec()
    out = {"reference": "FP64 +REPRO build (bit-identical to the oracle), every cell",
           "tight": "FP64 +REPRO build at rtol 1e-12, atol 1e-26, %d seeded cells (%d finished)" % (
               idx.size, int((s_t == 0).sum())),
           "units": "largest over a cell's species of |y - y_ref| / (rtol |y_ref| + atol), rtol 1e-6, atol 1e-20"}
    for name, (y, st) in list(results.items()) + [("fp64_reproducible", (y_ref, s_ref))]:
        ok = (st == 0) & (s_ref == 0)
        okt = (s_t == 0) & (st[idx] == 0)
        out[name] = {"err_vs_reference": error_summary(cell_errors(y[:, ok], y_ref[:, ok], w["rtol"], w["atol"])),
                     "err_vs_tight": error_summary(cell_errors(y[:, idx][:, okt], y_t[:, okt], w["rtol"], w["atol"])),
                     "drift": drift(w["y"][:, st == 0], y[:, st == 0])}
    return out


def batching(steps, warmup):
    """Batching form: one persistent launch (hipGraph replay) over the
    2,097,152-cell 40/30/30 ROBER/OREGO/PRIMORDIAL batch (seed 4), with the
    bundle, precisions and block from the autotuner's cached result for this
    device (odevec_amd/tuned.json), else all-FP64 at 256 threads."""
    import torch
    from odevec_amd import autotune as A
    from odevec_amd import workloads as W
    segs = W.batch_mix(2097152, seed=4)
    nets = [w["network"] for w in segs]
    t = A.lookup(nets)
    if t is None:
        print("batching: no tuned entry for %s in odevec_amd/tuned.json; all-FP64 at 256 threads"
              % A.cache_key(nets), file=sys.stderr)
        t = {"bundle": "MIX3_FP64", "block": 256, "precision": {n: 0 for n in nets}, "per_cell_seconds": {},
             "order": nets}
    d = A.cost_order(A._device_segments(segs, [t["precision"][n] for n in nets]),
                     {n: -k for k, n in enumerate(t["order"])})
    sec = A.time_batch(t["bundle"], d, t["block"], use_graph=True, reps=max(steps, 1),
                       fp64_fallback=t.get("fp64_fallback", 0))
    ok = sum(int((x["status"] == 0).sum()) for x in d)
    # each network's cells against an rtol-1e-12 solve of a sample of them
    errs = {}
    for x in d:
        w = next(v for v in segs if v["network"] == x["network"])
        errs[x["network"]] = err_vs_tight(w, x["y"].cpu().numpy(), x["status"].cpu().numpy(), 2)
    return {"value": 2097152 / sec, "unit": "cell-solves/s", "ms_per_launch": sec * 1e3,
            "err_vs_tight": errs,
            "workload": "2,097,152 cells, 40/30/30 ROBER/OREGO/PRIMORDIAL (seed 4), t_stop logU over 6 decades "
                        "per network, one persistent launch per step, hipGraph replay",
            "bundle": t["bundle"], "block": t["block"], "fp64_fallback": t.get("fp64_fallback", 0), "precision": t["precision"], "cells_ok": ok,
            "tuned": "per_cell_seconds" in t and bool(t["per_cell_seconds"]), "device": A.device_key()}


def synth64(steps, warmup, cells=65536):
    """SYNTH64 (64 species, 300 reactions, FF ordering, CSC; seed 7) through
    the cooperative kernel (one wavefront per cell, species on lanes), mixed
    and FP64: n_tot logU[1, 1e6], T logU[100, 1e4] K, abundances
    logU[1e-6, 1] x n_tot, integrated from 0 to 1e6 s at rtol 1e-6."""
    from odevec_amd import _lib as L
    from odevec_amd import workloads as W
    w = W.synth64(cells, seed=7)
    out = {"workload": "SYNTH64 %d cells (seed 7), t_stop 1e6 s, rtol %g, cooperative kernel; mixed form with the "
                       "reference's RMS corrector test" % (cells, w["rtol"])}
    for name, prec in (("mixed", 1), ("fp64", 0)):
        r = Runner(w, precision=prec, mode=L.COOP)
        if prec == 1:
            # the reference's RMS corrector test: SYNTH64's FP32 factors
            # contract every component (the max norm adds 35% iterations)
            r.ode.set_mixed_convergence_test("rms")
        wall, ms = r.timed(steps, warmup)
        out[name] = {"value": steps * cells / wall, "unit": "cell-solves/s", "kernel_ms": ms,
                     "cells_ok": int((r.st == 0).sum()),
                     "err_vs_tight": err_vs_tight(w, r.y.cpu().numpy(), r.st.cpu().numpy(), L.COOP, sample=1024)}
        r.ode.CloseOdevec()
    return out


def _server(nets):
    from odevec_amd import autotune as A
    from odevec_amd import serve as S
    tuned = A.lookup([n for n in nets if not S.is_coop(n)]) or {}
    srv = S.Server(tuned.get("bundle", "MIX3"), slots=2, precision=tuned.get("precision", {}),
                   block=tuned.get("block", 256), fp64_fallback=tuned.get("fp64_fallback", 0))
    # every kernel the trace uses loaded and run once, untimed
    S.run_trace(srv, S.make_trace(8.0, 1.0, seed=99, networks=nets, hi=2e4))
    return srv


def serving(rate, seconds, networks=None):
    """Serving form at sustained load: Poisson arrivals (seed 5) of blocks of
    logU[1e3, 2.56e5] cells of a network drawn uniformly from `networks`
    (default: ROBER, OREGO, PRIMORDIAL, SYNTH64), one dt target per block
    drawn logU over the network's span, rtol 1e-4 or 1e-6; two launch slots
    and the cooperative kernel's slot; latency = arrival to results on the
    host."""
    from odevec_amd import serve as S
    nets = networks or list(S.SERVE_SPANS)
    srv = _server(nets)
    res = S.run_trace(srv, S.make_trace(rate, seconds, seed=5, networks=nets))
    srv.close()
    return dict(rate_per_s=rate, seconds=seconds, networks=nets, **res)


def serving_sweep(rates, seconds):
    """The serving workload of serving() at each Poisson rate in `rates`
    (same seed, so a higher rate packs the same kinds of requests closer),
    one server: p50/p99 latency and cell-solves/s per rate, up to and past
    the rate the device sustains (the trace is replayed until it drains)."""
    from odevec_amd import serve as S
    nets = list(S.SERVE_SPANS)
    srv = _server(nets)
    out = []
    for rate in rates:
        r = S.run_trace(srv, S.make_trace(rate, seconds, seed=5, networks=nets))
        out.append(dict(rate_per_s=rate, seconds=seconds, **r))
    srv.close()
    return out


def sessions(nsess=8, ncells=65536, nadv=5):
    """Serving with warm restart: closed-loop clients each advance a block of
    PRIMORDIAL cells through five equal fractions of its span, as sessions
    (state and integrator history resident on the device, each advance
    continues the integration) and, for comparison, as ordinary requests
    that cold-start every advance from the returned state."""
    from odevec_amd import serve as S
    out = {}
    for warm in (True, False):
        srv = S.Server("MIX3", slots=2)
        r = S.run_sessions(srv, nsess=nsess, ncells=ncells, nadv=nadv, warm=warm, seed=7)
        srv.close()
        r.pop("final")
        r.pop("final_ok")
        out["warm" if warm else "cold"] = r
    out["warm_over_cold"] = out["warm"]["cell_solves_per_s"] / out["cold"]["cell_solves_per_s"]
    out["workload"] = ("%d closed-loop clients x %d PRIMORDIAL cells (seeds 7..), %d advances each through "
                       "equal fractions of t_stop logU[1e8,1e14] s, FP64, rtol 1e-6" % (nsess, ncells, nadv))
    return out


def timeline(w):
    """One more headline solve with the per-cell timeline attached: when the
    cells end, relative to the first cell's start (ms), and the attempts of
    the slowest cells -- the launch's tail."""
    import torch
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    dev = torch.device("cuda:0")
    ode = Odevec(w["network"], n, precision=1, mode=2, schedule=SCHEDULE)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    ode.set_fp64_fallback(FALLBACK)
    times, hz = ode.enable_timeline(n)
    y = torch.tensor(w["y"], device=dev)
    ts = torch.tensor(np.asarray(w["t_stop"], dtype=np.float64), device=dev)
    dt0 = torch.tensor(np.broadcast_to(w["dt0"], (n,)).copy(), device=dev)
    cnt = torch.zeros(8 * n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ode.solve_cells_async(n, ts.data_ptr(), y.data_ptr(), dt0_ptr=dt0.data_ptr(), counters_ptr=cnt.data_ptr(),
                          stream=stream)
    torch.cuda.synchronize()
    t = times.cpu().numpy().reshape(2, n).astype(np.float64)
    att = cnt.cpu().numpy().reshape(8, n)[0]
    ode.CloseOdevec()
    t0 = t[0].min()
    end = (t[1] - t0) / hz * 1e3
    start = (t[0] - t0) / hz * 1e3
    slow = np.argsort(-att)[:5]
    alone = _alone_ms(w, int(slow[0]))
    return {"cells_end_ms": {"p50": float(np.percentile(end, 50)), "p99": float(np.percentile(end, 99)),
                             "p99.9": float(np.percentile(end, 99.9)), "max": float(end.max())},
            "last_cell_start_ms": float(start.max()),
            "attempts": {"median": float(np.median(att)), "p99.9": float(np.percentile(att, 99.9)),
                         "max": int(att.max())},
            "slowest_cells": [{"attempts": int(att[k]), "start_ms": float(start[k]), "end_ms": float(end[k])}
                              for k in slow],
            "slowest_cell_alone_ms": alone,
            "note": "mixed persistent, counters on; ticks of the device wall clock; slowest_cell_alone_ms: "
                    "that cell solved by itself (HIP events), the floor of the launch time"}


def _alone_ms(w, k, reps=2):
    """Kernel time of cell k of workload w solved alone (mixed persistent),
    HIP events on the launch stream, `reps` launches."""
    import torch
    from odevec_amd.api import Odevec
    dev = torch.device("cuda:0")
    ode = Odevec(w["network"], 1, precision=1, mode=2)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    ode.set_fp64_fallback(FALLBACK)
    y0 = torch.tensor(np.ascontiguousarray(w["y"][:, k:k + 1]), device=dev)
    ts = torch.tensor([float(w["t_stop"][k])], dtype=torch.float64, device=dev)
    dt0 = torch.tensor([float(np.broadcast_to(w["dt0"], w["t_stop"].shape)[k])], dtype=torch.float64, device=dev)
    y = torch.empty_like(y0)
    stream = torch.cuda.current_stream()
    ms = []
    for _ in range(reps):
        y.copy_(y0)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        ode.solve_cells_async(1, ts.data_ptr(), y.data_ptr(), dt0_ptr=dt0.data_ptr(), stream=stream.cuda_stream)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    ode.CloseOdevec()
    return ms


def end_to_end(w, steps, kernel_ms):
    """The headline solve through the host-buffer entry point
    (odevec_solve_cells: the state, end times and dt hints go host -> device
    over PCIe, the state and status come back), mixed persistent form:
    cell-solves/s of the whole call, beside the device-resident kernel time
    measured for the headline (`kernel_ms`)."""
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    ode = Odevec(w["network"], n, precision=1, mode=2, schedule=SCHEDULE)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    ode.set_fp64_fallback(FALLBACK)
    ts = np.ascontiguousarray(w["t_stop"])
    dt0 = np.ascontiguousarray(np.broadcast_to(w["dt0"], (n,)))
    walls = []
    for k in range(steps + 1):
        y = np.array(w["y"], order="C", copy=True)
        t = time.perf_counter()
        ode.solve_cells(y, ts, dt0=dt0, raise_on_fail=False)
        if k > 0:          # the first call sizes the handle's buffers
            walls.append(time.perf_counter() - t)
    ode.CloseOdevec()
    wall = float(np.mean(walls))
    return {"value": n / wall, "unit": "cell-solves/s", "ms_per_call": wall * 1e3,
            "device_resident_kernel_ms": kernel_ms, "copies_and_host_ms": wall * 1e3 - kernel_ms,
            "entry": "odevec_solve_cells (host arrays), mixed persistent"}


def cpu_baseline(w, seconds=12.0):
    """The oracle (a C restatement of the reference solver) on the host cores,
    per-cell mode, on a seeded sample of the same workload."""
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    net = O.Network.load(w["network"])
    from odevec_amd import workloads as W
    n = 512 * threads
    done, t_used = 0, 0.0
    rng = np.random.default_rng(123)
    while t_used < seconds and done < 64 * n:
        idx = rng.choice(w["y"].shape[1], n, replace=False)
        s = W.subset(w, idx)
        t = time.perf_counter()
        y, st, _, _ = net.solve_cells(s["y"], s["t_stop"], t0=s["t0"], dt0=s["dt0"], rtol=s["rtol"], atol=s["atol"],
                                      dt_min=s["dt_min"], threads=threads)
        t_used += time.perf_counter() - t
        done += n
    return {"value": done / t_used, "unit": "cell-solves/s", "cores": threads, "kind": "port",
            "sample": "%d random cells of the same 1,048,576-cell PRIMORDIAL workload, oracle (C restatement of the "
                      "reference solver), per-cell mode, %d OpenMP threads, %.1f s" % (done, threads, t_used)}


def read_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprof summary."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    global FALLBACK
    a = parse()
    FALLBACK = a.fallback
    if a.gpus != 1:
        print("bench.py runs on one GPU: there is no multi-GPU path (cells are independent: each GPU "
              "would take a contiguous range of cells, one process per GPU, with no exchange during "
              "the solve).", file=sys.stderr)
        sys.exit(2)
    from odevec_amd import workloads as W
    import torch
    # a created stream: its handle is never the null stream, which the C ABI
    # reads as "the handle's own stream"
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    if a.only == "batching":
        print(json.dumps({"batching": batching(a.steps, a.warmup)}))
        return
    if a.only == "synth64":
        print(json.dumps({"synth64": synth64(a.steps, a.warmup)}))
        return
    if a.only == "sessions":
        print(json.dumps({"serving_sessions": sessions()}))
        return
    if a.only == "serving":
        print(json.dumps({"serving_sweep": serving_sweep(SERVE_RATES, 8.0)}))
        return
    if a.only == "tail":
        print(json.dumps({"timeline": timeline(W.primordial(a.cells, seed=3))}))
        return
    w = W.primordial(a.cells, seed=3)
    mixed = Runner(w, precision=1, mode=2, fallback=FALLBACK, schedule=SCHEDULE)
    if a.only == "headline":
        # the headline solve alone (for profiling its kernels)
        wall_m, ms_m = mixed.timed(a.steps, a.warmup)
        print(json.dumps({"headline": {"value": a.steps * a.cells / wall_m, "unit": "cell-solves/s",
                                       "ms_per_step": wall_m / a.steps * 1e3, "kernel_ms": ms_m}}))
        return
    fp64 = Runner(w, precision=0, mode=0)
    wall_m, ms_m = mixed.timed(a.steps, a.warmup)
    y_m, st_mixed = mixed.y.cpu().numpy(), mixed.st.cpu().numpy()
    wall_d, ms_d = fp64.timed(a.steps, a.warmup)
    y_d, st_fp64 = fp64.y.cpu().numpy(), fp64.st.cpu().numpy()
    # the FP64 form in the persistent kernel with the same cost-ordered
    # queue: the mixed form's gain from precision alone
    fp64p = Runner(w, precision=0, mode=2, schedule=SCHEDULE)
    wall_p, ms_p = fp64p.timed(a.steps, a.warmup)
    fp64p.ode.CloseOdevec()
    # the mixed form without the FP64 fallback, for comparison (not the value)
    wall_n, ms_n = Runner(w, precision=1, mode=2).timed(min(a.steps, 2), 0)
    fl_m, st_m = mixed.count_flops()
    fl_d, st_d = fp64.count_flops()
    ok_m, ok_d = int((st_m == 0).sum()), int((st_d == 0).sum())
    value = a.steps * a.cells / wall_m
    for name, wall, ms in (("mixed", wall_m, ms_m), ("fp64", wall_d, ms_d)):
        if not 0.8 * wall / a.steps <= float(np.mean(ms)) * 1e-3 <= 1.01 * wall / a.steps:
            print("warning: %s event time %.3f ms/launch disagrees with wall time %.3f ms/step"
                  % (name, float(np.mean(ms)), wall / a.steps * 1e3), file=sys.stderr)
    value_d = a.steps * a.cells / wall_d
    t_kernel = float(np.mean(ms_m)) * 1e-3
    achieved = (fl_m[0] + fl_m[1] * PEAK_FP64 / PEAK_FP32) / t_kernel
    frac = fl_m[0] / (PEAK_FP64 * t_kernel) + fl_m[1] / (PEAK_FP32 * t_kernel)
    res = {
        "metric": "PRIMORDIAL cell-solves/s, mixed vs FP64 GPU at rtol=1e-6; serving p50/p99 ms",
        "value": value,
        "unit": "cell-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall_m / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (seeded PRIMORDIAL initial states, seed 3)",
        "config": {"workload": "PRIMORDIAL 16-species, CSC+CMK, %d cells, per-cell t_stop logU[1e8,1e14] s, rtol 1e-6, "
                               "mixed precision (FP32 P/LU/solves, FP64 residual/history/control), persistent "
                               "work-compacting kernel with the queue in cost order; cells reaching %d "
                               "convergence failures continue in FP64 from where they stopped, in the same launch "
                               "(1 in 32 wavefronts starts as FP64 lanes)" % (a.cells, FALLBACK),
                   "cells": a.cells, "rtol": 1e-6, "atol": 1e-20, "network": w["network"],
                   "fp64_fallback_convfail": FALLBACK, "schedule": SCHEDULE},
        "fp64_plain": {"value": value_d, "unit": "cell-solves/s", "ms_per_step": wall_d / a.steps * 1e3,
                       "kernel": "one lane per cell, FP64, no compaction"},
        "mixed_over_fp64": value / value_d,
        "cells_ok": {"mixed": ok_m, "fp64": ok_d},
        "fp64_persistent": {"value": a.steps * a.cells / wall_p, "unit": "cell-solves/s",
                            "ms_per_step": wall_p / a.steps * 1e3,
                            "kernel": "persistent, queue in cost order, FP64 (the mixed form's scheduling)"},
        "mixed_over_fp64_persistent": wall_p / wall_m,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": frac, "traffic": read_traffic(),
                     "note": "FP64/FP32 vector-ALU bound (no MFMA: per-cell systems are tiny sparse LU "
                             "solves; neither hbm nor mfma applies); achieved = FP64 + FP32*peak64/peak32 algorithmic "
                             "flops per solve (op-count model; mixed and FP64 lanes of the launch) / mean solve "
                             "time (HIP events around both launches); traffic: the mixed kernel's HBM bytes per "
                             "launch; no MFMA is used"},
        "flops_per_launch": {"fp64": float(fl_m[0]), "fp32": float(fl_m[1])},
        "kernel_ms": {"mixed": ms_m, "fp64": ms_d},
        "mixed_without_fallback": {"value": min(a.steps, 2) * a.cells / wall_n, "unit": "cell-solves/s",
                                   "kernel_ms": ms_n},
    }
    res["accuracy"] = accuracy(w, {"mixed": (y_m, st_mixed), "fp64_plain": (y_d, st_fp64)})
    res["err_vs_fp64"] = res["accuracy"]["mixed"]["err_vs_reference"]
    res["err_vs_tight"] = res["accuracy"]["mixed"]["err_vs_tight"]
    res["end_to_end"] = end_to_end(w, a.steps, float(np.mean(ms_m)))
    res["timeline"] = timeline(w)
    if not a.no_batch:
        res["batching"] = batching(a.steps, a.warmup)
        res["synth64"] = synth64(a.steps, a.warmup)
        # all four networks (SYNTH64's 64-species cells dominate the cost)
        # swept over Poisson rates up to and past saturation; p50/p99 of the
        # bench line at a sustained 4 requests/s; and the three small
        # networks at 40/s
        sw = serving_sweep(SERVE_RATES, 8.0)
        sv = next(p for p in sw if p["rate_per_s"] == SERVE_SUSTAINED)
        res["serving"] = {"workload": "Poisson arrivals (seed 5, 8 s per rate) of blocks of logU[1e3, 2.56e5] "
                                      "cells of ROBER, OREGO, PRIMORDIAL or SYNTH64 (uniform), one dt target per "
                                      "block logU over the network's span, rtol 1e-4 or 1e-6",
                          "sustained_rate_per_s": SERVE_SUSTAINED, "sweep": sw}
        res["serving_p50_ms"], res["serving_p99_ms"] = sv["p50_ms"], sv["p99_ms"]
        res["serving_3net"] = serving(40.0, 8.0, ["ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC"])
        res["serving_sessions"] = sessions()
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
