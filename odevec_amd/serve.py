"""Serving form: a continuous-batching request loop over the batch kernel.

A request is a block of cells of one network with its own end time (dt
target) and tolerances. The server

  * admits waiting requests while their device footprint fits the HBM budget
    (state, end times, parameters, status: the bytes the request holds on the
    device until its results are returned);
  * keeps up to `slots` launches in flight, each on its own HIP stream with
    its own batch handle (segment table, queue counters, rate workspace) and
    a share of the device's workgroups, so a request that arrives while
    others run starts on a free slot instead of waiting for them: their work
    overlaps on the device;
  * runs requests of networks too large for one lane (SYNTH64, 64 species)
    through the cooperative kernel on a slot of their own, one chunk of at
    most `coop_chunk` cells per launch: the cooperative kernel fills the
    device for as long as a launch lasts, so a whole 256k-cell request in
    one launch would keep the other slots' launches off the device for
    seconds; between chunks they get it;
  * forms each launch on a free slot from the waiting requests that fit
    (continuous batching: requests join at the next free slot, not when the
    previous batch drains), ordered by cost per cell;
  * uploads a request's arrays on admission and keeps them resident until
    its launch completes, then copies the results back and reports
    latency = completion - arrival.

Every cell keeps its own step control, so a request's result does not
depend on what it was batched with (bitwise, for one bundle).

Sessions (warm restart; the reference cold-starts every SolveODE call): a
client opens a session for a block of cells, whose state, parameters and
integrator history then stay resident on the device, and sends `advance`
requests, each moving every cell on to a new end time. An advance continues
each cell's integration where the previous one stopped (Nordsieck history,
order, step size, LU factors), so it does not pay the cold start's order-1
steps, fresh Jacobian and initial step-size search again, and nothing but
the end times travels to the device. One session's advances run in order;
different sessions' advances overlap like other requests. Mixed-precision
sessions keep the FP64 fallback: a cell that FP64 lanes took over goes on
in FP64 in later advances, so advances equal one solve there too.

    python -m odevec_amd.serve --rates 4,8,16,32 --seconds 10
"""
from __future__ import annotations

import argparse
import json
import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L
from . import workloads as W

SERVE_SPANS = {"ROBER/NONE/DENSE": (1e-1, 1e5), "OREGO/NONE/DENSE": (3.6e-4, 360.0),
               "PRIMORDIAL/CMK/CSC": (1e8, 1e14), "SYNTH64/FF/CSC": (1e4, 1e10)}
COST_ORDER = ["PRIMORDIAL/CMK/CSC", "OREGO/NONE/DENSE", "ROBER/NONE/DENSE"]
# networks served by the cooperative kernel (one launch per request, on the
# same slot stream as the slot's batch launch) rather than the bundle
COOP_NETWORKS = {"SYNTH64/FF/CSC"}
# cooperative-kernel chunks of one request in flight at once, each on a
# stream of its own. Measured on the bench's serving workload at 4
# requests/s: with 2, p50/p99 45/959 ms (SYNTH64's p99 1.05 s, the per-lane
# networks' up to 317 ms); with 1, 27/1,255 ms (1.36 s; up to 112 ms): a
# second chunk overlaps the first one's last cells but takes the device
# from the batch slots, so the server keeps 1
COOP_INFLIGHT = 1
# networks whose mixed form runs the reference's RMS corrector test (their
# FP32 factors contract every component: the max norm only adds iterations)
MIXED_RMS = {"SYNTH64/FF/CSC"}


def _mixed_test(h, network, precision):
    if precision == L.MIXED and network.split("+")[0] in MIXED_RMS:
        h.set_mixed_convergence_test("rms")


def is_coop(network: str) -> bool:
    """Served by the cooperative kernel (the +REPRO build of one too)."""
    return network.split("+")[0] in COOP_NETWORKS


# Attempts per cell and launch in the batch slots, 0 = no limit. With a
# quantum q > 0 a cell that has not finished after q attempts pauses (its
# state in a history record) and goes on in the slot's next launch, which
# also takes the requests that arrived meanwhile: new work joins running
# work at every launch, with results equal to the solve in one launch. On
# the bench's serving workload it raised latencies at every q tried (256 to
# 4096, one or two slots): the slowest PRIMORDIAL cells need up to ~40,000
# attempts, i.e. tens to hundreds of relaunches, each of which rescans every
# cell of the slot's requests, while without a quantum a slot's requests
# finish in one launch and new ones take the other slot. Hence off.
QUANTUM = 0
L_MAXSEG = 1024   # segments per batch launch


@dataclass
class Session:
    """A block of cells resident on the device with its warm-restart history."""
    sid: int
    network: str
    precision: int
    ncells: int
    rtol: float
    atol: float
    dt_min: float
    dev: dict = field(default_factory=dict)   # y, t, dt0, params, hist, status
    busy: bool = False                        # an advance is admitted or in flight
    bytes: int = 0


@dataclass
class Request:
    rid: int
    network: str
    y: Optional[np.ndarray]         # (neq, ncells), float64 (None for a session's advance)
    t_stop: np.ndarray              # (ncells,)
    rtol: float
    atol: float
    dt_min: float
    dt0: np.ndarray                 # (ncells,)
    params: Optional[np.ndarray] = None
    precision: int = L.FP64
    arrival: float = 0.0            # seconds (server clock)
    done: float = 0.0
    status: Optional[np.ndarray] = None
    result: Optional[np.ndarray] = None
    dev: dict = field(default_factory=dict)
    session: Optional[Session] = None
    hist_planes: int = 0            # > 0: solved in quantum launches, with a history of this many planes
    launches: int = 0               # launches the request took part in

    @property
    def ncells(self) -> int:
        return self.session.ncells if self.session is not None else self.y.shape[1]

    def footprint(self) -> int:
        """Device bytes held from admission to completion (a session's
        resident arrays are counted once, when it opens)."""
        n = self.ncells
        if self.session is not None:
            return 8 * n
        p = 0 if self.params is None else self.params.shape[0]
        return n * (8 * self.y.shape[0] + 8 + 8 + 8 * p + 4 + 8 * self.hist_planes)


class Server:
    def __init__(self, bundle: str = "MIX3", slots: int = 2, hbm_budget: int = 16 << 30, block: int = 256,
                 precision: Optional[dict] = None, fp64_fallback: int = 0, coop_chunk: int = 16384,
                 quantum: int = QUANTUM, coop_inflight: int = COOP_INFLIGHT):
        import torch
        from .api import Batch
        self.torch = torch
        self.bundle = bundle
        self.entries = set(L.bundle_entries(bundle))
        dev = torch.cuda.get_device_properties(0)
        kinfo = L.KernelInfo()
        L.check(L.lib().odevec_batch_kernel_info(bundle.encode(), block, L.C.byref(kinfo)))
        full = dev.multi_processor_count * max(kinfo.blocks_per_cu, 1)
        self.grid = max(1, full // slots)
        self.slots = [{"stream": torch.cuda.Stream(),
                       "batch": Batch(bundle, block=block, grid=self.grid, fp64_fallback=fp64_fallback),
                       "lane": {}, "reqs": None, "event": None, "running": [], "finishing": []}
                      for _ in range(slots)]
        self.quantum = int(quantum)
        # the cooperative kernel's slot: one request at a time, chunk by chunk
        # the cooperative slot keeps coop_inflight chunks of its request in
        # flight, on streams of their own, so that one chunk's last cells can
        # overlap the next chunk (uploads on the first stream)
        if int(coop_inflight) < 1:
            raise ValueError("coop_inflight must be >= 1")
        streams = [torch.cuda.Stream() for _ in range(int(coop_inflight))]
        self.coop_slot = {"stream": streams[0], "streams": streams, "coop": {}, "req": None, "next": 0,
                          "inflight": [], "uploaded": None}
        self.coop_chunk = int(coop_chunk)
        self.coop_waiting: deque = deque()
        self.sessions = {}
        self._next_sid = 0
        self.budget = hbm_budget
        self.in_use = 0
        self.waiting: deque = deque()      # arrived, not yet admitted
        self.completed: List[Request] = []
        self.precision = precision or {}
        self.fp64_fallback = int(fp64_fallback)
        self.launches = 0
        self.max_inflight = 0      # most launches seen in flight at once
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    # -- sessions
    def open_session(self, network: str, y: np.ndarray, params: Optional[np.ndarray] = None, rtol: float = 1e-6,
                     atol: float = 1e-20, dt_min: float = 1e-20, dt0=1e-10, t0=0.0,
                     precision: Optional[int] = None) -> Session:
        """Upload a block of cells (y: (neq, ncells)) that later `advance`
        requests integrate further; its arrays and history stay on the device
        until close_session."""
        torch = self.torch
        n = y.shape[1]
        planes = L.history_planes(network)
        prec = self.precision.get(network, L.FP64) if precision is None else precision
        dev = torch.device("cuda:0")
        d = {"y": torch.tensor(np.ascontiguousarray(y), dtype=torch.float64, device=dev),
             "t": torch.tensor(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)).copy(), device=dev),
             "dt0": torch.tensor(np.broadcast_to(np.asarray(dt0, dtype=np.float64), (n,)).copy(), device=dev),
             "hist": torch.zeros(planes * n, dtype=torch.float64, device=dev),
             "status": torch.zeros(n, dtype=torch.int32, device=dev)}
        if params is not None:
            d["params"] = torch.tensor(np.ascontiguousarray(params), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ss = Session(self._next_sid, network, prec, n, rtol, atol, dt_min, d)
        ss.bytes = sum(v.numel() * v.element_size() for v in d.values())
        if self.in_use + ss.bytes > self.budget:
            raise ValueError("session of %d cells needs %d bytes, more than the HBM budget left" % (n, ss.bytes))
        self.in_use += ss.bytes
        self.sessions[ss.sid] = ss
        self._next_sid += 1
        return ss

    def advance(self, ss: Session, t_stop, rid: int = -1, arrival: Optional[float] = None) -> Request:
        """Queue an advance of every cell of the session to t_stop (scalar or
        per cell); the request's result is the state there."""
        ts = np.array(np.broadcast_to(np.asarray(t_stop, dtype=np.float64), (ss.ncells,)))
        r = Request(rid, ss.network, None, ts, ss.rtol, ss.atol, ss.dt_min, None, precision=ss.precision,
                    arrival=self.now() if arrival is None else arrival, session=ss)
        self.waiting.append(r)
        return r

    def close_session(self, ss: Session):
        while ss.busy or any(r.session is ss for r in self.waiting):
            self.poll()
            time.sleep(0.0002)
        self.in_use -= ss.bytes
        ss.dev = {}
        del self.sessions[ss.sid]

    def submit(self, r: Request):
        if is_coop(r.network):
            if r.precision is None:
                r.precision = self.precision.get(r.network, L.MIXED)
        else:
            if r.precision is None or (r.network, r.precision) not in self.entries:
                r.precision = self.precision.get(r.network, L.FP64)
            if (r.network, r.precision) not in self.entries:
                raise ValueError("network %s precision %d not in bundle %s" % (r.network, r.precision, self.bundle))
            r.hist_planes = L.history_planes(r.network) if self.quantum > 0 and r.session is None else 0
        if r.footprint() > self.budget:
            raise ValueError("request %d needs %d bytes, more than the HBM budget" % (r.rid, r.footprint()))
        (self.coop_waiting if is_coop(r.network) else self.waiting).append(r)

    def _admit(self, stream, room: int = L_MAXSEG) -> List[Request]:
        """Upload waiting requests that fit the budget (arrival order, at most
        `room` segments) on `stream`; returns them."""
        torch = self.torch
        new = []
        skipped = deque()
        while (self.waiting and len(new) < room and
               self.in_use + self.waiting[0].footprint() <= self.budget):
            r = self.waiting.popleft()
            dev = torch.device("cuda:0")
            if r.session is not None:
                if r.session.busy:           # one advance of a session at a time, in order
                    skipped.append(r)
                    continue
                r.session.busy = True
                with torch.cuda.stream(stream):
                    r.dev = {"t_stop": torch.from_numpy(r.t_stop).pin_memory().to(dev, non_blocking=True)}
                self.in_use += r.footprint()
                new.append(r)
                continue
            with torch.cuda.stream(stream):
                d = {"y": torch.from_numpy(r.y).pin_memory().to(dev, non_blocking=True),
                     "t_stop": torch.from_numpy(np.ascontiguousarray(r.t_stop)).pin_memory().to(dev, non_blocking=True),
                     "dt0": torch.from_numpy(np.ascontiguousarray(r.dt0)).pin_memory().to(dev, non_blocking=True),
                     "status": torch.empty(r.ncells, dtype=torch.int32, device=dev)}
                if r.params is not None:
                    d["params"] = torch.from_numpy(np.ascontiguousarray(r.params)).pin_memory().to(dev, non_blocking=True)
                if r.hist_planes:
                    # quantum launches: the cells' phase (all fresh) and room for
                    # their pause records (read only where a pause wrote one)
                    d["status"].fill_(L.FRESH)
                    d["hist"] = torch.empty(r.hist_planes * r.ncells, dtype=torch.float64, device=dev)
            r.dev = d
            self.in_use += r.footprint()
            new.append(r)
        self.waiting.extendleft(reversed(skipped))
        return new

    def _coop(self, network, precision):
        from .api import Odevec
        key = (network, precision)
        h = self.coop_slot["coop"].get(key)
        if h is None:
            h = Odevec(network, 1, precision=precision, mode=L.COOP)
            _mixed_test(h, network, precision)
            self.coop_slot["coop"][key] = h
        return h

    def _coop_admit(self) -> bool:
        """Upload the next cooperative-kernel request, one array set per
        chunk of at most coop_chunk cells; False if it does not fit yet."""
        torch = self.torch
        r = self.coop_waiting[0]
        if self.in_use + r.footprint() > self.budget:
            return False
        self.coop_waiting.popleft()
        dev = torch.device("cuda:0")
        chunks = []
        with torch.cuda.stream(self.coop_slot["stream"]):
            for c0 in range(0, r.ncells, self.coop_chunk):
                c1 = min(r.ncells, c0 + self.coop_chunk)
                up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)
                d = {"n": c1 - c0, "y": up(r.y[:, c0:c1]), "t_stop": up(r.t_stop[c0:c1]), "dt0": up(r.dt0[c0:c1]),
                     "status": torch.empty(c1 - c0, dtype=torch.int32, device=dev)}
                if r.params is not None:
                    d["params"] = up(r.params[:, c0:c1])
                chunks.append(d)
            up_ev = torch.cuda.Event()
            up_ev.record(self.coop_slot["stream"])
        r.dev = {"chunks": chunks}
        self.in_use += r.footprint()
        self.coop_slot["req"], self.coop_slot["next"], self.coop_slot["uploaded"] = r, 0, up_ev
        return True

    def _coop_launch(self):
        """Enqueue the current cooperative request's next chunk and the copy
        of its results to pinned host memory."""
        torch = self.torch
        cs = self.coop_slot
        r = cs["req"]
        d = r.dev["chunks"][cs["next"]]
        st = cs["streams"][cs["next"] % len(cs["streams"])]
        if st is not cs["stream"]:
            st.wait_event(cs["uploaded"])
        h = self._coop(r.network, r.precision)
        h.RelativeTolerance, h.AbsoluteTolerance, h.MinimalTimestep = r.rtol, r.atol, r.dt_min
        h.solve_cells_async(d["n"], d["t_stop"].data_ptr(), d["y"].data_ptr(), dt0_ptr=d["dt0"].data_ptr(),
                            params_ptr=d["params"].data_ptr() if "params" in d else None,
                            status_ptr=d["status"].data_ptr(), stream=st.cuda_stream)
        with torch.cuda.stream(st):
            d["y_host"] = torch.empty(tuple(d["y"].shape), dtype=torch.float64, pin_memory=True)
            d["st_host"] = torch.empty(d["n"], dtype=torch.int32, pin_memory=True)
            d["y_host"].copy_(d["y"], non_blocking=True)
            d["st_host"].copy_(d["status"], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        cs["inflight"].append(ev)
        cs["next"] += 1
        self.launches += 1

    def _coop_fill(self):
        cs = self.coop_slot
        while len(cs["inflight"]) < len(cs["streams"]) and cs["next"] < len(cs["req"].dev["chunks"]):
            self._coop_launch()

    def _coop_poll(self, done: List[Request]):
        cs = self.coop_slot
        if cs["req"] is not None:
            cs["inflight"] = [ev for ev in cs["inflight"] if not ev.query()]
            r = cs["req"]
            if cs["next"] < len(r.dev["chunks"]):
                self._coop_fill()
                return
            if cs["inflight"]:
                return
            r.result = np.concatenate([d["y_host"].numpy() for d in r.dev["chunks"]], axis=1)
            r.status = np.concatenate([d["st_host"].numpy() for d in r.dev["chunks"]])
            r.done = self.now()
            self.in_use -= r.footprint()
            r.dev = {}
            cs["req"], cs["uploaded"] = None, None
            done.append(r)
        if cs["req"] is None and self.coop_waiting and self._coop_admit():
            self._coop_fill()

    def _lane(self, slot, network, precision):
        """The slot's persistent-mode handle for (network, precision) (the
        cooperative mode's for a network too large for one lane): runs
        session advances, each with its session's history attached."""
        from .api import Odevec
        key = (network, precision)
        h = slot["lane"].get(key)
        if h is None and is_coop(network):
            h = Odevec(network, 1, precision=precision, mode=L.COOP)
            _mixed_test(h, network, precision)
            slot["lane"][key] = h
        if h is None:
            # the slot's share of the device's resident workgroups of this kernel
            k = L.kernel_info(network, L.PERSISTENT, precision)
            ncu = self.torch.cuda.get_device_properties(0).multi_processor_count
            grid = max(1, ncu * max(k["blocks_per_cu"], 1) // len(self.slots))
            h = Odevec(network, 1, precision=precision, mode=L.PERSISTENT, grid=grid)
            if self.fp64_fallback and precision == L.MIXED:
                h.set_fp64_fallback(self.fp64_fallback)
            slot["lane"][key] = h
        return h

    def _advance(self, slot, r: Request):
        ss, d = r.session, r.session.dev
        h = self._lane(slot, ss.network, ss.precision)
        h.RelativeTolerance, h.AbsoluteTolerance, h.MinimalTimestep = ss.rtol, ss.atol, ss.dt_min
        L.check(L.lib().odevec_set_history(h._h, L.C.c_void_p(d["hist"].data_ptr()), ss.ncells))
        st = slot["stream"]
        h.solve_cells_async(ss.ncells, r.dev["t_stop"].data_ptr(), d["y"].data_ptr(), t0_ptr=d["t"].data_ptr(),
                            dt0_ptr=d["dt0"].data_ptr(),
                            params_ptr=d["params"].data_ptr() if "params" in d else None,
                            status_ptr=d["status"].data_ptr(), stream=st.cuda_stream)
        L.check(L.lib().odevec_set_history(h._h, None, 0))
        with self.torch.cuda.stream(st):
            d["t"].copy_(r.dev["t_stop"], non_blocking=True)   # the next advance starts here
        r.dev["y"], r.dev["status"] = d["y"], d["status"]

    def _launch(self, slot, reqs: List[Request]):
        reqs = sorted(reqs, key=lambda r: COST_ORDER.index(r.network) if r.network in COST_ORDER else -1)
        st = slot["stream"]
        segs = []
        advanced = []
        for r in reqs:
            if r.session is not None:
                ss = r.session
                if (ss.network, ss.precision) not in self.entries:
                    self._advance(slot, r)      # not in the bundle: its own launch
                    continue
                d = ss.dev
                s = {"network": ss.network, "precision": ss.precision, "y": d["y"], "t0": d["t"],
                     "t_stop": r.dev["t_stop"], "dt0": d["dt0"], "status": d["status"], "hist": d["hist"],
                     "rtol": ss.rtol, "atol": ss.atol, "dt_min": ss.dt_min}
                if "params" in d:
                    s["params"] = d["params"]
                segs.append(s)
                r.dev["y"], r.dev["status"] = d["y"], d["status"]
                advanced.append(r)
                continue
            s = {"network": r.network, "precision": r.precision, "y": r.dev["y"], "t_stop": r.dev["t_stop"],
                 "dt0": r.dev["dt0"], "status": r.dev["status"], "rtol": r.rtol, "atol": r.atol, "dt_min": r.dt_min}
            if r.params is not None:
                s["params"] = r.dev["params"]
            if r.hist_planes:
                s["hist"], s["quantum"] = r.dev["hist"], self.quantum
            segs.append(s)
            r.launches += 1
        torch = self.torch
        if segs:
            slot["batch"].solve_async(segs, st.cuda_stream)
        with torch.cuda.stream(st):
            for r in advanced:     # the session's next advance starts where this one ends
                r.session.dev["t"].copy_(r.dev["t_stop"], non_blocking=True)
            for r in reqs:
                if "st_host" not in r.dev:
                    r.dev["st_host"] = torch.empty(r.ncells, dtype=torch.int32, pin_memory=True)
                r.dev["st_host"].copy_(r.dev["status"], non_blocking=True)
                if not r.hist_planes:
                    self._copy_y(r)
            ev = torch.cuda.Event()
            ev.record(st)
        slot["reqs"], slot["event"] = reqs, ev

    def _copy_y(self, r: Request):
        """Enqueue the copy of a request's results to pinned host memory (on
        the current stream)."""
        r.dev["y_host"] = self.torch.empty(tuple(r.dev["y"].shape), dtype=self.torch.float64, pin_memory=True)
        r.dev["y_host"].copy_(r.dev["y"], non_blocking=True)

    def _complete(self, r: Request, t: float, done: List[Request]):
        r.result = r.dev["y_host"].numpy()
        r.status = r.dev["st_host"].numpy()
        r.done = t
        self.in_use -= r.footprint()
        r.dev = {}
        if r.session is not None:
            r.session.busy = False
        done.append(r)

    def poll(self) -> List[Request]:
        """Complete finished launches, admit waiting requests, launch on free
        slots. Returns the requests completed in this call."""
        done = []
        torch = self.torch
        for slot in self.slots:
            # requests whose results are on their way to the host
            still = []
            for r, ev in slot["finishing"]:
                if ev.query():
                    self._complete(r, self.now(), done)
                else:
                    still.append((r, ev))
            slot["finishing"] = still
            if slot["reqs"] is not None and slot["event"].query():
                t = self.now()
                for r in slot["reqs"]:
                    if not r.hist_planes:
                        self._complete(r, t, done)
                    elif (r.dev["st_host"].numpy() == L.PAUSED).any():
                        slot["running"].append(r)      # continues in the slot's next launch
                    else:
                        with torch.cuda.stream(slot["stream"]):
                            self._copy_y(r)
                            ev = torch.cuda.Event()
                            ev.record(slot["stream"])
                        slot["finishing"].append((r, ev))
                slot["reqs"], slot["event"] = None, None
        for slot in self.slots:
            if slot["reqs"] is None and (self.waiting or slot["running"]):
                # the slot's unfinished requests and the waiting ones that fit:
                # new requests join the running work at every launch
                reqs = slot["running"] + self._admit(slot["stream"], L_MAXSEG - len(slot["running"]))
                slot["running"] = []
                if reqs:
                    self._launch(slot, reqs)
                    self.launches += 1
        self._coop_poll(done)
        self.max_inflight = max(self.max_inflight, sum(s["reqs"] is not None for s in self.slots) +
                                (self.coop_slot["req"] is not None))
        self.completed.extend(done)
        return done

    def busy(self) -> bool:
        return bool(self.waiting or self.coop_waiting or self.coop_slot["req"] is not None or
                    any(s["reqs"] is not None or s["running"] or s["finishing"] for s in self.slots))

    def drain(self):
        while self.busy():
            self.poll()
            time.sleep(0.0002)

    def close(self):
        self.drain()
        for s in self.slots:
            s["batch"].close()
            for h in s["lane"].values():
                h.CloseOdevec()
        for h in self.coop_slot["coop"].values():
            h.CloseOdevec()



def make_trace(rate: float, seconds: float, seed: int = 5, networks=None, lo: float = 1e3, hi: float = 2.56e5,
               rtols=(1e-4, 1e-6)) -> List[Request]:
    """Poisson arrivals (rate per second) over `seconds`; each request a block
    of logU[lo, hi] cells of a network drawn uniformly, one dt target drawn
    logU over the network's span, rtol drawn from `rtols`."""
    rng = np.random.default_rng(seed)
    networks = networks or list(SERVE_SPANS)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        net = networks[rng.integers(len(networks))]
        base = net.split("+")[0]            # the +REPRO build of a network: the same inputs
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        sd = int(rng.integers(1 << 30))
        w = {"ROBER/NONE/DENSE": W.rober, "OREGO/NONE/DENSE": W.orego, "PRIMORDIAL/CMK/CSC": W.primordial,
             "SYNTH64/FF/CSC": W.synth64}[base](n, seed=sd)
        lo_t, hi_t = SERVE_SPANS[base]
        ts = float(np.exp(rng.uniform(np.log(lo_t), np.log(hi_t))))
        rtol = float(rtols[rng.integers(len(rtols))])
        out.append(Request(rid, net, np.ascontiguousarray(w["y"]), np.full(n, ts), rtol,
                           w["atol"], w["dt_min"], np.ascontiguousarray(w["dt0"]),
                           None if w.get("params") is None else np.ascontiguousarray(w["params"]), None, t))
        rid += 1
    return out


def run_trace(server: Server, trace: List[Request]) -> dict:
    """Replay arrivals in real time; returns latency percentiles and throughput."""
    server.t0 = time.perf_counter()
    pending = deque(sorted(trace, key=lambda r: r.arrival))
    while pending or server.busy():
        now = server.now()
        while pending and pending[0].arrival <= now:
            server.submit(pending.popleft())
        server.poll()
        time.sleep(0.0002)
    wall = max(r.done for r in trace) if trace else 0.0
    lat = np.array([r.done - r.arrival for r in trace]) * 1e3
    cells = sum(r.ncells for r in trace)
    ok = sum(int((r.status == 0).sum()) for r in trace)
    per_net = {}
    for net in sorted({r.network for r in trace}):
        ln = np.array([(r.done - r.arrival) * 1e3 for r in trace if r.network == net])
        per_net[net] = {"requests": int(ln.size), "p50_ms": float(np.percentile(ln, 50)),
                        "p99_ms": float(np.percentile(ln, 99))}
    return {"requests": len(trace), "cells": cells, "cells_ok": ok, "wall_s": wall,
            "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_ms": float(lat.mean()) if len(lat) else None,
            "cell_solves_per_s": cells / wall if wall > 0 else None, "per_network": per_net}


def run_sessions(server: Server, nsess: int = 8, ncells: int = 65536, nadv: int = 5, warm: bool = True,
                 seed: int = 7, network: str = "PRIMORDIAL/CMK/CSC", precision: int = L.FP64) -> dict:
    """Closed-loop clients: each of `nsess` blocks of `ncells` seeded cells
    is advanced `nadv` times through equal fractions of its span (each
    client sends its next advance when the previous result is back). warm:
    sessions (state and history stay on the device, each advance continues
    the integration); cold: each advance is an ordinary request carrying the
    state returned by the previous one, integrated from a cold start over
    the same interval (the networks are autonomous). Returns latency
    percentiles and cell-solves/s over all advances."""
    gen = {"ROBER/NONE/DENSE": W.rober, "OREGO/NONE/DENSE": W.orego, "PRIMORDIAL/CMK/CSC": W.primordial}[network]
    blocks = [gen(ncells, seed=seed + k) for k in range(nsess)]
    spans = [np.broadcast_to(np.asarray(w["t_stop"], dtype=np.float64), (ncells,)).copy() for w in blocks]
    server.t0 = time.perf_counter()
    sess = [server.open_session(network, w["y"], w.get("params"), w["rtol"], w["atol"], w["dt_min"], w["dt0"],
                                precision=precision) if warm else None for w in blocks]
    state = [np.ascontiguousarray(w["y"]) for w in blocks]
    cell_ok = [np.ones(ncells, dtype=bool) for _ in blocks]   # OK in every advance so far
    step = [0] * nsess
    live = {}
    lat = []
    rid = 0

    def send(k):
        nonlocal rid
        frac0, frac1 = step[k] / nadv, (step[k] + 1) / nadv
        if warm:
            r = server.advance(sess[k], spans[k] * frac1, rid=rid)
        else:
            w = blocks[k]
            r = Request(rid, network, state[k], spans[k] * (frac1 - frac0), w["rtol"], w["atol"], w["dt_min"],
                        np.array(np.broadcast_to(w["dt0"], (ncells,))),
                        None if w.get("params") is None else np.ascontiguousarray(w["params"]), precision,
                        server.now())
            server.submit(r)
        live[r.rid] = k
        rid += 1

    t_begin = server.now()
    for k in range(nsess):
        send(k)
    ok = 0
    while live:
        for r in server.poll():
            k = live.pop(r.rid, None)
            if k is None:
                continue
            lat.append((r.done - r.arrival) * 1e3)
            ok += int((r.status == 0).sum())
            cell_ok[k] &= r.status == 0
            state[k] = np.ascontiguousarray(r.result)
            step[k] += 1
            if step[k] < nadv:
                send(k)
        time.sleep(0.0002)
    wall = server.now() - t_begin
    for ss in sess:
        if ss is not None:
            server.close_session(ss)
    lat = np.array(lat)
    cells = nsess * ncells * nadv
    return {"mode": "warm" if warm else "cold", "sessions": nsess, "cells": ncells, "advances": nadv,
            "cells_ok": ok, "wall_s": wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "cell_solves_per_s": cells / wall,
            "final": state, "final_ok": cell_ok}


def main(argv=None) -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--rates", default="4,8,16,32")
    p.add_argument("--seconds", type=float, default=10.0)
    p.add_argument("--slots", type=int, default=2)
    p.add_argument("--bundle", default="MIX3")
    p.add_argument("--quantum", type=int, default=QUANTUM, help="attempts per cell and batch launch (0: no limit)")
    p.add_argument("--coop_chunk", type=int, default=16384, help="cells per cooperative-kernel launch")
    p.add_argument("--coop_inflight", type=int, default=COOP_INFLIGHT, help="cooperative chunks in flight at once")
    a = p.parse_args(argv)
    from . import autotune as A
    tuned = A.lookup(list(SERVE_SPANS)) or {}
    prec = tuned.get("precision", {})
    for rate in [float(x) for x in a.rates.split(",")]:
        srv = Server(a.bundle, slots=a.slots, precision=prec, quantum=a.quantum, coop_chunk=a.coop_chunk,
                     coop_inflight=a.coop_inflight)
        res = run_trace(srv, make_trace(rate, a.seconds))
        srv.close()
        print(json.dumps(dict(rate_per_s=rate, **res)), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
