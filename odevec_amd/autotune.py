"""Autotuner of the batch launch: the bundle (register/LDS split), the
FP32/FP64 split per network and the threads per workgroup.

For a batch (a list of segments: dicts with host arrays as
``odevec_amd.workloads`` makes them) it times every candidate on a seeded
sample of each segment's cells (the same fraction of each, so the mix is
kept) through a bundle's batch kernel, with inputs resident on the GPU, and
returns the fastest. Results are cached in a JSON file keyed by device and
the batch's networks.

Candidates:
  * bundle (the register/LDS split): a bundle is one batch kernel compiled
    for a set of (network, precision) entries; its register allocation and
    spills cover all of them, and each precision places the lane state
    differently (mixed: Nordsieck history in registers, rates in LDS; FP64:
    history columns 2..5 in LDS, rates in the workspace). Every default-build
    bundle that covers the batch's networks is tried.
  * block (threads = cells per workgroup): 64, 128, 256. The LDS columns of
    the 16-species networks are laid out for 256 lanes per workgroup, so
    512 and 1024 are not offered (a 512-lane workgroup of the FP64
    PRIMORDIAL kernel would need 294 KiB of LDS; a CU has 160 KiB).
  * precision per network: FP64 or mixed (FP32 Jacobian/LU/solves), every
    combination the bundle has.
  * segment order: by decreasing measured cost per cell (always), so the
    cheap cells fill the end of the launch.

    python -m odevec_amd.autotune [--cells 2097152] [--sample 0.25]
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time
from typing import Dict, List

import numpy as np

from . import _lib as L
from . import workloads as W

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
BLOCKS = (64, 128, 256)


def _device_segments(segs, precisions):
    import torch
    dev = torch.device("cuda:0")
    out = []
    for w, prec in zip(segs, precisions):
        n = w["y"].shape[1]
        d = {"network": w["network"], "precision": int(prec), "ncells": n,
             "y0": torch.tensor(w["y"], dtype=torch.float64, device=dev),
             "t0": torch.tensor(np.broadcast_to(w["t0"], (n,)).copy(), dtype=torch.float64, device=dev),
             "t_stop": torch.tensor(np.broadcast_to(w["t_stop"], (n,)).copy(), dtype=torch.float64, device=dev),
             "dt0": torch.tensor(np.broadcast_to(w["dt0"], (n,)).copy(), dtype=torch.float64, device=dev),
             "status": torch.zeros(n, dtype=torch.int32, device=dev),
             "rtol": w["rtol"], "atol": w["atol"], "dt_min": w["dt_min"]}
        if w.get("params") is not None:
            d["params"] = torch.tensor(w["params"], dtype=torch.float64, device=dev)
        d["y"] = torch.empty_like(d["y0"])
        out.append(d)
    return out


# mixed segments are timed (and then run) with the FP64 fallback at this many
# convergence failures: without it a few cells set a mixed launch's length
FP64_FALLBACK = 128


def time_batch(bundle: str, dsegs, block: int, use_graph: bool = False, reps: int = 2, stream=None,
               fp64_fallback: int = FP64_FALLBACK) -> float:
    """Median seconds of one batch launch (with its FP64 passes) over device
    segments (state restored by a device copy before each launch, outside
    the timing)."""
    import torch
    from .api import Batch
    stream = stream or torch.cuda.Stream()
    b = Batch(bundle, block=block, use_graph=use_graph, fp64_fallback=fp64_fallback)
    ts = []
    with torch.cuda.stream(stream):
        for r in range(reps + 1):
            for d in dsegs:
                d["y"].copy_(d["y0"])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            b.solve_async(dsegs, stream.cuda_stream)
            e1.record()
            torch.cuda.synchronize()
            if r > 0:            # the first launch also allocates and (graph mode) captures
                ts.append(e0.elapsed_time(e1) * 1e-3)
    b.close()
    return float(np.median(ts))


def cost_order(dsegs, times_per_cell: Dict[str, float]):
    """Segments in decreasing cost per cell."""
    return sorted(dsegs, key=lambda d: -times_per_cell.get(d["network"], 0.0))


def candidate_bundles(networks: List[str]) -> List[str]:
    """Default-build bundles whose entries cover every network of a batch."""
    out = []
    for b in L.bundles():
        if b.endswith("+REPRO"):
            continue
        have = {n for n, _ in L.bundle_entries(b)}
        if all(n in have for n in networks):
            out.append(b)
    return out


def device_key() -> str:
    """The device a tuning result belongs to: its ISA and CU count (the
    marketing name varies with the driver's name tables)."""
    import torch
    p = torch.cuda.get_device_properties(0)
    return "%s/%dCU" % (str(getattr(p, "gcnArchName", "gpu")).split(":")[0], p.multi_processor_count)


def cache_key(networks: List[str]) -> str:
    return "%s|%s" % (device_key(), ",".join(sorted(set(networks))))


def tune(segs: List[dict], sample: float = 0.25, seed: int = 0, verbose: bool = True, cache: str = CACHE,
         bundles: List[str] = None) -> dict:
    """Fastest (bundle, precision per network, block) for the batch `segs`.
    The bundle is the register/LDS split: which precisions (and so which
    state placements) share one kernel's register allocation."""
    import torch
    rng = np.random.default_rng(seed)
    sub = []
    for w in segs:
        n = w["y"].shape[1]
        m = max(1, int(round(n * sample))) if n else 0
        sub.append(W.subset(w, rng.choice(n, m, replace=False)) if m else w)
    nets = [w["network"] for w in sub]
    bundles = bundles or candidate_bundles(nets)
    stream = torch.cuda.Stream()
    # per-network cost per cell (first bundle, FP64 where present) fixes the segment order
    per_cell = {}
    ref_bundle = bundles[0]
    ents = L.bundle_entries(ref_bundle)
    for w in sub:
        if w["y"].shape[1] == 0:
            continue
        precs = [p for n, p in ents if n == w["network"]]
        d = _device_segments([w], [min(precs)])
        per_cell[w["network"]] = time_batch(ref_bundle, d, 256, stream=stream) / w["y"].shape[1]
    results = []
    for bundle in bundles:
        ents = L.bundle_entries(bundle)
        choices = [sorted({p for n, p in ents if n == net}) for net in nets]
        for precs in itertools.product(*choices):
            dsegs = cost_order(_device_segments(sub, precs), per_cell)
            for block in BLOCKS:
                t = time_batch(bundle, dsegs, block, stream=stream)
                results.append({"bundle": bundle, "block": block, "seconds": t,
                                "precision": {n: int(p) for n, p in zip(nets, precs)},
                                "fp64_fallback": FP64_FALLBACK})
                if verbose:
                    print("autotune %-11s block %3d precision %s: %.2f ms" % (bundle, block, precs, t * 1e3),
                          flush=True)
    best = min(results, key=lambda r: r["seconds"])
    best = dict(best, order=sorted(per_cell, key=lambda n: -per_cell[n]), per_cell_seconds=per_cell,
                sample=sample, device=device_key(), device_name=torch.cuda.get_device_name(0), all=results)
    if cache:
        key = cache_key(nets)
        db = {}
        if os.path.exists(cache):
            try:
                db = json.load(open(cache))
            except (OSError, ValueError):
                db = {}
        db[key] = best
        with open(cache, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return best


def lookup(networks: List[str], cache: str = CACHE):
    """A cached tuning result for this device and set of networks, or None."""
    if not os.path.exists(cache):
        return None
    try:
        db = json.load(open(cache))
    except (OSError, ValueError):
        return None
    return db.get(cache_key(networks))


def main(argv=None) -> int:
    p = argparse.ArgumentParser()
    p.add_argument("--cells", type=int, default=2097152)
    p.add_argument("--sample", type=float, default=0.25)
    p.add_argument("--cache", default=CACHE)
    a = p.parse_args(argv)
    segs = W.batch_mix(a.cells, seed=4)
    best = tune(segs, a.sample, cache=a.cache)
    print(json.dumps({k: v for k, v in best.items() if k != "all"}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
