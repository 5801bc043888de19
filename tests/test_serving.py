"""Serving form: the continuous-batching request loop (odevec_amd/serve.py).

CPU: the arrival trace and the admission bookkeeping. GPU: requests that
overlap on the device give the same results (bitwise) as each request run
alone, and every request completes with latency recorded.
"""
import numpy as np
import pytest

from odevec_amd import serve as S


def test_trace_is_seeded_poisson_with_block_sizes_in_range():
    a = S.make_trace(20.0, 5.0, seed=5)
    b = S.make_trace(20.0, 5.0, seed=5)
    assert [r.rid for r in a] == list(range(len(a))) and len(a) == len(b)
    assert all(np.array_equal(x.y, y.y) and x.arrival == y.arrival for x, y in zip(a, b))
    arr = np.array([r.arrival for r in a])
    assert (np.diff(arr) > 0).all() and arr[-1] <= 5.0
    assert 40 <= len(a) <= 170                      # mean 100 arrivals
    assert all(1e3 * 0.999 <= r.ncells <= 2.56e5 for r in a)
    assert {r.network for r in a} == set(S.SERVE_SPANS)
    assert {r.rtol for r in a} == {1e-4, 1e-6}
    for r in a:
        lo, hi = S.SERVE_SPANS[r.network]
        assert lo <= r.t_stop[0] <= hi and (r.t_stop == r.t_stop[0]).all()


def test_request_footprint_counts_resident_arrays():
    r = S.make_trace(50.0, 1.0, seed=3)[0]
    n, neq = r.ncells, r.y.shape[0]
    p = 0 if r.params is None else r.params.shape[0]
    assert r.footprint() == n * (8 * neq + 8 + 8 + 8 * p + 4)


@pytest.mark.gpu
def test_serving_results_same_alone_and_in_mix():
    print("GPU serving mix invariance", flush=True)
    trace = S.make_trace(400.0, 0.05, seed=11, lo=1e3, hi=2e4)
    assert len(trace) >= 8
    srv = S.Server("MIX3", slots=2, precision={"PRIMORDIAL/CMK/CSC": 1})
    res = S.run_trace(srv, trace)
    srv.close()
    assert res["requests"] == len(trace) and res["p99_ms"] is not None
    assert srv.max_inflight == 2          # launches overlapped on the device
    for r in trace:
        alone = S.make_trace(400.0, 0.05, seed=11, lo=1e3, hi=2e4)[r.rid]
        one = S.Server("MIX3", slots=1, precision={"PRIMORDIAL/CMK/CSC": 1})
        one.submit(alone)
        one.drain()
        one.close()
        assert np.array_equal(alone.result, r.result), r.rid
        assert np.array_equal(alone.status, r.status), r.rid
        assert (r.status == 0).mean() > 0.99


@pytest.mark.gpu
def test_serving_admission_respects_hbm_budget():
    trace = S.make_trace(400.0, 0.03, seed=12, lo=1e3, hi=4e3)
    budget = max(r.footprint() for r in trace)       # one request at a time
    srv = S.Server("MIX3", slots=2, hbm_budget=budget)
    seen = []
    orig = srv._launch

    def spy(slot, reqs):
        seen.append(srv.in_use)
        return orig(slot, reqs)
    srv._launch = spy
    S.run_trace(srv, trace)
    srv.close()
    assert seen and max(seen) <= budget
    assert all(r.status is not None for r in trace)
