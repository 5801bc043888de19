"""odevec_solve_cells_noalloc: the per-cell solve whose device memory beyond
the caller's arrays is the library's fixed scratch array. Called through the
C ABI, it must equal odevec_solve_cells_async in persistent mode (cooperative
mode for SYNTH64) bit for bit, on two calls in a row, the second with fewer
cells (the scratch keeps the first call's contents), and reject sizes the
scratch cannot hold."""
import ctypes as C
import sys

import numpy as np
import pytest

from odevec_amd import workloads as W
from odevec_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _noalloc(w, precision, stream, with_params=True, grid=0):
    import torch
    n = w["y"].shape[1]
    y = _dev(w["y"])
    ts, t0, dt0 = _dev(w["t_stop"]), _dev(w["t0"]), _dev(np.broadcast_to(w["dt0"], (n,)))
    prm = _dev(w["params"]) if (with_params and w.get("params") is not None) else None
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    dto = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = _lib.lib().odevec_solve_cells_noalloc(
        w["network"].encode(), precision, n, t0.data_ptr(), ts.data_ptr(), dt0.data_ptr(), y.data_ptr(),
        None if prm is None else prm.data_ptr(), st.data_ptr(), dto.data_ptr(), w["rtol"], w["atol"],
        w["dt_min"], grid, C.c_void_p(stream.cuda_stream))
    assert rc == 0, _lib.last_error()
    stream.synchronize()
    return y.cpu().numpy(), st.cpu().numpy(), dto.cpu().numpy()


def _handle(w, precision, mode, stream, with_params=True):
    import torch
    from odevec_amd.api import Odevec
    n = w["y"].shape[1]
    ode = Odevec(w["network"], n, precision=precision, mode=mode)
    ode.RelativeTolerance, ode.AbsoluteTolerance, ode.MinimalTimestep = w["rtol"], w["atol"], w["dt_min"]
    y = _dev(w["y"])
    ts, t0, dt0 = _dev(w["t_stop"]), _dev(w["t0"]), _dev(np.broadcast_to(w["dt0"], (n,)))
    prm = _dev(w["params"]) if (with_params and w.get("params") is not None) else None
    st = torch.zeros(n, dtype=torch.int32, device="cuda")
    dto = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ode.solve_cells_async(n, ts.data_ptr(), y.data_ptr(), t0_ptr=t0.data_ptr(), dt0_ptr=dt0.data_ptr(),
                          params_ptr=None if prm is None else prm.data_ptr(), status_ptr=st.data_ptr(),
                          dt_out_ptr=dto.data_ptr(), stream=stream.cuda_stream)
    stream.synchronize()
    ode.CloseOdevec()
    return y.cpu().numpy(), st.cpu().numpy(), dto.cpu().numpy()


CASES = [
    ("PRIMORDIAL mixed", lambda n: W.primordial(n, seed=41), _lib.MIXED, _lib.PERSISTENT, True),
    ("PRIMORDIAL FP64", lambda n: W.primordial(n, seed=42), _lib.FP64, _lib.PERSISTENT, True),
    ("ROBER default params", lambda n: W.rober(n, seed=43), _lib.FP64, _lib.PERSISTENT, False),
    ("SYNTH64 mixed", lambda n: W.synth64(n, seed=44), _lib.MIXED, _lib.COOP, True),
]


@pytest.mark.parametrize("name,gen,precision,mode,with_params", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_async_entry(name, gen, precision, mode, with_params):
    import torch
    stream = torch.cuda.Stream()
    for n in ((3000, 700) if "SYNTH64" not in name else (300, 70)):
        w = gen(n)
        say("GPU no-alloc solve", name, "cells", n)
        a = _noalloc(w, precision, stream, with_params)
        say("GPU async solve", name, "cells", n)
        b = _handle(w, precision, mode, stream, with_params)
        assert (a[1] == 0).mean() > 0.99, np.bincount(a[1])
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_noalloc_rejects_sizes_it_cannot_hold():
    """A grid whose workspace exceeds the scratch, and 2^31 cells, are
    refused before anything is launched."""
    L = _lib.lib()
    one = C.c_void_p(1)   # never dereferenced: the call is refused first
    rc = L.odevec_solve_cells_noalloc(b"PRIMORDIAL/CMK/CSC", _lib.FP64, 1 << 24, None, one, None, one, None, None,
                                      None, 1e-6, 1e-20, 1e-20, 1 << 20, None)
    assert rc == _lib.EINVAL and "scratch" in _lib.last_error()
    rc = L.odevec_solve_cells_noalloc(b"PRIMORDIAL/CMK/CSC", _lib.FP64, 1 << 31, None, one, None, one, None, None,
                                      None, 1e-6, 1e-20, 1e-20, 0, None)
    assert rc == _lib.EINVAL
