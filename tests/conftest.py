import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)

HARNESS_SRC = os.path.join(ROOT, "tests", "harness", "lane_cpu.hip")
HARNESS_LIB = os.path.join(ROOT, "tests", "harness", "liblane_cpu.so")
CSRC = os.path.join(ROOT, "odevec_amd", "csrc")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")


def _newest_source_mtime():
    paths = [HARNESS_SRC]
    for d in (CSRC, os.path.join(CSRC, "generated")):
        paths += [os.path.join(d, f) for f in os.listdir(d) if f.endswith(".h")]
    return max(os.path.getmtime(p) for p in paths)


def build_harness():
    """The product's per-lane integrator compiled for the host (test only)."""
    if os.path.exists(HARNESS_LIB) and os.path.getmtime(HARNESS_LIB) >= _newest_source_mtime():
        return HARNESS_LIB
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-host-only", "-O2", "-ffp-contract=off", "-std=c++17",
           "-fPIC", "-shared", "-I", CSRC, "-o", HARNESS_LIB + ".tmp", HARNESS_SRC]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(HARNESS_LIB + ".tmp", HARNESS_LIB)
    return HARNESS_LIB


class Harness:
    _DP = C.POINTER(C.c_double)
    _IP = C.POINTER(C.c_int)

    def __init__(self, path):
        L = C.CDLL(path)
        DP, IP = self._DP, self._IP
        L.lane_solve_cells.argtypes = [C.c_char_p, C.c_int, C.c_long, DP, DP, DP, DP, DP, C.c_double, C.c_double,
                                       C.c_double, C.c_int, C.c_int, IP, DP, IP]
        L.lane_netfun.argtypes = [C.c_char_p, C.c_int, DP, DP, C.c_double, DP]
        self.L = L

    def solve_cells(self, net, w, mixed=0, strided=False, rtol=None, atol=None, dt_min=None, max_attempts=1000000,
                    check_negatives=False):
        y = np.array(w["y"], order="C", copy=True)
        n = y.shape[1]
        arr = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
        ts, t0, dt0 = arr(w["t_stop"]), arr(w["t0"]), arr(w["dt0"])
        prm = np.ascontiguousarray(w["params"] if w.get("params") is not None else np.zeros((1, n)))
        st = np.zeros(n, np.int32)
        dto = np.zeros(n)
        cnt = np.zeros((8, n), np.int32)
        dp = lambda a: a.ctypes.data_as(self._DP)
        rc = self.L.lane_solve_cells(net.encode(), int(mixed) | (2 if strided else 0), n, dp(t0), dp(ts), dp(dt0),
                                     dp(y), dp(prm), w["rtol"] if rtol is None else rtol,
                                     w["atol"] if atol is None else atol, w["dt_min"] if dt_min is None else dt_min,
                                     1 if check_negatives else 0, max_attempts, st.ctypes.data_as(self._IP), dp(dto),
                                     cnt.ctypes.data_as(self._IP))
        assert rc == 0, "network %s not in the harness" % net
        return y, st, dto, cnt

    def netfun(self, net, what, y, k, hb=0.0, nout=None):
        y = np.ascontiguousarray(y, dtype=np.float64)
        k = np.ascontiguousarray(k if len(k) else np.zeros(1), dtype=np.float64)
        out = np.zeros(nout if nout is not None else 4096)
        rc = self.L.lane_netfun(net.encode(), what, y.ctypes.data_as(self._DP), k.ctypes.data_as(self._DP),
                                float(hb), out.ctypes.data_as(self._DP))
        assert rc == 0
        return out


@pytest.fixture(scope="session")
def harness():
    return Harness(build_harness())


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
