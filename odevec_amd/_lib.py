"""ctypes binding of libodevec.so (include/odevec.h).

The library is loaded from this package directory; if it is missing the
import fails loudly (there is no other implementation to fall back to).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("ODEVEC_LIB", os.path.join(HERE, "libodevec.so"))

OK = 0
EINVAL, ENOMEM, EDEVICE, ENOTFOUND, ESOLVE = -1, -2, -3, -4, -5
STATUS = {0: "OK", 1: "DT_MIN", 2: "CONV_FAIL", 3: "ORDER", 4: "MAX_STEPS", 5: "SKIPPED", 8: "PAUSED"}
PAUSED, FRESH = 8, 9   # batch segments with a launch quantum
FP64, MIXED = 0, 1
PERCELL, LOCKSTEP, PERSISTENT, COOP = 0, 1, 2, 3

_I32 = C.c_int32
_I64 = C.c_int64
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_VP = C.c_void_p


class Opts(C.Structure):
    _fields_ = [("precision", _I32), ("mode", _I32), ("group", _I32), ("block", _I32), ("grid", _I32),
                ("max_attempts", _I32), ("device_ptrs", _I32), ("abort_on_error", _I32), ("schedule", _I32)]


class Stats(C.Structure):
    _fields_ = [(n, _I64) for n in ("attempts", "steps", "newton", "jac", "errfail", "convfail", "stepsel",
                                    "qsum", "cells_ok", "cells_failed", "cells_skipped")] + \
               [("flops64", C.c_double), ("flops32", C.c_double), ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NetInfo(C.Structure):
    _fields_ = [(n, _I32) for n in ("neq", "nnz", "nrate", "nparam", "rate_kind", "tgas_index", "flops_rhs",
                                    "flops_jac", "flops_lu", "flops_solve", "flops_lu_r", "flops_solve_r")]


class KernelInfo(C.Structure):
    _fields_ = [(n, _I32) for n in ("vgprs", "spill_bytes", "blocks_per_cu", "block")]


class Segment(C.Structure):
    _fields_ = [("network", C.c_char_p), ("precision", _I32), ("check_negatives", _I32), ("ncells", _I64),
                ("y", _VP), ("t0", _VP), ("t_stop", _VP), ("dt0", _VP), ("params", _VP), ("status", _VP),
                ("dt_out", _VP), ("counters", _VP), ("flops", _VP), ("rtol", C.c_double), ("atol", C.c_double),
                ("dt_min", C.c_double), ("max_attempts", _I32), ("quantum", _I32), ("hist", _VP)]


class OdevecError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("odevec error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError("libodevec.so not found at %s; build it with `python -m odevec_amd.build`" % LIBPATH)
    # PyTorch ships its own HIP runtime under the same soname; loading it first
    # lets this library bind to that one, so one process never holds two
    # runtimes (which leaves the second unable to see the GPU).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    sig = {
        "odevec_abi_version": (_I32, []),
        "odevec_last_error": (C.c_char_p, []),
        "odevec_network_count": (_I32, []),
        "odevec_network_name": (C.c_char_p, [_I32]),
        "odevec_network_info": (_I32, [C.c_char_p, C.POINTER(NetInfo)]),
        "odevec_kernel_info": (_I32, [C.c_char_p, _I32, _I32, _I32, C.POINTER(KernelInfo)]),
        "odevec_history_planes": (_I32, [C.c_char_p, C.POINTER(_I64)]),
        "odevec_set_history": (_I32, [_VP, _VP, _I64]),
        "odevec_set_timeline": (_I32, [_VP, _VP, _I64, C.POINTER(_I64)]),
        "odevec_get_rhs": (_I32, [_VP, _I64, _VP, _VP, _VP]),
        "odevec_get_jac": (_I32, [_VP, _I64, C.c_double, C.c_double, _VP, _VP, _VP]),
        "odevec_jac_entries": (_I32, [C.c_char_p, C.POINTER(_I64)]),
        "odevec_default_opts": (None, [C.POINTER(Opts)]),
        "odevec_create": (_I32, [C.c_char_p, _I64, C.POINTER(Opts), C.POINTER(_VP)]),
        "odevec_set_tolerances": (_I32, [_VP, C.c_double, C.c_double]),
        "odevec_set_dt_min": (_I32, [_VP, C.c_double]),
        "odevec_set_check_negatives": (_I32, [_VP, _I32]),
        "odevec_set_fp64_fallback": (_I32, [_VP, _I32]),
        "odevec_batch_set_fp64_fallback": (_I32, [_VP, _I32]),
        "odevec_set_invariant_test": (_I32, [_VP, C.c_double, _I32]),
        "odevec_set_mixed_convergence_test": (_I32, [_VP, _I32]),
        "odevec_batch_set_invariant_test": (_I32, [_VP, C.c_double, _I32]),
        "odevec_solve": (_I32, [_VP, _DP, _DP, C.c_double, _VP, _VP, _VP]),
        "odevec_solve_cells": (_I32, [_VP, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, C.POINTER(Stats)]),
        "odevec_solve_cells_async": (_I32, [_VP, _I64, _VP, _I64, _VP, _I64, _VP, _I64, _VP, _VP, _I64,
                                            _VP, _VP, _VP, _VP, _VP]),
        "odevec_destroy": (_I32, [_VP]),
        "odevec_solve_cells_noalloc": (_I32, [C.c_char_p, _I32, _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                                              C.c_double, C.c_double, C.c_double, _I32, _VP]),
        "odevec_bundle_count": (_I32, []),
        "odevec_bundle_name": (C.c_char_p, [_I32]),
        "odevec_bundle_entry": (_I32, [C.c_char_p, _I32, C.POINTER(C.c_char_p), C.POINTER(_I32)]),
        "odevec_batch_create": (_I32, [C.c_char_p, _I32, _I32, _I32, C.POINTER(_VP)]),
        "odevec_batch_solve_async": (_I32, [_VP, C.POINTER(Segment), _I32, _VP]),
        "odevec_batch_kernel_info": (_I32, [C.c_char_p, _I32, C.POINTER(KernelInfo)]),
        "odevec_batch_destroy": (_I32, [_VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


EXPORTED = [
    "odevec_abi_version", "odevec_last_error", "odevec_network_count", "odevec_network_name",
    "odevec_network_info", "odevec_kernel_info", "odevec_default_opts", "odevec_create",
    "odevec_set_tolerances", "odevec_set_dt_min", "odevec_set_check_negatives", "odevec_set_fp64_fallback", "odevec_solve",
    "odevec_solve_cells", "odevec_solve_cells_async", "odevec_destroy", "odevec_history_planes",
    "odevec_set_history", "odevec_set_timeline", "odevec_get_rhs", "odevec_get_jac", "odevec_jac_entries",
    "odevec_bundle_count",
    "odevec_bundle_name", "odevec_bundle_entry", "odevec_batch_create", "odevec_batch_solve_async",
    "odevec_batch_kernel_info", "odevec_batch_destroy", "odevec_batch_set_fp64_fallback",
    "odevec_set_invariant_test", "odevec_batch_set_invariant_test", "odevec_solve_cells_noalloc",
    "odevec_set_mixed_convergence_test",
]


def last_error() -> str:
    return lib().odevec_last_error().decode()


def check(rc: int) -> int:
    if rc != OK:
        raise OdevecError(rc, last_error())
    return rc


def networks():
    L = lib()
    return [L.odevec_network_name(i).decode() for i in range(L.odevec_network_count())]


def history_planes(name: str) -> int:
    """Doubles per cell of a warm-restart history for `name`."""
    v = _I64(0)
    check(lib().odevec_history_planes(name.encode(), C.byref(v)))
    return int(v.value)


def network_info(name: str) -> dict:
    info = NetInfo()
    check(lib().odevec_network_info(name.encode(), C.byref(info)))
    return {f: getattr(info, f) for f, _ in NetInfo._fields_}


def kernel_info(name: str, mode: int = PERCELL, precision: int = FP64, block: int = 0) -> dict:
    info = KernelInfo()
    check(lib().odevec_kernel_info(name.encode(), mode, precision, block, C.byref(info)))
    return {f: getattr(info, f) for f, _ in KernelInfo._fields_}


def bundles():
    L = lib()
    return [L.odevec_bundle_name(i).decode() for i in range(L.odevec_bundle_count())]


def bundle_entries(name: str):
    L = lib()
    net, prec = C.c_char_p(), _I32()
    n = L.odevec_bundle_entry(name.encode(), 0, C.byref(net), C.byref(prec))
    check(0 if n > 0 else n)
    out = []
    for k in range(n):
        check(0 if L.odevec_bundle_entry(name.encode(), k, C.byref(net), C.byref(prec)) > 0 else -1)
        out.append((net.value.decode(), prec.value))
    return out
