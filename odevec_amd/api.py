"""Python host interface mirroring the reference's Fortran module.

``Odevec`` mirrors ``type(odevec)`` of src/odevec.F90:33-74 with the
procedures that act on it:

=====================================  ====================================
reference (src/odevec.F90)             here
=====================================  ====================================
``call InitOdevec(ode)`` (:94)         ``ode = Odevec(network, nvector)`` /
                                       ``ode.InitOdevec()``
``ode%RelativeTolerance`` etc. (:48-52) same attribute names
``ode%y(nvector,neq)`` (:57)           ``ode.y`` (numpy, shape (nvector, neq))
``call SolveODE(ode,t,dt,t_stop,y,
GetRHS,GetJac,GetLU[,Mask])`` (:158)   ``t, dt = ode.SolveODE(t, dt, t_stop[, y, Mask])``
``call CloseOdevec(ode)`` (:1058)      ``ode.CloseOdevec()``
=====================================  ====================================

The network (RHS, Jacobian, LU) is fixed when the library is generated, so
the procedure arguments GetRHS/GetJac/GetLU have no counterpart; they are
accepted and ignored so that call sites read like the reference's.
Failures that make the reference print and STOP raise ``OdevecError`` here
(with per-cell status in ``ode.status``).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L


class Odevec:
    def __init__(self, network: str, nvector: int, precision: int = L.FP64, mode: int = L.PERCELL,
                 group: int = 0, block: int = 0, grid: int = 0, max_attempts: int = 0,
                 abort_on_error: bool = False, init: bool = True, schedule: int = 0):
        self.network = network
        self.nvector = int(nvector)
        self._opts = L.Opts(precision, mode, group, block, grid, max_attempts, 0, 1 if abort_on_error else 0,
                            int(schedule))
        self._h = None
        self._hist = None
        self._times = None
        info = L.network_info(network)
        self.neq = info["neq"]
        self.nnz = info["nnz"]
        self.nparam = info["nparam"]
        self.info = info
        self._rtol, self._atol = 1e-6, 1e-10
        self._dtmin = float(np.float32(1e-10))
        self._checkneg = False
        self.y = np.zeros((self.nvector, self.neq), order="F")
        self.status = np.zeros(self.nvector, dtype=np.int32)
        if init:
            self.InitOdevec()

    # -- reference procedures
    def InitOdevec(self):
        if self._h is not None:
            return
        h = C.c_void_p()
        L.check(L.lib().odevec_create(self.network.encode(), self.nvector, C.byref(self._opts), C.byref(h)))
        self._h = h
        self._push()

    def CloseOdevec(self):
        if self._h is not None:
            L.lib().odevec_destroy(self._h)
            self._h = None
        self._hist = None
        self._times = None

    # -- the generated plugin's network functions (src/odevec_commons.F90:32-95)
    def GetRHS(self, y: np.ndarray, params: Optional[np.ndarray] = None) -> np.ndarray:
        """f(y) for y of shape (neq, ncells) (rates from each cell's y and
        its parameters, shape (nparam, ncells), or the network's defaults);
        returns (neq, ncells)."""
        self._require()
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros_like(y)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        L.check(L.lib().odevec_get_rhs(self._h, y.shape[1], y.ctypes.data, None if p is None else p.ctypes.data,
                                       out.ctypes.data))
        return out

    def GetJac(self, beta: float, y: np.ndarray, dt: float, params: Optional[np.ndarray] = None) -> np.ndarray:
        """P = I - beta*dt*J in the packaging's storage: (entries, ncells)
        (DENSE: entry c*neq + r is P_order(r, c); CSC: the LU pattern)."""
        self._require()
        y = np.ascontiguousarray(y, dtype=np.float64)
        ne = C.c_int64(0)
        L.check(L.lib().odevec_jac_entries(self.network.encode(), C.byref(ne)))
        out = np.zeros((ne.value, y.shape[1]))
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        L.check(L.lib().odevec_get_jac(self._h, y.shape[1], float(beta), float(dt), y.ctypes.data,
                                       None if p is None else p.ctypes.data, out.ctypes.data))
        return out

    # -- warm restart (no reference counterpart: its SolveODE cold-starts every call)
    def enable_history(self, ncells: int):
        """Attach a zero-filled device history for `ncells` cells (a torch
        tensor this object keeps): later per-cell solves over exactly those
        cells continue each cell from where the previous call left it, when
        the call starts at that t_stop on the state it returned."""
        import torch
        planes = L.history_planes(self.network)
        self._hist = torch.zeros(planes * int(ncells), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()   # the fill runs on torch's stream, the solves on others
        L.check(L.lib().odevec_set_history(self._h, C.c_void_p(self._hist.data_ptr()), int(ncells)))
        return self._hist

    def enable_timeline(self, ncells: int):
        """Attach a device array of per-cell start / end ticks (2 x ncells);
        returns (tensor, ticks per second). Later solves over exactly those
        cells fill it."""
        import torch
        self._times = torch.zeros(2 * int(ncells), dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        hz = C.c_int64(0)
        L.check(L.lib().odevec_set_timeline(self._h, C.c_void_p(self._times.data_ptr()), int(ncells), C.byref(hz)))
        return self._times, int(hz.value)

    def set_fp64_fallback(self, convfail_limit: int):
        """Mixed form: cells that reach `convfail_limit` convergence failures
        continue in FP64 from where they stopped (FP64 lanes beside the
        launch in persistent mode, a second launch in per-cell mode); 0 turns
        it off (the default). On an FP64 handle results are unchanged."""
        L.check(L.lib().odevec_set_fp64_fallback(self._h, int(convfail_limit)))

    def set_invariant_test(self, tolerances: float = 1.0, fallback_fails: int = 32):
        """Mixed form: a corrector that passes the convergence test but leaves
        an element or charge invariant's residual above `tolerances`
        error-weight units counts as a convergence failure (0 turns the test
        off); with the FP64 fallback on, a cell is handed back after
        `fallback_fails` such failures (0: never)."""
        L.check(L.lib().odevec_set_invariant_test(self._h, float(tolerances), int(fallback_fails)))

    def set_mixed_convergence_test(self, test: str = "max"):
        """Mixed form's corrector test: "max" (the default) asks every
        weighted component of the last correction to be within the divisor;
        "rms" is the reference's test (RMS norm times min(1, 1.5 rate),
        src/odevec.F90:622-665), for networks whose FP32 factors contract
        every component (SYNTH64). The FP64 form is unaffected."""
        if test not in ("max", "rms"):
            raise ValueError("test must be 'max' or 'rms'")
        L.check(L.lib().odevec_set_mixed_convergence_test(self._h, 1 if test == "rms" else 0))

    def disable_history(self):
        L.check(L.lib().odevec_set_history(self._h, None, 0))
        self._hist = None

    def SolveODE(self, time: float, dt: float, t_stop: float, y: Optional[np.ndarray] = None,
                 GetRHS=None, GetJac=None, GetLU=None, Mask: Optional[np.ndarray] = None):
        """Integrate y (default: self.y) from time to t_stop; returns (t_stop, last dt)."""
        self._require()
        if y is None:
            y = self.y
        if y.shape != (self.nvector, self.neq):
            raise ValueError("y must have shape (nvector, neq) = (%d, %d)" % (self.nvector, self.neq))
        buf = np.asfortranarray(y, dtype=np.float64)
        t = C.c_double(time)
        d = C.c_double(dt)
        mask = None
        if Mask is not None:
            mask = np.ascontiguousarray(np.asarray(Mask, dtype=np.int8))
            if mask.shape != (self.nvector,):
                raise ValueError("Mask must have nvector entries")
        rc = L.lib().odevec_solve(self._h, C.byref(t), C.byref(d), float(t_stop), buf.ctypes.data,
                                  None if mask is None else mask.ctypes.data, self.status.ctypes.data)
        if buf is not y:
            y[...] = buf
        L.check(rc)
        return t.value, d.value

    # -- caller-set fields of type(odevec)
    @property
    def RelativeTolerance(self):
        return self._rtol

    @RelativeTolerance.setter
    def RelativeTolerance(self, v):
        self._rtol = float(v)
        self._push()

    @property
    def AbsoluteTolerance(self):
        return self._atol

    @AbsoluteTolerance.setter
    def AbsoluteTolerance(self, v):
        self._atol = float(v)
        self._push()

    @property
    def MinimalTimestep(self):
        return self._dtmin

    @MinimalTimestep.setter
    def MinimalTimestep(self, v):
        self._dtmin = float(v)
        self._push()

    @property
    def CheckNegatives(self):
        return self._checkneg

    @CheckNegatives.setter
    def CheckNegatives(self, v):
        self._checkneg = bool(v)
        self._push()

    # -- per-cell form
    def solve_cells(self, y: np.ndarray, t_stop, t0=None, dt0=None, params=None, stats: bool = False,
                    raise_on_fail: bool = True):
        """Per-cell solves of y (shape (neq, ncells), modified in place);
        t_stop/t0/dt0 scalars or arrays of ncells; params shape (nparam, ncells).
        Returns (status, dt_out, stats-dict or None)."""
        self._require()
        if y.ndim != 2 or y.shape[0] != self.neq or y.dtype != np.float64 or not y.flags.c_contiguous:
            raise ValueError("y must be a C-contiguous float64 array of shape (neq, ncells)")
        n = y.shape[1]

        def arr(v):
            if v is None:
                return None
            return np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))

        ts, t0a, d0 = arr(t_stop), arr(t0), arr(dt0)
        p = None
        if params is not None and self.nparam:
            p = np.ascontiguousarray(np.broadcast_to(np.asarray(params, dtype=np.float64), (self.nparam, n)))
        st = np.zeros(n, dtype=np.int32)
        dto = np.zeros(n)
        s = L.Stats() if stats else None
        ptr = lambda a: None if a is None else a.ctypes.data
        rc = L.lib().odevec_solve_cells(self._h, n, ptr(t0a), ptr(ts), ptr(d0), y.ctypes.data, ptr(p),
                                        st.ctypes.data, dto.ctypes.data, C.byref(s) if s is not None else None)
        if rc != L.OK and (raise_on_fail or rc != L.ESOLVE):
            L.check(rc)
        return st, dto, (s.as_dict() if s is not None else None)

    def solve_cells_async(self, ncells, t_stop_ptr, y_ptr, t0_ptr=None, dt0_ptr=None, params_ptr=None,
                          t_stop_stride=1, t0_stride=1, dt0_stride=1, params_stride=1, status_ptr=None,
                          dt_out_ptr=None, counters_ptr=None, flops_ptr=None, stream=None):
        """Enqueue a per-cell solve on device pointers (e.g. torch tensors' data_ptr())."""
        self._require()
        L.check(L.lib().odevec_solve_cells_async(
            self._h, int(ncells), t0_ptr, t0_stride, t_stop_ptr, t_stop_stride, dt0_ptr, dt0_stride, y_ptr,
            params_ptr, params_stride, status_ptr, dt_out_ptr, counters_ptr, flops_ptr, stream))

    # -- helpers
    def _require(self):
        if self._h is None:
            raise RuntimeError("call InitOdevec() first")

    def _push(self):
        if self._h is None:
            return
        lib = L.lib()
        L.check(lib.odevec_set_tolerances(self._h, self._rtol, self._atol))
        L.check(lib.odevec_set_dt_min(self._h, self._dtmin))
        L.check(lib.odevec_set_check_negatives(self._h, 1 if self._checkneg else 0))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.CloseOdevec()

    def __del__(self):
        try:
            self.CloseOdevec()
        except Exception:
            pass


def _ptr(v):
    """Device pointer of a torch tensor (or an int address, or None)."""
    if v is None:
        return None
    if isinstance(v, int):
        return v
    return v.data_ptr()


class Batch:
    """One launch over many blocks of cells of different networks and
    precisions (a bundle's batch kernel; no reference counterpart).

    ``solve_async(segments, stream)`` takes a list of dicts with keys
    ``network``, ``precision``, ``y`` (device, shape (neq, ncells)), ``t_stop``
    and optionally ``t0``, ``dt0``, ``params``, ``status``, ``dt_out``,
    ``counters``, ``flops``, ``rtol``, ``atol``, ``dt_min``,
    ``check_negatives``, ``max_attempts``. Array values are torch tensors on
    the GPU (or raw device addresses); ``ncells`` is taken from ``y``.
    """

    def __init__(self, bundle: str = "MIX3", block: int = 0, grid: int = 0, use_graph: bool = False,
                 fp64_fallback: int = 0):
        self.bundle = bundle
        self.entries = L.bundle_entries(bundle)
        h = C.c_void_p()
        L.check(L.lib().odevec_batch_create(bundle.encode(), int(block), int(grid), 1 if use_graph else 0,
                                            C.byref(h)))
        self._h = h
        self._keep = None
        if fp64_fallback:
            self.set_fp64_fallback(fp64_fallback)

    def set_fp64_fallback(self, convfail_limit: int):
        """Mixed segments: cells that reach `convfail_limit` convergence
        failures continue in FP64 from where they stopped, in FP64 lanes
        beside the batch launch (0: off)."""
        L.check(L.lib().odevec_batch_set_fp64_fallback(self._h, int(convfail_limit)))

    def set_invariant_test(self, tolerances: float = 1.0, fallback_fails: int = 32):
        """The invariant test of mixed segments (see Odevec.set_invariant_test)."""
        L.check(L.lib().odevec_batch_set_invariant_test(self._h, float(tolerances), int(fallback_fails)))

    def solve_async(self, segments, stream=None):
        if self._h is None:
            raise RuntimeError("batch is closed")
        n = len(segments)
        arr = (L.Segment * max(n, 1))()
        names = []
        for k, sg in enumerate(segments):
            y = sg["y"]
            ncells = int(sg.get("ncells", y.shape[-1] if hasattr(y, "shape") else 0))
            nm = sg["network"].encode()
            names.append(nm)
            arr[k] = L.Segment(nm, int(sg.get("precision", L.FP64)), 1 if sg.get("check_negatives") else 0, ncells,
                               _ptr(y), _ptr(sg.get("t0")), _ptr(sg["t_stop"]), _ptr(sg.get("dt0")),
                               _ptr(sg.get("params")), _ptr(sg.get("status")), _ptr(sg.get("dt_out")),
                               _ptr(sg.get("counters")), _ptr(sg.get("flops")), float(sg.get("rtol", 1e-6)),
                               float(sg.get("atol", 1e-10)), float(sg.get("dt_min", float(np.float32(1e-10)))),
                               int(sg.get("max_attempts", 0)), int(sg.get("quantum", 0)), _ptr(sg.get("hist")))
        self._keep = (arr, names)
        L.check(L.lib().odevec_batch_solve_async(self._h, arr, n, stream))

    def close(self):
        if self._h is not None:
            L.lib().odevec_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
