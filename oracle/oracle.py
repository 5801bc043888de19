"""ORACLE -- test infrastructure only.

ctypes wrapper of ``oracle/_build/liboracle.so``, the plain-C restatement of
the reference solver (see odevec_oracle.c for what it follows and how it is
pinned). Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
GEN = os.path.join(HERE, "..", "odevec_amd", "csrc", "generated")

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class ONet(C.Structure):
    _fields_ = [("neq", C.c_int), ("nrate", C.c_int), ("nparam", C.c_int), ("rate_kind", C.c_int),
                ("tgas", C.c_int), ("nterms", C.c_int),
                ("t_row", _IP), ("t_coef", _DP), ("t_rate", _IP), ("t_nfac", _IP), ("t_fac", _IP),
                ("perm", _IP), ("csc", C.c_int), ("nnz", C.c_int),
                ("row_index", _IP), ("u_col_start", _IP), ("l_col_start", _IP),
                ("arr_A", _DP), ("arr_Ea", _DP), ("lu_method", C.c_int),
                ("lc_nops", C.c_int), ("lc_code", _IP), ("lc_a", _IP), ("lc_b", _IP), ("lc_c", _DP),
                ("lc_nout", C.c_int), ("lc_out", _IP)]


class OCtl(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("dt_min", C.c_double),
                ("check_negatives", C.c_int), ("max_attempts", C.c_int)]


class OCnt(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("attempts", "steps", "newton", "jac", "errfail", "convfail", "stepsel", "qsum")]


RATE_KIND = {"none": 0, "param": 1, "primordial": 2, "arrhenius": 3}
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_solve_vector.restype = C.c_int
        L.oracle_solve_vector.argtypes = [C.POINTER(ONet), C.c_int, _DP, _DP, C.c_double, _DP, _DP,
                                          C.c_char_p, C.POINTER(OCtl), C.POINTER(OCnt)]
        L.oracle_solve_cells.restype = None
        L.oracle_solve_cells.argtypes = [C.POINTER(ONet), C.c_long, _DP, _DP, _DP, _DP, _DP,
                                         C.POINTER(OCtl), _IP, _DP, _IP, C.c_int]
        L.oracle_rhs.argtypes = [C.POINTER(ONet), _DP, _DP, _DP]
        L.oracle_p_dense.argtypes = [C.POINTER(ONet), _DP, _DP, C.c_double, _DP]
        L.oracle_factor_one.argtypes = [C.POINTER(ONet), _DP, _DP, C.c_double, _DP]
        L.oracle_solve_one.argtypes = [C.POINTER(ONet), _DP, _DP, _DP]
        L.oracle_fd_p_one.argtypes = [C.POINTER(ONet), _DP, _DP, _DP, C.c_double, C.c_double, _DP]
        L.oracle_rates_one.argtypes = [C.POINTER(ONet), _DP, _DP, _DP]
        L.oracle_primordial_rates.argtypes = [_DP, C.c_int, C.c_int, _DP]
        L.oracle_coeff_table.argtypes = [_DP]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_IP)


class Network:
    """A network IR (the preprocessor's JSON) turned into the oracle's tables."""

    def __init__(self, ir: dict):
        self.ir = ir
        self.neq = ir["neq"]
        self.nrate = ir["nrate"]
        self.nparam = ir["nparam"]
        terms = ir["terms"]
        nt = len(terms)
        self._row = np.array([t[0] for t in terms], dtype=np.int32)
        self._coef = np.array([t[1] for t in terms], dtype=np.float64)
        self._rate = np.array([t[2] for t in terms], dtype=np.int32)
        self._nfac = np.array([len(t[3]) for t in terms], dtype=np.int32)
        if nt and self._nfac.max() > 4:
            raise ValueError("oracle supports at most 4 factors per term")
        fac = np.zeros((max(nt, 1), 4), dtype=np.int32)
        for i, t in enumerate(terms):
            fac[i, :len(t[3])] = t[3]
        self._fac = np.ascontiguousarray(fac)
        self._perm = np.array(ir["perm"], dtype=np.int32)
        self._ri = np.array(ir["row_index"], dtype=np.int32)
        self._ucs = np.array(ir["u_col_start"], dtype=np.int32)
        self._lcs = np.array(ir["l_col_start"], dtype=np.int32)
        self._A = np.array(ir.get("arrhenius_A") or [0.0], dtype=np.float64)
        self._Ea = np.array(ir.get("arrhenius_Ea") or [0.0], dtype=np.float64)
        self.csc = ir["packaging"] == "CSC"
        self.nnz = ir["nnz"]
        self.param_defaults = np.array(ir.get("param_defaults") or [], dtype=np.float64)
        lc = ir.get("lu_closed") or {"ops": [], "out": []}
        ops = lc["ops"] or [[3, 0, 0, 0.0]]
        self._lc_code = np.array([o[0] for o in ops], dtype=np.int32)
        self._lc_a = np.array([o[1] for o in ops], dtype=np.int32)
        self._lc_b = np.array([o[2] for o in ops], dtype=np.int32)
        self._lc_c = np.array([o[3] for o in ops], dtype=np.float64)
        self._lc_out = np.array(lc["out"] or [[0, 0, 0]], dtype=np.int32).reshape(-1)
        self.s = ONet(self.neq, self.nrate, self.nparam, RATE_KIND[ir["rate_kind"]], ir["tgas_index"], nt,
                      _ip(self._row), _dp(self._coef), _ip(self._rate), _ip(self._nfac), _ip(self._fac),
                      _ip(self._perm), 1 if self.csc else 0, self.nnz, _ip(self._ri), _ip(self._ucs),
                      _ip(self._lcs), _dp(self._A), _dp(self._Ea), int(ir.get("lu_method", 1)),
                      len(lc["ops"]), _ip(self._lc_code), _ip(self._lc_a), _ip(self._lc_b), _dp(self._lc_c),
                      len(lc["out"]), _ip(self._lc_out))

    @classmethod
    def load(cls, net_id: str) -> "Network":
        base = net_id.replace("/", "_")
        with open(os.path.join(GEN, base + ".json")) as fh:
            return cls(json.load(fh))

    @property
    def lu_size(self) -> int:
        return self.nnz if self.csc else self.neq * self.neq

    # ---- single-cell network functions
    def rates(self, y0, params=None) -> np.ndarray:
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        p = np.ascontiguousarray(self.param_defaults if params is None else params, dtype=np.float64)
        if p.size == 0:
            p = np.zeros(1)
        k = np.zeros(max(self.nrate, 1))
        lib().oracle_rates_one(C.byref(self.s), _dp(y0), _dp(p), _dp(k))
        return k

    def rhs(self, y, k) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        if k.size == 0:
            k = np.zeros(1)
        f = np.zeros(self.neq)
        lib().oracle_rhs(C.byref(self.s), _dp(y), _dp(k), _dp(f))
        return f

    def p_dense(self, y, k, hb) -> np.ndarray:
        """P_order = I - hb*J (permuted), row-major."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        if k.size == 0:
            k = np.zeros(1)
        P = np.zeros(self.neq * self.neq)
        lib().oracle_p_dense(C.byref(self.s), _dp(y), _dp(k), float(hb), _dp(P))
        return P.reshape(self.neq, self.neq)

    def fd_p(self, y, k, w, dt, beta) -> np.ndarray:
        """CalcJac_dense's P = I - dt*beta*J (forward differences), row-major."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        if k.size == 0:
            k = np.zeros(1)
        P = np.zeros(self.neq * self.neq)
        lib().oracle_fd_p_one(C.byref(self.s), _dp(y), _dp(k), _dp(w), float(dt), float(beta), _dp(P))
        return P.reshape(self.neq, self.neq)

    def factor(self, y, k, hb) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        if k.size == 0:
            k = np.zeros(1)
        out = np.zeros(self.lu_size)
        lib().oracle_factor_one(C.byref(self.s), _dp(y), _dp(k), float(hb), _dp(out))
        return out

    def lin_solve(self, lu, res) -> np.ndarray:
        lu = np.ascontiguousarray(lu, dtype=np.float64)
        res = np.ascontiguousarray(res, dtype=np.float64)
        den = np.zeros(self.neq)
        lib().oracle_solve_one(C.byref(self.s), _dp(lu), _dp(res), _dp(den))
        return den

    # ---- solves
    def solve_vector(self, y, t0, t_stop, dt, rtol, atol, dt_min=float(np.float32(1e-10)),
                     params=None, mask=None, check_negatives=False, max_attempts=1000000):
        """SolveODE over one vector of G cells (lockstep). y: [neq, G]."""
        y = np.array(y, dtype=np.float64, order="C", copy=True)
        G = y.shape[1]
        if self.nparam:
            p = np.ascontiguousarray(np.broadcast_to(
                self.param_defaults[:, None] if params is None else params, (self.nparam, G)), dtype=np.float64)
        else:
            p = np.zeros(1)
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.int8)).tobytes()
        t = C.c_double(t0)
        d = C.c_double(dt)
        ctl = OCtl(rtol, atol, dt_min, 1 if check_negatives else 0, max_attempts)
        cnt = OCnt()
        st = lib().oracle_solve_vector(C.byref(self.s), G, C.byref(t), C.byref(d), float(t_stop), _dp(y),
                                       _dp(p), m, C.byref(ctl), C.byref(cnt))
        return y, st, t.value, d.value, {f: getattr(cnt, f) for f, _ in OCnt._fields_}

    def solve_cells(self, y, t_stop, t0=None, dt0=None, params=None, rtol=1e-6, atol=1e-10,
                    dt_min=float(np.float32(1e-10)), check_negatives=False, max_attempts=1000000,
                    threads=1):
        """Per-cell solves (each cell alone, G = 1). y: [neq, ncells]."""
        y = np.array(y, dtype=np.float64, order="C", copy=True)
        n = y.shape[1]
        ts = np.ascontiguousarray(np.broadcast_to(t_stop, (n,)), dtype=np.float64)
        t0a = np.ascontiguousarray(np.broadcast_to(0.0 if t0 is None else t0, (n,)), dtype=np.float64)
        d0 = np.ascontiguousarray(np.broadcast_to(1e-10 if dt0 is None else dt0, (n,)), dtype=np.float64)
        if self.nparam:
            p = np.ascontiguousarray(np.broadcast_to(
                self.param_defaults[:, None] if params is None else params, (self.nparam, n)), dtype=np.float64)
        else:
            p = np.zeros(1)
        st = np.zeros(n, dtype=np.int32)
        dto = np.zeros(n)
        cnt = np.zeros((8, n), dtype=np.int32)
        ctl = OCtl(rtol, atol, dt_min, 1 if check_negatives else 0, max_attempts)
        lib().oracle_solve_cells(C.byref(self.s), n, _dp(t0a), _dp(ts), _dp(d0), _dp(y), _dp(p), C.byref(ctl),
                                 _ip(st), _dp(dto), _ip(cnt), int(threads))
        return y, st, dto, cnt


def coeff_table() -> np.ndarray:
    out = np.zeros(42)
    lib().oracle_coeff_table(_dp(out))
    return out.reshape(6, 7)


def primordial_rates(y, tgas=14) -> np.ndarray:
    y = np.ascontiguousarray(y, dtype=np.float64)
    k = np.zeros(38)
    lib().oracle_primordial_rates(_dp(y), y.size, tgas, _dp(k))
    return k
