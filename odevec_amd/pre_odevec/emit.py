"""Emitters: straight-line HIP/C++ network code and the network IR (JSON).

The emitted header defines one struct per (network, ordering, packaging)
with static ``__host__ __device__`` members:

* ``rates(y0, prm, k)``  -- rate coefficients, once per cell-solve;
* ``rhs(y, k, f)``       -- f(y), the RHS (src/odevec_commons.F90:78-95);
* ``jac(y, k, hb, a)``   -- the entries of ``P_order = I - hb*J`` in LU storage
  order, fill entries zero (src/odevec_commons.F90:32-51,
  pre_odevec.py:144-180);
* ``lu(a)``              -- the pivot-free in-place factorization, unrolled for
  the pattern: left-looking Crout for CSC (src/odevec.F90:929-971), Doolittle
  for DENSE (src/odevec.F90:898-922);
* ``solve(a, x)``        -- the triangular solves in place, with the ordering's
  permutation folded into the indexing (src/odevec.F90:328-342, 797-891);
* ``lu_r`` / ``solve_r`` -- the same with reciprocal pivots stored in the
  diagonal slots (used by the FP32 form, where the pivot division is an
  approximation anyway).

Every evaluator in this project evaluates the network in one canonical
order: a term is ``coef * (k * y_f0 * y_f1 * ...)`` (left to right), a row
sums its terms left to right, a Jacobian entry sums ``(coef*e_j) * (k * y...)``
over the row's terms, and ``P = 1 - hb*J`` on the diagonal, ``-(hb*J)`` off
it.  The factorizations and solves follow the reference loops statement by
statement with the structurally-zero operations (exact no-ops) removed.
"""
from __future__ import annotations

import json
import math

import numpy as np
from typing import Dict, List, Tuple

from . import lu_closed
from .structure import Structure, Term, linear_invariants, reduce_frozen

RATE_KIND_ID = {"none": 0, "param": 1, "primordial": 2, "arrhenius": 3}


def _lit(v: float) -> str:
    """Exact C literal of a double."""
    if v == int(v) and abs(v) < 2 ** 53:
        return "%d.0" % int(v)
    r = repr(float(v))
    if "e" not in r and "." not in r:
        r += ".0"
    return r


def lu_suffix(lu_method: int, sep: str) -> str:
    """Variant suffix of a LUmethod: none for 1, ``/LU2`` or ``/LU3``."""
    return "" if lu_method == 1 else "%sLU%d" % (sep, lu_method)


def struct_name(st: Structure) -> str:
    return "odv_net_%s_%s_%s%s" % (st.net.name, st.ordering.upper(), st.packaging, lu_suffix(st.lu_method, "_"))


def network_id(st: Structure) -> str:
    return "%s/%s/%s%s" % (st.net.name, st.ordering.upper(), st.packaging, lu_suffix(st.lu_method, "/"))


class _Counter:
    def __init__(self):
        self.add = 0
        self.mul = 0
        self.div = 0

    @property
    def flops(self) -> int:
        return self.add + self.mul + self.div

    def as_dict(self) -> Dict[str, int]:
        return {"add": self.add, "mul": self.mul, "div": self.div, "flops": self.flops}


def _monomial(rate: int, factors: Tuple[int, ...], yname: str, cnt: _Counter) -> str:
    parts = []
    if rate >= 0:
        parts.append("k[%d * S]" % rate)
    parts += ["%s[%d]" % (yname, f) for f in factors]
    if not parts:
        return "1.0"
    cnt.mul += len(parts) - 1
    return " * ".join(parts)


def _scaled(coef: float, var: str, cnt: _Counter) -> str:
    if coef == 1.0:
        return var
    if coef == -1.0:
        return "(-%s)" % var
    cnt.mul += 1
    return "(%s * %s)" % (_lit(coef), var)


def _emit_rhs(st: Structure, cnt: _Counter) -> List[str]:
    """Scatter form: each distinct flux k*prod(y) is formed once, in
    (rate, factors) order, and added to every row that has it. Terms inside a
    row are in that same order, so each row is summed in its canonical order
    while only one flux is live at a time."""
    lines = []
    by_key: Dict[Tuple[int, Tuple[int, ...]], List[Term]] = {}
    for t in st.terms:
        by_key.setdefault((t.rate, t.factors), []).append(t)
    started = set()
    for n, key in enumerate(sorted(by_key)):
        lines.append("    {")
        lines.append("      const double F = %s;" % _monomial(key[0], key[1], "y", cnt))
        for t in by_key[key]:
            v = _scaled(t.coef, "F", cnt)
            if t.row not in started:
                started.add(t.row)
                lines.append("      f[%d] = %s;" % (t.row, v))
            else:
                cnt.add += 1
                if v.startswith("(-"):
                    lines.append("      f[%d] = f[%d] - %s;" % (t.row, t.row, v[2:-1]))
                else:
                    lines.append("      f[%d] = f[%d] + %s;" % (t.row, t.row, v))
        lines.append("    }")
    for i in range(st.neq):
        if i not in started:
            lines.insert(0, "    f[%d] = 0.0;" % i)
    return lines


def jacobian_contributions(st: Structure):
    """{(row, col): [(coef*e, rate, reduced factors), ...]} in canonical order."""
    contrib: Dict[Tuple[int, int], List[Tuple[float, int, Tuple[int, ...]]]] = {}
    for t in st.terms:
        for j in sorted(set(t.factors)):
            e = t.factors.count(j)
            red = list(t.factors)
            red.remove(j)
            contrib.setdefault((t.row, j), []).append((t.coef * e, t.rate, tuple(red)))
    return contrib


def _emit_jac(st: Structure, cnt: _Counter) -> List[str]:
    """Entries of P_order in storage order, each formed and written on its
    own: the derivative monomials of an entry are formed where the entry is
    (a monomial two entries share is formed twice, which keeps the live set to
    one entry's terms)."""
    lines = []
    contrib = jacobian_contributions(st)
    perm = [int(p) for p in st.perm]
    for s, (r, c) in enumerate(st.csc.entries):
        R, C = perm[r], perm[c]
        if (R, C) in contrib:
            vals = []
            for coef, rate, red in contrib[(R, C)]:
                mono = _monomial(rate, red, "y", cnt)
                vals.append(_scaled(coef, "(%s)" % mono if " " in mono else mono, cnt))
            cnt.add += len(vals) - 1
            expr = vals[0]
            for v in vals[1:]:
                expr = "%s + %s" % (expr, v) if not v.startswith("(-") else "%s - %s" % (expr, v[2:-1])
            jn = "J%d_%d" % (R, C)
            lines.append("    const double %s = %s;" % (jn, expr))
            cnt.mul += 1
            if R == C:
                cnt.add += 1
                val = "1.0 - hb * %s" % jn
            else:
                val = "-(hb * %s)" % jn
        elif R == C:
            val = "1.0"
        else:
            val = "0.0"      # fill entry (pre_odevec.py:175-177)
        lines.append("    a[%d] = (M)(%s);" % (s, val))
    return lines


def _emit_lu_crout(st: Structure, cnt: _Counter, rcp: bool) -> List[str]:
    cs = st.csc
    idx = cs.index()
    lines = []
    n = st.neq
    for k in range(n):
        for kk in range(cs.u_col_start[k], cs.l_col_start[k]):
            j = cs.row_index[kk]
            dj = cs.l_col_start[j]
            if rcp:
                cnt.mul += 1
                lines.append("    a[%d] = a[%d] * a[%d];" % (kk, kk, dj))
            else:
                cnt.div += 1
                lines.append("    a[%d] = a[%d] / a[%d];" % (kk, kk, dj))
            for jj in range(cs.l_col_start[j] + 1, cs.u_col_start[j + 1]):
                r = cs.row_index[jj]
                tgt = idx[(r, k)]
                cnt.mul += 1
                cnt.add += 1
                lines.append("    a[%d] = a[%d] - a[%d] * a[%d];" % (tgt, tgt, kk, jj))
        if rcp:
            cnt.div += 1
            d = cs.l_col_start[k]
            lines.append("    a[%d] = (M)1 / a[%d];" % (d, d))
    return lines


def _emit_lu_doolittle(st: Structure, cnt: _Counter, rcp: bool) -> List[str]:
    idx = st.csc.index()
    n = st.neq
    lines = []
    for k in range(n):
        dk = idx[(k, k)]
        if rcp:
            cnt.div += 1
            lines.append("    a[%d] = (M)1 / a[%d];" % (dk, dk))
        if k == n - 1:
            break
        for j in range(k + 1, n):
            if (j, k) in idx:
                s = idx[(j, k)]
                if rcp:
                    cnt.mul += 1
                    lines.append("    a[%d] = a[%d] * a[%d];" % (s, s, dk))
                else:
                    cnt.div += 1
                    lines.append("    a[%d] = a[%d] / a[%d];" % (s, s, dk))
        for j in range(k + 1, n):
            if (k, j) not in idx:
                continue
            for m in range(k + 1, n):
                if (m, k) in idx:
                    tgt = idx[(m, j)]
                    cnt.mul += 1
                    cnt.add += 1
                    lines.append("    a[%d] = a[%d] - a[%d] * a[%d];" % (tgt, tgt, idx[(m, k)], idx[(k, j)]))
    return lines


def _emit_solve_crout(st: Structure, cnt: _Counter, rcp: bool) -> List[str]:
    cs = st.csc
    perm = [int(p) for p in st.perm]
    X = lambda i: "x[%d]" % perm[i]
    lines = []
    n = st.neq
    for k in range(n):
        d = cs.l_col_start[k]
        if rcp:
            cnt.mul += 1
            lines.append("    %s = %s * a[%d];" % (X(k), X(k), d))
        else:
            cnt.div += 1
            lines.append("    %s = %s / a[%d];" % (X(k), X(k), d))
        for kk in range(cs.l_col_start[k] + 1, cs.u_col_start[k + 1]):
            r = cs.row_index[kk]
            cnt.mul += 1
            cnt.add += 1
            lines.append("    %s = %s - %s * a[%d];" % (X(r), X(r), X(k), kk))
    for k in range(n - 1, -1, -1):
        for kk in range(cs.u_col_start[k], cs.l_col_start[k]):
            r = cs.row_index[kk]
            cnt.mul += 1
            cnt.add += 1
            lines.append("    %s = %s - %s * a[%d];" % (X(r), X(r), X(k), kk))
    return lines


def _emit_solve_doolittle(st: Structure, cnt: _Counter, rcp: bool) -> List[str]:
    idx = st.csc.index()
    perm = [int(p) for p in st.perm]
    X = lambda i: "x[%d]" % perm[i]
    n = st.neq
    lines = []
    for k in range(n):
        prods = ["a[%d] * %s" % (idx[(k, j)], X(j)) for j in range(k) if (k, j) in idx]
        if prods:
            cnt.mul += len(prods)
            cnt.add += len(prods)     # (len-1) sums + the subtraction from res
            lines.append("    %s = %s - (%s);" % (X(k), X(k), " + ".join(prods)))
    for k in range(n - 1, -1, -1):
        prods = ["a[%d] * %s" % (idx[(k, j)], X(j)) for j in range(k + 1, n) if (k, j) in idx]
        d = idx[(k, k)]
        if prods:
            cnt.mul += len(prods)
            cnt.add += len(prods)
            inner = "(%s - (%s))" % (X(k), " + ".join(prods))
        else:
            inner = X(k)
        if rcp:
            cnt.mul += 1
            lines.append("    %s = %s * a[%d];" % (X(k), inner, d))
        else:
            cnt.div += 1
            lines.append("    %s = %s / a[%d];" % (X(k), inner, d))
    return lines


def _emit_rates(st: Structure, cnt: _Counter) -> List[str]:
    net = st.net
    kind = net.rate_kind
    if kind == "none":
        return ["    (void)y0; (void)prm; (void)k;"]
    if kind == "param":
        return ["    (void)y0;"] + ["    k[%d] = prm[%d];" % (i, i) for i in range(net.nrate)]
    if kind == "primordial":
        return ["    (void)prm;",
                "    odv_primordial_rates(y0, NEQ, TGAS, k);"]
    if kind == "arrhenius":
        lines = ["#pragma clang fp contract(off)", "    (void)y0;", "    const double invT = 1.0 / prm[0];"]
        for r, (A, Ea) in enumerate(zip(net.arrhenius_A, net.arrhenius_Ea)):
            lines.append("    k[%d] = %s * odv::m_exp(-(%s * invT));" % (r, _lit(A), _lit(Ea)))
        return lines
    raise ValueError(kind)


def opcounts(st: Structure) -> Dict[str, Dict[str, int]]:
    out = {}
    for name, fn in (("rhs", lambda c: _emit_rhs(st, c)), ("jac", lambda c: _emit_jac(st, c)),
                     ("lu", lambda c: (_emit_lu_crout if st.packaging == "CSC" else _emit_lu_doolittle)(st, c, False)),
                     ("solve", lambda c: (_emit_solve_crout if st.packaging == "CSC" else _emit_solve_doolittle)(st, c, False)),
                     ("lu_r", lambda c: (_emit_lu_crout if st.packaging == "CSC" else _emit_lu_doolittle)(st, c, True)),
                     ("solve_r", lambda c: (_emit_solve_crout if st.packaging == "CSC" else _emit_solve_doolittle)(st, c, True))):
        c = _Counter()
        fn(c)
        out[name] = c.as_dict()
    if st.lu_method == 2:
        # the closed-form factors: the program's operations, no runtime
        # factorization; the reciprocal form inverts the diagonal
        out["jac"] = lu_closed.opcount(lu_closed.closed_form_lu(st))
        out["lu"] = {"add": 0, "mul": 0, "div": 0, "flops": 0}
        out["lu_r"] = {"add": 0, "mul": 0, "div": st.neq, "flops": st.neq}
    if st.lu_method == 3:
        # forward differences: N+1 RHS evaluations; the weighted norm of f
        # (3N+2), r0 (4), per column 9 (|y|, two products, sqrt, quotient,
        # max, y+r, dt*beta, its quotient), per entry 2 (difference, product),
        # per diagonal 1
        n = st.neq
        fl = (n + 1) * out["rhs"]["flops"] + 3 * n + 6 + n * (9 + 2 * n) + n
        out["jac"] = {"add": 0, "mul": 0, "div": 0, "flops": fl}
    return out


COOP_MAX = 64   # species per cell of the cooperative kernels (one lane per species)
LANE_MAX = 32   # species per cell of the per-lane kernels (state in one lane's registers)


def _arr(ctype: str, name: str, vals, per_line: int = 24) -> List[str]:
    vals = list(vals) or [0]
    out = ["  static constexpr %s %s[%d] = {" % (ctype, name, len(vals))]
    for i in range(0, len(vals), per_line):
        out.append("      " + ", ".join(str(v) for v in vals[i:i + per_line]) + ",")
    out.append("  };")
    return out


def _emit_coop(st: Structure) -> List[str]:
    """Tables of the cooperative kernels (one lane per species, odevec_coop.h)
    in the oracle's evaluation order:
      RHS, per species row c: its terms in canonical order; f_c = v_first,
        then f_c + v for the rest, v = coef * ((k_rate * y_f0) * y_f1 ...);
      Jacobian, per permuted P entry (i, jj) (row-major, offsets JT_OFF): the
        contributions in term order, v = (coef*e) * (k_rate * reduced monomial);
      PERM / IPERM: permuted position <-> species; LU_MASK: bit jj of row i set
        when (i, jj) is in the factors' pattern (with fill);
      the same RHS terms and Jacobian contributions lane-major (RL_*, JL_*):
        entry q of row c at [q*n + c] (q < RL_N[c] / JL_N[c]; up to CO_MAXT /
        CO_MAXJ), so the lanes of a group read consecutive words; a term's
        rate index and factor count in one word (RL_W: rate & 0xffff | nf << 16),
        its factors in another (RL_F: byte p = factor p); a contribution's
        also carries its permuted column (JL_W: ... | col << 24), in the
        order of the permuted row's entries, then term order."""
    n = st.neq
    rows = [[] for _ in range(n)]
    for t in st.terms:
        rows[t.row].append(t)
    rt_off, rt_rate, rt_coef, rt_nf, rt_f = [0], [], [], [], []
    for c in range(n):
        for t in rows[c]:
            rt_rate.append(-1 if t.rate is None or t.rate < 0 else int(t.rate))
            rt_coef.append(_lit(float(t.coef)))
            rt_nf.append(len(t.factors))
            f = list(t.factors) + [0] * (4 - len(t.factors))
            rt_f += f[:4]
        rt_off.append(len(rt_rate))
    perm = [int(p) for p in st.perm]
    ip = [0] * n
    for i, p in enumerate(perm):
        ip[p] = i
    ent = [[[] for _ in range(n)] for _ in range(n)]
    for t in st.terms:
        fac = list(t.factors)
        for q, j in enumerate(fac):
            if q > 0 and fac[q - 1] == j:
                continue
            e = fac.count(j)
            red = list(fac)
            red.remove(j)
            ent[ip[t.row]][ip[j]].append((-1 if t.rate is None or t.rate < 0 else int(t.rate),
                                          float(t.coef) * e, red))
    jt_off, jt_rate, jt_ce, jt_nr, jt_r = [0], [], [], [], []
    for i in range(n):
        for jj in range(n):
            for rate, ce, red in ent[i][jj]:
                jt_rate.append(rate)
                jt_ce.append(_lit(ce))
                jt_nr.append(len(red))
                rr = red + [0] * (3 - len(red))
                jt_r += rr[:3]
            jt_off.append(len(jt_rate))
    lu = np.asarray(st.LU, dtype=bool)
    mask = []
    for i in range(n):
        m = 0
        for jj in range(n):
            if lu[i, jj]:
                m |= 1 << jj
        mask.append("0x%xull" % m)
    maxr = max((len(r) for r in rows), default=0)
    # lane-major copies
    rl_n = [len(rows[c]) for c in range(n)]
    rl_w, rl_f, rl_c = [0] * max(maxr * n, 1), [0] * max(maxr * n, 1), ["0.0"] * max(maxr * n, 1)
    for c in range(n):
        for q, t in enumerate(rows[c]):
            assert len(t.factors) <= 4
            rate = -1 if t.rate is None or t.rate < 0 else int(t.rate)
            i = q * n + c
            rl_w[i] = (rate & 0xffff) | (len(t.factors) << 16)
            rl_f[i] = sum(int(f) << (8 * p) for p, f in enumerate(t.factors))
            rl_c[i] = _lit(float(t.coef))
    jl_rows = [[(jj, rate, ce, red) for jj in range(n) for rate, ce, red in ent[i][jj]] for i in range(n)]
    maxj = max((len(r) for r in jl_rows), default=0)
    jl_n = [len(r) for r in jl_rows]
    jl_w, jl_f, jl_c = [0] * max(maxj * n, 1), [0] * max(maxj * n, 1), ["0.0"] * max(maxj * n, 1)
    for i in range(n):
        for q, (jj, rate, ce, red) in enumerate(jl_rows[i]):
            assert len(red) <= 3 and jj < 256
            k = q * n + i
            jl_w[k] = (rate & 0xffff) | (len(red) << 16) | (jj << 24)
            jl_f[k] = sum(int(f) << (8 * p) for p, f in enumerate(red))
            jl_c[k] = _lit(ce)
    out = ["", "  // cooperative kernels (species on lanes): tables in the oracle's order",
           "  static constexpr int CO_MAXT = %d;   // most RHS terms of one species" % maxr,
           "  static constexpr int CO_MAXJ = %d;   // most Jacobian contributions of one row" % maxj]
    out += _arr("short", "RT_OFF", rt_off)
    out += _arr("short", "RT_RATE", rt_rate)
    out += _arr("double", "RT_COEF", rt_coef, 8)
    out += _arr("unsigned char", "RT_NF", rt_nf)
    out += _arr("unsigned char", "RT_F", rt_f)
    out += _arr("short", "JT_OFF", jt_off)
    out += _arr("short", "JT_RATE", jt_rate)
    out += _arr("double", "JT_CE", jt_ce, 8)
    out += _arr("unsigned char", "JT_NR", jt_nr)
    out += _arr("unsigned char", "JT_R", jt_r)
    out += _arr("unsigned char", "PERM", perm)
    out += _arr("unsigned char", "IPERM", ip)
    out += _arr("unsigned long long", "LU_MASK", mask, 6)
    out += _arr("unsigned char", "RL_N", rl_n)
    out += _arr("unsigned", "RL_W", ["%du" % v for v in rl_w], 12)
    out += _arr("unsigned", "RL_F", ["%du" % v for v in rl_f], 12)
    out += _arr("double", "RL_C", rl_c, 8)
    out += _arr("unsigned char", "JL_N", jl_n)
    out += _arr("unsigned", "JL_W", ["%du" % v for v in jl_w], 12)
    out += _arr("unsigned", "JL_F", ["%du" % v for v in jl_f], 12)
    out += _arr("double", "JL_C", jl_c, 8)
    return out


def emit_header(full: Structure) -> str:
    """Header for the structure's active species (frozen species -- zero RHS,
    read by no term -- are carried unchanged by the caller's state)."""
    st = reduce_frozen(full)
    net = st.net
    sn = struct_name(st)
    crout = st.packaging == "CSC"
    lu_fn = _emit_lu_crout if crout else _emit_lu_doolittle
    sv_fn = _emit_solve_crout if crout else _emit_solve_doolittle
    c = _Counter()
    rates = _emit_rates(st, c)
    rhs = _emit_rhs(st, _Counter())
    jac = _emit_jac(st, _Counter())
    lu = lu_fn(st, _Counter(), False)
    sv = sv_fn(st, _Counter(), False)
    lur = lu_fn(st, _Counter(), True)
    if st.lu_method == 2:
        jac, lu, lur = _emit_lu_closed(st)
    svr = sv_fn(st, _Counter(), True)
    ops = opcounts(st)
    perm = ", ".join(str(int(p)) for p in st.perm)
    act = st.active
    hdr = []
    hdr.append("// Generated by odevec_amd.pre_odevec -- do not edit.")
    hdr.append("// network %s, ordering %s, packaging %s" % (net.name, st.ordering, st.packaging))
    hdr.append("// neq %d (integrated %d), nnz(J) %d, nnz(P) %d, nnz(LU) %d (full system %d), terms %d" % (
        full.neq, st.neq, int(full.J.sum()), int(full.P.sum()), st.csc.nnz, full.csc.nnz, len(st.terms)))
    hdr.append("// flops: " + ", ".join("%s %d" % (k, v["flops"]) for k, v in ops.items()))
    hdr.append("#pragma once")
    hdr.append('#include "odevec_netcommon.h"')
    hdr.append("")
    hdr.append("struct %s {" % sn)
    hdr.append('  static constexpr const char* ID = "%s";' % network_id(st))
    hdr.append("  static constexpr int NEQ = %d;   // species in the caller's state" % full.neq)
    hdr.append("  static constexpr int NACT = %d;  // species integrated; the others stay constant" % st.neq)
    hdr.append("  // ACT[a] = species index of integrated species a")
    hdr.append("  static constexpr int ACT[%d] = {%s};" % (len(act), ", ".join(str(a) for a in act)))
    inv = linear_invariants(st.neq, st.terms)
    hdr.append("  // linear invariants over the integrated species: sum_a INV[k][a] f_a(y) = 0 for all y")
    hdr.append("  static constexpr int NINV = %d;" % len(inv))
    hdr.append("  static constexpr signed char INV[%d][%d] = {%s};" % (
        max(len(inv), 1), st.neq,
        ", ".join("{%s}" % ", ".join(str(v) for v in row) for row in inv) if inv else "{}"))
    hdr.append("  ODV_HD static constexpr double inv(int k, int a) { return INV[k][a]; }")
    hdr.append("  static constexpr int NNZ = %d;" % st.csc.nnz)
    hdr.append("  static constexpr int NRATE = %d;" % max(net.nrate, 1))
    hdr.append("  static constexpr int NPARAM = %d;" % net.nparam)
    hdr.append("  static constexpr int RATE_KIND = %d;" % RATE_KIND_ID[net.rate_kind])
    hdr.append("  static constexpr int TGAS = %d;" % net.tgas_index)
    hdr.append("  static constexpr bool CROUT = %s;" % ("true" if crout else "false"))
    hdr.append("  // LUmethod 3: P by forward differences of the RHS (CalcJac_dense); AIDX[r*NACT + c]")
    hdr.append("  // = storage index of P_order(r, c), -1 outside the LU pattern (there the difference")
    hdr.append("  // is exactly zero: no RHS term of row r reads species c)")
    hdr.append("  static constexpr bool FD_JAC = %s;" % ("true" if st.lu_method == 3 else "false"))
    hdr.append("  // LUmethod 2: jac() writes the closed-form L and U factors of P_order (GetLU),")
    hdr.append("  // lu() does nothing and lu_r() only inverts the diagonal")
    hdr.append("  static constexpr int LU_METHOD = %d;" % st.lu_method)
    if st.lu_method == 3:
        pos = {tuple(map(int, e)): k for k, e in enumerate(st.csc.entries)}
        aidx = [pos.get((r, c), -1) for r in range(st.neq) for c in range(st.neq)]
        hdr.append("  static constexpr short AIDX[%d] = {%s};" % (len(aidx), ", ".join(str(a) for a in aidx)))
    hdr.append("  static constexpr int FLOPS_RHS = %d, FLOPS_JAC = %d, FLOPS_LU = %d, FLOPS_SOLVE = %d;" % (
        ops["rhs"]["flops"], ops["jac"]["flops"], ops["lu"]["flops"], ops["solve"]["flops"]))
    hdr.append("  static constexpr int FLOPS_LU_R = %d, FLOPS_SOLVE_R = %d;" % (
        ops["lu_r"]["flops"], ops["solve_r"]["flops"]))
    hdr.append("  // permutation Perm (0-based): P_order[i][j] = P[Perm[i]][Perm[j]]")
    hdr.append("  // Perm = {%s}" % perm)
    hdr.append("")
    hdr.append("  static inline void param_defaults(double* p) {")
    if net.nparam == 0:
        hdr.append("    (void)p;")
    for i, v in enumerate(net.param_defaults):
        hdr.append("    p[%d] = %s;" % (i, _lit(v)))
    hdr.append("  }")
    hdr.append("")
    hdr.append("  ODV_HD static inline void rates(const double* __restrict__ y0, const double* __restrict__ prm,")
    hdr.append("                                 double* __restrict__ k) {")
    hdr += rates
    hdr.append("  }")
    if st.packaging == "CSC" and st.neq <= COOP_MAX:
        hdr.append("  static constexpr bool COOP = true;   // cooperative kernels available")
        hdr += _emit_coop(st)
    else:
        hdr.append("  static constexpr bool COOP = false;")
    hdr.append("  static constexpr bool LANE = %s;   // per-lane kernels available" % (
        "true" if st.neq <= LANE_MAX else "false"))
    if st.neq > LANE_MAX:
        hdr.append("};")
        hdr.append("")
        return "\n".join(hdr)
    hdr += _emit_jmap(full, st)
    hdr.append("")
    hdr.append("  // k[r * S]: rate r of this cell (S = stride between a cell's rates)")
    hdr.append("  template <int S = 1>")
    hdr.append("  ODV_HD static inline void rhs(const double* __restrict__ y, const double* __restrict__ k,")
    hdr.append("                               double* __restrict__ f) {")
    # each flux is rounded once and added as that same value to every row
    # it feeds, so that the rounding noise of f cancels exactly along the
    # network's linear invariants (element and charge conservation); a
    # flux's last multiplication contracted into each row's sum (an FMA per
    # row, which this target forms for a product with several uses) puts
    # noise of size eps * |flux| into those directions, where P = I -
    # dt*beta*J does not damp it: on stiff cells (dt*|flux| far above the
    # trace species' tolerances) the corrector then stalls and the steps
    # stay small (PRIMORDIAL 1M cells, FP64: 592k -> 303k convergence
    # failures, slowest cell 46,132 -> 4,199 attempts)
    hdr.append("#pragma clang fp contract(off)")
    hdr.append("    (void)k;")
    hdr += rhs
    hdr.append("  }")
    hdr.append("")
    hdr.append("  template <class M, int S = 1>")
    hdr.append("  ODV_HD static inline void jac(const double* __restrict__ y, const double* __restrict__ k,")
    hdr.append("                               double hb, M* __restrict__ a) {")
    hdr.append("    (void)k; (void)y; (void)hb;")
    hdr += jac
    hdr.append("  }")
    for name, body, sig in (("lu", lu, "M* __restrict__ a"), ("lu_r", lur, "M* __restrict__ a")):
        hdr.append("")
        hdr.append("  template <class M>")
        hdr.append("  ODV_HD static inline void %s(%s) {" % (name, sig))
        hdr += body
        hdr.append("  }")
    for name, body in (("solve", sv), ("solve_r", svr)):
        hdr.append("")
        hdr.append("  template <class M, class V>")
        hdr.append("  ODV_HD static inline void %s(const M* __restrict__ a, V* __restrict__ x) {" % name)
        hdr += body
        hdr.append("  }")
    hdr.append("};")
    hdr.append("")
    return "\n".join(hdr)


def _emit_lu_closed(st: Structure):
    """LUmethod 2 bodies: (jac, lu, lu_r). jac() evaluates the closed-form
    program and writes every storage slot (zero where the symbolic factor is
    zero); lu() is empty; lu_r() turns the U diagonal into reciprocals."""
    idx = st.csc.index()
    prog = lu_closed.closed_form_lu(st)
    lit = lambda v: _lit(v) if v >= 0 else "(%s)" % _lit(v)
    body, _ = lu_closed.emit_c(prog, lit, lambda r, c, e: "    a[%d] = (M)(%s);" % (idx[(r, c)], e))
    written = {idx[(r, c)] for r, c, _ in prog["out"]}
    zero = ["    a[%d] = (M)0.0;" % s for s in range(st.csc.nnz) if s not in written]
    jac = body + zero
    lur = ["    a[%d] = (M)1 / a[%d];" % (idx[(k, k)], idx[(k, k)]) for k in range(st.neq)]
    return jac, ["    (void)a;"], lur


def _emit_jmap(full: Structure, st: Structure) -> List[str]:
    """The reference's GetJac storage (src/odevec_commons.F90:32-51,
    pre_odevec.py:144-180) from the generated code's: entry e of the full
    system's layout -- DENSE: P_order(r, c) of all neq x neq at e = c*neq + r
    (Fortran jac(:, r, c)); CSC: jac%sdata(:, e) over the full LU pattern --
    takes storage index JMAP[e] >= 0 of the integrated species' P, or is 0
    (-1: structurally zero, fill, or coupling a frozen species) or 1 (-2: a
    frozen species' diagonal)."""
    active = st.active if st.active is not None else list(range(full.neq))
    amap = {s: a for a, s in enumerate(active)}
    redpos = {int(a): p for p, a in enumerate(st.perm)}
    store = st.csc.index()
    fperm = [int(p) for p in full.perm]
    if full.packaging == "DENSE":
        cells = [(r, c) for c in range(full.neq) for r in range(full.neq)]
    else:
        cells = [tuple(map(int, e)) for e in full.csc.entries]
    jm = []
    for r, c in cells:
        sr, sc = fperm[r], fperm[c]
        if sr in amap and sc in amap:
            jm.append(store.get((redpos[amap[sr]], redpos[amap[sc]]), -1))
        else:
            jm.append(-2 if r == c else -1)
    out = ["  // the reference's GetJac layout from this storage (see odevec_get_jac)",
           "  static constexpr int JMAP_N = %d;" % len(jm)]
    out += _arr("short", "JMAP", jm)
    return out


def analyse_copy(st: Structure) -> Structure:
    import copy
    return copy.copy(st)


def ir_dict(st: Structure) -> dict:
    net = st.net
    return {
        "id": network_id(st),
        "name": net.name,
        "ordering": st.ordering,
        "packaging": st.packaging,
        "neq": st.neq,
        "nrate": net.nrate,
        "nparam": net.nparam,
        "rate_kind": net.rate_kind,
        "tgas_index": net.tgas_index,
        "species": list(net.species),
        "param_names": list(net.param_names),
        "param_defaults": list(net.param_defaults),
        "arrhenius_A": net.arrhenius_A,
        "arrhenius_Ea": net.arrhenius_Ea,
        "terms": [[t.row, t.coef, t.rate, list(t.factors)] for t in st.terms],
        "perm": [int(p) for p in st.perm],
        "nnz_J": int(st.J.sum()),
        "nnz_P": int(st.P.sum()),
        "lu_entries": [list(e) for e in st.csc.entries],
        "row_index": st.csc.row_index,
        "u_col_start": st.csc.u_col_start,
        "l_col_start": st.csc.l_col_start,
        "nnz": st.csc.nnz,
        "lu_method": st.lu_method,
        # LUmethod 2: the closed-form factors' program (row, col in P_order)
        "lu_closed": lu_closed.closed_form_lu(st) if st.lu_method == 2 else None,
        # the generated code integrates the active species only
        "active": reduce_frozen(analyse_copy(st)).active,
        # linear invariants of the full system (zero on frozen species)
        "invariants": _full_invariants(st),
        "opcounts": opcounts(st),
        "opcounts_active": opcounts(reduce_frozen(analyse_copy(st))),
    }


def _full_invariants(st: Structure) -> List[List[int]]:
    red = reduce_frozen(analyse_copy(st))
    out = []
    for row in linear_invariants(red.neq, red.terms):
        full = [0] * st.neq
        for a, v in zip(red.active, row):
            full[a] = v
        out.append(full)
    return out


def emit_ir_json(st: Structure) -> str:
    return json.dumps(ir_dict(st), indent=1)
