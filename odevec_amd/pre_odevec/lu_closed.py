"""LUmethod 2: the L and U factors of P_order = I - hb*J in closed form.

The reference's preprocessor factors P symbolically with SymPy's
``LUdecomposition_Simple`` (pre_odevec.py:474-494) and, for LUmethod 2,
writes every nonzero entry of the combined LU matrix as one expression in y,
the rate coefficients, dt and beta (pre_odevec.py:132-139; the solver calls
``GetLU`` instead of ``GetJac`` + ``LUDecompose``, src/odevec.F90:301-306,
src/odevec_commons.F90:55-72). The factors are then used by the ordinary
dense triangular solves (``SolveLU_dense``, src/odevec.F90:797-849: unit
lower L below the diagonal, U on and above it).

Here the same symbolic factorization is turned into one straight-line
program: a list of operations over the leaves y[i], k[r], hb (= dt*beta, the
product the solver forms once per factorization), constants, and the
results of earlier operations. Structurally identical subexpressions are
computed once and reused, which changes no value (the same operations on the
same operands). The HIP code generator and the oracle both execute exactly
this list, in this order, so the two agree bit for bit when neither contracts
a*b+c; what pins the list itself is that L*U reproduces P (tests).

Operation codes (``ops`` entries ``[code, a, b, c]``):
    0 y[a]   1 k[a]   2 hb   3 the constant c
    4 v[a] + v[b]   5 v[a] - v[b]   6 v[a] * v[b]   7 v[a] / v[b]   8 -v[a]
"""
from __future__ import annotations

from fractions import Fraction
from typing import Dict, List, Tuple

import sympy as sp

from .structure import Structure

Y, K, HB, CONST, ADD, SUB, MUL, DIV, NEG = range(9)
OPNAMES = ("y", "k", "hb", "const", "add", "sub", "mul", "div", "neg")


class _Program:
    def __init__(self, ysym, ksym, hb):
        self.ops: List[List] = []
        self.memo: Dict[tuple, int] = {}
        self.leaf = {s: (Y, i) for i, s in enumerate(ysym)}
        self.leaf.update({s: (K, r) for r, s in enumerate(ksym)})
        self.leaf[hb] = (HB, 0)

    def node(self, code: int, a: int = 0, b: int = 0, c: float = 0.0) -> int:
        key = (code, a, b, c)
        if key not in self.memo:
            self.memo[key] = len(self.ops)
            self.ops.append([code, a, b, c])
        return self.memo[key]

    def const(self, v) -> int:
        return self.node(CONST, c=float(v))

    def power(self, base, e: int) -> int:
        b = self.expr(base)
        acc = b
        for _ in range(e - 1):
            acc = self.node(MUL, acc, b)
        return acc

    def product(self, factors) -> int:
        acc = None
        for f in factors:
            if f.is_Pow and f.exp.is_Integer and int(f.exp) > 0:
                v = self.power(f.base, int(f.exp))
            else:
                v = self.expr(f)
            acc = v if acc is None else self.node(MUL, acc, v)
        return acc

    def expr(self, e) -> int:
        if e in self.leaf:
            code, i = self.leaf[e]
            return self.node(code, i)
        if e.is_Number:
            return self.const(e)
        if e.is_Add:
            acc = None
            for t in e.args:
                neg = t.could_extract_minus_sign() and not t.is_Add
                v = self.expr(-t if neg else t)
                if acc is None:
                    acc = self.node(NEG, v) if neg else v
                else:
                    acc = self.node(SUB if neg else ADD, acc, v)
            return acc
        if e.is_Mul:
            c, fs = e.as_coeff_mul()
            neg = c < 0
            c = -c if neg else c
            num, den = [], []
            for f in fs:
                if f.is_Pow and f.exp.is_Integer and int(f.exp) < 0:
                    den.append(f.base ** (-int(f.exp)))
                else:
                    num.append(f)
            acc = self.product(num) if num else None
            if c != 1:
                cv = self.const(c)
                acc = cv if acc is None else self.node(MUL, cv, acc)
            if acc is None:
                acc = self.const(1)
            if den:
                acc = self.node(DIV, acc, self.product(den))
            return self.node(NEG, acc) if neg else acc
        if e.is_Pow and e.exp.is_Integer:
            p = int(e.exp)
            if p > 0:
                return self.power(e.base, p)
            return self.node(DIV, self.const(1), self.power(e.base, -p))
        raise ValueError("closed-form LU: unsupported expression %s" % e)


def symbolic_p_order(st: Structure):
    """P_order = (I - hb*J)[perm][:, perm] from the structure's terms, with
    exact rational coefficients; returns (P_order, y symbols, k symbols, hb)."""
    n = st.neq
    ys = sp.symbols("y0:%d" % n)
    nrate = max([t.rate for t in st.terms] + [-1]) + 1
    ks = sp.symbols("k0:%d" % max(nrate, 1))
    hb = sp.Symbol("hb")
    f = [sp.Integer(0)] * n
    for t in st.terms:
        m = sp.Rational(Fraction(t.coef)) * (ks[t.rate] if t.rate >= 0 else 1)
        for i in t.factors:
            m = m * ys[i]
        f[t.row] = f[t.row] + m
    J = sp.Matrix(f).jacobian(ys)
    P = sp.eye(n) - hb * J
    perm = [int(p) for p in st.perm]
    Po = sp.Matrix(n, n, lambda i, j: P[perm[i], perm[j]])
    return Po, ys, ks, hb


def closed_form_lu(st: Structure) -> dict:
    """The straight-line program of the closed-form factors of the
    structure's P_order: ``{"ops": [...], "out": [[row, col, node], ...]}``,
    one output per entry of the LU pattern (row-major), every entry of the
    symbolic factors that is not identically zero inside the pattern."""
    Po, ys, ks, hb = symbolic_p_order(st)
    LU, piv = Po.LUdecomposition_Simple()
    if piv:
        raise ValueError("closed-form LU needs no row exchange; SymPy pivoted %s" % (piv,))
    n = st.neq
    pat = {tuple(map(int, e)) for e in st.csc.entries}
    prog = _Program(ys, ks, hb)
    out = []
    for r in range(n):
        for c in range(n):
            v = LU[r, c]
            if v == 0:
                continue
            if (r, c) not in pat:
                raise ValueError("closed-form LU entry (%d, %d) outside the LU pattern" % (r, c))
            out.append([r, c, prog.expr(v)])
    return {"ops": prog.ops, "out": out}


def opcount(prog: dict) -> Dict[str, int]:
    cnt = {"add": 0, "mul": 0, "div": 0}
    for code, _, _, _ in prog["ops"]:
        if code in (ADD, SUB):
            cnt["add"] += 1
        elif code == MUL:
            cnt["mul"] += 1
        elif code == DIV:
            cnt["div"] += 1
    cnt["flops"] = cnt["add"] + cnt["mul"] + cnt["div"]
    return cnt


def emit_c(prog: dict, lit, store) -> Tuple[List[str], Dict[str, int]]:
    """C statements computing the program: leaves inline (``y[i]``,
    ``k[r * S]``, ``hb``, literals), every other operation one ``const
    double`` temporary, then ``store(row, col, expr)`` for each output."""
    names: List[str] = []
    lines: List[str] = []
    for i, (code, a, b, c) in enumerate(prog["ops"]):
        if code == Y:
            names.append("y[%d]" % a)
        elif code == K:
            names.append("k[%d * S]" % a)
        elif code == HB:
            names.append("hb")
        elif code == CONST:
            names.append(lit(c))
        else:
            t = "u%d" % i
            if code == NEG:
                rhs = "-%s" % names[a]
            else:
                op = {ADD: "+", SUB: "-", MUL: "*", DIV: "/"}[code]
                rhs = "%s %s %s" % (names[a], op, names[b])
            lines.append("    const double %s = %s;" % (t, rhs))
            names.append(t)
    for r, c, node in prog["out"]:
        lines.append(store(r, c, names[node]))
    return lines, opcount(prog)

