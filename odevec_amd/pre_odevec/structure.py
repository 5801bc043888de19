"""Structural analysis of a network: term table, Jacobian and P patterns,
orderings (CMK / INVERT / FF), the LU fill pattern and the CSC index arrays.

The conventions follow the reference preprocessor:

* ``P = I - dt*beta*J`` (pre_odevec.py:474-475);
* orderings permute P symmetrically, ``P_order[i,j] = P[Perm[i],Perm[j]]``
  (pre_odevec.py:307-309, 328-330, 355-357), the RHS is not permuted
  (pre_odevec.py:305);
* CMK calls SciPy's reverse Cuthill-McKee on the transposed pattern built
  from the column-sorted entry list (pre_odevec.py:292-302);
* FF sorts by ``nnz(column) + nnz(row) - 1`` with NumPy's default argsort
  (pre_odevec.py:343-350);
* INVERT reverses the index order (pre_odevec.py:323);
* the LU pattern is stored column-major, rows ascending inside a column,
  with ``u_col_start`` (first entry of each column, plus nnz at the end) and
  ``l_col_start`` (the diagonal entry of each column) (pre_odevec.py:524-540).

The fill pattern is the structural fill of pivot-free Gaussian elimination.
The reference takes it from SymPy's symbolic ``LUdecomposition_Simple``
(pre_odevec.py:494), which can only be sparser where an entry cancels
exactly; ``sympy_lu_pattern`` recomputes that pattern for cross-checks.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np
import sympy as sp
from scipy import sparse
from scipy.sparse import csgraph

from .networks import Network

ORDERINGS = ("None", "CMK", "INVERT", "FF")
PACKAGINGS = ("DENSE", "CSC")


@dataclass(frozen=True)
class Term:
    """One monomial of a RHS row: ``coef * k[rate] * prod(y[f] for f in factors)``.

    ``rate`` is -1 when the term carries no rate symbol; ``factors`` lists the
    species with repetition for powers, ascending."""
    row: int
    coef: float
    rate: int
    factors: Tuple[int, ...]


def extract_terms(net: Network) -> List[Term]:
    """Split each (expanded, polynomial) RHS row into monomial terms.

    Terms inside a row are put in a canonical order (rate index, factors) so
    that every evaluator sums them in the same order."""
    yidx = {s: i for i, s in enumerate(net.y)}
    kidx = {s: i for i, s in enumerate(net.k)}
    out: List[Term] = []
    for row, expr in enumerate(net.rhs):
        expr = sp.expand(expr)
        if expr == 0:
            continue
        row_terms = []
        for t in sp.Add.make_args(expr):
            coef, facs = t.as_coeff_mul()
            rate = -1
            ys: List[int] = []
            for f in facs:
                base, ex = f.as_base_exp()
                if base in kidx:
                    if rate >= 0 or ex != 1:
                        raise ValueError("term %s: at most one rate symbol to the first power" % t)
                    rate = kidx[base]
                elif base in yidx:
                    if not (ex.is_Integer and int(ex) > 0):
                        raise ValueError("term %s: non-polynomial in y" % t)
                    ys.extend([yidx[base]] * int(ex))
                elif base.is_Number:
                    coef = coef * f
                else:
                    raise ValueError("term %s: unknown factor %s" % (t, f))
            row_terms.append(Term(row, float(coef), rate, tuple(sorted(ys))))
        row_terms.sort(key=lambda tm: (tm.rate, tm.factors, tm.coef))
        out.extend(row_terms)
    return out


def jacobian_pattern(neq: int, terms: List[Term]) -> np.ndarray:
    J = np.zeros((neq, neq), dtype=bool)
    for t in terms:
        for f in set(t.factors):
            J[t.row, f] = True
    return J


def p_pattern(neq: int, terms: List[Term]) -> np.ndarray:
    """Pattern of P = I - dt*beta*J: the Jacobian pattern plus the diagonal."""
    P = jacobian_pattern(neq, terms)
    P[np.arange(neq), np.arange(neq)] = True
    return P


def _size(v: List[int]) -> Tuple[int, int]:
    return sum(map(abs, v)), sum(1 for x in v if x)


def linear_invariants(neq: int, terms: List[Term]) -> List[List[int]]:
    """Integer basis of the network's linear invariants: the vectors c with
    sum_j c_j f_j(y) = 0 for every y (element and charge conservation of a
    mass-action network; PRIMORDIAL: H, He and D nuclei and charge; ROBER:
    y1 + y2 + y3; SYNTH64: the molecule count).

    f is a sum of monomials (rate x product of species), so c is an invariant
    exactly when it is orthogonal to every monomial's column of
    stoichiometric coefficients: the left null space of the species x
    monomial coefficient matrix, computed in exact rational arithmetic. The
    basis is the reduced row echelon form scaled to integers, then
    size-reduced (a vector is replaced by its sum or difference with another
    while that lowers its L1 norm, then its count of nonzeros), so that each
    vector involves few species (PRIMORDIAL: H, He and D nuclei, and charge
    combined with helium, E + H- + He - H+ - H2+ - D+ - He++)."""
    cols = {}
    for t in terms:
        col = cols.setdefault((t.rate, t.factors), {})
        col[t.row] = col.get(t.row, 0) + sp.nsimplify(t.coef, rational=True)
    M = sp.zeros(neq, max(len(cols), 1))
    for c, rows in enumerate(cols.values()):
        for r, v in rows.items():
            M[r, c] = v
    ns = M.T.nullspace()
    if not ns:
        return []
    R = sp.Matrix.hstack(*ns).T.rref()[0]
    basis = []
    for i in range(R.rows):
        row = list(R.row(i))
        den = sp.ilcm(*[sp.fraction(v)[1] for v in row])
        basis.append([int(v * den) for v in row])
    changed = True
    while changed:
        changed = False
        for a in range(len(basis)):
            for b in range(len(basis)):
                if a == b:
                    continue
                for sgn in (1, -1):
                    cand = [x + sgn * y for x, y in zip(basis[a], basis[b])]
                    if _size(cand) < _size(basis[a]):
                        basis[a] = cand
                        changed = True
    for v in basis:   # leading nonzero positive
        if next(x for x in v if x != 0) < 0:
            v[:] = [-x for x in v]
    return basis


def _colsorted_entries(P: np.ndarray) -> List[Tuple[int, int]]:
    n = P.shape[0]
    return [(i, j) for j in range(n) for i in range(n) if P[i, j]]


def perm_cmk(P: np.ndarray) -> np.ndarray:
    """pre_odevec.py:292-302: RCM on the pattern given as (col,row) pairs."""
    ent = _colsorted_entries(P)
    rows = np.array([e[0] for e in ent])
    cols = np.array([e[1] for e in ent])
    data = np.ones(len(ent), dtype=int)
    Psci = sparse.csc_matrix((data, (cols, rows)))
    return np.asarray(csgraph.reverse_cuthill_mckee(Psci, symmetric_mode=False), dtype=np.int64)


def perm_invert(P: np.ndarray) -> np.ndarray:
    return np.arange(P.shape[0])[::-1].copy()


def perm_ff(P: np.ndarray) -> np.ndarray:
    """pre_odevec.py:343-350: argsort of nnz(col)+nnz(row)-1 (default sort kind)."""
    n = P.shape[0]
    sorter = np.zeros(n)
    for i in range(n):
        sorter[i] = P[:, i].sum() + P[i, :].sum() - 1
    return np.arange(n)[sorter.argsort()]


def permutation(P: np.ndarray, ordering: str) -> np.ndarray:
    if ordering in (None, "None", "NONE", ""):
        return np.arange(P.shape[0])
    if ordering == "CMK":
        return perm_cmk(P)
    if ordering == "INVERT":
        return perm_invert(P)
    if ordering == "FF":
        return perm_ff(P)
    raise ValueError("unknown ordering %r (known: %s)" % (ordering, ", ".join(ORDERINGS)))


def permute_pattern(P: np.ndarray, perm: np.ndarray) -> np.ndarray:
    return P[np.ix_(perm, perm)]


def lu_fill(P: np.ndarray) -> np.ndarray:
    """Structural fill of pivot-free elimination (L and U share one pattern)."""
    A = P.copy()
    n = A.shape[0]
    for k in range(n):
        rows = [i for i in range(k + 1, n) if A[i, k]]
        cols = [j for j in range(k + 1, n) if A[k, j]]
        for i in rows:
            for j in cols:
                A[i, j] = True
    return A


def sympy_lu_pattern(P_sym: sp.Matrix) -> np.ndarray:
    """Pattern of SymPy's LUdecomposition_Simple of a symbolic matrix
    (the reference's own way of fixing nnz, pre_odevec.py:494)."""
    LU, piv = sp.Matrix(P_sym).LUdecomposition_Simple()
    if piv:
        raise ValueError("symbolic LU pivoted; pivot-free factorization not possible")
    n = LU.shape[0]
    return np.array([[LU[i, j] != 0 for j in range(n)] for i in range(n)], dtype=bool)


@dataclass
class CSC:
    entries: List[Tuple[int, int]]   # (row, col), column-major, rows ascending
    row_index: List[int]
    u_col_start: List[int]           # length neq+1
    l_col_start: List[int]           # length neq (diagonal position)

    @property
    def nnz(self) -> int:
        return len(self.entries)

    def index(self) -> dict:
        return {e: s for s, e in enumerate(self.entries)}


def csc_structure(LU: np.ndarray) -> CSC:
    """pre_odevec.py:524-540 (0-based)."""
    n = LU.shape[0]
    ent = _colsorted_entries(LU)
    u_col_start = [0] * (n + 1)
    l_col_start = [0] * n
    row_index = []
    j_old = 0
    for s, (i, j) in enumerate(ent):
        if j - 1 == j_old:
            u_col_start[j] = s
        if i == j:
            l_col_start[j] = s
        row_index.append(i)
        j_old = j
    u_col_start[n] = len(ent)
    for j in range(n):
        if not LU[j, j]:
            raise ValueError("missing diagonal in column %d" % j)
    return CSC(ent, row_index, u_col_start, l_col_start)


@dataclass
class Structure:
    net: Network
    ordering: str
    packaging: str
    terms: List[Term]
    J: np.ndarray
    P: np.ndarray
    perm: np.ndarray
    P_order: np.ndarray
    LU: np.ndarray
    csc: CSC

    # species integrated by the generated code (natural order); the others are
    # frozen: their RHS is identically zero and no RHS term reads them
    active: Optional[List[int]] = None

    # the reference's LUmethod: 1 symbolic Jacobian (GetJac) + LU; 3 the
    # Jacobian by forward differences of the RHS (CalcJac_dense,
    # src/odevec.F90:469-506) + LU
    lu_method: int = 1

    @property
    def neq(self) -> int:
        """Number of species in this structure's index space."""
        return self.P.shape[0]

    @property
    def neq_full(self) -> int:
        return self.net.neq

    @property
    def frozen(self) -> List[int]:
        act = set(self.active if self.active is not None else range(self.neq_full))
        return [i for i in range(self.neq_full) if i not in act]


def frozen_species(neq: int, terms: List[Term]) -> List[int]:
    """Species whose RHS row has no term and that no term reads. Their
    correction is exactly zero at every iteration (their row of P is the unit
    row, so the LU solve returns 0 for them and every update they feed is
    x - 0*a), hence y_NS(:,:,0) keeps their initial value and the higher
    Nordsieck columns stay zero."""
    has_row = {t.row for t in terms}
    read = {f for t in terms for f in t.factors}
    return [i for i in range(neq) if i not in has_row and i not in read]


def reduce_frozen(st: Structure) -> Structure:
    """The structure restricted to the active species: terms renumbered to
    active indices, the ordering restricted (relative order kept), and the LU
    pattern restricted to active rows and columns. Every elimination and solve
    operation that is dropped touches a frozen row or column and is an exact
    no-op (x - 0*a or x - a*0), so the remaining operations, in the same order,
    give the same values."""
    frozen = frozen_species(st.neq, st.terms)
    if not frozen:
        st.active = list(range(st.neq))
        return st
    active = [i for i in range(st.neq) if i not in frozen]
    amap = {s: a for a, s in enumerate(active)}
    terms = [Term(amap[t.row], t.coef, t.rate, tuple(amap[f] for f in t.factors)) for t in st.terms]
    keep_pos = [p for p in range(st.neq) if int(st.perm[p]) in amap]
    perm = np.array([amap[int(st.perm[p])] for p in keep_pos], dtype=np.int64)
    J = st.J[np.ix_(active, active)]
    P = st.P[np.ix_(active, active)]
    Po = st.P_order[np.ix_(keep_pos, keep_pos)]
    LU = st.LU[np.ix_(keep_pos, keep_pos)]
    red = Structure(st.net, st.ordering, st.packaging, terms, J, P, perm, Po, LU, csc_structure(LU))
    red.active = active
    red.lu_method = st.lu_method
    return red


def analyse(net: Network, ordering: str = "None", packaging: str = "CSC",
            lu_pattern: Optional[np.ndarray] = None) -> Structure:
    if packaging not in PACKAGINGS:
        raise ValueError("unknown packaging %r (known: %s)" % (packaging, ", ".join(PACKAGINGS)))
    ordering = "None" if ordering in (None, "", "NONE") else ordering
    terms = extract_terms(net)
    J = jacobian_pattern(net.neq, terms)
    P = p_pattern(net.neq, terms)
    perm = permutation(P, ordering)
    Po = permute_pattern(P, perm)
    LU = lu_fill(Po) if lu_pattern is None else lu_pattern
    return Structure(net, ordering, packaging, terms, J, P, perm, Po, LU, csc_structure(LU))
