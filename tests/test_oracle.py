"""Pins of the oracle (the C restatement of the reference solver) against
what fixes each part independently of it: closed forms, the reference's own
network definitions and rate text (golden vectors), finite differences,
textbook identities and an independent Radau integrator."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from odevec_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_bdf_coefficients_closed_form(oracle):
    """Nordsieck l-vector of order q: coefficients of prod_{i=1..q}(1 + x/i),
    normalised to l_1 = 1 (src/odevec.F90:127-134)."""
    tab = oracle.coeff_table()
    for q in range(1, 7):
        poly = [Fraction(1)]
        for i in range(1, q + 1):
            nxt = [Fraction(0)] * (len(poly) + 1)
            for d, c in enumerate(poly):
                nxt[d] += c
                nxt[d + 1] += c / i
            poly = nxt
        l = [c / poly[1] for c in poly]
        for kk in range(7):
            want = float(l[kk]) if kk < len(l) else 0.0
            assert tab[q - 1, kk] == pytest.approx(want, rel=1e-15, abs=0.0), (q, kk)


def _net(oracle, nid):
    return oracle.Network.load(nid)


@pytest.mark.parametrize("name,nid", [("ROBER", "ROBER/NONE/DENSE"), ("OREGO", "OREGO/NONE/DENSE"),
                                      ("PRIMORDIAL", "PRIMORDIAL/CMK/CSC")])
def test_rhs_matches_reference_definitions(oracle, name, nid):
    gold = json.load(open(os.path.join(GOLD, "reference_rhs.json")))["cases"][name]
    net = _net(oracle, nid)
    for pt in gold["points"]:
        k = pt["k"] if name == "PRIMORDIAL" else ([0.04, 3e7, 1e4] if name == "ROBER" else [])
        f = net.rhs(np.array(pt["y"]), np.array(k))
        ref = np.array(pt["rhs"])
        scale = np.max(np.abs(ref))
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-13 * scale), name


def test_primordial_rhs_conservation(oracle):
    net = _net(oracle, "PRIMORDIAL/CMK/CSC")
    H = np.array([0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0.0])
    He = np.array([0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0.0])
    D = np.array([0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0.0])
    Q = np.array([-1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 0, 0, 0, 0.0])
    rng = np.random.default_rng(3)
    for _ in range(20):
        y = np.exp(rng.uniform(-5, 5, 16))
        k = np.exp(rng.uniform(-5, 0, 38))
        f = net.rhs(y, k)
        for v in (H, He, D, Q):
            assert abs(v @ f) <= 1e-12 * np.abs(v[:, None] * f).sum()


def test_rates_match_reference_text(oracle):
    gold = json.load(open(os.path.join(GOLD, "reference_rates.json")))
    for c in gold["cases"]:
        y = np.zeros(16)
        y[14] = c["Tgas"]
        y[2] = c["nmax"] if c["nmax"] > c["Tgas"] else 0.0
        k = oracle.primordial_rates(y)
        assert np.allclose(k, c["k"], rtol=1e-12, atol=0.0), c["Tgas"]


def test_deterministic_math(oracle):
    L = oracle.lib()
    import ctypes as C
    for fn in ("oracle_exp", "oracle_log"):
        getattr(L, fn).restype = C.c_double
        getattr(L, fn).argtypes = [C.c_double]
    L.oracle_pow.restype = C.c_double
    L.oracle_pow.argtypes = [C.c_double, C.c_double]
    rng = np.random.default_rng(0)
    for x in rng.uniform(-700, 700, 2000):
        assert L.oracle_exp(x) == pytest.approx(math.exp(x), rel=4e-16)
    for x in np.exp(rng.uniform(-700, 700, 2000)):
        assert L.oracle_log(x) == pytest.approx(math.log(x), rel=4e-16, abs=4e-16)
    for x, yv in zip(np.exp(rng.uniform(-20, 20, 500)), rng.uniform(-3, 3, 500)):
        assert L.oracle_pow(x, yv) == pytest.approx(x ** yv, rel=1e-14)
    assert L.oracle_exp(-800.0) == 0.0 and math.isinf(L.oracle_exp(800.0))


@pytest.mark.parametrize("nid", ["ROBER/NONE/DENSE", "OREGO/NONE/CSC", "PRIMORDIAL/CMK/CSC", "PRIMORDIAL/NONE/DENSE"])
def test_jacobian_finite_differences(oracle, nid):
    net = _net(oracle, nid)
    rng = np.random.default_rng(1)
    n = net.neq
    y = np.exp(rng.uniform(-2, 2, n))
    k = np.exp(rng.uniform(-3, 0, max(net.nrate, 1)))
    hb = 0.37
    P = net.p_dense(y, k, hb)
    perm = net.ir["perm"]
    J = np.zeros((n, n))
    for j in range(n):
        h = 1e-6 * y[j]
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        J[:, j] = (net.rhs(yp, k) - net.rhs(ym, k)) / (2 * h)
    Pfd = np.eye(n) - hb * J
    Pfd = Pfd[np.ix_(perm, perm)]
    assert np.allclose(P, Pfd, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("nid", ["ROBER/NONE/DENSE/LU3", "PRIMORDIAL/NONE/DENSE/LU3"])
def test_calcjac_forward_differences(oracle, nid):
    """LUmethod 3 (CalcJac_dense): the forward-difference P against the
    analytic P of the same network. With steps r_k = 2^-26 |y_k| (the
    r0/sqrt(w_k) branch is far smaller at these weights) an entry's error
    is at most |dt*beta| * (truncation r_k |d2f_j/dy_k2| + rounding
    eps*T_j/r_k), T_j the sum of |terms| of row j; both are below
    |dt*beta| * T_j/|y_k| * 1e-6, about 1e-6 of the entry's natural scale,
    so a wrong step, sign or column fails."""
    net = _net(oracle, nid)
    sym = _net(oracle, nid[:-4])
    rng = np.random.default_rng(7)
    n = net.neq
    for trial in range(5):
        y = np.exp(rng.uniform(-3, 1, n))
        k = np.exp(rng.uniform(-3, 0, max(net.nrate, 1))) if net.nrate else np.zeros(1)
        if "PRIMORDIAL" in nid:
            y[net.ir["tgas_index"]] = 10.0 ** rng.uniform(1, 4)
            k = oracle.primordial_rates(y)
        w = 1.0 / (1e-6 * y + 1e-20) ** 2
        dt, beta = 10.0 ** rng.uniform(-8, -2), 6.0 / 11.0
        Pfd = net.fd_p(y, k, w, dt, beta)
        Pan = sym.p_dense(y, k, dt * beta)
        assert np.array_equal(np.diag(Pfd) != 0, np.ones(n, dtype=bool))
        # the analytic structural zeros are exact zeros of the differences too
        assert (Pfd[(Pan == 0)] == 0).all()
        T = np.zeros(n)
        for row, coef, rate, fac in net.ir["terms"]:
            T[row] += abs(coef * (k[rate] if rate >= 0 else 1.0) * np.prod([y[f] for f in fac]))
        bound = abs(dt * beta) * T[:, None] / y[None, :] * 1e-6
        err = np.abs(Pfd - Pan)
        assert (err <= bound + 4e-16 * np.abs(Pan)).all(), np.max(err / (bound + 1e-300))
        # the bound is tight enough to see the Jacobian: most off-diagonal
        # entries exceed it by orders of magnitude
        off = (Pan != 0) & ~np.eye(n, dtype=bool)
        assert np.median(np.abs(Pan[off]) / bound[off]) > 1e3


def test_lumethod3_solution_agrees_with_symbolic(oracle):
    """A ROBER solve with the difference Jacobian stays within the
    tolerance of the symbolic-Jacobian solve (only the Newton iteration
    matrix differs; converged steps satisfy the same equations)."""
    from odevec_amd import workloads as W
    w = W.rober(16, seed=3)
    a = _net(oracle, "ROBER/NONE/DENSE").solve_cells(w["y"], w["t_stop"], params=w["params"], rtol=1e-6,
                                                     atol=1e-10)
    b = _net(oracle, "ROBER/NONE/DENSE/LU3").solve_cells(w["y"], w["t_stop"], params=w["params"], rtol=1e-6,
                                                         atol=1e-10)
    assert (a[1] == 0).all() and (b[1] == 0).all()
    assert np.all(np.abs(a[0] - b[0]) <= 10 * 1e-6 * np.abs(a[0]) + 1e-9)
    assert not np.array_equal(a[0], b[0])


def _unpack(net, lu):
    """L and U from the factor storage (Crout: unit U; Doolittle: unit L)."""
    n = net.neq
    Lm, Um = np.zeros((n, n)), np.zeros((n, n))
    if net.csc:
        A = np.zeros((n, n))
        for s, (r, c) in enumerate(net.ir["lu_entries"]):
            A[r, c] = lu[s]
        Lm = np.tril(A)
        Um = np.triu(A, 1) + np.eye(n)
    else:
        A = lu.reshape(n, n)
        Lm = np.tril(A, -1) + np.eye(n)
        Um = np.triu(A)
    return Lm, Um


@pytest.mark.parametrize("nid", ["ROBER/NONE/DENSE", "OREGO/NONE/CSC", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC",
                                 "PRIMORDIAL/FF/CSC", "PRIMORDIAL/NONE/DENSE", "KROME/FF/CSC"])
def test_lu_identity_and_solve(oracle, nid):
    """L*U = P_order and P * den = res for the permuted system."""
    net = _net(oracle, nid)
    rng = np.random.default_rng(2)
    n = net.neq
    for _ in range(5):
        y = np.exp(rng.uniform(-3, 3, n))
        k = np.exp(rng.uniform(-3, 1, max(net.nrate, 1)))
        hb = float(np.exp(rng.uniform(-4, 1)))
        P = net.p_dense(y, k, hb)
        lu = net.factor(y, k, hb)
        Lm, Um = _unpack(net, lu)
        assert np.allclose(Lm @ Um, P, rtol=1e-12, atol=1e-12 * np.abs(P).max())
        res = rng.standard_normal(n)
        den = net.lin_solve(lu, res)
        perm = net.ir["perm"]
        assert np.allclose(P @ den[perm], res[perm], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("nid", ["ROBER/NONE/DENSE/LU2", "OREGO/NONE/DENSE/LU2"])
def test_closed_form_lu_reproduces_p(oracle, nid):
    """LUmethod 2 (GetLU): the oracle's closed-form factors -- unit L below
    the diagonal, U on and above it -- multiply back to P_order, evaluated
    independently term by term (p_dense), to within a few roundings of
    |L| |U|, over stiff and non-stiff points (hb from 1e-6 to 1e4, ROBER rate
    constants scattered over a factor e^3 around their defaults); a wrong
    sign, index or dropped term fails by orders of magnitude. The factors
    also agree with the numeric Doolittle factorization of the same P
    (LUmethod 1) to rounding."""
    net = _net(oracle, nid)
    sym = _net(oracle, nid[:-4])
    n = net.neq
    assert net.ir["lu_method"] == 2
    pat = {tuple(e) for e in net.ir["lu_entries"]}
    assert {(r, c) for r, c, _ in net.ir["lu_closed"]["out"]} <= pat
    rng = np.random.default_rng(9)
    for _ in range(500):
        y = 10.0 ** rng.uniform(-8, 0, n)
        prm = net.param_defaults * np.exp(rng.uniform(-1.5, 1.5, net.nparam)) if net.nparam else None
        k = net.rates(y, prm)
        hb = 10.0 ** rng.uniform(-6, 4)
        A = net.factor(y, k, hb).reshape(n, n)
        Lm, Um = np.tril(A, -1) + np.eye(n), np.triu(A)
        P = sym.p_dense(y, k, hb)
        scale = np.abs(Lm) @ np.abs(Um)
        assert (np.abs(Lm @ Um - P) <= 8 * 2.0 ** -52 * scale).all()
        B = sym.factor(y, k, hb).reshape(n, n)
        assert np.allclose(A, B, rtol=1e-13, atol=1e-15 * np.abs(B).max())
        res = rng.standard_normal(n)
        assert np.allclose(net.lin_solve(A.ravel(), res), sym.lin_solve(B.ravel(), res), rtol=1e-10, atol=1e-12)


def test_closed_form_lu_program_is_the_symbolic_factorization():
    """The program's expressions, evaluated in exact rational arithmetic at
    rational points, equal the exact Doolittle factors of P_order there
    (Fractions, no SymPy): the program is the factorization, not just close
    to it."""
    from fractions import Fraction
    from odevec_amd.pre_odevec import lu_closed, networks, structure
    for name in ("ROBER", "OREGO"):
        st = structure.analyse(networks.load(name), "None", "DENSE")
        prog = lu_closed.closed_form_lu(st)
        rng = np.random.default_rng(4)
        n = st.neq
        for _ in range(3):
            y = [Fraction(int(v), 97) for v in rng.integers(1, 500, n)]
            k = [Fraction(int(v), 13) for v in rng.integers(1, 50, 3)]
            hb = Fraction(int(rng.integers(1, 40)), 7)
            # P = I - hb*J from the structure's terms (derivative of each monomial)
            P = [[Fraction(int(r == c)) for c in range(n)] for r in range(n)]
            for t in st.terms:
                for j in set(t.factors):
                    red = list(t.factors)
                    red.remove(j)
                    v = Fraction(t.coef) * t.factors.count(j) * (k[t.rate] if t.rate >= 0 else 1)
                    for f in red:
                        v *= y[f]
                    P[t.row][j] -= hb * v
            A = [row[:] for row in P]
            for q in range(n - 1):
                for i in range(q + 1, n):
                    A[i][q] /= A[q][q]
                    for j in range(q + 1, n):
                        A[i][j] -= A[i][q] * A[q][j]
            val = []
            for code, a, b, c in prog["ops"]:
                val.append([lambda: y[a], lambda: k[a], lambda: hb, lambda: Fraction(c),
                            lambda: val[a] + val[b], lambda: val[a] - val[b], lambda: val[a] * val[b],
                            lambda: val[a] / val[b], lambda: -val[a]][code]())
            got = {(r, c): val[node] for r, c, node in prog["out"]}
            for r in range(n):
                for c in range(n):
                    exact = A[r][c]
                    if (r, c) in got:
                        # constants are doubles: the only inexact inputs
                        assert abs(got[(r, c)] - exact) <= Fraction(1, 10 ** 12) * (abs(exact) + 1), (name, r, c)
                    else:
                        assert exact == 0, (name, r, c)


def test_lumethod2_solution_agrees_with_lumethod1(oracle):
    """Solves with the closed-form factors stay within the tolerance of the
    LUmethod 1 solves (only the rounding of the iteration matrix differs)
    and converge to the Radau solutions like them."""
    gold = json.load(open(os.path.join(GOLD, "radau_known_answers.json")))
    for name in ("ROBER", "OREGO"):
        case = gold["cases"][name]
        net = _net(oracle, case["network"] + "/LU2")
        for c in case["cells"]:
            y0 = np.array(c["y0"])[:, None]
            prm = None if c["params"] is None else np.array(c["params"])[:, None]
            ref = np.array(c["y"])
            big = np.abs(ref) > 1e-6 * np.abs(ref).max()
            y, st, _, _ = net.solve_cells(y0, c["t_stop"], params=prm, rtol=1e-8, atol=1e-12, dt_min=1e-30)
            assert st[0] == 0
            assert np.max(np.abs(y[:, 0] - ref)[big] / np.abs(ref[big])) < 2e-6, (name, c["cell"])
    from odevec_amd import workloads as W
    w = W.rober(16, seed=3)
    a = _net(oracle, "ROBER/NONE/DENSE").solve_cells(w["y"], w["t_stop"], params=w["params"], rtol=1e-6,
                                                     atol=1e-10)
    b = _net(oracle, "ROBER/NONE/DENSE/LU2").solve_cells(w["y"], w["t_stop"], params=w["params"], rtol=1e-6,
                                                         atol=1e-10)
    assert (a[1] == 0).all() and (b[1] == 0).all()
    assert np.all(np.abs(a[0] - b[0]) <= 10 * 1e-6 * np.abs(a[0]) + 1e-9)


def test_known_answers_radau(oracle):
    """The oracle converges to SciPy Radau (rtol 1e-12) solutions."""
    gold = json.load(open(os.path.join(GOLD, "radau_known_answers.json")))
    for name, case in gold["cases"].items():
        net = _net(oracle, case["network"])
        for c in case["cells"]:
            y0 = np.array(c["y0"])[:, None]
            prm = None if c["params"] is None else np.array(c["params"])[:, None]
            ref = np.array(c["y"])
            big = np.abs(ref) > 1e-6 * np.abs(ref).max()
            atol = 1e-20 if name in ("PRIMORDIAL", "KROME_ISM") else 1e-12
            y, st, _, _ = net.solve_cells(y0, c["t_stop"], params=prm, rtol=1e-8, atol=atol, dt_min=1e-30)
            assert st[0] == 0
            err = np.max(np.abs(y[:, 0] - ref)[big] / np.abs(ref[big]))
            assert err < 2e-6, (name, c["cell"], err)


def test_rober_rtol_1e6_within_10rtol_of_radau(oracle):
    gold = json.load(open(os.path.join(GOLD, "radau_known_answers.json")))["cases"]["ROBER"]
    net = _net(oracle, gold["network"])
    for c in gold["cells"]:
        y, st, _, _ = net.solve_cells(np.array(c["y0"])[:, None], c["t_stop"], params=np.array(c["params"])[:, None],
                                      rtol=1e-6, atol=1e-10)
        ref = np.array(c["y"])
        assert np.all(np.abs(y[:, 0] - ref) <= 10 * 1e-6 * np.abs(ref) + 1e-10)


def test_lockstep_of_identical_cells_equals_percell(oracle):
    """With G identical cells every max-norm equals each cell's norm, so the
    vector solve reproduces the single-cell solve exactly."""
    net = _net(oracle, "PRIMORDIAL/CMK/CSC")
    w = W.primordial(1)
    y1, st1, _, _ = net.solve_cells(w["y"], w["t_stop"][0], dt0=w["dt0"], rtol=1e-6, atol=1e-20, dt_min=1e-20)
    yv, st, t, d, cnt = net.solve_vector(np.repeat(w["y"], 8, axis=1), 0.0, w["t_stop"][0], w["dt0"][0], 1e-6, 1e-20,
                                         1e-20)
    assert st == 0 and np.array_equal(yv, np.repeat(y1, 8, axis=1))


def test_zero_span_and_mask(oracle):
    net = _net(oracle, "ROBER/NONE/DENSE")
    w = W.rober(4)
    y, st, _, _ = net.solve_cells(w["y"], 0.0, params=w["params"])     # t_stop == t0
    assert np.array_equal(y, w["y"]) and (st == 5).all()
    yv, st, t, d, cnt = net.solve_vector(w["y"], 0.0, 10.0, 1e-10, 1e-6, 1e-10, params=w["params"],
                                         mask=[1, 0, 1, 0])
    assert st == 0 and np.array_equal(yv[:, 1], w["y"][:, 1]) and not np.array_equal(yv[:, 0], w["y"][:, 0])


def test_failure_paths_on_cpu(oracle):
    """Host-forced failures: a MinimalTimestep above every step fails with
    DT_MIN on the first convergence failure path, a tiny attempt budget with
    MAX_STEPS; failed cells return y_NS(:,:,0)."""
    net = _net(oracle, "PRIMORDIAL/CMK/CSC")
    w = W.primordial(8)
    y, st, _, cnt = net.solve_cells(w["y"], w["t_stop"], dt0=w["dt0"], rtol=1e-6, atol=1e-20, dt_min=1e-20,
                                    max_attempts=3)
    assert (st == 4).all() and (cnt[0] == 4).all()
    yneg = w["y"].copy()
    yneg[0, :] = -1.0
    y, st, _, _ = net.solve_cells(yneg, w["t_stop"], dt0=w["dt0"], rtol=1e-6, atol=1e-20, dt_min=1e300,
                                  check_negatives=True)
    assert (st == 1).all()


# --------------------------------------------------------------- SYNTH64
# SYNTH64 has no reference definition beyond its specification (64 species,
# 300 seeded two-body reactions A + B -> C + D, k = A exp(-Ea/T) with
# log10 A ~ U[-12,-8] and Ea/k ~ U[0, 5e4] K, FF ordering). These tests pin
# the oracle's SYNTH64 tables (which come from the preprocessor's term
# table) against that specification, evaluated here from the reaction list
# itself, and against the textbook identities the other networks are pinned
# by.

def _synth_list():
    from odevec_amd.pre_odevec.networks import synth_reactions
    reactions, A, Ea = synth_reactions(64, 300)
    ra = np.array([r[0][0] for r in reactions])
    rb = np.array([r[0][1] for r in reactions])
    pc = np.array([r[1][0] for r in reactions])
    pd = np.array([r[1][1] for r in reactions])
    return ra, rb, pc, pd, np.array(A), np.array(Ea)


def _synth_rhs(y, k):
    ra, rb, pc, pd, _, _ = _synth_list()
    v = k * y[ra] * y[rb]
    f = np.zeros(64)
    for idx, sgn in ((ra, -1.0), (rb, -1.0), (pc, 1.0), (pd, 1.0)):
        np.add.at(f, idx, sgn * v)
    return f


def test_synth64_generator_meets_its_specification():
    ra, rb, pc, pd, A, Ea = _synth_list()
    assert len(ra) == 300
    for r in zip(ra, rb, pc, pd):
        assert len(set(int(x) for x in r)) == 4 and all(0 <= x < 64 for x in r)
    assert np.all((np.log10(A) >= -12) & (np.log10(A) <= -8)) and np.all((Ea >= 0) & (Ea <= 5e4))
    # spread over the stated ranges (not a degenerate draw)
    assert np.ptp(np.log10(A)) > 3.9 and np.ptp(Ea) > 4.9e4


def test_synth64_rates_are_arrhenius(oracle):
    net = _net(oracle, "SYNTH64/FF/CSC")
    _, _, _, _, A, Ea = _synth_list()
    for T in (100.0, 1234.5, 1e4):
        k = net.rates(np.ones(64), np.array([T]))
        assert np.allclose(k, A * np.exp(-Ea / T), rtol=1e-13, atol=0.0), T


def test_synth64_rhs_is_mass_action_of_its_reactions(oracle):
    net = _net(oracle, "SYNTH64/FF/CSC")
    rng = np.random.default_rng(64)
    for _ in range(10):
        y = np.exp(rng.uniform(-10, 5, 64))
        k = np.exp(rng.uniform(-20, 0, 300))
        f = net.rhs(y, k)
        ref = _synth_rhs(y, k)
        scale = np.abs(k * y[_synth_list()[0]] * y[_synth_list()[1]]).max()
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-14 * scale)
        # every reaction conserves the molecule count
        assert abs(f.sum()) <= 1e-13 * np.abs(f).sum() + 1e-300


def test_synth64_jacobian_and_lu(oracle):
    """Central-difference Jacobian against the oracle's P, and L*U = P on the
    3,155-entry FF pattern, with random states and rates."""
    net = _net(oracle, "SYNTH64/FF/CSC")
    rng = np.random.default_rng(65)
    perm = net.ir["perm"]
    for _ in range(3):
        y = np.exp(rng.uniform(-2, 2, 64))
        k = np.exp(rng.uniform(-3, 0, 300))
        hb = float(np.exp(rng.uniform(-3, 0)))
        P = net.p_dense(y, k, hb)
        J = np.zeros((64, 64))
        for j in range(64):
            h = 1e-6 * y[j]
            yp, ym = y.copy(), y.copy()
            yp[j] += h
            ym[j] -= h
            J[:, j] = (_synth_rhs(yp, k) - _synth_rhs(ym, k)) / (2 * h)
        Pfd = (np.eye(64) - hb * J)[np.ix_(perm, perm)]
        assert np.allclose(P, Pfd, rtol=1e-6, atol=1e-8)
        lu = net.factor(y, k, hb)
        Lm, Um = _unpack(net, lu)
        assert np.allclose(Lm @ Um, P, rtol=1e-11, atol=1e-11 * np.abs(P).max())
        res = rng.standard_normal(64)
        den = net.lin_solve(lu, res)
        assert np.allclose(P @ den[perm], res[perm], rtol=1e-9, atol=1e-9)


def test_synth64_ff_ordering_and_fill_from_the_reaction_list(oracle):
    """The FF permutation is NumPy's default argsort of nnz(column) + nnz(row)
    - 1 of P's pattern, as the reference calls it (pre_odevec.py:343-350; not
    a stable sort, so ties keep NumPy's order), and the LU pattern is the
    structural fill of pivot-free elimination of the permuted pattern; both
    recomputed here from the reaction list (the Jacobian pattern: row r reads
    species s when s is a reactant of a reaction that changes r)."""
    net = _net(oracle, "SYNTH64/FF/CSC")
    ra, rb, pc, pd, _, _ = _synth_list()
    Pat = np.eye(64, dtype=bool)
    for a, b, c, d in zip(ra, rb, pc, pd):
        for row in (a, b, c, d):
            Pat[row, a] = Pat[row, b] = True
    score = Pat.sum(axis=0) + Pat.sum(axis=1) - 1
    perm = np.argsort(score)
    assert list(perm) == list(net.ir["perm"])
    assert np.all(np.diff(score[perm]) >= 0)
    F = Pat[np.ix_(perm, perm)].copy()
    for kk in range(64):
        rows = np.where(F[kk + 1:, kk])[0] + kk + 1
        cols = np.where(F[kk, kk + 1:])[0] + kk + 1
        F[np.ix_(rows, cols)] = True
    assert int(F.sum()) == net.ir["nnz"] == 3155
    assert sorted(map(tuple, net.ir["lu_entries"])) == sorted(zip(*np.where(F.T)[::-1]))


def _krome_file():
    """The KROME setup file executed as pre_odevec.py:97-105 executes it:
    its own y, k and rhs (SymPy), independent of the preprocessor."""
    import sympy as sp
    from odevec_amd.build import KROME_SETUPFILE
    scope = {}
    exec("from sympy import *", scope)
    with open(KROME_SETUPFILE) as fh:
        exec(fh.read(), scope)
    y, k, rhs = list(scope["y"]), list(scope["k"]), sp.Matrix(scope["rhs"])
    return (scope, sp.lambdify([y, k], list(rhs), "numpy"),
            sp.lambdify([y, k], rhs.jacobian(y).tolist(), "numpy"))


def test_krome_setupfile_rhs_jacobian_and_conservation(oracle):
    """KROME/FF/CSC: the oracle's RHS and P equal the setup file's own SymPy
    expressions at random states and rate coefficients (rate i is k(:,i) of
    the file, so all 21 parameters reach the network), and every RHS
    conserves H, C and O nuclei and charge (species e-, H, H2, H+, H3+, C,
    C+, O, OH, H2O, CO, HCO+)."""
    scope, f_sym, j_sym = _krome_file()
    assert (scope["neq"], scope["nrea"]) == (12, 21)
    net = _net(oracle, "KROME/FF/CSC")
    assert (net.neq, net.nrate, net.nparam) == (12, 21, 21)
    nuc = np.array([[0, 1, 2, 1, 3, 0, 0, 0, 1, 2, 0, 1],       # H
                    [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 1],       # C
                    [0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1],       # O
                    [-1, 0, 0, 1, 1, 0, 1, 0, 0, 0, 0, 1]], float)  # charge
    perm = net.ir["perm"]
    rng = np.random.default_rng(21)
    for _ in range(10):
        y = np.exp(rng.uniform(-8, 8, 12))
        k = np.exp(rng.uniform(-20, 0, 21))
        assert np.array_equal(net.rates(y, k), k)
        f = net.rhs(y, k)
        ref = np.array(f_sym(y, k), dtype=float)
        assert np.allclose(f, ref, rtol=1e-13, atol=1e-14 * np.abs(ref).max())
        for v in nuc:
            assert abs(v @ f) <= 1e-13 * (np.abs(v) @ np.abs(f)) + 1e-300
        hb = float(np.exp(rng.uniform(-4, 1)))
        P = np.eye(12) - hb * np.array(j_sym(y, k), dtype=float)
        assert np.allclose(net.p_dense(y, k, hb), P[np.ix_(perm, perm)], rtol=1e-13, atol=1e-13 * np.abs(P).max())
    # each rate coefficient moves exactly the species of its reaction
    y = np.exp(rng.uniform(-3, 3, 12))
    base = net.rhs(y, np.zeros(21))
    assert np.all(base == 0.0)
    for i in range(21):
        e = np.zeros(21)
        e[i] = 1.0
        assert np.allclose(net.rhs(y, e), np.array(f_sym(y, e), dtype=float), rtol=1e-14, atol=0.0), i


@pytest.mark.parametrize("t_stop", [2.0 ** 10, 2.0 ** 14, 3.0])
def test_step_controller_on_a_constant_solution(oracle, t_stop):
    """The step-size controller where its outcome is fixed in closed form:
    with every ROBER rate 0 the right-hand side is 0, so every predictor is
    exact, every correction 0, and the controller's candidates are all
    1/(factor * 0 + 1e-6) = 1e6. The reference's rules then fix the whole
    step sequence: the initial step sqrt(tol / (W0^-2 + 0)) with W0 =
    max(|t0|, |t_stop|) (src/odevec.F90:224-236, capped at the span), a step
    selection after every q + 1 = 2 successful steps (:447-450) that keeps
    order 1 (up does not beat same on a tie: :570-600) and scales the step
    by the cap, 1e4 at the first selection and 10 after (:605-611), the
    last step overshooting t_stop (no clamping). A mis-stated cap, cadence,
    tie rule or initial-step formula changes the steps, the selections or
    the last step size."""
    net = _net(oracle, "ROBER/NONE/DENSE")
    rtol = 2.0 ** -40                       # powers of two: the initial step is exact
    y0 = np.array([[1.0], [0.0], [0.0]])
    y, st, dto, cnt = net.solve_cells(y0, t_stop, params=np.zeros((3, 1)), rtol=rtol, atol=1e-10,
                                      dt_min=1e-300)
    # the sequence those rules give
    W0 = max(0.0, t_stop)
    h = min(math.sqrt(rtol / W0 ** -2.0), t_stop)
    t, steps, sel, swu, first, jac = 0.0, 0, 0, 0, True, 1
    while t < t_stop:
        t += h
        steps += 1
        swu += 1
        if swu == 2:
            h = h * (1e4 if first else 10.0)
            first, swu, sel = False, 0, sel + 1
            # the step change (>= 10, far past the 0.3 threshold, :454-458)
            # refreshes the Jacobian at the next attempt, if there is one
            jac += 1 if t < t_stop else 0
    assert st[0] == 0 and np.array_equal(y, y0)
    # attempts, steps, Jacobian refreshes, selections
    assert (cnt[0][0], cnt[1][0], cnt[3][0], cnt[6][0]) == (steps, steps, jac, sel)
    # W0^-2 goes through the restated pow (exp(-2 log W0)), within an ulp of 2^-2k
    assert dto[0] == pytest.approx(h, rel=1e-14, abs=0.0)
