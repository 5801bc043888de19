"""Preprocessor: network structure, orderings, LU fill, CSC conventions,
frozen species, emitted code and the command line (pre_odevec.py)."""
import json
import os

import numpy as np
import pytest
import sympy as sp
from scipy import sparse
from scipy.sparse import csgraph

from odevec_amd.pre_odevec import cli, emit, networks as N, structure as S

GEN = os.path.join(os.path.dirname(__file__), "..", "odevec_amd", "csrc", "generated")


@pytest.mark.parametrize("ordering,nnz", [("None", 143), ("CMK", 93), ("FF", 88), ("INVERT", 90)])
def test_primordial_structure_counts(ordering, nnz):
    st = S.analyse(N.load("PRIMORDIAL"), ordering, "CSC")
    assert int(st.J.sum()) == 81
    assert int(st.P.sum()) == 85
    assert len(st.terms) == 122
    assert st.csc.nnz == nnz


def test_small_networks_structure():
    rober = S.analyse(N.load("ROBER"), "None", "DENSE")
    assert (int(rober.P.sum()), rober.csc.nnz) == (8, 8)
    orego = S.analyse(N.load("OREGO"), "None", "DENSE")
    assert (int(orego.P.sum()), orego.csc.nnz) == (7, 8)
    assert orego.LU[2, 1] and not orego.P[2, 1]      # the one fill entry (3,2)


@pytest.mark.parametrize("name", ["ROBER", "OREGO", "PRIMORDIAL"])
@pytest.mark.parametrize("ordering", ["None", "CMK", "FF", "INVERT"])
def test_fill_equals_sympy_symbolic_lu(name, ordering):
    """pre_odevec.py:494 fixes nnz(LU) with SymPy's symbolic LU; the structural
    fill used here must give the same pattern."""
    net = N.load(name)
    st = S.analyse(net, ordering, "CSC")
    J = sp.Matrix(net.rhs).jacobian(net.y)
    dt, beta = sp.symbols("dt beta")
    P = sp.eye(net.neq) - dt * beta * J
    Po = sp.Matrix(net.neq, net.neq, lambda i, j: P[int(st.perm[i]), int(st.perm[j])])
    assert (S.sympy_lu_pattern(Po) == st.LU).all()


def test_orderings_follow_library_routines():
    st = S.analyse(N.load("PRIMORDIAL"), "None", "CSC")
    P = st.P
    # CMK: SciPy reverse Cuthill-McKee of the pattern (symmetrised by SciPy)
    rows, cols = np.nonzero(P)
    M = sparse.csc_matrix((np.ones(len(rows), dtype=int), (rows, cols)), shape=P.shape)
    assert list(S.perm_cmk(P)) == list(csgraph.reverse_cuthill_mckee(M.T.tocsc(), symmetric_mode=False))
    # FF: stable order of nnz(col) + nnz(row) - 1 for n <= 16
    key = P.sum(axis=0) + P.sum(axis=1) - 1
    assert list(S.perm_ff(P)) == list(np.argsort(key, kind="stable"))
    assert list(S.perm_invert(P)) == list(range(16))[::-1]


def test_csc_conventions():
    """u_col_start: first entry of each column (+ nnz), l_col_start: the
    diagonal, rows ascending within a column (pre_odevec.py:524-540)."""
    st = S.analyse(N.load("PRIMORDIAL"), "CMK", "CSC")
    cs = st.csc
    assert cs.u_col_start[0] == 0 and cs.u_col_start[-1] == cs.nnz
    for j in range(16):
        col = cs.entries[cs.u_col_start[j]:cs.u_col_start[j + 1]]
        assert all(c == j for _, c in col)
        assert [r for r, _ in col] == sorted(r for r, _ in col)
        assert cs.entries[cs.l_col_start[j]] == (j, j)


def test_frozen_species_primordial():
    st = S.analyse(N.load("PRIMORDIAL"), "CMK", "CSC")
    assert S.frozen_species(16, st.terms) == [12, 13, 14, 15]   # CR, g, Tgas, dummy
    red = S.reduce_frozen(S.analyse(N.load("PRIMORDIAL"), "CMK", "CSC"))
    assert red.neq == 12 and red.csc.nnz == 89 and red.active == list(range(12))
    assert S.frozen_species(3, S.analyse(N.load("ROBER"), "None", "DENSE").terms) == []


def test_primordial_reactions_conserve_elements_and_charge():
    """Mass action of the reaction list conserves H, He, D nuclei and charge."""
    H = [0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0, 0, 0, 0, 0]
    He = [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0]
    D = [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0]
    Q = [-1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 0, 0, 0, 0]
    for reac, prod in N.PRIMORDIAL_REACTIONS:
        for vec in (H, He, D, Q):
            assert sum(vec[s] for s in reac) == sum(vec[s] for s in prod), (reac, prod)


def test_emitted_files_are_current(tmp_path):
    """The committed generated headers/IR are what the preprocessor emits."""
    from odevec_amd.build import KROME_SETUPFILE
    for base in ("ROBER_NONE_DENSE", "OREGO_NONE_DENSE", "PRIMORDIAL_CMK_CSC", "PRIMORDIAL_FF_CSC", "KROME_FF_CSC"):
        name, ordering, packaging = base.split("_")
        cli.generate(name, "None" if ordering == "NONE" else ordering, packaging, str(tmp_path),
                     krome_setupfile=KROME_SETUPFILE if name == "KROME" else None)
        for f in ("net_%s.h" % base, "%s.json" % base, "inst_%s.hip" % base):
            with open(os.path.join(GEN, f)) as a, open(os.path.join(str(tmp_path), f)) as b:
                assert a.read() == b.read(), f


def test_opcounts_primordial_cmk():
    ir = json.load(open(os.path.join(GEN, "PRIMORDIAL_CMK_CSC.json")))
    oc = ir["opcounts_active"]
    # one multiply-subtract per elimination update, one division per U entry
    assert oc["lu"]["div"] == sum(1 for r, c in ir["lu_entries"] if r < c and r < 12 and c < 12)
    assert oc["solve"]["div"] == 12
    assert oc["rhs"]["flops"] < 383      # flux sharing beats the naive 383 flops


def test_cli_flags(tmp_path):
    assert cli.main(["--example=ROBER", "--nvector=32", "--outdir", str(tmp_path)]) == 0
    assert os.path.exists(os.path.join(str(tmp_path), "net_ROBER_NONE_DENSE.h"))
    with pytest.raises(SystemExit):   # closed-form LU: DENSE only
        cli.main(["--example=ROBER", "--LUmethod=2", "--packaging=CSC", "--outdir", str(tmp_path)])
    with pytest.raises(SystemExit):
        cli.main(["--example=ROBER", "--LUmethod=4", "--outdir", str(tmp_path)])
    assert cli.main(["--example=ROBER", "--LUmethod=2", "--outdir", str(tmp_path)]) == 0
    hdr = open(os.path.join(str(tmp_path), "net_ROBER_NONE_DENSE_LU2.h")).read()
    assert "LU_METHOD = 2" in hdr and '"ROBER/NONE/DENSE/LU2"' in hdr
    ir = json.load(open(os.path.join(str(tmp_path), "ROBER_NONE_DENSE_LU2.json")))
    assert ir["lu_method"] == 2 and len(ir["lu_closed"]["out"]) == 8 and ir["opcounts"]["lu"]["flops"] == 0
    with pytest.raises(SystemExit):
        cli.main(["--example=ROBER", "--packaging=COO", "--outdir", str(tmp_path)])
    with pytest.raises(SystemExit):   # pre_odevec.py:455-456
        cli.main(["--example=ROBER", "--equilibrium_H=1", "--packaging=CSC", "--outdir", str(tmp_path)])
    with pytest.raises(SystemExit):   # pre_odevec.py:459-460
        cli.main(["--example=ROBER", "--equilibrium_H=1", "--outdir", str(tmp_path)])
    with pytest.raises(SystemExit):   # LUmethod 3 needs DENSE (src/odevec.F90:509-523)
        cli.main(["--example=ROBER", "--LUmethod=3", "--packaging=CSC", "--outdir", str(tmp_path)])
    with pytest.raises(SystemExit):   # ... and no ordering
        cli.main(["--example=PRIMORDIAL", "--LUmethod=3", "--ordering=CMK", "--outdir", str(tmp_path)])
    assert cli.main(["--example=OREGO", "--LUmethod=3", "--equilibrium_H=1", "--outdir", str(tmp_path)]) == 0
    hdr = open(os.path.join(str(tmp_path), "net_OREGO_NONE_DENSE_LU3.h")).read()
    assert "FD_JAC = true" in hdr and '"OREGO/NONE/DENSE/LU3"' in hdr


def test_krome_setupfile(tmp_path):
    f = tmp_path / "net.py"
    f.write_text("neq = 2\nnrea = 1\ny = symbols('y(\\\\:\\\\,1:3)')\nk = symbols('k(\\\\:\\\\,1:2)')\n"
                 "rhs = Matrix([-k[0]*y[0]*y[1], k[0]*y[0]*y[1]])\n")
    net = N.load("KROME", str(f))
    assert net.neq == 2 and net.nparam == 1 and net.rate_kind == "param"
    st = S.analyse(net, "None", "DENSE")
    assert len(st.terms) == 2
    # parameters follow the file's k (KROME's reaction indices), used or not
    g = tmp_path / "net12.py"
    g.write_text("neq = 2\nnrea = 12\ny = symbols('y(\\\\:\\\\,1:3)')\nk = symbols('k(\\\\:\\\\,1:13)')\n"
                 "rhs = Matrix([-k[9]*y[0] + k[1]*y[1], k[9]*y[0] - k[1]*y[1]])\n")
    net = N.load("KROME", str(g))
    assert net.nparam == 12 and net.param_names[9] == "k(:,10)" and net.param_names[1] == "k(:,2)"
    st = S.analyse(net, "None", "DENSE")
    assert sorted(t.rate for t in st.terms) == [1, 1, 9, 9]
    # without k: the other symbols in natural order of their names
    h = tmp_path / "netx.py"
    h.write_text("neq = 1\nnrea = 2\ny = [Symbol('a')]\nrhs = [Symbol('k10')*y[0] - Symbol('k2')]\n")
    assert N.load("KROME", str(h)).param_names == ["k2", "k10"]


def test_synth64_seeded():
    a = N.synth_reactions(64, 300, 6)
    b = N.synth_reactions(64, 300, 6)
    assert a == b
    reac, A, Ea = a
    assert len(reac) == 300 and all(len(set(r + p)) == 4 for r, p, _ in reac)
    assert all(1e-12 <= v <= 1e-8 for v in A) and all(0 <= v <= 5e4 for v in Ea)


def test_linear_invariants():
    """The invariants the generator emits (the mixed form's corrector test):
    ROBER's total, none for OREGO, SYNTH64's molecule count, and for
    PRIMORDIAL a basis of the same space as H, He and D nuclei and charge;
    each satisfies sum_j c_j f_j = 0 at random states and rates."""
    def inv(name):
        st = S.reduce_frozen(S.analyse(N.load(name), "None", "CSC"))
        return st, S.linear_invariants(st.neq, st.terms)
    assert inv("ROBER")[1] == [[1, 1, 1]]
    assert inv("OREGO")[1] == []
    assert inv("SYNTH64")[1] == [[1] * 64]
    st, B = inv("PRIMORDIAL")
    phys = np.array([[0, 1, 1, 0, 2, 0, 1, 1, 0, 2, 0, 0], [0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1],
                     [0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0], [-1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2]])
    B = np.array(B)
    assert B.shape == (4, 12)
    assert np.linalg.matrix_rank(np.vstack([B, phys])) == 4 == np.linalg.matrix_rank(B)
    rng = np.random.default_rng(8)
    for _ in range(5):
        y = np.exp(rng.uniform(-5, 5, 12))
        k = np.exp(rng.uniform(-5, 0, 38))
        f = np.zeros(12)
        terms = np.zeros(12)
        for t in st.terms:
            v = t.coef * (k[t.rate] if t.rate >= 0 else 1.0) * np.prod([y[i] for i in t.factors])
            f[t.row] += v
            terms[t.row] += abs(v)
        for c in B:
            assert abs(c @ f) <= 1e-12 * (np.abs(c) @ terms)
    hdr = emit.emit_header(S.analyse(N.load("PRIMORDIAL"), "CMK", "CSC"))
    assert "static constexpr int NINV = 4;" in hdr


def _coop_tables(name, ordering):
    """The cooperative-kernel tables the generator emits, as lists of the
    literals in the header (unsigned suffixes dropped)."""
    st = S.reduce_frozen(S.analyse(N.load(name), ordering, "CSC"))
    out, cur = {}, None
    for ln in emit._emit_coop(st):
        t = ln.strip()
        if t.startswith("static constexpr") and t.endswith("{"):
            cur = t.split("[")[0].split()[-1]
            out[cur] = []
        elif t == "};":
            cur = None
        elif cur is not None:
            out[cur] += [v.strip().rstrip("u") for v in t.split(",") if v.strip()]
    return out


@pytest.mark.parametrize("name,ordering", [("SYNTH64", "FF"), ("PRIMORDIAL", "CMK"), ("ROBER", "None")])
def test_coop_lane_major_tables_restate_the_row_major_ones(name, ordering):
    """RL_* / JL_* (term q of row c at [q*n + c], packed words) hold exactly
    the RHS terms and Jacobian contributions of RT_* / JT_*, row by row in
    the same order, with zero padding past each row's count."""
    T = _coop_tables(name, ordering)
    n = len(T["RL_N"])
    rt_off = [int(v) for v in T["RT_OFF"]]
    maxt = max(int(v) for v in T["RL_N"])
    assert len(rt_off) == n + 1
    for c in range(n):
        cnt = rt_off[c + 1] - rt_off[c]
        assert int(T["RL_N"][c]) == cnt
        for q in range(maxt):
            i = q * n + c
            w, fw, coef = int(T["RL_W"][i]), int(T["RL_F"][i]), T["RL_C"][i]
            if q >= cnt:
                assert w == 0 and fw == 0 and float(coef) == 0.0
                continue
            t = rt_off[c] + q
            rate = w & 0xffff
            rate = rate - 0x10000 if rate >= 0x8000 else rate
            nf = w >> 16
            assert rate == int(T["RT_RATE"][t]) and nf == int(T["RT_NF"][t]), (c, q)
            assert [(fw >> (8 * p)) & 0xff for p in range(nf)] == [int(v) for v in T["RT_F"][4 * t:4 * t + nf]]
            assert coef == T["RT_COEF"][t]
    jt_off = [int(v) for v in T["JT_OFF"]]
    maxj = max(int(v) for v in T["JL_N"])
    for r in range(n):
        lo, hi = jt_off[r * n], jt_off[r * n + n]
        assert int(T["JL_N"][r]) == hi - lo
        for q in range(maxj):
            k = q * n + r
            w, fw, ce = int(T["JL_W"][k]), int(T["JL_F"][k]), T["JL_C"][k]
            if q >= hi - lo:
                assert w == 0 and fw == 0 and float(ce) == 0.0
                continue
            c = lo + q
            col = w >> 24
            assert jt_off[r * n + col] <= c < jt_off[r * n + col + 1], (r, q, col)
            rate = w & 0xffff
            rate = rate - 0x10000 if rate >= 0x8000 else rate
            nr = (w >> 16) & 0xff
            assert rate == int(T["JT_RATE"][c]) and nr == int(T["JT_NR"][c])
            assert [(fw >> (8 * p)) & 0xff for p in range(nr)] == [int(v) for v in T["JT_R"][3 * c:3 * c + nr]]
            assert ce == T["JT_CE"][c]
