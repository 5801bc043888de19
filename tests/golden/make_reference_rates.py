"""Write tests/golden/reference_rates.json: the reference's KROME primordial
rate coefficients (function coe, src/odevec_commons.F90:101-384) evaluated on a
temperature grid, by translating the Fortran text mechanically (no hand
transcription): continuation lines are joined, `where (cond)` blocks become
conditions, literals keep their Fortran kind (a real literal with a `d`
exponent is double precision; one without is default real, i.e. single
precision, promoted to double), `**` with an integer literal exponent stays an
integer power.

coe fixes Tgas = 1e3 for every cell (src/odevec_commons.F90:120-124); here the
same expressions are evaluated for each Tgas of the grid, with `small` from a
given max(y). The fixture holds only the evaluated vectors (inputs and the 38
rates); the translated statements exist only while this script runs.

Runs only in the build container, where the reference is mounted.
"""
import json
import math
import os
import re

import numpy as np

REF = os.environ.get("ODEVEC_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_rates.json")

TOKEN = re.compile(r"(?P<id>[A-Za-z_][A-Za-z_0-9]*)|(?P<num>(\d+\.\d*|\.\d+|\d+)([dDeE][+-]?\d+)?)|(?P<op>\*\*|\.(?:LE|GT|LT|GE|EQ|NE)\.|.)", re.I)
CMP = {".LE.": "<=", ".GT.": ">", ".LT.": "<", ".GE.": ">=", ".EQ.": "==", ".NE.": "!="}
FUNCS = {"exp": "math.exp", "dexp": "math.exp", "log": "math.log", "log10": "math.log10", "sqrt": "math.sqrt"}


def lit(tok):
    m = re.fullmatch(r"(\d+\.\d*|\.\d+|\d+)([dDeE][+-]?\d+)?", tok)
    mant, exp = m.group(1), m.group(2)
    if exp is None and "." not in mant:
        return mant                      # integer literal
    if exp is not None and exp[0] in "dD":
        return repr(float(mant + "e" + exp[1:]))
    val = float(mant + ("e" + exp[1:] if exp else ""))
    return repr(float(np.float32(val)))  # default real


def translate(expr):
    out = []
    for m in TOKEN.finditer(expr):
        if m.group("id"):
            name = m.group("id")
            out.append(FUNCS.get(name.lower(), name))
        elif m.group("num"):
            out.append(lit(m.group("num")))
        else:
            op = m.group("op")
            out.append(CMP.get(op.upper(), op))
    return "".join(out)


def statements(lines):
    buf = ""
    for raw in lines:
        line = raw.split("!")[0].rstrip()
        if not line.strip():
            continue
        if line.endswith("&"):
            buf += line[:-1]
            continue
        buf += line
        yield buf.strip()
        buf = ""


def load_coe():
    src = open(os.path.join(REF, "src", "odevec_commons.F90")).read().split("\n")
    start = next(i for i, l in enumerate(src) if "Te = Tgas*" in l)
    end = next(i for i, l in enumerate(src) if "coe(:,:) = vec_k(:,:)" in l)
    prog = []
    cond = None
    for st in statements(src[start:end]):
        low = st.lower()
        if low.startswith("do ") or low.startswith("end do") or "vec_k(i,:)" in low:
            continue
        if low.startswith("where"):
            cond = translate(st[st.index("(") + 1:st.rindex(")")])
            continue
        if low.startswith("end where"):
            cond = None
            continue
        lhs, rhs = st.split("=", 1)
        m = re.match(r"\s*vec_k\(:,(\d+)\)", lhs)
        target = "k[%d]" % (int(m.group(1)) - 1) if m else lhs.strip()
        stmt = "%s = %s" % (target, translate(rhs))
        prog.append(("if %s: " % cond if cond else "") + stmt)
    return prog


def rates(prog, Tgas, nmax):
    env = {"math": math, "Tgas": float(Tgas), "small": 1e-40 / (nmax * nmax * nmax), "k": [0.0] * 38}
    env["k"] = [env["small"]] * 38
    for st in prog:
        exec(st, env)
    return env["k"]


def main():
    prog = load_coe()
    grid = sorted(set([10.0, 20.0, 49.9, 50.0, 50.1, 100.0, 199.9, 200.0, 200.1, 299.9, 300.0, 300.1, 500.0, 616.9,
                       617.0, 617.1, 1000.0, 1159.9, 1160.0, 1160.1, 1999.9, 2000.0, 2000.1, 3479.9, 3480.0, 3480.1,
                       5499.9, 5500.0, 5500.1, 6699.9, 6700.0, 6700.1, 8000.0, 9279.9, 9280.0, 9280.1, 1e4]
                      + [float(v) for v in np.exp(np.random.default_rng(5).uniform(np.log(10), np.log(1e4), 24))]))
    cases = []
    for T in grid:
        for nmax in (max(T, 1.0), 1e6):   # a state holding only Tgas, or one species at 1e6
            cases.append({"Tgas": T, "nmax": nmax, "k": rates(prog, T, nmax)})
    with open(OUT, "w") as fh:
        # only the evaluated vectors are kept: the translated statements stay in memory
        json.dump({"source": "src/odevec_commons.F90 coe", "cases": cases}, fh, indent=1)
    print("wrote", OUT, len(cases), "cases,", len(prog), "statements")


if __name__ == "__main__":
    main()
