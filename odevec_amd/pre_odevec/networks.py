"""Network definitions for the preprocessor.

A network is a set of species ``y[0..neq-1]``, a set of rate symbols
``k[0..nrate-1]`` and a right-hand side ``rhs[i](y, k)`` that is a polynomial in
``y`` (mass-action chemistry).  How the rate symbols get their values per cell
is the network's ``rate_kind``:

* ``"param"``      -- ``k`` are per-cell input parameters (defaults given);
* ``"none"``       -- no rate symbols, constants are folded into the RHS;
* ``"primordial"`` -- the 38 KROME primordial rates, a function of the gas
  temperature stored in the ``Tgas`` slot of the state vector;
* ``"arrhenius"``  -- ``k_r = A_r * exp(-Ea_r / T)`` with ``T = param[0]``.

Networks follow the reference definitions:

* ROBER  -- pre_odevec/example_systems.py:3-17 (rates k1=0.04, k2=3e7, k3=1e4
  made per-cell parameters so that configurations can scale them per cell);
* OREGO  -- pre_odevec/example_systems.py:20-37 (k1=77.27, k2=8.375e-6,
  k3=0.161 folded in exactly as the reference's Python expression folds them);
* PRIMORDIAL -- pre_odevec/example_systems.py:40-191, written here as the list
  of its 38 reactions (mass action); rate formulas follow
  src/odevec_commons.F90:101-384;
* SYNTH64 -- a seeded synthetic 64-species network of 300 random two-body
  reactions A + B -> C + D with Arrhenius rates (no reference counterpart);
* KROME  -- a KROME-style setup file that defines ``y``, ``rhs``, ``neq`` and
  ``nrea`` in Python syntax (pre_odevec.py:97-105).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import sympy as sp


@dataclass
class Network:
    name: str
    species: List[str]
    y: List[sp.Symbol]
    k: List[sp.Symbol]
    rhs: List[sp.Expr]
    rate_kind: str = "none"
    param_names: List[str] = field(default_factory=list)
    param_defaults: List[float] = field(default_factory=list)
    arrhenius_A: Optional[List[float]] = None
    arrhenius_Ea: Optional[List[float]] = None
    tgas_index: int = -1          # state slot holding Tgas (primordial)
    reactions: Optional[List[Tuple[Tuple[int, ...], Tuple[int, ...], int]]] = None

    @property
    def neq(self) -> int:
        return len(self.y)

    @property
    def nrate(self) -> int:
        return len(self.k)

    @property
    def nparam(self) -> int:
        return len(self.param_names)


def _ysyms(neq: int) -> List[sp.Symbol]:
    return [sp.Symbol("y%d" % i) for i in range(neq)]


def _ksyms(n: int) -> List[sp.Symbol]:
    return [sp.Symbol("k%d" % i) for i in range(n)]


def mass_action_rhs(neq: int, reactions, y, k) -> List[sp.Expr]:
    """RHS of a mass-action network: d y_s/dt = sum_r nu_{s r} k_r prod(reactants)."""
    rhs = [sp.Integer(0)] * neq
    for reac, prod, ridx in reactions:
        flux = k[ridx]
        for s in reac:
            flux = flux * y[s]
        for s in reac:
            rhs[s] = rhs[s] - flux
        for s in prod:
            rhs[s] = rhs[s] + flux
    return [sp.expand(e) for e in rhs]


# --------------------------------------------------------------------------- ROBER
ROBER_K0 = (0.04, 3e7, 1e4)


def load_rober() -> Network:
    """Robertson problem, pre_odevec/example_systems.py:3-17.

    The reference hard-codes k = (0.04, 3e7, 1e4) in the expression
    ``-0.04*y1 + 1e4*y2*y3``; here they are per-cell parameters whose defaults
    are those constants (a per-cell lognormal scale is a workload input).
    """
    y = _ysyms(3)
    k = _ksyms(3)
    rhs = [-k[0] * y[0] + k[2] * y[1] * y[2],
           k[0] * y[0] - k[1] * y[1] * y[1] - k[2] * y[1] * y[2],
           k[1] * y[1] * y[1]]
    return Network("ROBER", ["A", "B", "C"], y, k, [sp.expand(e) for e in rhs],
                   rate_kind="param", param_names=["k1", "k2", "k3"],
                   param_defaults=list(ROBER_K0))


# --------------------------------------------------------------------------- OREGO
def load_orego() -> Network:
    """Oregonator, pre_odevec/example_systems.py:20-37 (constants folded in
    Python float arithmetic in the same order as the reference expression)."""
    y = _ysyms(3)
    k1 = 77.27
    k2 = 8.375e-6
    k3 = 0.161
    rhs = [k1 * y[0] + k1 * y[1] - k1 * k2 * y[0] ** 2 - k1 * y[0] * y[1],
           1.0 / k1 * y[2] - 1.0 / k1 * y[1] - 1.0 / k1 * y[0] * y[1],
           k3 * y[0] - k3 * y[2]]
    return Network("OREGO", ["X", "Y", "Z"], y, [], [sp.expand(e) for e in rhs],
                   rate_kind="none")


# ---------------------------------------------------------------------- PRIMORDIAL
PRIMORDIAL_SPECIES = ["E", "H-", "H", "HE", "H2", "D", "HD", "H+", "HE+", "H2+",
                      "D+", "HE++", "CR", "g", "Tgas", "dummy"]
(E, HM, H, HE, H2, D, HD, HP, HEP, H2P, DP, HEPP, CR, G, TGAS, DUMMY) = range(16)

# (reactants, products) for k[0..37], the KROME primordial network
# (src/odevec_commons.F90:143-371 names each reaction in a comment; the RHS in
# pre_odevec/example_systems.py:69-189 is the mass-action form of this list).
PRIMORDIAL_REACTIONS = [
    ((H, E), (HP, E, E)),          # 0  H + E -> H+ + E + E
    ((HP, E), (H,)),               # 1  H+ + E -> H          (T <= 5.5e3)
    ((HP, E), (H,)),               # 2  H+ + E -> H          (T >  5.5e3)
    ((HE, E), (HEP, E, E)),        # 3  HE + E -> HE+ + E + E
    ((HEP, E), (HE,)),             # 4  HE+ + E -> HE        (T <= 9.28e3)
    ((HEP, E), (HE,)),             # 5  HE+ + E -> HE        (T >  9.28e3)
    ((HEP, E), (HEPP, E, E)),      # 6  HE+ + E -> HE++ + E + E
    ((HEPP, E), (HEP,)),           # 7  HE++ + E -> HE+
    ((H, E), (HM,)),               # 8  H + E -> H-
    ((HM, H), (H2, E)),            # 9  H- + H -> H2 + E     (T < 1160)
    ((HM, H), (H2, E)),            # 10 H- + H -> H2 + E     (T > 1160)
    ((H, HP), (H2P,)),             # 11 H + H+ -> H2+        (T <= 6.7e3)
    ((H, HP), (H2P,)),             # 12 H + H+ -> H2+        (T >  6.7e3)
    ((H2P, H), (H2, HP)),          # 13 H2+ + H -> H2 + H+
    ((H2, HP), (H2P, H)),          # 14 H2 + H+ -> H2+ + H   (T > 3.48e3)
    ((H2, E), (H, H, E)),          # 15 H2 + E -> H + H + E
    ((H2, H), (H, H, H)),          # 16 H2 + H -> H + H + H
    ((HM, E), (H, E, E)),          # 17 H- + E -> H + E + E
    ((HM, H), (H, H, E)),          # 18 H- + H -> H + H + E  (T <= 1.16e3)
    ((HM, H), (H, H, E)),          # 19 H- + H -> H + H + E  (T >  1.16e3)
    ((HM, HP), (H, H)),            # 20 H- + H+ -> H + H
    ((HM, HP), (H2P, E)),          # 21 H- + H+ -> H2+ + E
    ((H2P, E), (H, H)),            # 22 H2+ + E -> H + H     (T <= 617)
    ((H2P, E), (H, H)),            # 23 H2+ + E -> H + H     (T >  617)
    ((H2P, HM), (H, H2)),          # 24 H2+ + H- -> H + H2
    ((H, H, H), (H2, H)),          # 25 H + H + H -> H2 + H  (T <= 300)
    ((H, H, H), (H2, H)),          # 26 H + H + H -> H2 + H  (T >  300)
    ((H2, H, H), (H2, H2)),        # 27 H2 + H + H -> H2 + H2 (T <= 300)
    ((H2, H, H), (H2, H2)),        # 28 H2 + H + H -> H2 + H2 (T >  300)
    ((HP, D), (H, DP)),            # 29 H+ + D -> H + D+     (T >= 50)
    ((H, DP), (HP, D)),            # 30 H + D+ -> H+ + D     (T >= 50)
    ((H2, DP), (HD, HP)),          # 31 H2 + D+ -> HD + H+
    ((HD, HP), (H2, DP)),          # 32 HD + H+ -> H2 + D+
    ((H2, D), (HD, H)),            # 33 H2 + D -> HD + H     (T <= 2e3)
    ((H2, D), (HD, H)),            # 34 H2 + D -> HD + H     (T >  2e3)
    ((HD, H), (H2, D)),            # 35 HD + H -> H2 + D     (T > 200)
    ((D, HM), (HD, E)),            # 36 D + H- -> HD + E
    ((DP, E), (D,)),               # 37 D+ + E -> D
]


def load_primordial() -> Network:
    """16-species primordial H/He/D network, pre_odevec/example_systems.py:40-191."""
    y = _ysyms(16)
    k = _ksyms(38)
    reactions = [(r, p, i) for i, (r, p) in enumerate(PRIMORDIAL_REACTIONS)]
    rhs = mass_action_rhs(16, reactions, y, k)
    return Network("PRIMORDIAL", list(PRIMORDIAL_SPECIES), y, k, rhs,
                   rate_kind="primordial", tgas_index=TGAS, reactions=reactions)


# ------------------------------------------------------------------------- SYNTH64
SYNTH64_SEED = 6


def synth_reactions(nspec: int = 64, nreac: int = 300, seed: int = SYNTH64_SEED):
    """Seeded random two-body reactions A + B -> C + D (four distinct species)
    with Arrhenius rates, log10 A ~ U[-12,-8], Ea/k ~ U[0, 5e4] K."""
    rng = np.random.default_rng(seed)
    reactions = []
    for r in range(nreac):
        a, b, c, d = (int(v) for v in rng.choice(nspec, size=4, replace=False))
        reactions.append(((min(a, b), max(a, b)), (min(c, d), max(c, d)), r))
    log10A = rng.uniform(-12.0, -8.0, size=nreac)
    Ea = rng.uniform(0.0, 5e4, size=nreac)
    return reactions, [float(10.0 ** v) for v in log10A], [float(v) for v in Ea]


def load_synth(nspec: int = 64, nreac: int = 300, seed: int = SYNTH64_SEED) -> Network:
    reactions, A, Ea = synth_reactions(nspec, nreac, seed)
    y = _ysyms(nspec)
    k = _ksyms(nreac)
    rhs = mass_action_rhs(nspec, reactions, y, k)
    return Network("SYNTH%d" % nspec, ["S%d" % i for i in range(nspec)], y, k, rhs,
                   rate_kind="arrhenius", param_names=["T"], param_defaults=[1000.0],
                   arrhenius_A=A, arrhenius_Ea=Ea, reactions=reactions)


def load_synth64(seed: int = SYNTH64_SEED) -> Network:
    return load_synth(64, 300, seed)


def load_synth40(seed: int = SYNTH64_SEED) -> Network:
    """40 species, 180 reactions: a cell fills 40 of a 64-lane group (the
    cooperative kernel's padding lanes)."""
    return load_synth(40, 180, seed)


# --------------------------------------------------------------------------- KROME
def load_krome_setupfile(path: str) -> Network:
    """Network from a KROME-style setup file (pre_odevec.py:97-105).

    The file is executed with SymPy's names in scope and must define ``y``
    (a sequence of symbols), ``rhs`` (a sequence or Matrix of expressions),
    ``neq`` and ``nrea``. Every symbol of the RHS that is not a species is a
    rate and becomes a per-cell parameter (KROME computes its rate
    coefficients ``k(:,i)`` itself, outside the solver): if the file defines
    ``k`` (a sequence of symbols), parameter i is ``k[i]``, used or not, so
    that parameter indices are KROME's reaction indices; any other symbols
    follow in natural order of their names (``k(:,2)`` before ``k(:,10)``)."""
    scope: Dict[str, object] = {}
    exec("from sympy import *", scope)
    with open(path) as fh:
        exec(compile(fh.read(), path, "exec"), scope)
    ys = list(scope["y"])
    rhs_in = list(scope["rhs"])
    neq = int(scope["neq"])
    if len(ys) != neq or len(rhs_in) != neq:
        raise ValueError("KROME setup file: len(y) and len(rhs) must equal neq")
    yn = _ysyms(neq)
    sub = {ys[i]: yn[i] for i in range(neq)}
    rhs = [sp.expand(sp.sympify(e).xreplace(sub)) for e in rhs_in]
    declared = [sp.sympify(s) for s in scope["k"]] if "k" in scope else []
    extra = sorted({s for e in rhs for s in e.free_symbols if s not in yn and s not in declared},
                   key=lambda s: [int(t) if t.isdigit() else t for t in re.split(r"(\d+)", s.name)])
    ksyms = declared + extra
    kn = _ksyms(len(ksyms))
    rhs = [e.xreplace(dict(zip(ksyms, kn))) for e in rhs]
    return Network("KROME", [str(s) for s in ys], yn, kn, rhs, rate_kind="param",
                   param_names=[s.name for s in ksyms], param_defaults=[0.0] * len(ksyms))


BUILTIN = {
    "ROBER": load_rober,
    "OREGO": load_orego,
    "PRIMORDIAL": load_primordial,
    "SYNTH64": load_synth64,
    "SYNTH40": load_synth40,
}


def load(example: str, krome_setupfile: Optional[str] = None) -> Network:
    if krome_setupfile is not None:
        return load_krome_setupfile(krome_setupfile)
    try:
        return BUILTIN[example.upper()]()
    except KeyError:
        raise ValueError("unknown network %r (known: %s)" % (example, ", ".join(BUILTIN)))
