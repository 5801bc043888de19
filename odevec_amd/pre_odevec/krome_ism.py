# KROME-style setup file for the preprocessor (python -m odevec_amd.pre_odevec
# --krome_setupfile <this file>): the right-hand side of a network as KROME
# hands it to the solver, in Python syntax. The preprocessor executes the file
# with SymPy's names in scope and reads neq, nrea, y and rhs; the rate
# coefficients k(:,i) are computed per cell by the caller (KROME's own rate
# code) and passed to the solver as parameters, in the order of k.
#
# Network: a small carbon/oxygen chemistry of diffuse and translucent clouds,
# 12 species and 21 reactions (cosmic-ray ionisation, photo processes,
# ion-molecule and neutral-neutral reactions, recombination and H2 formation
# on grains). Reactions, with the rate law that each term of rhs uses:
#    1  H + H       -> H2               k1 H H      (formation on grains)
#    2  H2 + H2     -> H3+ + H + e-     k2 H2       (cosmic-ray ionisation of H2,
#                                                     H2+ + H2 folded in)
#    3  H           -> H+ + e-          k3 H        (cosmic-ray ionisation)
#    4  H+ + e-     -> H                k4 H+ e-
#    5  H3+ + e-    -> H2 + H           k5 H3+ e-
#    6  H3+ + O     -> OH + H + H+      k6 H3+ O    (OH+ chain folded in)
#    7  OH + H2     -> H2O + H          k7 OH H2
#    8  H2O         -> OH + H           k8 H2O      (photodissociation)
#    9  OH          -> O + H            k9 OH       (photodissociation)
#   10  C + O       -> CO               k10 C O     (radiative association)
#   11  C + OH      -> CO + H           k11 C OH
#   12  C+ + OH     -> CO + H+          k12 C+ OH   (CO+ chain folded in)
#   13  C           -> C+ + e-          k13 C       (photoionisation)
#   14  C+ + e-     -> C                k14 C+ e-
#   15  CO          -> C + O            k15 CO      (photodissociation)
#   16  H3+ + CO    -> HCO+ + H2        k16 H3+ CO
#   17  HCO+ + e-   -> CO + H           k17 HCO+ e-
#   18  C+ + H2O    -> HCO+ + H         k18 C+ H2O
#   19  H2          -> H + H            k19 H2      (photodissociation)
#   20  O + H2      -> OH + H           k20 O H2
#   21  H + OH      -> O + H2           k21 H OH
# Every reaction conserves H, C and O nuclei and charge.
from sympy import Matrix, symbols

neq = 12
nrea = 21

# species:      1   2  3   4   5    6  7   8  9   10   11  12
#               e-  H  H2  H+  H3+  C  C+  O  OH  H2O  CO  HCO+
y = symbols('y(\\:\\,1:13)')
k = symbols('k(\\:\\,1:22)')

E, H, H2, HP, H3P, C, CP, O, OH, H2O, CO, HCOP = y
r = [None,
     k[0] * H * H,
     k[1] * H2,
     k[2] * H,
     k[3] * HP * E,
     k[4] * H3P * E,
     k[5] * H3P * O,
     k[6] * OH * H2,
     k[7] * H2O,
     k[8] * OH,
     k[9] * C * O,
     k[10] * C * OH,
     k[11] * CP * OH,
     k[12] * C,
     k[13] * CP * E,
     k[14] * CO,
     k[15] * H3P * CO,
     k[16] * HCOP * E,
     k[17] * CP * H2O,
     k[18] * H2,
     k[19] * O * H2,
     k[20] * H * OH]

rhs = Matrix([
    # e-
    r[2] + r[3] - r[4] - r[5] + r[13] - r[14] - r[17],
    # H
    -2 * r[1] + r[2] - r[3] + r[4] + r[5] + r[6] + r[7] + r[8] + r[9] + r[11] + r[17] + r[18]
    + 2 * r[19] + r[20] - r[21],
    # H2
    r[1] - 2 * r[2] + r[5] - r[7] + r[16] - r[19] - r[20] + r[21],
    # H+
    r[3] - r[4] + r[6] + r[12],
    # H3+
    r[2] - r[5] - r[6] - r[16],
    # C
    -r[10] - r[11] - r[13] + r[14] + r[15],
    # C+
    -r[12] + r[13] - r[14] - r[18],
    # O
    -r[6] + r[9] - r[10] + r[15] - r[20] + r[21],
    # OH
    r[6] - r[7] + r[8] - r[9] - r[11] - r[12] + r[20] - r[21],
    # H2O
    r[7] - r[8] - r[18],
    # CO
    r[10] + r[11] + r[12] - r[15] - r[16] + r[17],
    # HCO+
    r[16] - r[17] + r[18],
])
