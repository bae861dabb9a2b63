"""Basis-set data for the chemistry front end.

cc-pVDZ for H and O is the set the reference's own saved run used
(psi3/output.c1_fzc:45-80: exponents and contraction coefficients, with
spherical d).  STO-3G for H, O and F is the standard minimal basis (Hehre,
Stewart and Pople, J. Chem. Phys. 51, 2657 (1969)); it serves the HOF
known-answer run (psi4/hof/output.c1_nofzc, "Spherical Harmonics?: true").

A shell is (l, [(exponent, coefficient), ...]) with coefficients that refer
to normalised primitives; contractions are renormalised by the integral code.
"""

CC_PVDZ = {
    "O": [
        (0, [(11720.0, 0.00071), (1759.0, 0.00547), (400.8, 0.027837), (113.7, 0.1048), (37.03, 0.283062),
             (13.27, 0.448719), (5.025, 0.270952), (1.013, 0.015458), (0.3023, -0.002585)]),
        (0, [(11720.0, -0.00016), (1759.0, -0.001263), (400.8, -0.006267), (113.7, -0.025716), (37.03, -0.070924),
             (13.27, -0.165411), (5.025, -0.116955), (1.013, 0.557368), (0.3023, 0.572759)]),
        (0, [(0.3023, 1.0)]),
        (1, [(17.7, 0.043018), (3.854, 0.228913), (1.046, 0.508728), (0.2753, 0.460531)]),
        (1, [(0.2753, 1.0)]),
        (2, [(1.185, 1.0)]),
    ],
    "H": [
        (0, [(13.01, 0.019685), (1.962, 0.137977), (0.4446, 0.478148), (0.122, 0.50124)]),
        (0, [(0.122, 1.0)]),
        (1, [(0.727, 1.0)]),
    ],
}

_STO3G_C1S = (0.15432897, 0.53532814, 0.44463454)
_STO3G_C2S = (-0.09996723, 0.39951283, 0.70011547)
_STO3G_C2P = (0.15591627, 0.60768372, 0.39195739)


def _sto3g_first_row(e1s, e2sp):
    return [
        (0, list(zip(e1s, _STO3G_C1S))),
        (0, list(zip(e2sp, _STO3G_C2S))),
        (1, list(zip(e2sp, _STO3G_C2P))),
    ]


STO_3G = {
    "H": [(0, list(zip((3.42525091, 0.62391373, 0.16885540), _STO3G_C1S)))],
    "O": _sto3g_first_row((130.7093200, 23.8088610, 6.4436083), (5.0331513, 1.1695961, 0.3803890)),
    "F": _sto3g_first_row((166.6791300, 30.3608120, 8.2168207), (6.4648032, 1.5022812, 0.4885885)),
}

BASIS_SETS = {"cc-pvdz": CC_PVDZ, "sto-3g": STO_3G}

Z = {"H": 1, "O": 8, "F": 9}


def shells_for(atoms, basis):
    """[(center_xyz, l, [(exp, coef)])] for atoms [(symbol, (x, y, z) bohr)]."""
    table = BASIS_SETS[basis.lower()]
    out = []
    for sym, xyz in atoms:
        for l, prims in table[sym]:
            out.append((tuple(xyz), l, prims))
    return out
