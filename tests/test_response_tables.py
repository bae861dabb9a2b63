"""The contraction tables of the GPU linear response (kPertBar, kPertX and
kPertYSource in ugacc_amd/csrc/ugx_ccsd.cpp: pertbar, build_X1/X2 and the
Y-independent part of build_Y1/Y2 of ccpert.cc) evaluated term by term with
numpy on the CPU, in the order and with the coefficients the library runs
them, against the oracle's restatement: the right-hand residual at random
amplitudes, and the left-hand residual (the Lambda operator of the oracle
plus omega Y plus the table's source) at random amplitudes.  This checks the
einsum strings of every term without a GPU; the GPU tests then compare the
library's results with the same oracle."""
import os
import re

import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd.synthetic import random_hamiltonian

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ugacc_amd", "csrc", "ugx_ccsd.cpp")
TERM = re.compile(r'\{"([^"]+)", ([-0-9.]+), "(\w+)", (?:nullptr|"(\w+)"), ([-0-9.]+), "(\w+)"\}')


def _table(name):
    src = open(SRC).read()
    start = src.index("const PTerm %s[] = {" % name)
    body = src[start:src.index("};", start)]
    return [(m.group(1), float(m.group(2)), m.group(3), m.group(4), float(m.group(5)), m.group(6))
            for m in TERM.finditer(body)]


def _run(table, env):
    for spec, alpha, a, b, beta, c in table:
        lhs, out = spec.split("->")
        if b is None:
            val = alpha * np.einsum(lhs + "->" + out, env[a])
        else:
            val = alpha * np.einsum(spec, env[a], env[b], optimize=True)
        env[c] = val if beta == 0.0 or c not in env else val + beta * env[c]


def _setup(no, nv, seed, omega):
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=400, conv=1e-13)["converged"]
    o.lambda_init()
    assert o.lambda_solve(maxiter=400, conv=1e-13)["converged"]
    A = np.random.default_rng(seed + 7).standard_normal((n, n))
    A = 0.5 * (A + A.T)
    P = o.pert(A, omega)
    t1, t2 = o.amps()
    l1, l2 = o.lambdas()
    L = o.Lints()
    env = dict(t1=t1, t2=t2, l1=l1, l2=l2, Loovv=L[:no, :no, no:, no:], oovv=ints[:no, :no, no:, no:],
               Poo=A[:no, :no], Pov=A[:no, no:], Pvo=A[no:, :no], Pvv=A[no:, no:])
    for k, v in o.hbar().items():
        env[k] = v
    _run(_table("kPertBar"), env)
    return o, P, env


def _sym(Z):
    return Z + Z.transpose(1, 0, 3, 2)


@pytest.mark.parametrize("no,nv,seed,omega", [(2, 4, 3, 0.0), (3, 5, 4, 0.15)])
def test_right_hand_table_matches_the_oracle(no, nv, seed, omega):
    o, P, env = _setup(no, nv, seed, omega)
    rng = np.random.default_rng(seed)
    X1 = 0.1 * rng.standard_normal((no, nv))
    X2 = 0.05 * rng.standard_normal((no, no, nv, nv))
    X2 = X2 + X2.transpose(1, 0, 3, 2)
    env.update(X1o=X1, X2o=X2)
    _run(_table("kPertX"), env)
    R1 = env["R1"] - omega * X1
    R2 = _sym(env["Z"] - 0.5 * omega * X2)
    R1o, R2o = P.residual(0, X1, X2)
    assert np.max(np.abs(R1 - R1o)) < 1e-12 * max(1.0, np.max(np.abs(R1o)))
    assert np.max(np.abs(R2 - R2o)) < 1e-12 * max(1.0, np.max(np.abs(R2o)))
    P.close(); o.close()


@pytest.mark.parametrize("no,nv,seed,omega", [(2, 4, 5, 0.0), (3, 5, 6, -0.1)])
def test_left_hand_source_table_matches_the_oracle(no, nv, seed, omega):
    o, P, env = _setup(no, nv, seed, omega)
    assert P.solve(0, conv=1e-12)["converged"]
    X1, X2, _, _ = P.amplitudes()
    env.update(X1=X1, X2=X2)
    _run(_table("kPertYSource"), env)
    rng = np.random.default_rng(seed)
    Y1 = 0.1 * rng.standard_normal((no, nv))
    Y2 = 0.05 * rng.standard_normal((no, no, nv, nv))
    Y2 = Y2 + Y2.transpose(1, 0, 3, 2)
    # the Lambda operator: the Lambda residual is affine in (l1, l2)
    opY = [a - b for a, b in zip(o.lambda_residual(Y1, Y2)[:2],
                                 o.lambda_residual(np.zeros_like(Y1), np.zeros_like(Y2))[:2])]
    R1 = opY[0] + omega * Y1 + env["S1"]
    R2 = opY[1] + omega * Y2 + _sym(env["S2"])
    R1o, R2o = P.residual(1, Y1, Y2)
    assert np.max(np.abs(R1 - R1o)) < 1e-11 * max(1.0, np.max(np.abs(R1o)))
    assert np.max(np.abs(R2 - R2o)) < 1e-11 * max(1.0, np.max(np.abs(R2o)))
    P.close(); o.close()
