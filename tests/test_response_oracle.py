"""Pins of the oracle's linear response (orc_pert_*, orc_linresp: a
restatement of ccpert.cc and ccresp.cc) by what mathematics fixes, on
synthetic Hamiltonians with a random symmetric one-electron perturbation A:

- at omega = 0 the right-hand perturbed amplitudes are the derivatives of
  the CCSD amplitudes with respect to a static field F added as F A to the
  Fock matrix, X = dt/dF, and the left-hand ones those of the Lambda
  amplitudes, Y = dl/dF (five-point differences of full oracle solves to
  1e-14).  Every term of build_X1/X2 (ccpert.cc:382-470) and of the long
  build_Y1/Y2 (:516-882, including the X, Lambda and perturbation terms)
  enters these derivatives;
- at omega = 0 with both hands solved, linresp(A, B) = -d2E/dF_A dF_B, the
  static response of the CCSD energy (E = 2 f.t1 + tau.L, all terms), for
  A = B and for two different perturbations;
- at omega != 0 the right-hand equations are (J - omega) X + Abar = 0, with
  the Jacobian J X and Abar = dR/dF taken as finite differences of the
  oracle's own CCSD residual;
- the guesses X = Y = Abar / (D + omega) with D from the HBAR diagonals
  (ccpert.cc:44-70), and linresp with the right-hand amplitudes only (the
  reference's default MYHAND=RIGHT) uses those Y guesses.
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd.synthetic import random_hamiltonian


def _sym(n, seed, scale=1.0):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return scale * 0.5 * (A + A.T)


def _solved(fock, ints, no, conv=1e-14):
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=500, conv=conv)["converged"]
    o.lambda_init()
    assert o.lambda_solve(maxiter=500, conv=conv)["converged"]
    return o


def _at(fock, ints, no):
    o = _solved(fock, ints, no)
    t1, t2 = o.amps()
    l1, l2 = o.lambdas()
    e = o.energy_full()
    o.close()
    return e, t1, t2, l1, l2


@pytest.mark.parametrize("no,nv,seed", [(2, 4, 3), (3, 5, 4)])
def test_static_perturbed_amplitudes_are_derivatives(no, nv, seed):
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3)
    A = _sym(n, seed + 10)
    o = _solved(fock, ints, no)
    P = o.pert(A, 0.0)
    assert P.solve(0, conv=1e-13)["converged"]
    assert P.solve(1, conv=1e-13)["converged"]
    X1, X2, Y1, Y2 = P.amplitudes()
    h = 1e-3
    pts = [_at(fock + k * h * A, ints, no) for k in (-2, -1, 1, 2)]
    e0 = _at(fock, ints, no)[0]

    def d1(i):
        v = [p[i] for p in pts]
        return (v[0] - 8 * v[1] + 8 * v[2] - v[3]) / (12 * h)

    for got, idx in ((X1, 1), (X2, 2), (Y1, 3), (Y2, 4)):
        ref = d1(idx)
        assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), idx
    d2 = (-pts[0][0] + 16 * pts[1][0] - 30 * e0 + 16 * pts[2][0] - pts[3][0]) / (12 * h * h)
    assert abs(P.linresp(P) + d2) < 1e-8 * max(1.0, abs(d2))
    P.close()
    o.close()


def test_static_response_of_two_perturbations():
    no, nv = 2, 5
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=8, scale=0.3)
    A, B = _sym(n, 21), _sym(n, 22)
    o = _solved(fock, ints, no)
    PA, PB = o.pert(A, 0.0), o.pert(B, 0.0)
    for P in (PA, PB):
        assert P.solve(0, conv=1e-13)["converged"] and P.solve(1, conv=1e-13)["converged"]
    h = 1e-3
    e0 = _at(fock, ints, no)[0]

    def d2(C):  # five-point second derivative along C
        E = [_at(fock + k * h * C, ints, no)[0] for k in (-2, -1, 1, 2)]
        return (-E[0] + 16 * E[1] - 30 * e0 + 16 * E[2] - E[3]) / (12 * h * h)

    # the mixed derivative by polarization: (d2(A + B) - d2(A) - d2(B)) / 2
    d2 = 0.5 * (d2(A + B) - d2(A) - d2(B))
    for val in (PA.linresp(PB), PB.linresp(PA)):
        assert abs(val + d2) < 1e-8 * max(1.0, abs(d2)), (val, d2)
    PA.close(); PB.close(); o.close()


@pytest.mark.parametrize("omega", [0.05, -0.05, 0.2])
def test_dynamic_right_hand_equations(omega):
    no, nv = 2, 4
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=5, scale=0.3)
    A = _sym(n, 31)
    o = _solved(fock, ints, no)
    t1, t2 = o.amps()
    P = o.pert(A, omega)
    assert P.solve(0, conv=1e-13)["converged"]
    X1, X2, _, _ = P.amplitudes()
    h = 1e-4

    def R(f, a1, a2):
        q = Oracle(f, ints, no)
        q.set_amps(a1, a2)
        r = q.residual_rows(no)
        q.close()
        return r

    def cd(fn):
        p1, m1 = fn(h), fn(-h)
        p2, m2 = fn(2 * h), fn(-2 * h)
        return [(m2[k] - 8 * m1[k] + 8 * p1[k] - p2[k]) / (12 * h) for k in range(2)]

    JX = cd(lambda e: R(fock, t1 + e * X1, t2 + e * X2))
    Abar = cd(lambda e: R(fock + e * A, t1, t2))
    r1 = JX[0] + Abar[0] - omega * X1
    r2 = JX[1] + Abar[1] - omega * X2
    assert np.max(np.abs(r1)) < 1e-9 and np.max(np.abs(r2)) < 1e-9, (np.max(np.abs(r1)), np.max(np.abs(r2)))
    # the frequency enters: the omega = 0 amplitudes do not solve these equations
    Q = o.pert(A, 0.0)
    assert Q.solve(0, conv=1e-13)["converged"]
    assert np.max(np.abs(Q.amplitudes()[0] - X1)) > 1e-4
    P.close(); Q.close(); o.close()


def test_guesses_and_right_hand_only_response():
    no, nv = 2, 4
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=6, scale=0.3)
    A = _sym(n, 41)
    o = _solved(fock, ints, no)
    omega = 0.1
    P = o.pert(A, omega)
    X1, X2, Y1, Y2 = P.amplitudes()
    np.testing.assert_array_equal(X1, Y1)
    np.testing.assert_array_equal(X2, Y2)
    Hb = o.hbar()
    D1 = np.diag(Hb["Hoo"])[:, None] - np.diag(Hb["Hvv"])[None, :]
    # with zero X the X1 residual is Abar_vo: the guess is that over D1 + omega
    R1, _ = P.residual(0, np.zeros((no, nv)), np.zeros((no, no, nv, nv)))
    np.testing.assert_allclose(X1, R1 / (D1 + omega), rtol=1e-13, atol=1e-15)
    # right hand only: linresp reads the Y guess; solving Y changes it
    assert P.solve(0, conv=1e-13)["converged"]
    before = P.linresp(P)
    assert P.solve(1, conv=1e-13)["converged"]
    assert abs(P.linresp(P) - before) > 1e-6
    P.close(); o.close()
