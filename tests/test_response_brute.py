"""Pins of the oracle's linear response for general one-electron
perturbations (orc_pert_*, orc_linresp: ccpert.cc, ccresp.cc), where the
Fock-matrix derivative pins cannot reach: a field that is not symmetric
(the magnetic dipole "L" of the optical rotation is antisymmetric) is no
Fock matrix the closed-shell CCSD equations accept.  The checker is CCSD in
the determinant basis (tests/brute_cc.py), whose Hamiltonian takes any
one-electron operator V = sum_pq A_pq E_pq:

- the determinant CCSD reproduces the oracle's amplitudes and energy;
- at omega = 0 the oracle's X for a general A is dt/dlambda of the
  determinant CCSD of H + lambda V (five-point differences), and with both
  hands solved linresp(A, B) = linresp(B, A) = -d2E/dlambda_A dlambda_B for
  general A and B (E the determinant CCSD energy);
- at omega != 0 the oracle's X solves the right-hand equations
  <mu| [Hbar, X] + Abar - omega X |0> = 0 in the determinant basis, with
  Hbar = exp(-T) H exp(T) and Abar = exp(-T) V exp(T) built there.
"""
import numpy as np
import pytest

from oracle import Oracle
from tests.brute_cc import BruteCC
from ugacc_amd.synthetic import random_hamiltonian


def _setup(no, nv, seed):
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=500, conv=1e-14)["converged"]
    o.lambda_init()
    assert o.lambda_solve(maxiter=500, conv=1e-14)["converged"]
    return fock, ints, o, BruteCC(fock, ints, no)


def test_determinant_ccsd_matches_the_oracle():
    no, nv = 2, 3
    fock, ints, o, B = _setup(no, nv, 3)
    t1, t2 = o.amps()
    r1, r2 = B.residual(t1, t2)
    assert max(np.max(np.abs(r1)), np.max(np.abs(r2))) < 1e-14
    s1, s2 = B.solve(np.zeros_like(t1), np.zeros_like(t2), conv=1e-14)
    assert np.max(np.abs(s1 - t1)) < 1e-13 and np.max(np.abs(s2 - t2)) < 1e-13
    e_ref = B.S.ref @ B.H @ B.S.ref
    assert abs(B.energy(t1, t2) - e_ref - o.energy_full()) < 1e-13
    o.close()


@pytest.mark.parametrize("no,nv,seed", [(2, 3, 3), (2, 4, 6)])
def test_static_response_to_general_perturbations(no, nv, seed):
    n = no + nv
    fock, ints, o, B = _setup(no, nv, seed)
    t1, t2 = o.amps()
    B.keep_jacobian(t1, t2)
    rng = np.random.default_rng(seed + 40)
    A = rng.standard_normal((n, n))
    C = rng.standard_normal((n, n))
    C = 0.5 * (C - C.T)  # antisymmetric, as the magnetic dipole
    VA, VC = B.S.one_body(A), B.S.one_body(C)
    PA, PC = o.pert(A, 0.0), o.pert(C, 0.0)
    for P in (PA, PC):
        assert P.solve(0, conv=1e-13)["converged"] and P.solve(1, conv=1e-13)["converged"]
    h = 2e-3
    E0 = B.energy(t1, t2)

    def at(V):
        return [B.solve(t1, t2, V=k * h * V, conv=1e-14) for k in (-2, -1, 1, 2)]

    def d2(V, pts):
        E = [B.energy(*p, V=k * h * V) for k, p in zip((-2, -1, 1, 2), pts)]
        return (-E[0] + 16 * E[1] - 30 * E0 + 16 * E[2] - E[3]) / (12 * h * h)

    pA, pC = at(VA), at(VC)
    for P, pts in ((PA, pA), (PC, pC)):
        X1, X2, _, _ = P.amplitudes()
        for got, k in ((X1, 0), (X2, 1)):
            ref = (pts[0][k] - 8 * pts[1][k] + 8 * pts[2][k] - pts[3][k]) / (12 * h)
            assert np.max(np.abs(got - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref))), k
    dA, dC = d2(VA, pA), d2(VC, pC)
    mixed = 0.5 * (d2(VA + VC, at(VA + VC)) - dA - dC)
    assert abs(PA.linresp(PA) + dA) < 1e-8 * max(1.0, abs(dA))
    assert abs(PC.linresp(PC) + dC) < 1e-8 * max(1.0, abs(dC))
    for val in (PA.linresp(PC), PC.linresp(PA)):
        assert abs(val + mixed) < 1e-8 * max(1.0, abs(mixed)), (val, -mixed)
    # the antisymmetric field has a visible mixed response here
    assert abs(mixed) > 1e-3
    PA.close(); PC.close(); o.close()


@pytest.mark.parametrize("omega", [0.1, -0.07])
def test_dynamic_right_hand_equations_for_a_general_perturbation(omega):
    no, nv = 2, 3
    n = no + nv
    fock, ints, o, B = _setup(no, nv, 8)
    t1, t2 = o.amps()
    A = np.random.default_rng(9).standard_normal((n, n))
    P = o.pert(A, omega)
    assert P.solve(0, conv=1e-13)["converged"]
    X1, X2, _, _ = P.amplitudes()
    # <mu|[Hbar, X]|0> as the derivative of the residual along X (a
    # polynomial in the amplitudes: the five-point difference is exact to
    # rounding at this step)
    h = 1e-3
    r = {k: B.residual(t1 + k * h * X1, t2 + k * h * X2) for k in (-2, -1, 1, 2)}
    JX = [(r[-2][m] - 8 * r[-1][m] + 8 * r[1][m] - r[2][m]) / (12 * h) for m in (0, 1)]
    r0 = B.residual(t1, t2)
    rA = B.residual(t1, t2, V=B.S.one_body(A))
    SX = B.excitation(X1, X2)
    for m in (0, 1):
        lhs = JX[m] + (rA[m] - r0[m]) - omega * SX[m]
        assert np.max(np.abs(lhs)) < 1e-10, (m, np.max(np.abs(lhs)))
    # the frequency term is not negligible
    assert np.max(np.abs(omega * SX[0])) > 1e-3
    P.close(); o.close()
