"""Pins of the oracle's (T) gradient terms (orc_tgrad, a restatement of
ccwfn.cc:854-1098) with the CCSD(T) Lambda (s1, s2/2 added to the l1/l2
residuals, cclambda.cc:181, :243) and densities ((T) parts copied in first,
ccdensity.cc:43-74), by what mathematics fixes:

- Hellmann-Feynman: the one- and two-particle densities of CCSD(T) are the
  derivatives of the CCSD(T) energy with respect to the Hamiltonian (the
  amplitudes relaxed, the orbitals fixed).  Along a random perturbation of
  one class of integrals ((oo|oo), (vv|vv), (oo|ov), (vv|vo), (oo|vv),
  (ov|ov), with their permutational partners), of the Fock diagonal, or of
  the f_ov block, the five-point difference of E_CCSD + E_(T) (each point a
  full CCSD solve to 1e-14 and the oracle's tcorr) equals the contraction of
  the densities with the perturbation, in the weights of the density energy
  (ccdensity.cc:138-339).  Every block of the (T) parts and both sources
  s1/s2 enter these derivatives: with the CCSD Lambda and densities alone
  each check misses by the (T) share, asserted to be large;
- stationarity: with (l1, l2) the converged CCSD(T) Lambda,
      Lag(t) = E(t) + E_(T)(t) + l1.R1(t) + 1/2 l2.R2(t)
  has zero gradient at the converged t along t1 elements and t2 pair
  directions; E_(T)(t) is the oracle's tcorr (pinned against spin-orbital
  (T) and the reference's known answers);
- at convergence the density energy E_one + E_two is E_CCSD + E_(T), the
  check the reference prints for WFN=CCSD_T (ugacc.cc:129-137), and with
  f_ov != 0 once the f_ov blocks are included.
The (T) terms use diagonal denominators, so these identities hold for Fock
matrices whose oo and vv blocks are diagonal (canonical orbitals), the
reference's setting.
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd.synthetic import dirac_from_chemist, random_hamiltonian
from tests.test_density_oracle import PARTS, _class_mask, _random_chemist


def _solve_t(fock, ints, no, with_t=True, conv=1e-14):
    """CCSD (conv), then tgrad, the CCSD(T) Lambda and the densities; with_t
    False: the CCSD Lambda and densities instead."""
    o = Oracle(fock, ints, no)
    r = o.ccsd(maxiter=500, conv=conv)
    assert r["converged"]
    o.build_tstar()
    et = o.tcorr()
    tg = o.tgrad(enable=with_t)
    o.lambda_init()
    assert o.lambda_solve(maxiter=500, conv=conv)["converged"]
    d = o.density()
    return o, r, et, tg, d


def _energy(fock, ints, no):
    o = Oracle(fock, ints, no)
    r = o.ccsd(maxiter=500, conv=1e-14)
    assert r["converged"]
    e = o.energy_full() + o.tcorr()
    o.close()
    return e


def _contract(B, no, df, dg):
    """sum_pq f_pq D_pq + the two-electron density energy, of a perturbation
    (df, dg) (dg Dirac, nact^4), in the weights of ccdensity.cc:138-339."""
    n = df.shape[0]
    O, V = slice(0, no), slice(no, n)
    e = (np.sum(df[O, O] * B["Doo"]) + np.sum(df[V, V] * B["Dvv"]) + np.sum(df[O, V] * B["Dov"]) +
         np.sum(df[V, O] * B["Dvo"]))
    e += 0.5 * np.sum(dg[O, O, O, O] * B["Goooo"]) + 0.5 * np.sum(dg[V, V, V, V] * B["Gvvvv"])
    e += np.sum(dg[O, O, O, V] * B["Gooov"]) + np.sum(dg[V, V, V, O] * B["Gvvvo"])
    e += np.sum(dg[O, V, O, V] * B["Govov"]) + 0.5 * np.sum(dg[O, O, V, V] * B["Goovv"])
    return e


@pytest.mark.parametrize("no,nv,seed", [(3, 5, 4), (2, 6, 9)])
def test_densities_are_the_derivatives_of_the_ccsd_t_energy(no, nv, seed):
    n = no + nv
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3)
    o, _, _, _, d = _solve_t(fock, ints, no)
    o.close()
    o, _, _, _, d0 = _solve_t(fock, ints, no, with_t=False)
    o.close()
    h = 5e-4  # the truncation error of the stencil (~h^4) stays below 1e-12

    def deriv(df, dg):
        E = [_energy(fock + k * h * df, ints + k * h * dg, no) for k in (-2, -1, 1, 2)]
        return (E[0] - 8 * E[1] + 8 * E[2] - E[3]) / (12 * h)

    rng = np.random.default_rng(seed)
    zf, zg = np.zeros((n, n)), np.zeros((n, n, n, n))
    cases = [("fock diagonal", np.diag(rng.standard_normal(n)), zg)]
    Y = 0.2 * rng.standard_normal((no, nv))  # t1 follows f_ov: a smaller step keeps the stencil exact
    dfov = np.zeros((n, n))
    dfov[:no, no:] = Y
    dfov[no:, :no] = Y.T
    cases.append(("f_ov", dfov, zg))
    g = _random_chemist(n, seed + 50)
    for part in PARTS:
        cases.append((part, zf, dirac_from_chemist(np.where(_class_mask(no, nv, part), g, 0.0))))
    for name, df, dg in cases:
        dE = deriv(df, dg)
        pred = _contract(d["blocks"], no, df, dg)
        pred0 = _contract(d0["blocks"], no, df, dg)
        assert abs(dE - pred) < 5e-12 * max(1.0, abs(dE)) + 2e-12, (name, dE, pred)
        # the (T) parts carry a visible share of every derivative
        assert abs(dE - pred0) > 1e-6, (name, dE, pred0)


def test_ccsd_t_lagrangian_is_stationary():
    no, nv = 2, 4
    fock, ints = random_hamiltonian(no, nv, seed=3, scale=0.3)
    o, _, _, _, _ = _solve_t(fock, ints, no)
    l1, l2 = o.lambdas()
    t1, t2 = o.amps()

    def lag(a1, a2):
        o.set_amps(a1, a2)
        R1, R2 = o.residual_rows(no)
        return o.energy_full() + o.tcorr() + np.sum(l1 * R1) + 0.5 * np.sum(l2 * R2)

    h = 1e-3

    def deriv(fn):
        return (-fn(2 * h) + 8 * fn(h) - 8 * fn(-h) + fn(-2 * h)) / (12 * h)

    g1, g2 = [], []
    for i in range(no):
        for a in range(nv):
            def f(x, i=i, a=a):
                p = t1.copy()
                p[i, a] += x
                return lag(p, t2)
            g1.append(deriv(f))
    for i in range(no):
        for j in range(i, no):
            for a in range(nv):
                for b in range(nv):
                    def f(x, i=i, j=j, a=a, b=b):
                        p = t2.copy()
                        p[i, j, a, b] += x
                        if (i, a) != (j, b):
                            p[j, i, b, a] += x
                        return lag(t1, p)
                    g2.append(deriv(f))
    o.close()
    # E_(T) alone has a gradient of order 1e-3 here; the Lagrangian's vanishes
    assert np.max(np.abs(g1)) < 1e-10 and np.max(np.abs(g2)) < 1e-10, (np.max(np.abs(g1)), np.max(np.abs(g2)))


@pytest.mark.parametrize("kw", [{}, dict(fov=0.05)])
def test_converged_density_energy_is_the_ccsd_t_energy(kw):
    no, nv = 3, 5
    fock, ints = random_hamiltonian(no, nv, seed=4, scale=0.3, **kw)
    o, r, et, tg, d = _solve_t(fock, ints, no)
    e_cc = o.energy_full()
    o.close()
    b = d["blocks"]
    e_one = d["e_one"] + np.sum(fock[:no, no:] * b["Dov"]) + np.sum(fock[no:, :no] * b["Dvo"])
    assert abs(e_one + d["e_two"] - e_cc - et) < 1e-13
    if not kw:
        assert abs(d["e_one"] + d["e_two"] - r["eccsd"] - et) < 1e-12  # the reference's printed check
    # s2 has the pair symmetry of the l2 residual it enters
    assert np.max(np.abs(tg["s2"] - tg["s2"].transpose(1, 0, 3, 2))) < 1e-15


@pytest.mark.parametrize("no,nv,seed", [(3, 5, 4), (2, 6, 9)])
def test_out_of_core_form_is_the_in_core_form_with_diagonal_doo_dvv(no, nv, seed):
    """tgrad_ooc (ccwfn.cc:1150-1316, which compute_energy runs with
    OOC=true) forms W and V one (i,j,k) and one (a,b,c) at a time and
    accumulates only the diagonals of Doo and Dvv: for a canonical Fock
    matrix every other block equals tgrad's to rounding, and Doo, Dvv equal
    the diagonals of tgrad's (whose off-diagonals are not small).  With some
    f_ov != 0 the two forms differ in s2 (tgrad_ooc's V carries t2 f_ov
    terms that tgrad's l3 carries in another form), linearly in f_ov: the
    case ugx_tgrad_ooc refuses beyond |f_ov| = 1e-8."""
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=500, conv=1e-13)["converged"]
    o.build_tstar()
    a = o.tgrad(enable=False)
    b = o.tgrad_ooc(enable=False)
    for k in a:
        if k in ("Doo", "Dvv"):
            assert np.max(np.abs(b[k] - np.diag(np.diag(a[k])))) < 1e-15
            assert np.max(np.abs(a[k] - np.diag(np.diag(a[k])))) > 1e-7
        else:
            assert np.max(np.abs(a[k] - b[k])) < 1e-14 * max(1.0, np.max(np.abs(a[k]))), k
    o.close()
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=0.3, fov=0.03)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=500, conv=1e-13)["converged"]
    o.build_tstar()
    a, b = o.tgrad(enable=False), o.tgrad_ooc(enable=False)
    assert np.max(np.abs(a["s2"] - b["s2"])) > 1e-6
    o.close()
