"""Linear response on the GPU (ugx_pert_*, ugx_linresp: FP64 engine
contractions from the term tables of ugx_ccsd.cpp) against the oracle's
restatement of ccpert.cc / ccresp.cc.

- synthetic o=2..4 at omega = 0 and omega != 0, one perturbation symmetric
  and one general: at the library's converged CCSD and Lambda amplitudes
  (given to the oracle), the right- and left-hand
  residuals at random amplitudes to 1e-12 of their scale; the solved X and Y
  to 1e-10; the pseudoresponses and linresp of pairs of perturbations, with
  the right-hand amplitudes only (the reference's MYHAND=RIGHT, Y at its
  guess) and with both, to 1e-10 relative;
- H2O/cc-pVDZ dipole perturbations ("Mu" of the front end): the oracle's
  residuals vanish at the library's converged X and Y, and the static
  polarizability <<mu_z;mu_z>> equals -d2E/dF2 from CCSD solves of the
  library at F = +-h, +-2h (the identity the oracle tests pin), also for
  (H2O)2 where no oracle runs;
- the plugin mirror's DERTYPE=SECOND (ugacc.cc:143-241, both MYHAND): the
  polarizability and optical rotation tensors of H2O rotate with the
  molecule.
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from tests.gpu_util import announce, context, loaded

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def _sym(n, seed):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return 0.5 * (A + A.T)


def _ground(c, conv=1e-12):
    assert c.ccsd(maxiter=300, r_convergence=conv)["converged"]
    assert c.lambda_solve(maxiter=300, r_convergence=conv)["converged"]


def _oracle_like(c, fock, ints, no):
    o = Oracle(fock, ints, no)
    o.set_amps(*c.amplitudes())
    o.lambda_init()
    o.set_lambdas(*c.lambdas())
    return o


@pytest.mark.parametrize("no,nv,seed,omega", [(2, 4, 51, 0.0), (3, 6, 52, 0.12), (4, 7, 53, -0.08)])
def test_response_matches_oracle(no, nv, seed, omega):
    announce("response o=%d v=%d omega=%g" % (no, nv, omega))
    n = no + nv
    c, fock, ints = loaded(no, nv, seed, scale=0.3)
    _ground(c)
    o = _oracle_like(c, fock, ints, no)
    # B general: an antisymmetric part as the magnetic dipole ("L") has
    A, B = _sym(n, seed + 1), np.random.default_rng(seed + 2).standard_normal((n, n))
    ga, gb = c.pert(A, omega), c.pert(B, omega)
    oa, ob = o.pert(A, omega), o.pert(B, omega)
    # the guesses and the right-hand residual at random amplitudes
    for g, q in ((ga, oa), (gb, ob)):
        for x, y in zip(g.amplitudes(), q.amplitudes()):
            assert _rel(x, y) < 1e-12
    rng = np.random.default_rng(seed)
    R1 = 0.1 * rng.standard_normal((no, nv))
    R2 = 0.05 * rng.standard_normal((no, no, nv, nv))
    R2 = R2 + R2.transpose(1, 0, 3, 2)
    for x, y in zip(ga.residual(0, R1, R2), oa.residual(0, R1, R2)):
        assert _rel(x, y) < 1e-12
    # right-hand solves, then linresp with Y at its guess (MYHAND=RIGHT)
    for g, q in ((ga, oa), (gb, ob)):
        rg = g.solve(0, maxiter=200, r_convergence=1e-12)
        rq = q.solve(0, conv=1e-12)
        assert rg["converged"] and rq["converged"]
        assert abs(rg["pseudoresponse"] - rq["pseudoresponse"]) < 1e-10 * max(1.0, abs(rq["pseudoresponse"]))
        assert _rel(g.amplitudes()[0], q.amplitudes()[0]) < 1e-10
        assert _rel(g.amplitudes()[1], q.amplitudes()[1]) < 1e-10
    for x, y in ((ga, gb), (gb, ga), (ga, ga)):
        ox = oa if x is ga else ob
        oy = oa if y is ga else ob
        assert abs(x.linresp(y) - ox.linresp(oy)) < 1e-10 * max(1.0, abs(ox.linresp(oy)))
    # the left-hand residual (with the solved X) and the left-hand solves
    for x, y in zip(ga.residual(1, R1, R2), oa.residual(1, R1, R2)):
        assert _rel(x, y) < 1e-12
    for g, q in ((ga, oa), (gb, ob)):
        assert g.solve(1, maxiter=200, r_convergence=1e-12)["converged"]
        assert q.solve(1, conv=1e-12)["converged"]
        assert _rel(g.amplitudes()[2], q.amplitudes()[2]) < 1e-10
        assert _rel(g.amplitudes()[3], q.amplitudes()[3]) < 1e-10
    for x, y in ((ga, gb), (gb, ga)):
        ox = oa if x is ga else ob
        oy = oa if y is ga else ob
        assert abs(x.linresp(y) - ox.linresp(oy)) < 1e-10 * max(1.0, abs(ox.linresp(oy)))
    for g in (ga, gb):
        g.release()
    oa.close(); ob.close(); o.close()


def test_left_hand_needs_the_right_hand_amplitudes():
    announce("response state rules")
    c, fock, ints = loaded(2, 4, 54, scale=0.3)
    _ground(c)
    g = c.pert(_sym(6, 3), 0.0)
    with pytest.raises(ugx.UgxError):
        g.solve(1)
    g.release()
    with pytest.raises(ugx.UgxError):
        g.solve(0)


def _field_check(ref, comp, conv, oracle):
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    _ground(c, conv)
    mu = ref.perturbation("Mu")[comp]
    g = c.pert(mu, 0.0)
    assert g.solve(0, maxiter=200, r_convergence=conv)["converged"]
    assert g.solve(1, maxiter=200, r_convergence=conv)["converged"]
    alpha = g.linresp(g)
    if oracle is not None:
        o = _oracle_like(c, ref.fock, oracle, ref.no)
        q = o.pert(mu, 0.0)
        X1, X2, Y1, Y2 = g.amplitudes()
        for hand, a1, a2 in ((0, X1, X2), (1, Y1, Y2)):
            if hand == 1:
                q.set_amplitudes(X1=X1, X2=X2)
            r1, r2 = q.residual(hand, a1, a2)
            assert max(np.max(np.abs(r1)), np.max(np.abs(r2))) < 50 * conv, hand
        q.close(); o.close()
    g.release()
    # -d2E/dF2 from the library's own CCSD solves in the field F mu
    h = 2e-3
    E = {}
    for k in (-2, -1, 0, 1, 2):
        c.load_hamiltonian(ref.no, ref.nv, ref.fock + k * h * mu, ref.blocks)
        r = c.ccsd(maxiter=300, r_convergence=1e-12)
        assert r["converged"]
        E[k] = c.energy(form=ugx.ENERGY_FULL)
    d2 = (-E[-2] + 16 * E[-1] - 30 * E[0] + 16 * E[1] - E[2]) / (12 * h * h)
    print("<<mu;mu>> %.10f   -d2E/dF2 %.10f" % (alpha, -d2))
    return alpha, -d2


def test_h2o_static_polarizability():
    announce("response h2o cc-pvdz")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    alpha, fd = _field_check(ref, 2, 1e-11, ref.full_ints())
    assert abs(alpha - fd) < 2e-6 * abs(fd)


def test_h2o2_static_polarizability():
    announce("response (h2o)2 seed 200")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.random_dimer(200), "cc-pvdz")
    alpha, fd = _field_check(ref, 0, 1e-10, None)
    assert abs(alpha - fd) < 2e-6 * abs(fd)


def _plugin_second(atoms, hand, omega):
    from ugacc_amd import plugin
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(atoms), "cc-pvdz")
    opts = dict(plugin.read_options(), DERTYPE="SECOND", MYHAND=hand, MY_OMEGA=omega, R_CONVERGENCE=1e-10)
    plugin.ugacc(ref, opts)
    ax = plugin.CART
    P = np.array([[ref.polarizabilities["<<Mu_%s;Mu_%s>>" % (p, q)] for q in ax] for p in ax])
    G = np.array([[ref.rotations["<<Mu_%s;L_%s>>" % (p, q)] for q in ax] for p in ax])
    return P, G


@pytest.mark.parametrize("hand", ["LEFT", "RIGHT"])
def test_plugin_second_derivative_tensors_rotate_with_the_molecule(hand):
    """The plugin mirror's DERTYPE=SECOND (ugacc.cc:143-241) at 589 nm: the
    polarizability and the optical rotation tensor (gauge origin at O, the
    rotation centre) of a rotated H2O are R P R^T and R G R^T; in its own
    frame (C2v, molecule in xz) the polarizability is diagonal."""
    announce("plugin DERTYPE=SECOND h2o, MYHAND=%s" % hand)
    from ugacc_amd.chem import molecules
    omega = 0.0773571
    R = molecules.random_rotation(np.random.default_rng(11))
    w0 = molecules.water()
    w1 = [(s, R @ np.asarray(x)) for s, x in w0]
    P0, G0 = _plugin_second(w0, hand, omega)
    P1, G1 = _plugin_second(w1, hand, omega)
    print("P0", np.diag(P0), "G0 max", np.max(np.abs(G0)))
    assert np.max(np.abs(P0 - P0.T)) < 1e-12
    assert np.max(np.abs(P0 - np.diag(np.diag(P0)))) < 1e-8 * np.max(np.abs(P0))
    assert np.max(np.abs(P1 - R @ P0 @ R.T)) < 1e-7 * np.max(np.abs(P0))
    assert np.max(np.abs(G0)) > 1e-3
    assert np.max(np.abs(G1 - R @ G0 @ R.T)) < 1e-7 * np.max(np.abs(G0))
