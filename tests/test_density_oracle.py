"""Pins of the oracle's CCSD densities (orc_density, a restatement of
ccdensity.cc:89-339) by what mathematics fixes:

- the density energy is the CCSD Lagrangian: for ANY amplitudes t and
  multipliers l (not only converged ones),
      sum_pq f_pq D_pq + sum G.<pq|rs> = E(t) + l1.R1(t) + 1/2 l2.R2(t)
  with E = 2 f.t1 + tau.L and R the oracle's raw CCSD residuals, the
  one-electron sum taken over all four Fock blocks (the reference's E_one
  keeps only f_oo and f_vv, which is all of it for f_ov = 0).  The densities
  do not depend on the Hamiltonian and both sides are linear in it, so
  Hamiltonians that hold only the Fock matrix, or only one class of
  integrals ((oo|oo), (vv|vv), (oo|ov), (vv|vo), (oo|vv), (ov|ov) with their
  permutational partners), pin E_one and each of the six parts of E_two
  separately, at random t and l;
- at converged t and l the density energy equals E_CCSD, the check the
  reference prints (ugacc.cc:133-139).
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd.synthetic import dirac_from_chemist, random_hamiltonian

PARTS = Oracle.DENSITY_PARTS
# chemist class (pq|rs) of each part of E_two, by its occupied/virtual pattern
CLASS_OF = {"oooo": ("oo", "oo"), "vvvv": ("vv", "vv"), "ooov": ("oo", "ov"), "vvvo": ("vv", "vo"),
            "ovov": ("oo", "vv"), "oovv": ("ov", "ov")}


def _random_chemist(n, seed):
    g = np.random.default_rng(seed).standard_normal((n, n, n, n))
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    return 0.05 * (g + g.transpose(2, 3, 0, 1))


def _class_mask(no, nv, part):
    """Mask of the chemist elements (pq|rs) in the class of `part`: its pair
    patterns up to the order inside each pair and the order of the pairs."""
    n = no + nv
    kind = np.array(["o"] * no + ["v"] * nv)
    x, y = CLASS_OF[part]
    want = {tuple(sorted(x)) + tuple(sorted(y)), tuple(sorted(y)) + tuple(sorted(x))}
    p, q, r, s = np.meshgrid(*[np.arange(n)] * 4, indexing="ij")
    keys = np.stack([kind[p], kind[q], kind[r], kind[s]], axis=-1)
    mask = np.zeros((n, n, n, n), dtype=bool)
    for a in want:
        m1 = (np.sort(keys[..., :2], axis=-1) == np.array(a[:2])).all(-1)
        m2 = (np.sort(keys[..., 2:], axis=-1) == np.array(a[2:])).all(-1)
        mask |= m1 & m2
    return mask


def _fock(no, nv, seed, fov):
    rng = np.random.default_rng(seed)
    n = no + nv
    f = np.diag(np.concatenate([-1.0 - rng.random(no), 0.5 + rng.random(nv)]))
    off = 0.02 * rng.standard_normal((n, n))
    off = off + off.T
    off[:no, no:] *= fov
    off[no:, :no] *= fov
    return f + off


def _random_amps(no, nv, seed):
    rng = np.random.default_rng(seed)
    out = []
    for s1, s2 in ((0.1, 0.05), (0.2, 0.1)):
        a1 = s1 * rng.standard_normal((no, nv))
        a2 = s2 * rng.standard_normal((no, no, nv, nv))
        out += [a1, a2 + a2.transpose(1, 0, 3, 2)]
    return out


def _lagrangian_and_density(fock, ints, no, t1, t2, l1, l2):
    o = Oracle(fock, ints, no)
    o.set_amps(t1, t2)
    o.set_lambdas(l1, l2)
    R1, R2 = o.residual_rows(no)
    lag = o.energy_full() + np.sum(l1 * R1) + 0.5 * np.sum(l2 * R2)
    d = o.density()
    o.close()
    return lag, d


def _one_electron(fock, no, d):
    b = d["blocks"]
    return d["e_one"] + np.sum(fock[:no, no:] * b["Dov"]) + np.sum(fock[no:, :no] * b["Dvo"])


@pytest.mark.parametrize("no,nv,seed", [(2, 4, 1), (3, 4, 2)])
def test_each_density_block_by_its_part_of_the_lagrangian(no, nv, seed):
    n = no + nv
    t1, t2, l1, l2 = _random_amps(no, nv, seed)
    fock = _fock(no, nv, seed, fov=1.0)
    # Fock only: E_one (+ the f_ov blocks), no two-electron part
    lag, d = _lagrangian_and_density(fock, np.zeros((n, n, n, n)), no, t1, t2, l1, l2)
    assert abs(_one_electron(fock, no, d) - lag) < 1e-13 * max(1.0, abs(lag))
    assert all(abs(d["parts"][k]) == 0.0 for k in PARTS)
    lag0 = lag
    g = _random_chemist(n, seed + 100)
    for part in PARTS:
        ints = dirac_from_chemist(np.where(_class_mask(no, nv, part), g, 0.0))
        lag, d = _lagrangian_and_density(fock, ints, no, t1, t2, l1, l2)
        assert abs(d["parts"][part] - (lag - lag0)) < 1e-13 * max(1.0, abs(lag)), part
        assert abs(d["parts"][part]) > 1e-4, part
        for other in PARTS:
            if other != part:
                assert abs(d["parts"][other]) < 1e-15, (part, other)


@pytest.mark.parametrize("kw", [{}, dict(offdiag=0.03), dict(offdiag=0.03, fov=0.05)])
def test_converged_density_energy_is_the_ccsd_energy(kw):
    no, nv = 3, 5
    fock, ints = random_hamiltonian(no, nv, seed=4, scale=0.3, **kw)
    o = Oracle(fock, ints, no)
    r = o.ccsd(maxiter=400, conv=1e-13)
    assert r["converged"]
    e_cc = o.energy_full()
    o.lambda_init()
    assert o.lambda_solve(maxiter=400, conv=1e-13)["converged"]
    d = o.density()
    e_dens = _one_electron(fock, no, d) + d["e_two"]
    assert abs(e_dens - e_cc) < 1e-12
    if not kw.get("fov"):
        # the reference's own check (f_ov = 0)
        assert abs(d["e_one"] + d["e_two"] - r["eccsd"]) < 1e-12
    o.close()
