"""Pins of the FP64 CPU oracle (oracle/ugacc_oracle.c) by what mathematics
fixes, independent of the reference's spin-adapted equations:

- textbook spin-orbital CCSD and (T) (Stanton-Gauss / Raghavachari, in
  tests/spinorb_ref.py) on tiny systems give the same energies;
- for two electrons CCSD is exact: it equals brute-force FCI, and (T) is 0;
- E_MP2 has a closed form; one Jacobi step from t = 0 gives R2 = <ij|ab>;
- E_CCSD is invariant under occupied-occupied and virtual-virtual rotations;
- E_CCSD and E_(T) are size extensive (direct-sum Hamiltonians);
- the three (T) algorithms (tcorr, tcorr_ooc, tcorr_ooc_TJL) agree and the
  i = j = k triplets contribute nothing.
A plausible slip in the restatement (a dropped term, a wrong sign or index, a
transposed operand) breaks at least one of these.
"""
import numpy as np
import pytest

from oracle import Oracle
from tests import spinorb_ref as sr
from ugacc_amd.synthetic import direct_sum, random_hamiltonian, rotate


def _ccsd(fock, ints, no, conv=1e-12, **kw):
    o = Oracle(fock, ints, no)
    r = o.ccsd(maxiter=300, conv=conv, **kw)
    assert r["converged"]
    return o, r


@pytest.mark.parametrize("no,nv,seed,scale", [(2, 4, 1, 0.3), (3, 5, 2, 0.3), (2, 6, 5, 0.2)])
def test_ccsd_and_triples_match_spin_orbital_textbook(no, nv, seed, scale):
    fock, ints = random_hamiltonian(no, nv, seed=seed, scale=scale)
    o, r = _ccsd(fock, ints, no)
    e_so, et_so = sr.triples_spinorb(fock, ints, no)
    assert abs(r["eccsd"] - e_so) < 1e-11
    assert abs(o.tcorr_tjl() - et_so) < 1e-12


def test_ccsd_with_noncanonical_fock_matches_spin_orbital():
    # off-diagonal oo/vv Fock and a nonzero f_ov exercise every Fock term
    fock, ints = random_hamiltonian(2, 5, seed=7, scale=0.3, offdiag=0.05, fov=0.03)
    o, r = _ccsd(fock, ints, 2)
    e_so, _, _, _ = sr.ccsd_spinorb(fock, ints, 2)
    # the oracle reports the reference's energy form (only the last f_ia t1_ia
    # term of ccwfn.cc:272 survives); the full form is the textbook energy
    assert abs(o.energy_full() - e_so) < 1e-11
    assert abs(o.energy() - o.energy_full()) > 1e-8  # the two forms differ here


@pytest.mark.parametrize("nv,seed", [(4, 3), (6, 11)])
def test_two_electron_ccsd_is_fci(nv, seed):
    fock, ints = random_hamiltonian(1, nv, seed=seed, scale=0.3)
    o, r = _ccsd(fock, ints, 1, conv=1e-13)
    assert abs(r["eccsd"] - sr.fci_two_electron(fock, ints)) < 1e-11
    assert abs(o.tcorr_tjl()) < 1e-15


def test_mp2_closed_form_and_first_residual():
    no, nv = 3, 6
    fock, ints = random_hamiltonian(no, nv, seed=9, scale=0.2)
    o = Oracle(fock, ints, no)
    e = np.diag(fock)
    g = ints[:no, :no, no:, no:]
    D = e[:no, None, None, None] + e[None, :no, None, None] - e[None, None, no:, None] - e[None, None, None, no:]
    emp2 = np.sum((2 * g - g.transpose(0, 1, 3, 2)) * g / D)
    assert abs(o.energy() - emp2) < 1e-14
    o.set_amps(np.zeros((no, nv)), np.zeros((no, no, nv, nv)))
    _, R1, R2 = o.iterate()
    np.testing.assert_allclose(R2, g, atol=1e-15)
    np.testing.assert_allclose(R1, fock[:no, no:], atol=1e-15)


def test_orbital_rotation_invariance():
    no, nv = 3, 5
    fock, ints = random_hamiltonian(no, nv, seed=13, scale=0.25)
    _, r0 = _ccsd(fock, ints, no)
    f2, g2 = rotate(fock, ints, no, seed=3, angle=0.2)
    _, r1 = _ccsd(f2, g2, no)
    assert abs(r0["eccsd"] - r1["eccsd"]) < 1e-10


def test_size_extensivity():
    hA = random_hamiltonian(2, 4, seed=21, scale=0.25)
    hB = random_hamiltonian(1, 5, seed=22, scale=0.25)
    oA, rA = _ccsd(*hA, 2)
    oB, rB = _ccsd(*hB, 1)
    f, g = direct_sum(hA, 2, hB, 1)
    oAB, rAB = _ccsd(f, g, 3)
    assert abs(rAB["eccsd"] - (rA["eccsd"] + rB["eccsd"])) < 1e-11
    assert abs(oAB.tcorr_tjl() - (oA.tcorr_tjl() + oB.tcorr_tjl())) < 1e-13


def test_three_triples_algorithms_agree_and_ijk_diagonal_vanishes():
    no, nv = 4, 6
    fock, ints = random_hamiltonian(no, nv, seed=31, scale=0.25)
    o, _ = _ccsd(fock, ints, no)
    a, b, c = o.tcorr(), o.tcorr_ooc(), o.tcorr_tjl()
    assert abs(a - b) < 1e-13 and abs(b - c) < 1e-13
    per = o.tjl_per_triplet()
    assert len(per) == no * (no + 1) * (no + 2) // 6
    k = 0
    for i in range(no):
        for j in range(i + 1):
            for kk in range(j + 1):
                if i == j == kk:
                    assert abs(per[k]) < 1e-18
                k += 1
    assert abs(per.sum() - c) < 1e-14


def test_amplitude_symmetry():
    fock, ints = random_hamiltonian(3, 5, seed=41, scale=0.25)
    o, _ = _ccsd(fock, ints, 3)
    _, t2 = o.amps()
    assert np.max(np.abs(t2 - t2.transpose(1, 0, 3, 2))) < 1e-12


@pytest.mark.parametrize("no,nv,seed,fov,offdiag", [(3, 7, 5, 0.01, 0.0), (3, 6, 9, 0.02, 0.01), (4, 5, 2, 0.03, 0.0)])
def test_triples_fov_terms_fold_into_the_bracket(no, nv, seed, fov, offdiag):
    """With f_ov != 0 the reference returns tcorr (OOC false) or tcorr_ooc
    (OOC true), whose X2 carries f_kc t3 terms (ccwfn.cc:706-709, 725) that
    tcorr_ooc_TJL lacks.  The TJL bracket with t2_pq^xy f_r^z terms added to
    V (the form the GPU kernel evaluates) reproduces both reference values;
    the plain TJL value visibly differs."""
    fock, ints = random_hamiltonian(no, nv, seed=seed, fov=fov, offdiag=offdiag)
    o, _ = _ccsd(fock, ints, no, conv=1e-11)
    a, b = o.tcorr(), o.tcorr_ooc()
    per = o.tjl_fov_per_triplet()
    scale = abs(a)
    assert abs(a - b) < 1e-12 * scale
    assert abs(per.sum() - a) < 1e-12 * scale
    assert abs(o.tcorr_tjl() - a) > 1e-3 * scale
    # with f_ov = 0 the folded and plain per-triplet energies coincide
    f0 = fock.copy()
    f0[:no, no:] = 0.0
    f0[no:, :no] = 0.0
    o0 = Oracle(f0, ints, no)
    o0.set_amps(*o.amps())
    np.testing.assert_allclose(o0.tjl_fov_per_triplet(), o0.tjl_per_triplet(), rtol=0, atol=1e-18)
