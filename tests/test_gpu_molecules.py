"""GPU parity on real molecules built by the front end.

- H2O/cc-pVDZ all-electron at O-H 0.9572 A, 104.52 deg (the single-water
  configuration): full CCSD(T) against the oracle (1e-10 Eh, 1e-9 amplitudes).
- The reference's saved runs: H2O/cc-pVDZ with one frozen core
  (psi3/output.c1_fzc) and HOF/STO-3G (psi4/hof/output.*), GPU energies
  against the printed known answers.
- (H2O)4/cc-pVDZ (o=20, v=76) at full size: the oracle is too slow for a
  whole solve there, so the GPU residual is compared with the oracle's on the
  first t2 rows (and all of R1) at the GPU's converged amplitudes, and the
  GPU per-triplet (T) energies with the oracle's on a spread of triplets.
"""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from ugacc_amd.chem import molecules
from ugacc_amd.chem.scf import Reference
from tests.gpu_util import announce, context

pytestmark = pytest.mark.gpu

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def _solve(ref, **kw):
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    opts = dict(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    opts.update(kw)
    r = c.ccsd(**opts)
    assert r["converged"], r
    return c, r


def test_h2o_single_water_configuration():
    announce("h2o cc-pvdz all-electron parity")
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    assert (ref.no, ref.nv) == (5, 19)
    c, r = _solve(ref)
    o = Oracle(ref.fock, ref.full_ints(), ref.no)
    ro = o.ccsd(maxiter=100, conv=1e-10, mode=0)
    assert abs(r["e_mp2"] - ro["emp2"]) < 1e-12
    assert abs(r["e_ccsd"] - ro["eccsd"]) < 1e-10
    t1o, t2o = o.amps()
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - t1o)) < 1e-9 and np.max(np.abs(t2 - t2o)) < 1e-9
    assert abs(r["e_t"] - o.tcorr_tjl()) < 1e-10


def test_printed_iteration_table_matches_oracle(capfd):
    """The iteration table ugx_ccsd prints with opt.print (ccwfn.cc:194-196,
    224): row 0 holds E_MP2, row k the energy, the T1 norm (t1norm,
    ccwfn.cc:541-553) and the residual norm after iteration k's DIIS, and
    the converged iteration prints no row.  Every row against the oracle
    stopped after exactly k iterations (printed precision: 15 decimals,
    6 decimals, 7 significant digits)."""
    announce("h2o printed iteration table vs oracle")
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    capfd.readouterr()
    c, r = _solve(ref, print=1, wfn_ccsd_t=0)
    ctypes.CDLL(None).fflush(None)
    out = capfd.readouterr().out
    rows = [l.split() for l in out.splitlines() if re.match(r"^\t +\d+ +-?\d", l)]
    assert len(rows) == r["niter"], (len(rows), r["niter"], out[-2000:])
    assert rows[0][0] == "0" and len(rows[0]) == 2
    o = Oracle(ref.fock, ref.full_ints(), ref.no)
    ro = o.ccsd(maxiter=100, conv=1e-10, mode=0)
    assert abs(float(rows[0][1]) - ro["emp2"]) < 1e-12
    assert ro["niter"] == r["niter"]
    norms = []
    for k in range(1, r["niter"]):
        it, e, t1n, rms = rows[k]
        assert int(it) == k
        assert abs(float(e) - ro["trace"][k]) < 1e-11, (k, e, ro["trace"][k])
        ok = Oracle(ref.fock, ref.full_ints(), ref.no)
        rk = ok.ccsd(maxiter=k, conv=1e-10, mode=0)
        assert abs(float(t1n) - ok.t1norm()) <= 6e-7, (k, t1n, ok.t1norm())
        assert abs(float(rms) - rk["rms"]) <= 1e-6 * rk["rms"], (k, rms, rk["rms"])
        norms.append(ok.t1norm())
        ok.close()
    # the column is not constant: a dropped or misplaced term would show
    assert max(norms) - min(norms) > 1e-4
    o.close()


def test_psi3_h2o_known_answers_on_gpu():
    announce("psi3 h2o fc known answers")
    ka = KA["h2o_ccpvdz_fc1"]
    ref = Reference(molecules.PSI3_H2O, "cc-pvdz", nfzc=1)
    c, r = _solve(ref, r_convergence=1e-12)
    assert abs(r["e_mp2"] - ka["ccsd_trace"][0]) < 1e-11
    assert abs(ref.escf + r["e_ccsd"] - ka["e_ccsd_total"]) < 1e-10
    assert abs(ref.escf + r["e_ccsd"] + r["e_t"] - ka["e_ccsd_t_total"]) < 1e-10


@pytest.mark.parametrize("key,nfzc", [("hof_sto3g_c1_allelec", 0), ("hof_sto3g_fc2", 2)])
def test_hof_known_answers_on_gpu(key, nfzc):
    announce("hof " + key)
    ka = KA[key]
    ref = Reference(molecules.HOF, "sto-3g", nfzc=nfzc)
    c, r = _solve(ref, r_convergence=1e-12)
    assert abs(r["e_mp2"] - ka["e_mp2"]) < 1e-11
    assert abs(r["e_ccsd"] - ka["e_ccsd"]) < 1e-10
    assert abs(r["e_t"] - ka["e_t"]) < 1e-12


def test_h2o4_full_size_sampled_parity():
    announce("(h2o)4 full-size sampled parity")
    ref = Reference(molecules.water_cluster(4, seed=1), "cc-pvdz")
    assert (ref.no, ref.nv) == (20, 76)
    c, r = _solve(ref, r_convergence=1e-8, wfn_ccsd_t=0)
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t2 - t2.transpose(1, 0, 3, 2))) < 1e-12
    o = Oracle(ref.fock, ref.full_ints(), ref.no, omp=True)
    o.set_amps(t1, t2)
    rows = 2
    R1o, R2o = o.residual_rows(rows)
    R1, R2 = c.residual(t1, t2)
    assert np.max(np.abs(R1 - R1o)) < 1e-12
    assert np.max(np.abs(R2[:rows] - R2o)) < 1e-12
    per = c.triples_per_triplet()
    nt = len(per)
    for t in range(0, nt, nt // 7):
        assert abs(per[t] - o.tcorr_tjl_range(t, t + 1)) < 1e-13, t


@pytest.mark.parametrize("form,ladder", [("bf16x3", "packed"), ("fp32", "packed"), ("bf16x3", "full")])
def test_refinement_polish_reaches_the_fp64_fixed_point(form, ladder, monkeypatch):
    """Low-precision iterations to ||R|| < 1e-5, then the refinement polish:
    one FP64 residual at the switch point T_s, then residuals R(T_s) + [R(T) -
    R(T_s)] whose large contractions are low-precision products with a
    difference from T_s as one operand.  At the converged amplitudes the FP64
    residual, evaluated independently by ugx_residual, must be below the
    threshold; the energy is the FP64 form's and the iteration count the
    plain FP64 polish's (+-1).  Both the pair-packed tau terms and the full
    contractions (UGX_LADDER=full, read when a context is created)."""
    announce("refinement polish (h2o)2 %s %s" % (form, ladder))
    ref = Reference(molecules.water_cluster(2, seed=3), "cc-pvdz")
    if ladder == "full":
        monkeypatch.setenv("UGX_LADDER", "full")
        c = ugx.Context(0)
    else:
        c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r64 = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=0)
    common = dict(maxiter=100, r_convergence=1e-9, wfn_ccsd_t=0, precision=ugx.PREC_NAMES[form],
                  polish_threshold=1e-5)
    rp = c.ccsd(**common)
    rr = c.ccsd(polish_mode=ugx.POLISH_REFINE, **common)
    t1, t2 = c.amplitudes()
    R1, R2 = c.residual(t1, t2, ugx.PREC_FP64)
    assert r64["converged"] and rp["converged"] and rr["converged"]
    assert rr["niter_low"] == rp["niter_low"] > 0
    assert abs(rr["niter"] - rp["niter"]) <= 1, (rr["niter"], rp["niter"])
    assert np.sqrt((R1 ** 2).sum() + (R2 ** 2).sum()) < 2e-9
    assert abs(rr["e_ccsd_final"] - r64["e_ccsd_final"]) < 1e-9
    if ladder == "full":
        c.close()


@pytest.mark.parametrize("form", ["fp32", "bf16x3"])
def test_mixed_precision_forms_reach_the_fp64_fixed_point(form):
    """Low-precision iterations to ||R|| < 1e-5, then FP64 polish: the same
    CCSD fixed point as the FP64 form (|dE| <= 1e-9).  (T) evaluated in the
    low precision (FP64 energy sums and denominators) stays within 1e-7
    (FP32) / 5e-7 (BF16x3, whose dropped lo*lo term is ~2^-16 per product)
    relative, far inside the 1 microhartree budget of the mixed form."""
    announce("mixed form " + form + " (h2o)2")
    ref = Reference(molecules.water_cluster(2, seed=3), "cc-pvdz")
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r64 = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    rm = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1, precision=ugx.PREC_NAMES[form],
                polish_threshold=1e-5)
    assert r64["converged"] and rm["converged"]
    assert rm["niter_low"] > 0
    assert abs(rm["e_ccsd"] - r64["e_ccsd"]) < 1e-9
    rel = 1e-7 if form == "fp32" else 5e-7
    assert abs(rm["e_t"] - r64["e_t"]) < rel * abs(r64["e_t"])
