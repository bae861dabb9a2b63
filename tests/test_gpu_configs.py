"""The benchmark configurations at full size on the GPU.

- (H2O)8/cc-pVDZ, seed 8 (o=40, v=152), the headline workload: the FP64
  form to 1e-10 and the mixed form (BF16x3, learned guess, refinement
  polish, 1e-7) with |dE_total| <= 1 uEh; at the FP64 amplitudes, R1 and the
  first row of R2 against the oracle's residual, and the (T) energies of 8
  triplets against the oracle's (the value tcorr returns, per triplet), to
  1e-12.
- The batch workload: 32 distorted H2O monomers and 32 (H2O)2 dimers in one
  ugx_batch_solve, FP64 and mixed: every member against its own single FP64
  solve, and 2 monomers and 2 dimers against complete oracle solves
  (1e-10 Eh; amplitudes of the single solves to 1e-9).
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from ugacc_amd.chem import molecules
from ugacc_amd.chem.scf import Reference
from ugacc_amd.hamiltonian import full_from_blocks
from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def test_h2o8_mixed_learned_vs_fp64_and_sampled_oracle():
    announce("(h2o)8 seed 8: front end")
    ref = Reference(molecules.water_cluster(8, seed=8), "cc-pvdz")
    no, nv = ref.no, ref.nv
    assert (no, nv) == (40, 152)
    from ugacc_amd.learned import WEIGHTS, load_network
    net = load_network(WEIGHTS)
    c = ugx.Context(0)
    c.set_autotune(True)
    c.set_guess_network(net)
    c.load_hamiltonian(no, nv, ref.fock, ref.blocks)
    announce("(h2o)8 seed 8: FP64 to 1e-10")
    r64 = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    t1, t2 = c.amplitudes()
    announce("(h2o)8 seed 8: mixed + learned, refinement polish, 1e-7")
    rm = c.ccsd(maxiter=100, r_convergence=1e-7, wfn_ccsd_t=1, precision=ugx.PREC_BF16X3, polish_threshold=1e-5,
                polish_mode=ugx.POLISH_REFINE, guess=ugx.GUESS_MLP)
    assert r64["converged"] and rm["converged"] and rm["niter_low"] > 0
    d_ccsd = abs(rm["e_ccsd_final"] - r64["e_ccsd_final"])
    d_t = abs(rm["e_t"] - r64["e_t"])
    d_tot = abs(rm["e_ccsd_final"] + rm["e_t"] - r64["e_ccsd_final"] - r64["e_t"])
    print("(H2O)8: |dE_CCSD| %.2e  |dE_(T)| %.2e  |dE_total| %.2e  (FP64 %d iterations, mixed %d)"
          % (d_ccsd, d_t, d_tot, r64["niter"], rm["niter"]))
    assert d_tot <= 1e-6 and d_ccsd <= 1e-6 and d_t <= 1e-6
    # residual at the FP64 amplitudes (R ~ 1e-10 there: compare scales of
    # the terms, i.e. against the oracle's own values to 1e-12 absolute)
    R1, R2 = c.residual(t1, t2, ugx.PREC_FP64)
    per = c.triples_per_triplet()
    c.close()
    announce("(h2o)8 seed 8: oracle row sample and 8 triplets")
    o = Oracle(ref.fock, full_from_blocks(no, nv, ref.blocks), no, omp=True)
    o.set_amps(t1, t2)
    R1o, R2o = o.residual_rows(1)
    assert np.max(np.abs(R1 - R1o)) < 1e-12
    assert np.max(np.abs(R2[:1] - R2o)) < 1e-12
    nt = len(per)
    idx = np.array([0, 1, nt // 7, nt // 3, nt // 2, (2 * nt) // 3, nt - 2, nt - 1])
    eo = o.tjl_fov_list(idx)
    o.close()
    assert np.max(np.abs(per[idx] - eo)) < 1e-12, np.max(np.abs(per[idx] - eo))


def test_batch64_workload_vs_single_solves_and_oracle():
    announce("batch64: front end")
    geoms = [molecules.distorted_monomer(100 + k) for k in range(32)]
    geoms += [molecules.random_dimer(200 + k) for k in range(32)]
    refs = [Reference(a, "cc-pvdz") for a in geoms]
    problems = [(r.no, r.nv, r.fock, r.blocks) for r in refs]
    c = ugx.Context(0)
    c.set_autotune(True)
    announce("batch64: single FP64 solves")
    single, amps = [], {}
    checked = (0, 1, 32, 33)
    for q, (no, nv, fock, blocks) in enumerate(problems):
        c.load_hamiltonian(no, nv, fock, blocks)
        r = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
        assert r["converged"]
        single.append(r)
        if q in checked:
            amps[q] = c.amplitudes()
    announce("batch64: batch FP64 and mixed (graph replays, 8 workers)")
    res64, _ = c.batch_solve(problems, nworkers=8, wfn_ccsd_t=1, maxiter=100, r_convergence=1e-10, graph=1)
    resm, _ = c.batch_solve(problems, nworkers=8, wfn_ccsd_t=1, maxiter=100, r_convergence=1e-7, graph=1,
                            precision=ugx.PREC_BF16X3, polish_threshold=1e-5)
    c.close()
    for q in range(len(problems)):
        a, b, m = single[q], res64[q], resm[q]
        assert b["converged"] and m["converged"]
        assert b["niter"] == a["niter"], q
        assert abs(b["e_ccsd_final"] - a["e_ccsd_final"]) < 1e-13 and abs(b["e_t"] - a["e_t"]) < 1e-13, q
        assert abs(m["e_ccsd_final"] + m["e_t"] - a["e_ccsd_final"] - a["e_t"]) < 1e-6, q
        assert b["device_allocs"] == 0 and m["device_allocs"] == 0, q
    announce("batch64: oracle solves of 2 monomers and 2 dimers")
    for q in checked:
        no, nv, fock, blocks = problems[q]
        o = Oracle(fock, full_from_blocks(no, nv, blocks), no, omp=True)
        ro = o.ccsd(maxiter=100, conv=1e-10, mode=0)
        et = o.tjl_fov_range(0, o.ntriplets())
        t1o, t2o = o.amps()
        o.close()
        assert abs(res64[q]["e_ccsd"] - ro["eccsd"]) < 1e-10, q
        assert abs(res64[q]["e_t"] - et) < 1e-10, q
        t1, t2 = amps[q]
        assert np.max(np.abs(t1 - t1o)) < 1e-9 and np.max(np.abs(t2 - t2o)) < 1e-9, q
