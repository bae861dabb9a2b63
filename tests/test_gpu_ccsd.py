"""Parity of the gfx950 CCSD / (T) path with the FP64 oracle on seeded
synthetic Hamiltonians: intermediates, residuals, converged energies and
amplitudes, W3 and the per-triplet (T) energies.

Bars (FP64 form): energies 1e-10 Eh, amplitudes 1e-9 max-abs; per-function
outputs 1e-11 relative to their scale."""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from tests.gpu_util import loaded, announce

pytestmark = pytest.mark.gpu

CASES = [
    dict(no=2, nv=5, seed=1),
    dict(no=3, nv=7, seed=2, offdiag=0.02),
    dict(no=4, nv=9, seed=3, offdiag=0.01, fov=0.01),
    dict(no=5, nv=19, seed=4, scale=0.05),
    dict(no=6, nv=22, seed=5, offdiag=0.01),  # long-K terms take the split-K path
]


def _case_id(c):
    return "o%d_v%d_s%d" % (c["no"], c["nv"], c["seed"])


def _load(case):
    kw = {k: v for k, v in case.items() if k not in ("no", "nv", "seed")}
    return loaded(case["no"], case["nv"], case["seed"], **kw)


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_residual_and_intermediates(case):
    announce("residual " + _case_id(case))
    c, fock, ints = _load(case)
    no = case["no"]
    orc = Oracle(fock, ints, no)
    rng = np.random.default_rng(case["seed"] + 100)
    t1 = 0.02 * rng.standard_normal((no, c.nv))
    t2 = 0.02 * rng.standard_normal((no, no, c.nv, c.nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    orc.set_amps(t1, t2)
    _, R1o, R2o = orc.iterate()
    mo = orc.intermediates()
    R1, R2 = c.residual(t1, t2)
    mg = c.intermediates()
    for k in mo:
        s = max(1.0, np.max(np.abs(mo[k])))
        assert np.max(np.abs(mg[k] - mo[k])) < 1e-11 * s, k
    assert np.max(np.abs(R1 - R1o)) < 1e-11 * max(1.0, np.max(np.abs(R1o)))
    assert np.max(np.abs(R2 - R2o)) < 1e-11 * max(1.0, np.max(np.abs(R2o)))


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_ccsd_and_triples_parity(case):
    announce("ccsd+triples " + _case_id(case))
    c, fock, ints = _load(case)
    no = case["no"]
    orc = Oracle(fock, ints, no)
    ro = orc.ccsd(maxiter=100, conv=1e-10, diis=True, mode=0, diis_policy=0, min_vecs=3)
    rg = c.ccsd(maxiter=100, r_convergence=1e-10, diis=1, diis_policy=0, diis_min_vecs=3)
    assert ro["converged"] and rg["converged"]
    assert abs(rg["e_mp2"] - ro["emp2"]) < 1e-12
    assert abs(rg["e_ccsd"] - ro["eccsd"]) < 1e-10
    t1o, t2o = orc.amps()
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - t1o)) < 1e-9
    assert np.max(np.abs(t2 - t2o)) < 1e-9
    # (T) from the same (oracle) amplitudes: per-triplet and total.  The
    # value the reference returns is tcorr (OOC false) or tcorr_ooc (OOC
    # true), f_ov terms included (the fov case); per triplet it is the TJL
    # bracket with those terms folded in (equal to plain TJL when f_ov = 0)
    c.set_amplitudes(t1o, t2o)
    et_ref = orc.tcorr()
    per_o = orc.tjl_fov_per_triplet()
    per_g = c.triples_per_triplet()
    assert np.max(np.abs(per_g - per_o)) < 1e-13
    et_g = c.triples()
    assert abs(et_g - et_ref) < 1e-12
    assert abs(et_g - orc.tcorr_ooc()) < 1e-12
    W_o = orc.w3(no - 1, no // 2, 0)
    W_g = c.w3(no - 1, no // 2, 0)
    assert np.max(np.abs(W_g - W_o)) < 1e-12 * max(1.0, np.max(np.abs(W_o)))


@pytest.mark.parametrize("no,nv", [(3, 70), (2, 130)])
def test_triples_large_virtual_space(no, nv):
    """(T) with v above 64 and 128, where the W3 GEMMs use the 80- and
    160-wide tiles: per-triplet energies and one W3 against the oracle, from
    seeded amplitudes."""
    announce("triples o=%d v=%d" % (no, nv))
    c, fock, ints = loaded(no, nv, seed=nv)
    orc = Oracle(fock, ints, no, omp=True)
    rng = np.random.default_rng(nv)
    t1 = 0.01 * rng.standard_normal((no, nv))
    t2 = 0.01 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    orc.set_amps(t1, t2)
    c.set_amplitudes(t1, t2)
    per_o = orc.tjl_per_triplet()
    per_g = c.triples_per_triplet()
    assert np.max(np.abs(per_g - per_o)) < 1e-12 * max(1e-6, np.max(np.abs(per_o))) / 1e-6
    W_o = orc.w3(no - 1, 0, 0)
    W_g = c.w3(no - 1, 0, 0)
    assert np.max(np.abs(W_g - W_o)) < 1e-13


@pytest.mark.parametrize("no,nv", [(8, 40), (5, 19)])
def test_triples_low_precision_forms(no, nv):
    """(T) with the W3 build in FP32 and in BF16x3 (pre-split pair operands;
    o, v multiples of 8 take the 16-byte load path) against the FP64 build:
    per-triplet energies to 2e-6 relative of the largest, W3 to 1e-5
    relative."""
    announce("triples low precision o=%d v=%d" % (no, nv))
    c, fock, ints = loaded(no, nv, seed=no * 100 + nv)
    rng = np.random.default_rng(no + nv)
    t1 = 0.01 * rng.standard_normal((no, nv))
    t2 = 0.02 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    c.set_amplitudes(t1, t2)
    ref = c.triples_per_triplet(ugx.PREC_FP64)
    W64 = c.w3(no - 1, 1, 0, ugx.PREC_FP64)
    scale = np.max(np.abs(ref))
    for prec in (ugx.PREC_FP32, ugx.PREC_BF16X3):
        got = c.triples_per_triplet(prec)
        assert np.max(np.abs(got - ref)) < 2e-6 * scale, prec
        W = c.w3(no - 1, 1, 0, prec)
        assert np.max(np.abs(W - W64)) < 1e-5 * np.max(np.abs(W64)), prec


@pytest.mark.parametrize("no,nv", [(5, 19), (1, 9), (6, 33)])
def test_packed_ladder_matches_full_contraction(no, nv, monkeypatch):
    """The ladder as two packed GEMMs over the pairs i>=j, a>=b, e>=f (the
    default for real-orbital integrals) and as the full contraction
    (UGX_LADDER=full, read when a context is created): R2 in FP64 to
    rounding, and in BF16x3 to its precision.  o = 1 has no antisymmetric
    part."""
    announce("packed vs full ladder o=%d v=%d" % (no, nv))
    c, fock, ints = loaded(no, nv, seed=no * 13 + nv)
    rng = np.random.default_rng(no + 7 * nv)
    t1 = 0.02 * rng.standard_normal((no, nv))
    t2 = 0.02 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    R1p, R2p = c.residual(t1, t2)
    R1b, R2b = c.residual(t1, t2, ugx.PREC_BF16X3)
    monkeypatch.setenv("UGX_LADDER", "full")
    full = ugx.Context(0)
    full.load_hamiltonian(no, nv, fock, ugx.split_blocks(ints, no))
    R1f, R2f = full.residual(t1, t2)
    R1fb, R2fb = full.residual(t1, t2, ugx.PREC_BF16X3)
    full.close()
    s = max(1.0, np.max(np.abs(R2f)))
    assert np.max(np.abs(R2p - R2f)) < 1e-13 * s
    np.testing.assert_array_equal(R1p, R1f)
    assert np.max(np.abs(R2b - R2fb)) < 2e-6 * s
    np.testing.assert_allclose(R2p, R2p.transpose(1, 0, 3, 2), atol=1e-15 * s)


@pytest.mark.parametrize("no,nv", [(8, 32), (6, 40), (4, 24), (5, 32)])
def test_bf16x3_panel_operands_match_pair_arrays(no, nv, monkeypatch):
    """BF16x3 contractions whose K is whole 32-slices take their operands as
    MFMA panels (the LDS-copy kernel); UGX_CCSD_PANELS=0 (read when a context
    is created) keeps the hi/lo pair arrays and the register-staged kernel.
    Same split values, same MFMA order: the residuals agree to rounding."""
    announce("panel vs pair BF16x3 residual o=%d v=%d" % (no, nv))
    c, fock, ints = loaded(no, nv, seed=no * 7 + nv)
    rng = np.random.default_rng(no + nv)
    t1 = 0.02 * rng.standard_normal((no, nv))
    t2 = 0.02 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    R1p, R2p = c.residual(t1, t2, ugx.PREC_BF16X3)
    monkeypatch.setenv("UGX_CCSD_PANELS", "0")
    pairs = ugx.Context(0)
    pairs.load_hamiltonian(no, nv, fock, ugx.split_blocks(ints, no))
    R1q, R2q = pairs.residual(t1, t2, ugx.PREC_BF16X3)
    pairs.close()
    s = max(1.0, np.max(np.abs(R2q)))
    assert np.max(np.abs(R2p - R2q)) < 1e-12 * s
    assert np.max(np.abs(R1p - R1q)) < 1e-12 * s


@pytest.mark.gpu
@pytest.mark.parametrize("no,nv,env,val", [(8, 40, "UGX_T_DMA", "3"), (5, 19, "UGX_T_DMA", "0"),
                                           (3, 70, "UGX_T_DMA", "0"), (6, 38, "UGX_T_MERGE", "0"),
                                           (4, 130, "UGX_T_MERGE", "0"), (5, 38, "UGX_T_ZAFIRST", "0"),
                                           (4, 152, "UGX_T_ZAFIRST", "0")])
def test_triples_bf16x3_lds_copy_matches_register_path(no, nv, env, val, monkeypatch):
    """The default BF16x3 W3 build (panel operands copied straight into LDS,
    two stages, the z = a group's launch writing W, then the z = c and z = b
    permutation groups merged into one accumulating launch per (triplet, a))
    against the other W3 kernels, selected when a
    context is created: UGX_T_DMA=3 (three LDS stages) runs the same MFMA
    products in the same order, so per-triplet energies and W3 agree
    bitwise; UGX_T_DMA=0 (register-staged operands) and UGX_T_MERGE=0 (three
    launches) round the z = c group's FP32 partial sums into W before the
    z = b group adds to them, and UGX_T_ZAFIRST=0 (the merged launch writes W,
    the z = a launch accumulates) rounds in the other order, so they agree to
    FP32 rounding (1e-6 of the largest value).  v = 19, 38, 70 and 130 pad the panels' rows and k (and
    the merged launch's a-slabs); o = 3 and 5 leave the hole terms' k mostly
    padding."""
    announce("triples bf16x3 default vs %s=%s o=%d v=%d" % (env, val, no, nv))
    c, fock, ints = loaded(no, nv, seed=no * 31 + nv)
    rng = np.random.default_rng(no * nv)
    t1 = 0.01 * rng.standard_normal((no, nv))
    t2 = 0.02 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    c.set_amplitudes(t1, t2)
    e_copy = c.triples_per_triplet(ugx.PREC_BF16X3)
    w_copy = c.w3(no - 1, 1, 0, ugx.PREC_BF16X3)
    monkeypatch.setenv(env, val)
    other = ugx.Context(0)
    other.load_hamiltonian(no, nv, fock, ugx.split_blocks(ints, no))
    other.set_amplitudes(t1, t2)
    e_other = other.triples_per_triplet(ugx.PREC_BF16X3)
    w_other = other.w3(no - 1, 1, 0, ugx.PREC_BF16X3)
    other.close()
    if env == "UGX_T_DMA" and val == "3":
        np.testing.assert_array_equal(e_copy, e_other)
        np.testing.assert_array_equal(w_copy, w_other)
    else:
        assert np.max(np.abs(e_copy - e_other)) <= 1e-6 * np.max(np.abs(e_copy))
        assert np.max(np.abs(w_copy - w_other)) <= 1e-6 * np.max(np.abs(w_copy))
    ref = c.triples_per_triplet(ugx.PREC_FP64)
    assert np.max(np.abs(e_copy - ref)) < 2e-6 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_autotuned_tiles_give_the_same_results():
    # the autotuner only changes which tile computes each contraction
    announce("autotune on vs off, o=6 v=40, CCSD(T) FP64 and BF16x3")
    c, fock, ints = loaded(6, 40, seed=61)
    ref = c.ccsd(r_convergence=1e-10, wfn_ccsd_t=1)
    ref3 = c.ccsd(r_convergence=1e-9, wfn_ccsd_t=1, precision=ugx.PREC_BF16X3)
    c.set_autotune(True)
    try:
        r = c.ccsd(r_convergence=1e-10, wfn_ccsd_t=1)
        r3 = c.ccsd(r_convergence=1e-9, wfn_ccsd_t=1, precision=ugx.PREC_BF16X3)
        table = c.tile_table()
    finally:
        c.set_autotune(False)
    assert abs(r["e_ccsd_final"] - ref["e_ccsd_final"]) < 1e-11
    assert abs(r["e_t"] - ref["e_t"]) < 1e-12
    assert abs(r3["e_ccsd_final"] - ref3["e_ccsd_final"]) < 1e-9
    assert abs(r3["e_t"] - ref3["e_t"]) < 1e-10
    assert table.count("->") >= 10, table


@pytest.mark.gpu
@pytest.mark.parametrize("prec,polish", [(ugx.PREC_FP64, ugx.POLISH_PLAIN), (ugx.PREC_BF16X3, ugx.POLISH_PLAIN),
                                         (ugx.PREC_BF16X3, ugx.POLISH_REFINE)])
def test_graph_replayed_iterations_are_bitwise_eager(prec, polish):
    # every iteration after the first of each precision (and polish mode)
    # replays captured residual graphs, with the eager update / DIIS / energy
    # work and host synchronisations between replays; the trajectory must be
    # identical.  The refinement polish's base residual always runs eagerly.
    announce("graph vs eager iterations o=5 v=19 prec=%d polish=%d" % (prec, polish))
    c, fock, ints = loaded(5, 19, seed=71)
    conv = 1e-10 if polish == ugx.POLISH_PLAIN else 1e-9
    kw = dict(r_convergence=conv, wfn_ccsd_t=1, precision=prec, polish_mode=polish)
    ra = c.ccsd(**kw)
    t1a, t2a = c.amplitudes()
    rb = c.ccsd(graph=1, **kw)
    t1b, t2b = c.amplitudes()
    rc = c.ccsd(graph=1, **kw)  # replays the graphs of rb
    t1c, t2c = c.amplitudes()
    assert ra["niter"] == rb["niter"] == rc["niter"] and ra["niter"] > 4
    for r in (rb, rc):
        assert r["e_ccsd_final"] == ra["e_ccsd_final"] and r["e_t"] == ra["e_t"] and r["rms"] == ra["rms"]
    np.testing.assert_array_equal(t2b, t2a)
    np.testing.assert_array_equal(t2c, t2a)
    np.testing.assert_array_equal(t1c, t1a)


@pytest.mark.gpu
def test_larger_diis_subspaces_reach_the_same_fixed_point():
    # subspaces above the reference's 8 vectors (up to 16): the DIIS dots
    # then write 512 partials for each of up to 16 stored vectors
    announce("diis subspace 8 / 12 / 16 o=5 v=19")
    c, fock, ints = loaded(5, 19, seed=71)
    e = {}
    for n in (8, 12, 16):
        r = c.ccsd(r_convergence=1e-10, diis_nvec=n)
        assert r["converged"], n
        e[n] = r["e_ccsd_final"]
    assert abs(e[12] - e[8]) < 1e-10 and abs(e[16] - e[8]) < 1e-10
    with pytest.raises(ugx.UgxError):
        c.ccsd(diis_nvec=17)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,polish", [(ugx.PREC_FP64, ugx.POLISH_PLAIN), (ugx.PREC_BF16X3, ugx.POLISH_PLAIN),
                                         (ugx.PREC_BF16X3, ugx.POLISH_REFINE)])
def test_batch_matches_single_solves(prec, polish):
    # uneven sizes in one batch; each problem's result must be the one its
    # own single solve gives (the per-problem kernels are the same, and the
    # batched (T) builds each triplet's W3 with the same GEMM shapes)
    from ugacc_amd.synthetic import random_hamiltonian
    announce("batch of 7 uneven problems, prec=%d polish=%d" % (prec, polish))
    shapes = [(3, 8, 81), (4, 13, 82), (3, 8, 83), (2, 20, 84), (4, 13, 85), (3, 8, 86), (5, 10, 87)]
    problems = []
    for no, nv, seed in shapes:
        fock, ints = random_hamiltonian(no, nv, seed=seed)
        problems.append((no, nv, fock, ugx.split_blocks(ints, no)))
    c, _, _ = loaded(3, 6, seed=80)
    kw = dict(r_convergence=1e-10 if polish == ugx.POLISH_PLAIN else 1e-9, wfn_ccsd_t=1, precision=prec,
              polish_mode=polish)
    single = []
    for no, nv, fock, blocks in problems:
        c.load_hamiltonian(no, nv, fock, blocks)
        single.append(c.ccsd(**kw))
    batch, wall = c.batch_solve(problems, nworkers=4, graph=1, **kw)
    assert wall > 0
    for s, b in zip(single, batch):
        assert b["converged"] and s["niter"] == b["niter"]
        assert abs(s["e_ccsd_final"] - b["e_ccsd_final"]) < 1e-13
        assert abs(s["e_t"] - b["e_t"]) < 1e-13
        assert b["seconds_ccsd"] > 0 and b["seconds_triples"] > 0
