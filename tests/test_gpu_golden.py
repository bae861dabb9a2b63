"""The GPU path against the committed golden fixtures (tests/golden/*.ugxd,
written by tests/golden/make_golden.py from the oracle; format in
ugacc_amd/dump.py).  No oracle call here: the stored vectors are the
expected values.

- synthetic o=2..4, v=4..8 (canonical, non-canonical, f_ov != 0): one
  residual from the MP2 guess (intermediates, R1, R2), the converged
  E_CCSD / t1 / t2, and the (T) value the reference returns (tcorr =
  tcorr_ooc) in total and per triplet;
- H2O/cc-pVDZ all-electron (the single-water configuration), integrals
  rebuilt by the front end and checked against the fixture's checksums;
- (H2O)4/cc-pVDZ (o=20, v=76), converged to 1e-10: E_MP2, E_CCSD and the
  (T) energy against the oracle's full solve, t1 and the first t2 rows.

Bars: energies 1e-10 Eh (FP64 form), amplitudes 1e-9 max-abs, per-function
outputs 1e-11 relative to their scale, per-triplet energies 1e-13.
"""
import os

import numpy as np
import pytest

from ugacc_amd import dump, ugx
from tests.gpu_util import announce, context

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SYN = ["syn_o2v4", "syn_o3v6_offdiag", "syn_o4v8_fov"]


def _read(name):
    return dump.Dump(os.path.join(GOLDEN, name + ".ugxd"))


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("name", SYN)
def test_synthetic_fixture(name):
    announce("golden " + name)
    d = _read(name)
    no, nv = d.sizes["no"], d.sizes["nv"]
    fock, blocks = d.hamiltonian()
    c = context()
    c.load_hamiltonian(no, nv, fock, blocks)
    R1, R2 = c.residual(d["t1_guess"], d["t2_guess"])
    assert _rel(R1, d["R1_iter1"]) < 1e-11 and _rel(R2, d["R2_iter1"]) < 1e-11
    inter = c.intermediates()
    for k, v in inter.items():
        assert _rel(v, d[k + "_iter1"]) < 1e-11, k
    r = c.ccsd(maxiter=200, r_convergence=d.scalars["r_convergence"], wfn_ccsd_t=1)
    assert r["converged"]
    assert abs(r["e_mp2"] - d.scalars["e_mp2"]) < 1e-12
    assert abs(r["e_ccsd"] - d.scalars["eccsd"]) < 1e-10
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - d["t1"])) < 1e-9 and np.max(np.abs(t2 - d["t2"])) < 1e-9
    assert abs(r["e_t"] - d.scalars["e_t_tcorr"]) < 1e-10
    # (T) from the fixture's amplitudes: per triplet and the reference's total
    c.set_amplitudes(d["t1"], d["t2"])
    assert np.max(np.abs(c.triples_per_triplet() - d["e_triplet"])) < 1e-13
    assert abs(c.triples() - d.scalars["e_t_tcorr"]) < 1e-12
    assert abs(c.triples() - d.scalars["e_t_tcorr_ooc"]) < 1e-12


def _rebuild(d):
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    g = d.generator
    if g["builder"].endswith("water_cluster"):
        atoms = molecules.water_cluster(g["n"], seed=g["seed"])
    else:
        atoms = molecules.to_bohr(molecules.water(g["r_oh"], g["angle"]))
    ref = Reference(atoms, g["basis"], nfzc=g["nfzc"])
    assert (ref.no, ref.nv) == (d.sizes["no"], d.sizes["nv"])
    assert abs(ref.escf - d.scalars["escf"]) < 1e-10
    # the integrals the fixture was made from (sum and norm of each block)
    for k, b in ref.blocks.items():
        assert d.verify(k, b, rtol=1e-11), k
    if "fock" in d.checksums:
        assert d.verify("fock", ref.fock, rtol=1e-11)
    return ref


def test_h2o_fixture():
    announce("golden h2o_ccpvdz")
    d = _read("h2o_ccpvdz")
    ref = _rebuild(d)
    assert np.max(np.abs(ref.fock - d["fock"])) < 1e-11
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    assert r["converged"]
    assert abs(r["e_mp2"] - d.scalars["e_mp2"]) < 1e-12
    assert abs(r["e_ccsd"] - d.scalars["eccsd"]) < 1e-10
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - d["t1"])) < 1e-9 and np.max(np.abs(t2 - d["t2"])) < 1e-9
    assert abs(r["e_t"] - d.scalars["e_t_tcorr"]) < 1e-10
    assert np.max(np.abs(c.triples_per_triplet() - d["e_triplet"])) < 1e-12


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "h2o4_seed1.ugxd")), reason="fixture not generated")
def test_h2o4_fixture_converged_to_1e10():
    """(H2O)4 at the configuration's R_CONVERGENCE of 1e-10 against the
    oracle's complete solve (E_MP2, E_CCSD, E_(T) to 1e-10; t1 and the t2
    row i = 0 to 1e-9)."""
    announce("golden h2o4_seed1 (FP64, 1e-10)")
    d = _read("h2o4_seed1")
    ref = _rebuild(d)
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    assert r["converged"] and r["rms"] < 1e-10
    assert abs(r["e_mp2"] - d.scalars["e_mp2"]) < 1e-11
    assert abs(r["e_ccsd"] - d.scalars["eccsd"]) < 1e-10
    assert abs(r["e_t"] - d.scalars["e_t_fov_tjl"]) < 1e-10
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - d["t1"])) < 1e-9
    rows = d["t2_rows1"]
    assert np.max(np.abs(t2[: rows.shape[0]] - rows)) < 1e-9
    print("(H2O)4: E_CCSD %.12f (oracle %.12f), E_(T) %.12f (oracle %.12f), %d iterations (oracle %d)"
          % (r["e_ccsd"], d.scalars["eccsd"], r["e_t"], d.scalars["e_t_fov_tjl"], r["niter"],
             d.scalars["niter"]))


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "h2o4_seed1.ugxd")), reason="fixture not generated")
def test_h2o4_fixture_with_library_fp64_gemms():
    """With autotuning on, the engine's large single FP64 GEMMs (at least
    2^30 multiply-adds, unit-stride operands) run on rocBLAS dgemm where it
    times faster than every tile of the library's kernel (at (H2O)4 the ring
    and ladder GEMMs qualify): the fixture's energies and amplitudes to the
    same bars; rocBLAS runs with atomics off, so a second eager solve and a
    solve with graph-replayed residuals (rocBLAS captured with its preset
    workspace) give bitwise-equal amplitudes."""
    announce("golden h2o4_seed1 with rocBLAS FP64 GEMMs, eager and graph")
    d = _read("h2o4_seed1")
    ref = _rebuild(d)
    c = ugx.Context(0)
    c.set_autotune(True)
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1)
    assert r["converged"]
    assert abs(r["e_ccsd"] - d.scalars["eccsd"]) < 1e-10
    assert abs(r["e_t"] - d.scalars["e_t_fov_tjl"]) < 1e-10
    t1, t2 = c.amplitudes()
    assert np.max(np.abs(t1 - d["t1"])) < 1e-9
    lib = [l for l in c.tile_table().splitlines() if "rocblas" in l]
    print("\n".join(lib))
    assert lib, "no FP64 shape tuned to rocBLAS"
    for graph in (0, 1):
        r2 = c.ccsd(maxiter=100, r_convergence=1e-10, wfn_ccsd_t=1, graph=graph)
        u1, u2 = c.amplitudes()
        assert r2["niter"] == r["niter"] and r2["e_ccsd"] == r["e_ccsd"], graph
        np.testing.assert_array_equal(u1, t1)
        np.testing.assert_array_equal(u2, t2)
    c.close()
