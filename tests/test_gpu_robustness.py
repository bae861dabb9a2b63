"""Failure paths, allocation behaviour and graph safety of the C ABI.

- MAXITER reached: UGX_NOT_CONVERGED with the reference's message
  "Computation has not converged." (ccwfn.cc:230-231), results still filled;
- a NaN in the Fock matrix or in any integral block: UGX_NONFINITE at load;
- once ugx_prepare has run (ugx_ccsd runs it first), the iterations make no
  device allocation, in the FP64, mixed and refinement forms, eagerly and
  with graph replays, and in every member of a batch;
- a fresh context that solves in mixed precision with graph replays twice
  (the polish's FP64 residuals need larger workspaces than the BF16x3 ones
  captured before them) gives results bitwise equal to eager solves;
- an autotuned shape first timed eagerly keeps its tuned tile when the
  same shape is later captured into a graph (bitwise-equal results).
"""
import numpy as np
import pytest

from ugacc_amd import ugx
from ugacc_amd.synthetic import random_hamiltonian
from tests.gpu_util import announce, context, loaded

pytestmark = pytest.mark.gpu


def test_maxiter_not_converged():
    announce("maxiter=1 not converged")
    c, _, _ = loaded(3, 8, seed=61)
    r = c.ccsd(maxiter=1, r_convergence=1e-10)
    assert not r["converged"] and r["niter"] == 1 and r["rms"] > 1e-10
    assert c.last_error() == "Computation has not converged."
    assert r["e_ccsd"] != 0.0 and np.isfinite(r["e_ccsd"])


@pytest.mark.parametrize("where", ["fock", "oooo", "ooov", "oovv", "ovov", "ovvv", "vvvv"])
def test_nonfinite_input_rejected(where):
    announce("non-finite " + where)
    no, nv = 3, 7
    fock, ints = random_hamiltonian(no, nv, seed=62)
    blocks = ugx.split_blocks(ints, no)
    if where == "fock":
        fock = fock.copy()
        fock[1, no + 2] = np.nan
    else:
        b = blocks[where].copy()
        b.reshape(-1)[b.size // 2] = np.inf if where == "vvvv" else np.nan
        blocks[where] = b
    c = context()
    with pytest.raises(ugx.UgxError) as e:
        c.load_hamiltonian(no, nv, fock, blocks)
    assert e.value.code == ugx.NONFINITE
    # the context stays usable
    f2, g2 = random_hamiltonian(no, nv, seed=62)
    c.load_hamiltonian(no, nv, f2, ugx.split_blocks(g2, no))
    assert c.ccsd(r_convergence=1e-9)["converged"]


FORMS = [
    dict(precision=ugx.PREC_FP64),
    dict(precision=ugx.PREC_BF16X3, polish_threshold=1e-5),
    dict(precision=ugx.PREC_BF16X3, polish_threshold=1e-5, polish_mode=ugx.POLISH_REFINE),
    dict(precision=ugx.PREC_FP32, polish_threshold=1e-5, polish_mode=ugx.POLISH_REFINE),
]


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("form", range(len(FORMS)))
def test_iterations_allocate_nothing(form, graph):
    announce("allocations form=%d graph=%d" % (form, graph))
    no, nv = 8, 40  # o v < 1024: the t1 terms stay FP64; large enough for split-K
    fock, ints = random_hamiltonian(no, nv, seed=63, scale=0.05)
    c = ugx.Context(0)
    c.set_autotune(True)
    c.load_hamiltonian(no, nv, fock, ugx.split_blocks(ints, no))
    opts = dict(maxiter=100, r_convergence=1e-9, graph=graph, **FORMS[form])
    r1 = c.ccsd(**opts)
    r2 = c.ccsd(**opts)
    c.close()
    assert r1["converged"] and r2["converged"]
    assert r1["device_allocs"] == 0 and r2["device_allocs"] == 0, (r1["device_allocs"], r2["device_allocs"])
    assert r1["e_ccsd"] == r2["e_ccsd"]


def test_mixed_graph_solves_on_a_fresh_context_match_eager():
    """Regression: BF16x3 residuals captured into graphs, then the FP64
    polish's residuals (larger workspaces), then a second solve replaying
    the first solve's graphs.  Workspaces are sized before the first
    iteration and never freed while graphs may use them."""
    announce("fresh-context mixed graph solves vs eager")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.water_cluster(2, seed=3), "cc-pvdz")
    opts = dict(maxiter=100, r_convergence=1e-9, wfn_ccsd_t=0, precision=ugx.PREC_BF16X3, polish_threshold=1e-5)
    out = {}
    for graph in (1, 0):
        c = ugx.Context(0)
        c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
        a = c.ccsd(graph=graph, **opts)
        ta = c.amplitudes()
        b = c.ccsd(graph=graph, **opts)
        tb = c.amplitudes()
        c.close()
        out[graph] = (a, ta, b, tb)
    for k in (0, 2):
        rg, re_ = out[1][k], out[0][k]
        assert rg["converged"] and rg["niter"] == re_["niter"]
        assert rg["e_ccsd"] == re_["e_ccsd"] and rg["rms"] == re_["rms"]
        for x, y in zip(out[1][k + 1], out[0][k + 1]):
            assert np.array_equal(x, y)


def test_tuned_tiles_are_used_under_capture(tmp_path):
    """Autotuning on, one eager solve (tunes every shape), then a solve with
    graph replays: the same tiles, so bitwise-equal amplitudes.  The tuned
    table survives a save / load round trip."""
    announce("tuned tiles under capture")
    c, _, _ = loaded(6, 30, seed=64)
    c.set_autotune(True)
    r0 = c.ccsd(r_convergence=1e-10)
    t0 = c.amplitudes()
    r1 = c.ccsd(r_convergence=1e-10, graph=1)
    t1 = c.amplitudes()
    assert r0["niter"] == r1["niter"] and r0["e_ccsd"] == r1["e_ccsd"]
    assert np.array_equal(t0[1], t1[1])
    table = c.tile_table()
    assert table.count("\n") > 5 and "gfx" in table
    p = str(tmp_path / "tiles.txt")
    c.save_tile_table(p)
    lines = open(p).read().splitlines()
    assert len(lines) == table.count("\n") and all(len(l.split("\t")) == 3 for l in lines)
    c.load_tile_table(p)
    assert c.tile_table() == table
    c.set_autotune(False)
