"""Writes the golden fixtures in tests/golden/*.ugxd (format: ugacc_amd/dump.py)
from the FP64 CPU oracle, which the CPU tests pin against the reference's
known answers and the mathematical invariants.

    python tests/golden/make_golden.py              # every fixture
    python tests/golden/make_golden.py h2o4_seed1   # one of them

Fixtures (each records its generator and seed):
- syn_o2v4, syn_o3v6_offdiag, syn_o4v8_fov: seeded synthetic Hamiltonians
  (ugacc_amd/synthetic.py), stored whole.  Per reference function: the MP2
  guess (CCWfn ctor, ccwfn.cc:75-87) and E_MP2 (:196); from it one
  iteration's intermediates (build_F / build_W, :311-430) and raw residuals
  R1, R2 (build_t1 / build_t2, :432-539); the converged t1, t2 and E_CCSD of
  compute_energy (:188-258, R_CONVERGENCE 1e-11, DIIS); the three (T) values
  tcorr, tcorr_ooc, tcorr_ooc_TJL (:647-852) and the per-triplet energies of
  the bracket with tcorr's f_ov terms folded in.
- h2o_ccpvdz: the single-water configuration (O-H 0.9572 A, 104.52 deg,
  cc-pVDZ, all electrons) from the chemistry front end; fock, t1, t2 and the
  energies stored, the integral blocks described by checksums (rebuilt by
  the front end).
- h2o4_seed1: the seeded (H2O)4 cluster (o=20, v=76), solved to 1e-10 by the
  oracle (OpenMP): energies, t1 and the t2 row i = 0 stored, fock and the
  blocks described by checksums.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Oracle  # noqa: E402
from ugacc_amd import dump  # noqa: E402
from ugacc_amd.hamiltonian import blocks_from_full  # noqa: E402
from ugacc_amd.synthetic import random_hamiltonian  # noqa: E402

SYNTHETIC = {
    "syn_o2v4": dict(no=2, nv=4, seed=101, scale=0.25),
    "syn_o3v6_offdiag": dict(no=3, nv=6, seed=102, scale=0.2, offdiag=0.02),
    "syn_o4v8_fov": dict(no=4, nv=8, seed=103, scale=0.15, fov=0.01),
}
CONV = 1e-11


def _path(name):
    return os.path.join(HERE, name + ".ugxd")


def _oracle_results(fock, ints, no, conv, omp=False, rows=None):
    """Iteration-1 outputs from the MP2 guess, then the converged solve and
    the (T) values."""
    o1 = Oracle(fock, ints, no, omp=omp)
    t1g, t2g = o1.amps()
    emp2 = o1.energy()
    out = {}
    if rows is None:
        _, R1, R2 = o1.iterate()
        inter = o1.intermediates()
        out.update(t1_guess=t1g, t2_guess=t2g, R1_iter1=R1, R2_iter1=R2,
                   **{k + "_iter1": v for k, v in inter.items()})
    o1.close()
    o = Oracle(fock, ints, no, omp=omp)
    r = o.ccsd(maxiter=200, conv=conv, diis=True, mode=0, diis_policy=0, min_vecs=3)
    assert r["converged"], r
    t1, t2 = o.amps()
    sc = dict(e_mp2=emp2, eccsd=r["eccsd"], e_ccsd_final=o.energy(), rms=r["rms"], niter=r["niter"],
              e_t_tjl=o.tcorr_tjl(), e_t_fov_tjl=o.tjl_fov_range(0, o.ntriplets()))
    if rows is None:
        sc.update(e_t_tcorr=o.tcorr(), e_t_tcorr_ooc=o.tcorr_ooc())
        out.update(t1=t1, t2=t2, e_triplet=o.tjl_fov_per_triplet())
    else:
        out.update(t1=t1, **{"t2_rows%d" % rows: t2[:rows]})
    o.close()
    return sc, out


def make_synthetic(name):
    g = SYNTHETIC[name]
    no, nv = g["no"], g["nv"]
    kw = {k: v for k, v in g.items() if k not in ("no", "nv", "seed")}
    fock, ints = random_hamiltonian(no, nv, seed=g["seed"], **kw)
    sc, out = _oracle_results(fock, ints, no, CONV)
    blocks = blocks_from_full(ints, no)
    arrays = dict(fock=fock, **blocks, **out)
    sizes = dict(nmo=no + nv, nact=no + nv, no=no, nv=nv, nfzc=0, nfzv=0)
    gen = dict(kind="synthetic", module="ugacc_amd.synthetic.random_hamiltonian", **g)
    sc["r_convergence"] = CONV
    dump.write(_path(name), sizes, sc, arrays, gen)


def _molecule(atoms, basis, nfzc=0):
    from ugacc_amd.chem.scf import Reference
    return Reference(atoms, basis, nfzc=nfzc)


def make_h2o():
    from ugacc_amd.chem import molecules
    ref = _molecule(molecules.to_bohr(molecules.water()), "cc-pvdz")
    sc, out = _oracle_results(ref.fock, ref.full_ints(), ref.no, 1e-10)
    sc.update(escf=ref.escf, efzc=ref.efzc, enuc=ref.enuc, r_convergence=1e-10)
    arrays = dict(fock=ref.fock, t1=out["t1"], t2=out["t2"], e_triplet=out["e_triplet"],
                  R1_iter1=out["R1_iter1"], R2_iter1=out["R2_iter1"])
    sizes = dict(nmo=ref.nmo, nact=ref.no + ref.nv, no=ref.no, nv=ref.nv, nfzc=0, nfzv=0)
    gen = dict(kind="molecule", builder="ugacc_amd.chem.molecules.water", r_oh=0.9572, angle=104.52,
               basis="cc-pvdz", nfzc=0)
    dump.write(_path("h2o_ccpvdz"), sizes, sc, arrays, gen, described=ref.blocks)


def make_h2o4():
    from ugacc_amd.chem import molecules
    ref = _molecule(molecules.water_cluster(4, seed=1), "cc-pvdz")
    t0 = time.time()
    sc, out = _oracle_results(ref.fock, ref.full_ints(), ref.no, 1e-10, omp=True, rows=1)
    sc.update(escf=ref.escf, efzc=ref.efzc, enuc=ref.enuc, r_convergence=1e-10, oracle_seconds=time.time() - t0)
    sizes = dict(nmo=ref.nmo, nact=ref.no + ref.nv, no=ref.no, nv=ref.nv, nfzc=0, nfzv=0)
    gen = dict(kind="molecule", builder="ugacc_amd.chem.molecules.water_cluster", n=4, seed=1, basis="cc-pvdz",
               nfzc=0)
    dump.write(_path("h2o4_seed1"), sizes, sc, out, gen, described=dict(fock=ref.fock, **ref.blocks))


MAKERS = dict({k: (lambda k=k: make_synthetic(k)) for k in SYNTHETIC}, h2o_ccpvdz=make_h2o, h2o4_seed1=make_h2o4)

if __name__ == "__main__":
    names = sys.argv[1:] or list(MAKERS)
    for n in names:
        t = time.time()
        MAKERS[n]()
        print("%s: %.1f s" % (n, time.time() - t), flush=True)
