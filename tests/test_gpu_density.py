"""CCSD densities on the GPU (ugx_density: FP64 contractions) against the
oracle's restatement of ccdensity.cc:89-339.

- synthetic o=2..4 (canonical, non-canonical, f_ov != 0) and H2O/cc-pVDZ:
  after the library's CCSD and Lambda solves, every density block against
  the oracle's from the same amplitudes and multipliers to 1e-12 of the
  block's scale, and the eight energies to 1e-12;
- the check the reference prints (ugacc.cc:129-137): E_one + E_two equals
  the CCSD correlation energy at convergence (f_ov = 0), for H2O and for
  (H2O)4/cc-pVDZ (o=20, v=76), where no oracle runs.
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from tests.gpu_util import announce, context, loaded

pytestmark = pytest.mark.gpu

CASES = [dict(no=2, nv=4, seed=91), dict(no=3, nv=6, seed=92, offdiag=0.02),
         dict(no=4, nv=8, seed=93, offdiag=0.01, fov=0.02)]


def _solve(c, conv=1e-12):
    r = c.ccsd(maxiter=200, r_convergence=conv)
    assert r["converged"]
    assert c.lambda_solve(maxiter=200, r_convergence=conv)["converged"]
    return r


def _compare(c, fock, ints, no):
    d = c.density()
    o = Oracle(fock, ints, no)
    o.set_amps(*c.amplitudes())
    o.set_lambdas(*c.lambdas())
    ref = o.density()
    o.close()
    for k, r in ref["blocks"].items():
        err = np.max(np.abs(d["blocks"][k] - r)) / max(1.0, np.max(np.abs(r)))
        assert err < 1e-12, (k, err)
    assert abs(d["e_one"] - ref["e_one"]) < 1e-12 and abs(d["e_two"] - ref["e_two"]) < 1e-12
    for k in ref["parts"]:
        assert abs(d["parts"][k] - ref["parts"][k]) < 1e-12, k
    return d


@pytest.mark.parametrize("case", CASES, ids=lambda c: "o%d_v%d_s%d" % (c["no"], c["nv"], c["seed"]))
def test_density_matches_oracle(case):
    announce("density o=%d v=%d" % (case["no"], case["nv"]))
    kw = {k: v for k, v in case.items() if k not in ("no", "nv", "seed")}
    no = case["no"]
    c, fock, ints = loaded(no, case["nv"], case["seed"], scale=0.3, **kw)
    _solve(c)
    d = _compare(c, fock, ints, no)
    # the density energy with the f_ov blocks is the CCSD energy
    b = d["blocks"]
    e = d["e_one"] + d["e_two"] + np.sum(fock[:no, no:] * b["Dov"]) + np.sum(fock[no:, :no] * b["Dvo"])
    assert abs(e - c.energy(form=ugx.ENERGY_FULL)) < 1e-11


def test_density_h2o_matches_oracle_and_ccsd_energy():
    announce("density h2o cc-pvdz")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = _solve(c)
    d = _compare(c, ref.fock, ref.full_ints(), ref.no)
    assert abs(d["e_one"] + d["e_two"] - r["e_ccsd"]) < 1e-10


def test_density_energy_h2o4():
    announce("density (h2o)4 seed 1")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.water_cluster(4, seed=1), "cc-pvdz")
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = _solve(c, conv=1e-10)
    d = c.density(which=("Doo", "Dvv"))
    print("(H2O)4: E_one %.12f  E_two %.12f  sum %.12f  E_CCSD %.12f"
          % (d["e_one"], d["e_two"], d["e_one"] + d["e_two"], r["e_ccsd"]))
    assert abs(d["e_one"] + d["e_two"] - r["e_ccsd"]) < 1e-9
    # trace of the one-particle density change: tr Doo + tr Dvv = 0
    b = d["blocks"]
    assert abs(np.trace(b["Doo"]) + np.trace(b["Dvv"])) < 1e-10
