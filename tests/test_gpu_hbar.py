"""HBAR on the GPU (ugx_hbar, FP64 contractions) against the oracle's
restatement of hbar.cc:8-228, block by block, to 1e-12 of each block's
scale: at seeded amplitudes (canonical, non-canonical and f_ov != 0
Hamiltonians), at converged CCSD amplitudes, and for H2O/cc-pVDZ."""
import numpy as np
import pytest

from oracle import Oracle
from tests.gpu_util import announce, loaded

pytestmark = pytest.mark.gpu

CASES = [dict(no=2, nv=5, seed=71), dict(no=3, nv=7, seed=72, offdiag=0.02),
         dict(no=4, nv=9, seed=73, offdiag=0.01, fov=0.02), dict(no=5, nv=14, seed=74)]


def _cmp(got, ref):
    for k, r in ref.items():
        scale = max(1.0, np.max(np.abs(r)))
        err = np.max(np.abs(got[k] - r))
        assert err < 1e-12 * scale, (k, err)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "o%d_v%d_s%d" % (c["no"], c["nv"], c["seed"]))
def test_hbar_matches_oracle(case):
    announce("hbar o=%d v=%d" % (case["no"], case["nv"]))
    kw = {k: v for k, v in case.items() if k not in ("no", "nv", "seed")}
    no, nv = case["no"], case["nv"]
    c, fock, ints = loaded(no, nv, case["seed"], **kw)
    o = Oracle(fock, ints, no)
    rng = np.random.default_rng(case["seed"])
    t1 = 0.05 * rng.standard_normal((no, nv))
    t2 = 0.03 * rng.standard_normal((no, no, nv, nv))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    o.set_amps(t1, t2)
    c.set_amplitudes(t1, t2)
    _cmp(c.hbar(), o.hbar())
    # at the converged amplitudes of the library's own solve
    r = c.ccsd(maxiter=100, r_convergence=1e-10)
    assert r["converged"]
    o.set_amps(*c.amplitudes())
    _cmp(c.hbar(), o.hbar())
    o.close()


def test_hbar_h2o_converged():
    announce("hbar h2o cc-pvdz")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    from tests.gpu_util import context
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    assert c.ccsd(maxiter=100, r_convergence=1e-10)["converged"]
    o = Oracle(ref.fock, ref.full_ints(), ref.no)
    o.set_amps(*c.amplitudes())
    _cmp(c.hbar(), o.hbar())
    got = c.hbar(which=("Hov", "Hvvvo"))
    assert set(got) == {"Hov", "Hvvvo"}
    o.close()
