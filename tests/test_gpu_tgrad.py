"""(T) gradient terms on the GPU (ugx_tgrad: FP64 engine contractions over
one occupied slice of t3 and l3 at a time) against the oracle's restatement
of ccwfn.cc:854-1098, and the CCSD(T) Lambda and densities built on them.

- synthetic o=2..4 (canonical; f_ov != 0): at the oracle's converged
  amplitudes, s1, s2 and the six (T) density parts against the oracle to
  1e-12 of each block's scale; then the CCSD(T) Lambda (wfn_ccsd_t=1 after
  ugx_tgrad: s1 and s2/2 in the residuals) against the oracle's Lambda to
  1e-10, and every density block against the oracle's CCSD(T) densities to
  1e-11 of its scale;
- the check the reference prints for WFN=CCSD_T (ugacc.cc:129-137): E_one +
  E_two equals E_CCSD + E_(T) at convergence, for H2O/cc-pVDZ (with the
  oracle) and for (H2O)4/cc-pVDZ (o=20, v=76), where no oracle runs;
- ugx_tgrad needs amplitudes, and a Lambda solve with wfn_ccsd_t=1 stays the
  CCSD Lambda until ugx_tgrad has run on the current amplitudes;
- the plugin mirror with WFN=CCSD_T, DERTYPE=FIRST (ugacc.cc:117-141) prints
  the CCSD(T) energy from the density equal to the one from ccwfn.
"""
import numpy as np
import pytest

from oracle import Oracle
from ugacc_amd import ugx
from tests.gpu_util import announce, context, loaded

pytestmark = pytest.mark.gpu

CASES = [dict(no=2, nv=5, seed=71), dict(no=3, nv=6, seed=72), dict(no=4, nv=9, seed=73),
         dict(no=3, nv=7, seed=74, fov=0.02)]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("case", CASES, ids=lambda c: "o%d_v%d_s%d" % (c["no"], c["nv"], c["seed"]))
def test_tgrad_lambda_and_density_match_oracle(case):
    announce("tgrad o=%d v=%d" % (case["no"], case["nv"]))
    kw = {k: v for k, v in case.items() if k not in ("no", "nv", "seed")}
    no = case["no"]
    c, fock, ints = loaded(no, case["nv"], case["seed"], scale=0.3, **kw)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=400, conv=1e-13)["converged"]
    o.build_tstar()
    t1, t2 = o.amps()
    c.set_amplitudes(t1, t2)
    ref = o.tgrad(enable=True)
    got = c.tgrad()
    for k in ugx.Context.TGRAD_BLOCKS:
        assert _rel(got[k], ref[k]) < 1e-12, (k, _rel(got[k], ref[k]))
    # the CCSD(T) Lambda and densities
    lr = c.lambda_solve(maxiter=300, r_convergence=1e-12, wfn_ccsd_t=1)
    assert lr["converged"]
    o.lambda_init()
    lo = o.lambda_solve(maxiter=300, conv=1e-12)
    assert lo["converged"]
    l1, l2 = c.lambdas()
    l1o, l2o = o.lambdas()
    assert np.max(np.abs(l1 - l1o)) < 1e-10 and np.max(np.abs(l2 - l2o)) < 1e-10
    o.set_lambdas(l1, l2)
    d, do = c.density(), o.density()
    for k, r in do["blocks"].items():
        assert _rel(d["blocks"][k], r) < 1e-11, (k, _rel(d["blocks"][k], r))
    assert abs(d["e_one"] - do["e_one"]) < 1e-12 and abs(d["e_two"] - do["e_two"]) < 1e-12
    # the density energy (with the f_ov blocks) is E_CCSD + E_(T)
    b = d["blocks"]
    e = d["e_one"] + d["e_two"] + np.sum(fock[:no, no:] * b["Dov"]) + np.sum(fock[no:, :no] * b["Dvo"])
    et = o.tcorr()
    assert abs(e - c.energy(form=ugx.ENERGY_FULL) - et) < 1e-11
    o.close()


def test_lambda_stays_ccsd_without_tgrad_and_tgrad_needs_amplitudes():
    announce("tgrad state rules")
    no, nv = 2, 5
    c, fock, ints = loaded(no, nv, 75, scale=0.3)
    with pytest.raises(ugx.UgxError):
        c.tgrad()
    r = c.ccsd(maxiter=200, r_convergence=1e-12, wfn_ccsd_t=1)
    assert r["converged"]
    assert c.lambda_solve(maxiter=200, r_convergence=1e-12, wfn_ccsd_t=1)["converged"]
    l_ccsd = c.lambdas()
    o = Oracle(fock, ints, no)
    o.set_amps(*c.amplitudes())
    o.lambda_init()
    assert o.lambda_solve(maxiter=300, conv=1e-12)["converged"]
    assert np.max(np.abs(l_ccsd[1] - o.lambdas()[1])) < 1e-10  # the CCSD Lambda
    c.tgrad(which=())
    assert c.lambda_solve(maxiter=200, r_convergence=1e-12, wfn_ccsd_t=0)["converged"]
    assert np.max(np.abs(c.lambdas()[1] - l_ccsd[1])) < 1e-12  # wfn_ccsd_t=0: still CCSD
    assert c.lambda_solve(maxiter=200, r_convergence=1e-12, wfn_ccsd_t=1)["converged"]
    assert np.max(np.abs(c.lambdas()[1] - l_ccsd[1])) > 1e-6  # now with the (T) sources
    o.close()


def _molecule_check(ref, conv, oracle):
    c = context()
    c.load_hamiltonian(ref.no, ref.nv, ref.fock, ref.blocks)
    r = c.ccsd(maxiter=200, r_convergence=conv, wfn_ccsd_t=1)
    assert r["converged"]
    g = c.tgrad(which=("s1",) if oracle is None else None)
    assert c.lambda_solve(maxiter=200, r_convergence=conv, wfn_ccsd_t=1)["converged"]
    d = c.density(which=("Doo", "Dvv"))
    print("E_one %.12f  E_two %.12f  sum %.12f  E_CCSD + E_(T) %.12f"
          % (d["e_one"], d["e_two"], d["e_one"] + d["e_two"], r["e_ccsd"] + r["e_t"]))
    if oracle is not None:
        o = Oracle(ref.fock, oracle, ref.no)
        o.set_amps(*c.amplitudes())
        o.build_tstar()
        rg = o.tgrad()
        for k in ugx.Context.TGRAD_BLOCKS:
            assert _rel(g[k], rg[k]) < 1e-12, k
        o.close()
    return r, d


def test_ccsd_t_density_energy_h2o():
    announce("tgrad h2o cc-pvdz")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    r, d = _molecule_check(ref, 1e-12, ref.full_ints())
    assert abs(d["e_one"] + d["e_two"] - r["e_ccsd"] - r["e_t"]) < 1e-11


def test_ccsd_t_density_energy_h2o4():
    announce("tgrad (h2o)4 seed 1")
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.water_cluster(4, seed=1), "cc-pvdz")
    r, d = _molecule_check(ref, 1e-10, None)
    assert abs(d["e_one"] + d["e_two"] - r["e_ccsd"] - r["e_t"]) < 1e-9
    b = d["blocks"]
    assert abs(np.trace(b["Doo"]) + np.trace(b["Dvv"])) < 1e-10


def test_plugin_first_derivative_ccsd_t_h2o():
    announce("plugin CCSD_T DERTYPE=FIRST h2o")
    from ugacc_amd import plugin
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    opts = dict(plugin.read_options(), WFN="CCSD_T", DERTYPE="FIRST", R_CONVERGENCE=1e-11)
    r = plugin.ugacc(ref, opts)
    assert abs(r.density_energy - r.cc_energy) < 1e-10
    assert abs(r.current_energy - r.cc_energy - ref.escf) < 1e-12


def test_tgrad_ooc_matches_oracle_and_refuses_fov():
    """ugx_tgrad_ooc (the OOC=true form, ccwfn.cc:1150-1316) against the
    oracle's restatement of tgrad_ooc on a canonical synthetic Hamiltonian
    (every block to 1e-12 of its scale; Doo and Dvv diagonal), the CCSD(T)
    Lambda and density built on it, and UGX_BAD_INPUT for |f_ov| beyond a
    converged SCF's residual (here 0.02)."""
    announce("tgrad_ooc o=3 v=6")
    no, nv = 3, 6
    c, fock, ints = loaded(no, nv, 76, scale=0.3)
    o = Oracle(fock, ints, no)
    assert o.ccsd(maxiter=400, conv=1e-13)["converged"]
    o.build_tstar()
    c.set_amplitudes(*o.amps())
    ref = o.tgrad_ooc(enable=True)
    got = c.tgrad(ooc=True)
    for k in ugx.Context.TGRAD_BLOCKS:
        assert _rel(got[k], ref[k]) < 1e-12, (k, _rel(got[k], ref[k]))
    assert np.max(np.abs(got["Doo"] - np.diag(np.diag(got["Doo"])))) == 0.0
    assert c.lambda_solve(maxiter=300, r_convergence=1e-12, wfn_ccsd_t=1)["converged"]
    o.lambda_init()
    assert o.lambda_solve(maxiter=300, conv=1e-12)["converged"]
    assert np.max(np.abs(c.lambdas()[1] - o.lambdas()[1])) < 1e-10
    o.set_lambdas(*c.lambdas())
    d, do = c.density(), o.density()
    for k, r in do["blocks"].items():
        assert _rel(d["blocks"][k], r) < 1e-11, k
    o.close()
    c2, _, _ = loaded(no, nv, 77, scale=0.3, fov=0.02)
    assert c2.ccsd(maxiter=300, r_convergence=1e-12)["converged"]
    with pytest.raises(ugx.UgxError):
        c2.tgrad(ooc=True)


def test_plugin_first_derivative_ccsd_t_ooc_h2o():
    announce("plugin CCSD_T DERTYPE=FIRST OOC=true h2o")
    from ugacc_amd import plugin
    from ugacc_amd.chem import molecules
    from ugacc_amd.chem.scf import Reference
    ref = Reference(molecules.to_bohr(molecules.water()), "cc-pvdz")
    opts = dict(plugin.read_options(), WFN="CCSD_T", DERTYPE="FIRST", OOC=True, R_CONVERGENCE=1e-11)
    r = plugin.ugacc(ref, opts)
    # canonical orbitals: the diagonal-only (T) Doo / Dvv give the same density energy
    assert abs(r.density_energy - r.cc_energy) < 1e-10
