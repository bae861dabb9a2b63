/*
 * ugacc_mi355x.cc -- the ugacc Psi4 plugin entry with its CCSD / CCSD(T)
 * hot path on the MI355X library (include/ugx.h, libugx.so).
 *
 * It exports the two symbols the reference plugin exports (ugacc.cc:54-76),
 * from a plugin of the same name (ugacc.so), so psi4.core.plugin_load /
 * plugin("ugacc.so", scf_wfn) call it unchanged:
 *   read_options  -- the same option names and defaults (ugacc.cc:58-68)
 *   ugacc         -- the same header, the same RHF error trapping
 *                    (ugacc.cc:95-99), then the Hamiltonian, CCWfn and
 *                    compute_energy steps (ugacc.cc:104-115) on the GPU:
 *                    the six Dirac integral blocks and the active Fock
 *                    matrix are built on the host and handed to ugx_ccsd.
 * Energies are printed in the layout of ccwfn.cc:234-247.  DERTYPE=FIRST
 * runs the post-CC steps of ugacc.cc:117-141 on the GPU: for CCSD_T first
 * the (T) gradient terms (ugx_tgrad, or ugx_tgrad_ooc with OOC=true, as
 * compute_energy runs tgrad or tgrad_ooc, ccwfn.cc:252-255), then Lambda
 * over HBAR (ugx_lambda, with the (T) sources for CCSD_T) and the density
 * energies (ugx_density, with the (T) parts),
 * printed as the reference prints them, with CURRENT ENERGY set.  Every
 * other DERTYPE (SECOND) adds the linear response of ugacc.cc:143-241 on the
 * GPU (ugx_pert_create / ugx_pert_solve / ugx_linresp): the dipole ("Mu",
 * MintsHelper::ao_dipole) and magnetic-dipole ("L", ao_angular_momentum
 * scaled by -1/2, perturbation.cc:83-86) perturbed wave functions per
 * Cartesian component at MY_OMEGA, left-hand ones too with MYHAND=LEFT and
 * Mu at -MY_OMEGA with MYHAND=RIGHT and MY_OMEGA != 0, then the dipole
 * polarizabilities and the length-gauge optical rotation tensor, printed in
 * the reference's layout.  WFN=CCSD_T with SECOND is refused: the
 * reference's CCDensity reads (T) density parts that CCWfn allocates for
 * DERTYPE=FIRST only (ccdensity.cc:44-69, ccwfn.cc:108-118).
 *
 * Integrals: MintsHelper::mo_eri over the active occupied / virtual AO
 * coefficients gives chemist (pq|rs); the blocks are <pq|rs> = (pr|qs).
 * Orbitals are ordered by energy within the occupied and virtual spaces
 * (the reference uses QT order; every energy is invariant to the order).
 */
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "psi4/libmints/matrix.h"
#include "psi4/libmints/mintshelper.h"
#include "psi4/libmints/wavefunction.h"
#include "psi4/liboptions/liboptions.h"
#include "psi4/libpsi4util/PsiOutStream.h"
#include "psi4/libpsi4util/exception.h"
#include "psi4/libpsi4util/process.h"
#include "psi4/psi4-dec.h"

#include "ugx.h"

namespace psi {
namespace ugacc {

extern "C" PSI_API int read_options(std::string name, Options &options) {
  if (name == "UGACC" || options.read_globals()) {
    options.add_int("PRINT", 1);
    options.add_str("REFERENCE", "RHF");
    options.add_str("WFN", "CCSD");
    options.add_str("DERTYPE", "NONE");
    options.add_str("MYHAND", "RIGHT");
    options.add_int("MAXITER", 100);
    options.add_bool("DIIS", true);
    options.add_double("R_CONVERGENCE", 1e-7);
    options.add_double("MY_OMEGA", 0.00);
    options.add_bool("OOC", false);
    options.add_str("PROPERTY", "POLARIZABILITY", "POLARIZABILITY ROTATION ROA ROA_TENSOR ALL");
    // MI355X forms (no counterpart in the reference): arithmetic of the
    // iterations before the FP64 polish, and the GPU to run on
    options.add_str("UGX_PRECISION", "FP64", "FP64 FP32 BF16X3");
    options.add_int("UGX_DEVICE", 0);
  }
  return true;
}

namespace {

// Dirac block X[p][q][r][s] = <pq|rs> = (pr|qs) from the chemist matrix
// G[(p r)][(q s)] of mo_eri(C_p, C_r, C_q, C_s).
std::vector<double> dirac_block(const SharedMatrix &G, int np, int nq, int nr, int ns) {
  std::vector<double> X((size_t)np * nq * nr * ns);
  double **g = G->pointer();
  for (int p = 0; p < np; p++)
    for (int q = 0; q < nq; q++)
      for (int r = 0; r < nr; r++)
        for (int s = 0; s < ns; s++)
          X[(((size_t)p * nq + q) * nr + r) * ns + s] = g[(size_t)p * nr + r][(size_t)q * ns + s];
  return X;
}

// An AO one-electron operator in the active MO basis (occupied first),
// row-major nact x nact, times scale.
std::vector<double> mo_property(const SharedMatrix &X, const SharedMatrix &Ca, double scale) {
  SharedMatrix M = linalg::triplet(Ca, X, Ca, true, false, false);
  const int n = M->rowspi()[0];
  std::vector<double> out((size_t)n * n);
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) out[(size_t)p * n + q] = scale * M->get(p, q);
  return out;
}

struct Context {
  ugx_ctx *c = nullptr;
  ~Context() {
    if (c) ugx_destroy(c);
  }
};

}  // namespace

extern "C" PSI_API SharedWavefunction ugacc(SharedWavefunction ref, Options &options) {
  outfile->Printf("\t**************************\n");
  outfile->Printf("\t*                        *\n");
  outfile->Printf("\t*    UGA-CC (MI355X)     *\n");
  outfile->Printf("\t*                        *\n");
  outfile->Printf("\t**************************\n\n");
  outfile->Printf("\tWave function  = %s\n", options.get_str("WFN").c_str());
  outfile->Printf("\tMaxiter        = %d\n", options.get_int("MAXITER"));
  outfile->Printf("\tConvergence    = %3.1e\n", options.get_double("R_CONVERGENCE"));
  outfile->Printf("\tDIIS           = %s\n", options.get_bool("DIIS") ? "Yes" : "No");
  outfile->Printf("\tOut-of-core    = %s\n", options.get_bool("OOC") ? "Yes" : "No");
  outfile->Printf("\tDertype        = %s\n", options.get_str("DERTYPE").c_str());
  outfile->Printf("\tOMEGA          = %3.1e\n", options.get_double("MY_OMEGA"));
  outfile->Printf("\tHAND           = %s\n", options.get_str("MYHAND").c_str());
  outfile->Printf("\tPrecision      = %s\n", options.get_str("UGX_PRECISION").c_str());

  // the reference's error trapping (closed-shell RHF in place)
  if (!ref) throw PSIEXCEPTION("SCF has not been run yet!");
  if (options.get_str("REFERENCE") != "RHF") throw PSIEXCEPTION("Only for use with RHF references.");
  for (int h = 0; h < ref->nirrep(); h++)
    if (ref->soccpi()[h]) throw PSIEXCEPTION("UGACC is for closed-shell systems only.");
  const std::string dertype = options.get_str("DERTYPE");
  const bool second = dertype != "NONE" && dertype != "FIRST";  // ugacc.cc:119-143
  if (second && options.get_str("WFN") == "CCSD_T")
    throw PSIEXCEPTION("ugacc (MI355X): WFN=CCSD_T supports DERTYPE=NONE and FIRST only");

  // active orbitals (frozen core and frozen virtuals removed), AO basis
  SharedMatrix Co = ref->Ca_subset("AO", "ACTIVE_OCC");
  SharedMatrix Cv = ref->Ca_subset("AO", "ACTIVE_VIR");
  const int no = Co->colspi()[0], nv = Cv->colspi()[0], nact = no + nv;

  // active Fock matrix, occupied first (hamiltonian.cc:44-56)
  SharedMatrix Fao = ref->Fa_subset("AO");
  std::vector<SharedMatrix> cs = {Co, Cv};
  SharedMatrix Ca = linalg::horzcat(cs);
  SharedMatrix Fmo = linalg::triplet(Ca, Fao, Ca, true, false, false);
  std::vector<double> fock((size_t)nact * nact);
  for (int p = 0; p < nact; p++)
    for (int q = 0; q < nact; q++) fock[(size_t)p * nact + q] = Fmo->get(p, q);

  // the six blocks (include/ugx.h layouts)
  auto mints = std::make_shared<MintsHelper>(ref->basisset(), options, 0);
  const std::vector<double> oooo = dirac_block(mints->mo_eri(Co, Co, Co, Co), no, no, no, no);
  const std::vector<double> ooov = dirac_block(mints->mo_eri(Co, Co, Co, Cv), no, no, no, nv);
  const std::vector<double> oovv = dirac_block(mints->mo_eri(Co, Cv, Co, Cv), no, no, nv, nv);
  const std::vector<double> ovov = dirac_block(mints->mo_eri(Co, Co, Cv, Cv), no, nv, no, nv);
  const std::vector<double> ovvv = dirac_block(mints->mo_eri(Co, Cv, Cv, Cv), no, nv, nv, nv);
  const std::vector<double> vvvv = dirac_block(mints->mo_eri(Cv, Cv, Cv, Cv), nv, nv, nv, nv);

  Context ctx;
  int rc = ugx_create(&ctx.c, options.get_int("UGX_DEVICE"));
  if (rc != UGX_OK) throw PSIEXCEPTION("ugacc (MI355X): no usable MI355X device");
  rc = ugx_load_hamiltonian(ctx.c, no, nv, fock.data(), oooo.data(), ooov.data(), oovv.data(), ovov.data(),
                            ovvv.data(), vvvv.data());
  if (rc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));

  ugx_options opt;
  ugx_default_options(&opt);
  opt.wfn_ccsd_t = options.get_str("WFN") == "CCSD_T";
  opt.maxiter = options.get_int("MAXITER");
  opt.diis = options.get_bool("DIIS");
  opt.r_convergence = options.get_double("R_CONVERGENCE");
  opt.ooc = options.get_bool("OOC");
  opt.print = options.get_int("PRINT") > 1;
  const std::string prec = options.get_str("UGX_PRECISION");
  opt.precision = prec == "BF16X3" ? UGX_PREC_BF16X3 : (prec == "FP32" ? UGX_PREC_FP32 : UGX_PREC_FP64);
  if (opt.precision != UGX_PREC_FP64) opt.polish_mode = UGX_POLISH_REFINE;

  ugx_result res;
  rc = ugx_ccsd(ctx.c, &opt, &res);
  if (rc == UGX_NOT_CONVERGED) throw PSIEXCEPTION("Computation has not converged.");
  if (rc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));

  const double eref = ref->energy();
  outfile->Printf("\n\tMP2 Correlation Energy     = %20.14f\n", res.e_mp2);
  outfile->Printf("\tCCSD Correlation Energy    = %20.14f\n", res.e_ccsd);
  outfile->Printf("\tCCSD Total Energy          = %20.14f\n", res.e_ccsd + eref);
  if (opt.wfn_ccsd_t) {
    outfile->Printf("\t(T) Correction             = %20.14f\n", res.e_t);
    outfile->Printf("\tCCSD(T) Correlation Energy = %20.14f\n", res.e_ccsd + res.e_t);
    outfile->Printf("\tCCSD(T) Total Energy       = %20.14f\n", res.e_ccsd + res.e_t + eref);
  }
  outfile->Printf("\t%d iterations (%d before the FP64 polish), CCSD %.3f s, (T) %.3f s on the GPU\n", res.niter,
                  res.niter_low, res.seconds_ccsd, res.seconds_triples);
  if (dertype == "NONE") return ref;

  // the (T) gradient terms (ccwfn.cc:252-255), then HBAR, Lambda and the
  // densities (ugacc.cc:118-137), in FP64
  if (opt.wfn_ccsd_t) {
    // tgrad, or tgrad_ooc with OOC=true (ccwfn.cc:252-255)
    rc = (opt.ooc ? ugx_tgrad_ooc : ugx_tgrad)(ctx.c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                               nullptr);
    if (rc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
  }
  double pseudo = 0.0, lrms = 0.0;
  int lniter = 0;
  std::vector<double> trace((size_t)opt.maxiter + 1, 0.0);
  ugx_options lopt = opt;
  lopt.precision = UGX_PREC_FP64;
  rc = ugx_lambda(ctx.c, &lopt, &pseudo, &lniter, &lrms, trace.data());
  outfile->Printf("\n\tThe Coupled-Cluster Lambda Iteration:\n");
  outfile->Printf("\t-------------------------------------\n");
  outfile->Printf("\t Iter   PseudoEnergy\n");
  for (int k = 0; k < lniter; k++) outfile->Printf("\t  %3d  %20.15f\n", k, trace[k]);
  if (rc == UGX_NOT_CONVERGED) throw PSIEXCEPTION("Computation has not converged.");
  if (rc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
  double E[8];
  rc = ugx_density(ctx.c, E, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                   nullptr);
  if (rc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
  const char *part[6] = {"OOOO", "VVVV", "OOOV", "VVVO", "OVOV", "OOVV"};
  for (int k = 0; k < 6; k++) outfile->Printf("\t%s Energy = %20.14f\n", part[k], E[2 + k]);
  const double eone = E[0], etwo = E[1], ecc = res.e_ccsd + (opt.wfn_ccsd_t ? res.e_t : 0.0);
  const char *wfn = opt.wfn_ccsd_t ? "CCSD(T)" : "CCSD";
  outfile->Printf("\tOne-Electron Energy        = %20.14f\n", eone);
  outfile->Printf("\tTwo-Electron Energy        = %20.14f\n", etwo);
  outfile->Printf("\t%s Correlation Energy    = %20.14f (from density)\n", wfn, eone + etwo);
  outfile->Printf("\t%s Correlation Energy    = %20.14f (from ccwfn)\n", wfn, ecc);
  outfile->Printf("\t%s Total Energy          = %20.14f (from density)\n", wfn, eone + etwo + eref);
  outfile->Printf("\t%s Total Energy          = %20.14f (from ccwfn)\n", wfn, ecc + eref);
  Process::environment.globals["CURRENT ENERGY"] = ecc + eref;
  if (!second) return ref;

  // the linear response, ugacc.cc:143-241
  std::vector<SharedMatrix> dip = mints->ao_dipole(), ang = mints->ao_angular_momentum();
  std::vector<double> mu[3], am[3];
  for (int k = 0; k < 3; k++) {
    mu[k] = mo_property(dip[k], Ca, 1.0);
    am[k] = mo_property(ang[k], Ca, -0.5);
  }
  const double omega = options.get_double("MY_OMEGA");
  const bool right = options.get_str("MYHAND") == "RIGHT";
  ugx_options popt = lopt;
  std::map<std::string, int> perts;
  auto create = [&](const std::string &key, const std::vector<double> &A, double w) {
    int h = -1;
    if (ugx_pert_create(ctx.c, A.data(), w, &h) != UGX_OK)
      throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
    perts[key] = h;
  };
  auto solve = [&](const std::string &key, int hand) {
    outfile->Printf("\n\tCoupled-Cluster %s Perturbed Wfn Iteration:\n", hand ? "LH" : "RH");
    outfile->Printf("\t-------------------------------------\n");
    outfile->Printf("\t Iter   Pseudoresponse      RMS   \n");
    outfile->Printf("\t-------------------------------------\n");
    double pr = 0.0, prms = 0.0;
    int pn = 0;
    const int prc = ugx_pert_solve(ctx.c, perts.at(key), hand, &popt, &pr, &pn, &prms);
    if (prc == UGX_OK || prc == UGX_NOT_CONVERGED) outfile->Printf("\t  %3d  %20.15f  %5.3e\n", pn, pr, prms);
    if (prc == UGX_NOT_CONVERGED) throw PSIEXCEPTION("Computation has not converged.");
    if (prc != UGX_OK) throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
  };
  auto linresp = [&](const std::string &x, const std::string &y) {
    double v = 0.0;
    if (ugx_linresp(ctx.c, perts.at(x), perts.at(y), &v) != UGX_OK)
      throw PSIEXCEPTION(std::string("ugacc (MI355X): ") + ugx_last_error(ctx.c));
    return v;
  };
  const char *cart[3] = {"X", "Y", "Z"};
  outfile->Printf("\n\tSolving right hand perturbed CC amplitudes\n");
  for (int k = 0; k < 3; k++) {
    std::string entry = std::string("Mu") + cart[k] + std::to_string(omega);
    const std::string entry_1 = std::string("L") + cart[k] + std::to_string(omega);
    outfile->Printf("\n\tCC RH Perturbed Wavefunction: %s\n", entry.c_str());
    create(entry, mu[k], omega);
    solve(entry, 0);
    outfile->Printf("\n\tCC RH Perturbed Wavefunction: %s\n", entry_1.c_str());
    create(entry_1, am[k], omega);
    solve(entry_1, 0);
    if (omega != 0.0 && right) {
      entry = std::string("Mu") + cart[k] + std::to_string(-omega);
      outfile->Printf("\n\tCC RH Perturbed Wavefunction: %s\n", entry.c_str());
      create(entry, mu[k], -omega);
      solve(entry, 0);
    }
    if (!right) {
      outfile->Printf("\n\tSolving left hand perturbed CC amplitudes\n");
      outfile->Printf("\n\tCC LH Perturbed Wavefunction: %s\n", entry.c_str());
      solve(entry, 1);
      outfile->Printf("\n\tCC LH Perturbed Wavefunction: %s\n", entry_1.c_str());
      solve(entry_1, 1);
    }
  }
  std::map<std::string, double> polars, rots;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q <= p; q++) {
      const std::string mp = std::string("Mu") + cart[p] + std::to_string(omega);
      const std::string mq = std::string("Mu") + cart[q] + std::to_string(omega);
      const double polar = p == q ? linresp(mp, mq) : 0.5 * linresp(mp, mq) + 0.5 * linresp(mq, mp);
      polars[std::string("<<Mu_") + cart[p] + ";Mu_" + cart[q] + ">>"] = polar;
      polars[std::string("<<Mu_") + cart[q] + ";Mu_" + cart[p] + ">>"] = polar;
    }
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      const std::string mp = std::string("Mu") + cart[p] + std::to_string(omega);
      const std::string lq = std::string("L") + cart[q] + std::to_string(omega);
      const double rotation = 0.5 * linresp(mp, lq) - 0.5 * linresp(lq, mp);
      rots[std::string("<<Mu_") + cart[p] + ";L_" + cart[q] + ">>"] = -rotation;
    }
  outfile->Printf("\n\tDipole Polarizabilities (au)\n\n");
  for (const auto &e : polars) outfile->Printf("\t%s : %20.14lf \n\n", e.first.c_str(), e.second);
  outfile->Printf("\n\tOptical rotation (length gauge) (a.u)\n\n");
  for (const auto &e : rots) outfile->Printf("\t%s : %20.14lf \n\n", e.first.c_str(), e.second);
  return ref;
}

}  // namespace ugacc
}  // namespace psi
