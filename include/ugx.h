/*
 * ugx.h -- C ABI of the MI355X (gfx950) closed-shell CCSD / CCSD(T) library.
 *
 * This is the boundary a Psi4 plugin (or any other host) binds to replace
 * ugacc's CPU hot path.  Signatures use plain pointers and sizes only; no
 * exception crosses the ABI; every call returns a status code, and the text
 * of the last error is available from ugx_last_error().
 *
 * Reference interfaces replaced (paths relative to the ugacc source tree):
 *   ugx_default_options   <- read_options, ugacc.cc:54-73 (WFN, MAXITER,
 *                            DIIS, R_CONVERGENCE, OOC with their defaults)
 *   ugx_load_hamiltonian  <- Hamiltonian::Hamiltonian, hamiltonian.cc:19-103
 *                            (fock_, ints_, L_; here six Dirac blocks that
 *                            stay resident on the device)
 *   ugx_ccsd              <- CCWfn::CCWfn (denominators and MP2 guess,
 *                            ccwfn.cc:56-94) + CCWfn::compute_energy's
 *                            iterations, ccwfn.cc:188-231
 *   ugx_triples           <- CCWfn::tcorr / tcorr_ooc, ccwfn.cc:647-793 (the
 *                            value compute_energy returns, :239-248),
 *                            evaluated as tcorr_ooc_TJL's i>=j>=k bracket
 *                            (:795-852, W3_ijk :1388-1418) with the f_ov t3
 *                            terms of tcorr's X2 (:706-709) folded into V
 *   ugx_energy            <- CCWfn::energy, ccwfn.cc:260-282
 *   ugx_residual          <- build_F/build_W/build_t1/build_t2,
 *                            ccwfn.cc:311-539
 *   ugx_intermediates     <- the members Fvv_, Foo_, Fov_, Woooo_, Wovov_,
 *                            Wovvo_, Wovoo_ (ccwfn.h) after build_F/build_W
 *   ugx_w3                <- CCWfn::W3_ijk, ccwfn.cc:1388-1418
 *   ugx_hbar              <- HBAR::HBAR, hbar.cc:8-228
 *   ugx_lambda            <- CCLambda::compute_lambda, cclambda.cc:93-135
 *   ugx_tgrad             <- CCWfn::tgrad, ccwfn.cc:854-1098
 *   ugx_tgrad_ooc         <- CCWfn::tgrad_ooc, ccwfn.cc:1150-1316
 *   ugx_density           <- CCDensity::onepdm / twopdm, ccdensity.cc:89-339
 *   ugx_pert_create       <- CCPert::CCPert + pertbar, ccpert.cc:11-176
 *   ugx_pert_solve        <- CCPert::solve, ccpert.cc:178-225
 *   ugx_linresp           <- CCResp::linresp, ccresp.cc:27-124
 *   UGX_NOT_CONVERGED     <- PSIEXCEPTION("Computation has not converged."),
 *                            ccwfn.cc:230-231
 *
 * Orbital order is the reference's QT order restricted to active orbitals:
 * occupied (no) first, then virtual (nv).  Integrals are real Dirac
 * <pq|rs> = (pr|qs); each block is row-major in the index order its name
 * gives:
 *   oooo[m][n][i][j] = <mn|ij>      no^4
 *   ooov[m][n][i][e] = <mn|ie>      no^3 nv
 *   oovv[m][n][e][f] = <mn|ef>      no^2 nv^2
 *   ovov[m][b][j][e] = <mb|je>      no^2 nv^2
 *   ovvv[m][b][e][f] = <mb|ef>      no nv^3
 *   vvvv[a][b][e][f] = <ab|ef>      nv^4
 * The Fock matrix is nact x nact (nact = no + nv), row-major.
 * Amplitudes: t1[i][a] (no*nv), t2[i][j][a][b] (no^2 nv^2), b fastest.
 *
 * Threading: one context = one HIP stream on one device; calls on a context
 * must come from one host thread at a time.  Host buffers are read during
 * the call only and never retained.
 */
#ifndef UGX_H
#define UGX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  UGX_OK = 0,
  UGX_NOT_CONVERGED = 1,
  UGX_BAD_INPUT = 2,
  UGX_OOM = 3,
  UGX_HIP_ERROR = 4,
  UGX_NO_DEVICE = 5,
  UGX_NOT_READY = 6,
  UGX_NONFINITE = 7
};

/* Arithmetic of the dense contractions. */
enum {
  UGX_PREC_FP64 = 0,   /* FP64 MFMA, FP64 accumulate: the parity form */
  UGX_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulate */
  UGX_PREC_BF16X3 = 2, /* split-BF16x3 MFMA (hi*hi + hi*lo + lo*hi), FP32 accumulate */
  UGX_PREC_BF16X3_PAIR = 3, /* ugx_gemm only: BF16x3 on operands already split into BF16 hi/lo
                               halves (hi array, then lo array), A[m][k] and B[n][k] */
  UGX_PREC_BF16 = 4,         /* ugx_mlp_forward only: BF16 MFMA inference, FP32 accumulate */
  UGX_PREC_BF16X3_PANEL = 5  /* ugx_gemm only: FP64 operands packed into BF16 hi/lo MFMA panels,
                                multiplied by the LDS-copy BF16x3 kernel of the (T) W3 build */
};

/* Initial amplitudes: MP2 (ccwfn.cc:75-87), the caller's (ugx_set_amplitudes),
 * or the learned guess: the MP2 doubles corrected by the network set with
 * ugx_set_guess_network, with the first-order singles at those doubles,
 * t1 = R1(t1 = 0, t2) / D1 (one Jacobi step of build_t1, ccwfn.cc:432-469,
 * from t1 = 0). */
enum { UGX_GUESS_MP2 = 0, UGX_GUESS_USER = 1, UGX_GUESS_MLP = 2 };

/* Polish of the mixed forms, after the switch to FP64 accuracy:
 *  PLAIN:  every residual evaluated in FP64;
 *  REFINE: one FP64 residual at the switch point T_s, then residuals
 *          R(T) = R(T_s) + [R(T) - R(T_s)], where each o^2v^4 / o^3v^3
 *          contraction A.B of the difference is evaluated as
 *          A.(B - B_s) + (A - A_s).B_s in the low precision and every other
 *          term in FP64.  The low precision's error then scales with the
 *          size of T - T_s instead of T, well below R_CONVERGENCE = 1e-7. */
enum { UGX_POLISH_PLAIN = 0, UGX_POLISH_REFINE = 1 };

/* Energy expression: the reference's (ccwfn.cc:272 assigns instead of
 * accumulating, so only the last (i,a) term of 2 f_ia t1_ia survives; zero
 * for canonical RHF orbitals) or the full sum. */
enum { UGX_ENERGY_REFERENCE = 0, UGX_ENERGY_FULL = 1 };

typedef struct ugx_ctx ugx_ctx;
typedef struct ugx_mlp ugx_mlp;

typedef struct {
  int wfn_ccsd_t;          /* WFN: 0 = "CCSD", 1 = "CCSD_T" (ugacc.cc:60) */
  int maxiter;             /* MAXITER, default 100 (ugacc.cc:63) */
  int diis;                /* DIIS, default 1 (ugacc.cc:64) */
  double r_convergence;    /* R_CONVERGENCE, default 1e-7 (ugacc.cc:65) */
  int ooc;                 /* OOC, default 0 (ugacc.cc:67); accepted for the
                              drop-in: the reference returns tcorr (0) or
                              tcorr_ooc (1), equal values, which ugx_triples
                              computes either way */
  int diis_nvec;           /* subspace size, 8 (ccwfn.cc:205); at most 16 */
  int diis_min_vecs;       /* extrapolate once this many vectors are stored;
                              3 = "subspace_size() > 2" (ccwfn.cc:220) */
  int diis_policy;         /* 0: replace the oldest vector once full;
                              1: replace the one with the largest error */
  int precision;           /* UGX_PREC_* of the low-precision phase */
  double polish_threshold; /* mixed forms: switch to FP64 once ||R|| drops
                              below this (e.g. 1e-5); ignored for FP64 */
  int guess;               /* UGX_GUESS_MP2 (ccwfn.cc:81-87), UGX_GUESS_USER or UGX_GUESS_MLP */
  int energy_form;         /* UGX_ENERGY_REFERENCE or UGX_ENERGY_FULL */
  int print;               /* PRINT: 0 silent, >=1 iteration table on stdout */
  int polish_keep_diis;    /* mixed forms: keep the DIIS subspace when switching
                              to the FP64 polish (default 1); 0 restarts it */
  int graph;               /* 1: replay each iteration's residual kernels from a
                              captured hipGraph (launch-bound small systems);
                              results are bitwise those of eager execution */
  int polish_mode;         /* mixed forms: UGX_POLISH_PLAIN (default) or
                              UGX_POLISH_REFINE */
} ugx_options;

typedef struct {
  double e_mp2;            /* correlation energy of the initial amplitudes (ccwfn.cc:196) */
  double e_ccsd;           /* CCSD correlation energy as the reference reports it (ccwfn.cc:233) */
  double e_ccsd_final;     /* energy of the final amplitudes */
  double e_t;              /* (T) correction, if computed by ugx_ccsd (WFN=CCSD_T) */
  double rms;              /* final ||R||_2 (ccwfn.cc:568-581) */
  int niter;               /* iterations */
  int niter_low;           /* iterations in the low-precision phase */
  double seconds_ccsd;     /* wall time of the iterations */
  double seconds_triples;  /* wall time of (T) */
  double flops_ccsd;       /* algorithmic flops executed by the contractions */
  int device_allocs;       /* device allocations made by the iterations: 0 when
                              ugx_prepare (which ugx_ccsd calls first) left every
                              workspace and cached conversion in place */
} ugx_result;

/* Fills *opt with the reference's defaults (ugacc.cc:58-68). */
void ugx_default_options(ugx_options *opt);

int ugx_create(ugx_ctx **ctx, int device);
void ugx_destroy(ugx_ctx *ctx);
const char *ugx_last_error(const ugx_ctx *ctx);

/* Copies the Fock matrix and the six Dirac blocks to the device, where they
 * stay resident (hamiltonian.cc:19-103). */
int ugx_load_hamiltonian(ugx_ctx *ctx, int no, int nv, const double *fock, const double *oooo,
                         const double *ooov, const double *oovv, const double *ovov, const double *ovvv,
                         const double *vvvv);

/* Initial amplitudes for UGX_GUESS_USER (for example a learned guess). */
int ugx_set_amplitudes(ugx_ctx *ctx, const double *t1, const double *t2);

/* Allocates and converts, for the loaded Hamiltonian, everything a
 * ugx_ccsd with these options needs before its first iteration: DIIS
 * vectors, refinement and guess buffers, the contraction workspaces, the
 * cached integral conversions of every precision the solve uses and (with
 * autotuning on) the tuned tiles.  ugx_ccsd calls it itself when the
 * options differ from the last prepared ones; calling it beforehand moves
 * that one-time cost out of the solve.  (The reference allocates in the
 * CCWfn constructor, ccwfn.cc:96-106.) */
int ugx_prepare(ugx_ctx *ctx, const ugx_options *opt);

/* Solves the CCSD equations (ccwfn.cc:188-231), then (T) when
 * opt->wfn_ccsd_t.  Returns UGX_OK or UGX_NOT_CONVERGED; res is filled in
 * either case. */
int ugx_ccsd(ugx_ctx *ctx, const ugx_options *opt, ugx_result *res);

/* (T) correction from the current amplitudes: the value of tcorr and
 * tcorr_ooc (ccwfn.cc:647-793), f_ov terms included. */
int ugx_triples(ugx_ctx *ctx, int precision, double *e_t);

/* Per-triplet (T) energies in the i>=j>=k loop order of ccwfn.cc:810-812
 * (no(no+1)(no+2)/6 values). */
int ugx_triples_per_triplet(ugx_ctx *ctx, int precision, double *e);

/* W3 of one triplet (ccwfn.cc:1388-1418), nv^3 values in [a][b][c] order. */
int ugx_w3(ugx_ctx *ctx, int i, int j, int k, int precision, double *W3);

/* Copies the current amplitudes to the host (either pointer may be NULL). */
int ugx_get_amplitudes(ugx_ctx *ctx, double *t1, double *t2);

/* Correlation energy of the current amplitudes (ccwfn.cc:260-282). */
int ugx_energy(ugx_ctx *ctx, int energy_form, double *e);

/* One residual evaluation from amplitudes t1/t2 (device copies are set from
 * them): raw residuals R1 (build_t1, ccwfn.cc:432-469) and R2 (build_t2,
 * :471-539).  Either output may be NULL. */
int ugx_residual(ugx_ctx *ctx, const double *t1, const double *t2, int precision, double *R1, double *R2);

/* The F and W intermediates of the last ugx_residual call in the
 * reference's layouts: Fvv[a][e], Foo[m][i], Fov[m][e], Woooo[m][n][i][j],
 * Wovov[m][b][j][e], Wovvo[m][b][e][j], Wovoo[m][b][i][j] (ccwfn.cc:311-430).
 * Any pointer may be NULL. */
int ugx_intermediates(ugx_ctx *ctx, double *Fvv, double *Foo, double *Fov, double *Woooo, double *Wovov,
                      double *Wovvo, double *Wovoo);

/* HBAR::HBAR (hbar.cc:8-228): the one- and two-body blocks of the
 * similarity-transformed Hamiltonian built on the device in FP64 from the
 * context's current amplitudes (after ugx_ccsd, or ugx_set_amplitudes) and
 * kept there; the blocks asked for (non-NULL) are copied out in the
 * reference's layouts: Hov[m][e], Hoo[m][i], Hvv[a][e], Hoooo[m][n][i][j],
 * Hvvvv[a][b][e][f], Hvovv[a][m][e][f], Hooov[m][n][i][e], Hovvo[m][b][e][j],
 * Hovov[m][b][j][e], Hvvvo[a][b][e][i], Hovoo[m][b][i][j].  Hoooo and Hvvvv
 * carry the reference's doubled single-t1 term (hbar.cc:80, :93). */
int ugx_hbar(ugx_ctx *ctx, double *Hov, double *Hoo, double *Hvv, double *Hoooo, double *Hvvvv, double *Hvovv,
             double *Hooov, double *Hovvo, double *Hovov, double *Hvvvo, double *Hovoo);

/* CCLambda::CCLambda + CCLambda::compute_lambda (cclambda.cc:13-46,
 * :93-135): the CCSD Lambda equations over the HBAR of the context's current
 * amplitudes (after ugx_ccsd), in FP64 on the device.  Guess l1 = 2 t1,
 * l2 = 2 (2 t2 - t2^T); each iteration builds G (cclambda.cc:137-156), the
 * raw l1 residual (:158-210) and l2 residual Z + Z^T (:212-268), updates
 * l = l_old + R/D with the CCSD denominators and stops when the 2-norm of
 * the raw residual is below r_convergence (before DIIS, as the reference);
 * DIIS as in ugx_ccsd.  Options read: maxiter, diis, r_convergence,
 * diis_nvec, diis_min_vecs, diis_policy, print (NULL: the defaults).
 * Outputs (each may be NULL): the pseudoenergy 1/2 <ij|ab> l2_ij^ab of the
 * final l2 (cclambda.cc:270-288), the iteration count, the final residual
 * norm, and trace (maxiter + 1 values): the pseudoenergy of the guess, after
 * each iteration (the values the reference prints) and of the final l2 at
 * index niter.  UGX_NOT_CONVERGED with "Computation has not converged."
 * as the reference's exception (cclambda.cc:133-134).  With
 * opt->wfn_ccsd_t = 1 and ugx_tgrad run on the current amplitudes, the
 * residuals include the (T) sources s1 and s2/2 that CCSD_T with
 * DERTYPE=FIRST adds (cclambda.cc:181, :243): the CCSD(T) Lambda. */
int ugx_lambda(ugx_ctx *ctx, const ugx_options *opt, double *pseudoenergy, int *niter, double *rms,
               double *trace);

/* The Lambda amplitudes of the last ugx_lambda, l1[i][a] and l2[i][j][a][b]
 * (either pointer may be NULL). */
int ugx_get_lambda(ugx_ctx *ctx, double *l1, double *l2);

/* One evaluation of build_G, build_l1 and build_l2 (cclambda.cc:137-268) at
 * the given l1/l2 over the HBAR of the current amplitudes: the raw residuals
 * R1[i][a] and R2[i][j][a][b] (before the denominators) and Goo[m][i],
 * Gvv[a][e].  Outputs may be NULL. */
int ugx_lambda_residual(ugx_ctx *ctx, const double *l1, const double *l2, double *R1, double *R2, double *Goo,
                        double *Gvv);

/* CCWfn::tgrad (ccwfn.cc:854-1098; compute_energy runs it for WFN=CCSD_T
 * with DERTYPE=FIRST, :252-255): the (T) gradient terms of CCSD(T) from the
 * context's current amplitudes (after ugx_ccsd), in FP64 on the device, with
 * t3 and l3 (:917-956) formed one occupied slice at a time (the peak extra
 * device memory is about 3 no^2 nv^3 doubles).  They stay on the device for
 * the next ugx_lambda (wfn_ccsd_t = 1) and ugx_density, until the amplitudes
 * change.  Outputs asked for (non-NULL): the Lambda sources s1[i][a] and
 * s2[i][j][a][b], and the (T) density parts Doo[i][j], Dvv[a][b], Dov[i][a],
 * Gooov[i][j][k][a], Gvvvo[a][b][c][i], Goovv[i][j][a][b]. */
int ugx_tgrad(ugx_ctx *ctx, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
              double *Gvvvo, double *Goovv);

/* CCWfn::tgrad_ooc (ccwfn.cc:1150-1316; compute_energy runs it instead of
 * tgrad when OOC=true, :252-255): the same terms and outputs as ugx_tgrad,
 * except that the (T) parts of Doo and Dvv keep only their diagonals (the
 * reference accumulates Doo[i][i] and Dvv[a][a] only, :1197, :1242).  For
 * canonical Fock matrices that is tgrad_ooc's result.  tgrad_ooc's s2
 * takes a disconnected f_ov term in another form than tgrad's, a difference
 * linear in f_ov: up to |f_ov| = 1e-8 (a converged SCF's residual) it stays
 * below ~1e-8 of s2's scale; beyond that the call returns UGX_BAD_INPUT. */
int ugx_tgrad_ooc(ugx_ctx *ctx, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
                  double *Gvvvo, double *Goovv);

/* CCDensity::onepdm + twopdm (ccdensity.cc:89-339): the CCSD one- and
 * two-particle density blocks, built on the device in FP64 from the
 * context's amplitudes and the Lambda amplitudes of the last ugx_lambda, and
 * their energies.  energies (8 values, or NULL): E_one (f_oo.Doo + f_vv.Dvv,
 * ccdensity.cc:138-146), E_two, then the OOOO, VVVV, OOOV, VVVO, OVOV and
 * OOVV parts of E_two (the values the reference prints).  Blocks asked for
 * (non-NULL) are copied out in the reference's layouts: Doo[i][j], Dvv[a][b],
 * Dov[i][a], Dvo[a][i], Goooo[i][j][k][l], Gvvvv[a][b][c][d],
 * Gooov[i][j][k][a], Gvvvo[a][b][c][i], Govov[i][a][j][b], Goovv[i][j][a][b].
 * When the last ugx_lambda was the CCSD(T) Lambda, the blocks include the
 * (T) parts of ugx_tgrad, as CCSD_T copies them in first
 * (ccdensity.cc:43-74).  At converged amplitudes and multipliers E_one +
 * E_two is the CCSD (or CCSD(T)) correlation energy when f_ov = 0
 * (ugacc.cc:129-137). */
int ugx_density(ugx_ctx *ctx, double *energies, double *Doo, double *Dvv, double *Dov, double *Dvo, double *Goooo,
                double *Gvvvv, double *Gooov, double *Gvvvo, double *Govov, double *Goovv);

/* ---------------------------------------------------------------------
 * Linear response (DERTYPE=SECOND, ugacc.cc:143-241).  A perturbation is a
 * one-electron MO operator over the active orbitals (nact x nact, row-major,
 * occupied first), such as Perturbation::prop_p (perturbation.cc): the
 * dipole -(r - O) ("Mu") or the magnetic dipole -1/2 (r - O) x grad ("L").
 * Each perturbation lives in a slot of the context, named by a handle,
 * until ugx_pert_release or the next ugx_load_hamiltonian.
 * --------------------------------------------------------------------- */

/* CCPert::CCPert (ccpert.cc:11-75): the perturbation at frequency omega over
 * the HBAR and Lambda amplitudes of the last ugx_lambda: its similarity
 * transform (pertbar, :112-176), the HBAR-diagonal denominators and the
 * guesses X = Y = Abar / (D + omega).  *handle receives the slot. */
int ugx_pert_create(ugx_ctx *ctx, const double *pert, double omega, int *handle);

/* CCPert::solve (ccpert.cc:178-225): hand 0 solves the right-hand perturbed
 * amplitudes X (build_X1/X2, :382-470), hand 1 the left-hand Y (build_G,
 * build_Y1/Y2, :489-882; needs the slot's X); x = x_old + R/(D + omega),
 * break on the 2-norm of the raw residual below r_convergence before DIIS,
 * DIIS as in ugx_ccsd.  Options read: maxiter, diis, r_convergence,
 * diis_nvec, diis_min_vecs, diis_policy, print (NULL: the defaults).
 * Outputs (may be NULL): the pseudoresponse (:229-249), iterations, final
 * residual norm.  UGX_NOT_CONVERGED with "Computation has not converged.". */
int ugx_pert_solve(ugx_ctx *ctx, int handle, int hand, const ugx_options *opt, double *pseudoresponse, int *niter,
                   double *rms);

/* The slot's perturbed amplitudes X1[i][a], X2[i][j][a][b], Y1, Y2 (any
 * pointer may be NULL); Y holds its guess until a left-hand solve. */
int ugx_pert_amplitudes(ugx_ctx *ctx, int handle, double *X1, double *X2, double *Y1, double *Y2);

/* One residual evaluation of the slot's equations at amplitudes (A1, A2):
 * hand 0 build_X1/X2, hand 1 build_G/Y1/Y2 (with the slot's X); raw R1 and
 * R2 (before the denominators).  Outputs may be NULL. */
int ugx_pert_residual(ugx_ctx *ctx, int handle, int hand, const double *A1, const double *A2, double *R1, double *R2);

/* CCResp::linresp (ccresp.cc:27-124): <<A;B>> with A the perturbation of
 * slot hx and B the amplitudes of slot hy (its X, and its Y as they are:
 * the guess after a right-hand solve only, as with MYHAND=RIGHT). */
int ugx_linresp(ugx_ctx *ctx, int hx, int hy, double *value);

/* Frees a perturbation slot. */
int ugx_pert_release(ugx_ctx *ctx, int handle);

/* Device time per contraction label of the last ugx_ccsd / ugx_triples
 * call when profiling is on, as "label launches ms gflop\n" lines.  Returns
 * the number of bytes needed (including the terminating zero). */
void ugx_set_profiling(ugx_ctx *ctx, int on);
size_t ugx_kernel_stats(ugx_ctx *ctx, char *buf, size_t buflen);

/* General GEMM on device buffers with the library's MFMA kernel:
 * C = alpha op(A) op(B) + beta C (row-major, op 'N' or 'T'); pointers are
 * device addresses; A/B are double for UGX_PREC_FP64 and UGX_PREC_BF16X3_PANEL
 * and float otherwise.
 * tile = -1 picks the tile by size.  For kernel tests and benchmarks. */
int ugx_gemm(ugx_ctx *ctx, char opa, char opb, int M, int N, int K, double alpha, const void *A, long lda,
             const void *B, long ldb, double beta, double *C, long ldc, int precision, int tile);

/* ugx_gemm with UGX_PREC_BF16X3 (FP32 operands A, B in any layout), without
 * any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The BF16 hi/lo staging of A and B lives
 * in a 64 MiB device array of the library, one per GPU, held until the
 * process ends: (M + N) K must stay within 16 Mi (2^24) elements, otherwise
 * UGX_BAD_INPUT.  The work is queued on the context's stream (ugx_synchronize
 * waits for it); every call on one GPU shares that array, so calls that
 * overlap on one GPU (several contexts or threads) race.  Results are
 * bitwise those of ugx_gemm(..., UGX_PREC_BF16X3, tile). */
int ugx_gemm_bf16x3_static(ugx_ctx *ctx, char opa, char opb, int M, int N, int K, double alpha, const float *A,
                           long lda, const float *B, long ldb, double beta, double *C, long ldc, int tile);

/* Synchronises the context's stream. */
int ugx_synchronize(ugx_ctx *ctx);

/* One closed-shell problem of a batch: sizes, Fock matrix and the six
 * Dirac blocks, in the layouts of ugx_load_hamiltonian. */
typedef struct {
  int no, nv;
  const double *fock, *oooo, *ooov, *oovv, *ovov, *ovvv, *vvvv;
} ugx_problem;

/* Solves nprob independent problems of any sizes (the reference solves one
 * at a time through CCWfn::compute_energy, ccwfn.cc:188-258, and
 * tcorr_ooc_TJL, :795-852).  Each problem gets its own stream and device
 * buffers; nworkers host threads first load and prepare (ugx_prepare)
 * every problem, then take problems largest first and run their CCSD
 * iterations concurrently, allocating no device memory (opt->graph replays
 * each iteration's residual from a hipGraph).  With opt->wfn_ccsd_t, (T) then runs once per
 * (no, nv) class: grouped W3 GEMMs whose pointer tables span every triplet
 * of every problem of the class, and one persistent energy kernel.
 * results[p] are those of ugx_ccsd for problem p; seconds_ccsd is the
 * problem's CCSD wall time on its worker, seconds_triples its share of the
 * class's (T) time.  *wall_seconds receives the whole batch's wall time.
 * The context supplies the device, the autotuning and profiling settings
 * and the guess network; with profiling on, its ugx_kernel_stats then
 * covers every problem of the batch.  Returns UGX_NOT_CONVERGED when any problem did not converge
 * (its (T) is then skipped). */
int ugx_batch_solve(ugx_ctx *ctx, int nprob, const ugx_problem *probs, const ugx_options *opt, int nworkers,
                    ugx_result *results, double *wall_seconds);

/* GEMM tile autotuning (off by default): when on, the first launch of each
 * contraction shape times every tile of the kernel's tile table on the
 * real operands (into scratch) and keeps the fastest for the context's
 * lifetime.  No counterpart in the reference (its loops have no tiles). */
int ugx_set_autotune(ugx_ctx *ctx, int on);
/* The tuned shapes as "shape -> tile ms" lines; returns the bytes needed. */
size_t ugx_tile_table(ugx_ctx *ctx, char *buf, size_t buflen);
/* The process's tuned tiles to / from a text file, keyed by GPU
 * architecture, CU count and contraction shape, so a later process skips
 * the tuning (loading merges; entries of other GPUs are kept, unused).
 * ugx_create loads the file named by the UGX_TILE_FILE environment
 * variable, once per process. */
int ugx_save_tile_table(ugx_ctx *ctx, const char *path);
int ugx_load_tile_table(ugx_ctx *ctx, const char *path);

/* ---------------------------------------------------------------------
 * Learned guess.  Replaces the MP2 doubles of CCWfn::CCWfn
 * (ccwfn.cc:81-87) by t2 = t2(MP2) + y_scale (f(x_ijab) + f(x_jiba)) / 2,
 * and the singles f_ia / D_ia (:75-79) by the first-order singles at that
 * t2 (see UGX_GUESS_MLP), where f is a network 8 -> 128 -> 128 -> 128 -> 1 with GELU (tanh form)
 * after the hidden layers and x are the amplitude's features (computed by
 * ugx_features): asinh(1e3 t_MP2), log(-D), asinh(1e2 <ij|ab>),
 * asinh(1e2 <ij|ba>), f_ii, f_jj, f_aa, f_bb.  The reference has no such
 * guess; the fixed point of the iterations does not depend on it.
 * Parameters are FP32 in the order W1[128][8], b1, W2[128][128], b2,
 * W3[128][128], b3, W4[128], b4 (layers z = W h + b); the features are
 * standardised as (x - shift) / scale before the first layer.
 * --------------------------------------------------------------------- */

/* Number of FP32 parameters (34305). */
long ugx_mlp_num_params(void);
/* New network on a device, initialised from seed (U(+-1/sqrt(fan_in))). */
int ugx_mlp_create(ugx_mlp **net, int device, unsigned long long seed);
void ugx_mlp_destroy(ugx_mlp *net);
const char *ugx_mlp_last_error(const ugx_mlp *net);
int ugx_mlp_set_params(ugx_mlp *net, const float *params);
int ugx_mlp_get_params(ugx_mlp *net, float *params);
/* Feature standardisation (8 shifts, 8 positive scales) and output scale. */
int ugx_mlp_set_scaling(ugx_mlp *net, const float *x_shift, const float *x_scale, double y_scale);
/* y[r] = f(X[r]) for raw feature rows X[n][8]; precision UGX_PREC_FP32 (the
 * training forward, n <= 65536) or UGX_PREC_BF16 (the fused inference
 * kernel the guess uses). */
int ugx_mlp_forward(ugx_mlp *net, const float *X, long n, int precision, float *y);
/* Loss (1/n) sum (f(X[r]) - y[r])^2 and its gradient with respect to the
 * parameters (FP32 forward and backward on the device), n <= 65536. */
int ugx_mlp_loss_grad(ugx_mlp *net, const float *X, const float *y, long n, double *loss, float *grad);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over `epochs` passes of shuffled
 * mini-batches (seeded) of the n rows; epoch_loss (may be NULL) receives
 * the mean batch loss of each epoch. */
int ugx_mlp_train(ugx_mlp *net, const float *X, const float *y, long n, int epochs, int batch, double lr,
                  unsigned long long seed, double *epoch_loss);
/* Raw features X[r][8] of the amplitudes idx[r] (flattened ijab, b fastest;
 * idx NULL: the first n) of the loaded Hamiltonian. */
int ugx_features(ugx_ctx *ctx, long n, const long *idx, float *X);
/* Network used by UGX_GUESS_MLP (same device; it must outlive its use). */
int ugx_set_guess_network(ugx_ctx *ctx, const ugx_mlp *net);

#ifdef __cplusplus
}
#endif
#endif /* UGX_H */
