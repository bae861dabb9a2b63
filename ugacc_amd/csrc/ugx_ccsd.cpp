// C-ABI layer and host driver of the gfx950 CCSD / CCSD(T) library.
//
// The host owns the iteration (ccwfn.cc:188-258): each iteration is a fixed
// sequence of device work on the context's stream (tau, the F and W
// intermediates and the T1/T2 residuals as MFMA contractions, the fused
// update/denominator/norm kernel, the DIIS vectors), and only scalars come
// back to the host: the residual norm, the energy partials and the DIIS dot
// products.  The DIIS linear solve is done on the host in FP64.
//
// The T2 residual is assembled from its half X: every term of build_t2
// (ccwfn.cc:471-539) comes in a pair related by (i,a) <-> (j,b), and
// R2_ijab = X_ijab + X_jiba, so each pair is computed once.  Terms that are
// symmetric on their own (<ij|ab>, tau.Wmnij, the ladder tau.<ab|ef>) enter
// X with a factor 1/2.
#include <hip/hip_runtime.h>

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ugx.h"
#include "ugx_engine.h"
#include "ugx_host.h"
#include "ugx_internal.h"

using namespace ugx;


// One perturbed wave function of the linear response (ccpert.cc): the MO
// perturbation's blocks P*, its similarity transform A*, and the right- and
// left-hand perturbed amplitudes with their previous values, by name;
// aliases point names at caller arrays for one residual evaluation.
struct PertSlot {
  double omega = 0.0;
  bool x_solved = false;
  std::map<std::string, std::unique_ptr<ugx::DBuf>> bufs;
  std::map<std::string, const double *> aliases;
  double *buf(const char *n) {
    auto it = bufs.find(n);
    if (it == bufs.end()) throw std::runtime_error(std::string("ugx: response slot has no ") + n);
    return it->second->d();
  }
  void make(const char *n, size_t elems) {
    auto &b = bufs[n];
    if (!b) b.reset(new ugx::DBuf());
    b->alloc(elems * 8);
  }
  void swap(const char *a, const char *b) { std::swap(bufs[a], bufs[b]); }
  void alias(const char *n, const double *p) { aliases[n] = p; }
};

struct ugx_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::unique_ptr<Engine> eng;
  int no = 0, nv = 0;
  bool loaded = false;
  bool have_amps = false;

  // Hamiltonian (device, constant)
  DBuf fock, eps, fov, fvo, foo, fvv;
  std::vector<double> h_fov;  // host copy of f_ov for the one-electron energy
  bool has_fov = false;       // some f_ov element is non-zero
  DBuf oooo, ooov, oovv, ovov, ovvv, vvvv;
  DBuf Loovv, Looov, Lovvv;
  // the integral parts of Wmbje and Wmbej in their [m][e][j][b] order:
  // -<mb|je> and <mj|eb>, copied into WJ / WB at the start of each residual
  DBuf WJ0, WB0;
  // amplitudes and intermediates
  DBuf t1, t1o, t2, t2o, tau, ttau, U, Y;
  DBuf Fvv, Foo, Fov, FaeP, FmiP;
  DBuf Woooo, WJ, WB, Z, Zx, P2;
  DBuf R1, X;
  DBuf part;
  // pinned landing buffer of the per-iteration scalar readbacks (partial
  // sums, DIIS dots, t1 for the energy)
  HBuf hpin;
  // DIIS
  std::vector<std::unique_ptr<DBuf>> diis_err, diis_amp;
  std::vector<double> diis_B;  // nvec x nvec, slot order
  std::vector<double> diis_norm;
  int diis_count = 0, diis_next = 0, diis_nvec = 0;
  // (T)
  DBuf tP, tQ, tWbuf, ttrip, tptrs, tptrs2, tE, tPart;
  DBuf tOff;  // blocked-W offset tables of the three (T) launches
  DBuf tOT2, tT1F;  // f_ov energy kernel: (<ij|ab>, t2) and (t1, f_ov) pair arrays
  DBuf tProb, tCounter;  // energy kernel: problem table (one entry), work-queue counter
  int tOff_blk = 0, tOff_v = 0;
  // per launch g: the column table of the old orientation, used as the row
  // table of the transposed BF16x3 launch, holds aligned runs of 4
  bool tOff_row4[3] = {false, false, false};
  bool tOff_merge4 = false;  // the same for the merged launch's tables (rows c, columns b, offsets a)
  int tW_f32 = 0;  // the (T) W buffer holds FP32 values
  DBuf f_tP, f_tQ, f_t2, f_ooov;  // FP32 operand copies for the mixed forms
  // BF16x3 (T): BF16 hi/lo pair operands, all k contiguous
  DBuf b_P, b_Q, b_OVt, b_t2, b_T2t, b_T2tq;
  // the same operands as MFMA panels, for the LDS-copy W3 kernel
  DBuf p_OV, p_t2, p_T2t, p_T2tq;
  // the big panels again with rows (x, y) at x * VP + y, VP = v rounded up
  // to 16, so every x-slab starts on a panel block (the merged launch)
  DBuf q_P, q_Q, q_T2t, q_T2tq;
  int t_merge = 1;   // BF16x3 W3: the z = c and z = b groups as one launch per (triplet, a)
  int t_zgroup_m = 4;  // consecutive triplets per group in the merged launch
  // merged W3: the z = a launch writes W first and the merged launch (K four
  // times as long) accumulates, so the read of the old W hides behind more
  // MFMA work
  int t_za_first = 1;
  // packed symmetric / antisymmetric ladder (launch_ladder_*): V+, V- once
  // per Hamiltonian; S, A and the two GEMM outputs per residual
  bool lad_packed = false;
  // Zmbij = sum_ef <mb|ef> tau_ij^ef on the same pairs: <mb|ef> +- <mb|fe>
  // once per Hamiltonian, the two GEMM outputs per residual
  DBuf zm_vp, zm_vm, zm_P, zm_M;
  // the tau term of Woooo, sum_ef <mn|ef> tau_ij^ef, likewise (rows (m,n))
  DBuf wo_vp, wo_vm, wo_P, wo_M;
  bool lad_full = false;  // UGX_LADDER=full: always the full contraction
  DBuf lad_vp, lad_vm, lad_S, lad_A, lad_LP, lad_LM;
  // refinement polish (UGX_POLISH_REFINE): the operands of the large
  // contractions at the base point (rf_*), each output's sum of its large
  // contractions there (rs_*), and the differences from the base (rd_*)
  DBuf rf_t2, rf_tau, rf_ttau, rf_U, rf_Y, rf_S, rf_A, rf_WB, rf_WJ, rf_WO, rf_FaeP, rf_FmiP;
  DBuf rs_Fvv, rs_Foo, rs_WO, rs_WJ, rs_WB, rs_Z, rs_R1, rs_X;
  DBuf rd_t2, rd_tau, rd_ttau, rd_U, rd_Y, rd_S, rd_A, rd_WB, rd_WJ, rd_WO, rd_FaeP, rd_FmiP;
  bool tables_ready = false;
  std::vector<int> trip_host;
  // learned guess: network (owned by the caller) and its output f(x_ijab)
  const ugx_mlp *mlp = nullptr;
  DBuf G;
  int t_batch = 0;  // cap on triplets per (T) batch (0: by the W budget)
  long t_wbytes = 48L << 30;  // (T) W buffer budget, bytes (at most a quarter of the free HBM)
  int t_zgroup = 4;  // consecutive triplets per output tile in the BF16x3 W3 launches
  int t_dma = 2;     // > 1: BF16x3 W3 launches copy panels into LDS, with that many stages (0: registers)
  int t_sort = 1;    // BF16x3 W3 launches: each launch's entries ordered by its own panel pair
  int t_align = 1;   // (T) batches end at a triplet-block boundary (all triplets of one i)
  // hipGraphs of the residual, keyed by (amplitude buffer parity, precision
  // + 4 for the refinement polish's difference residuals)
  struct ResidualGraph {
    const void *t1o = nullptr;
    int prec = -1;
    hipGraphExec_t exec = nullptr;
    double flops = 0.0;
  };
  std::vector<ResidualGraph> graphs;
  int graph_eager[8] = {};  // eager residuals run per (precision, mode) since the last reset
  // options signature the context's workspaces, cached conversions and
  // buffers were prepared for (prepare_solve); empty after a load
  std::string prepared_for;
  // HBAR blocks (hbar.cc:8-228), built by ugx_hbar, and its intermediates
  DBuf hb_ov, hb_oo, hb_vv, hb_oooo, hb_vvvv, hb_vovv, hb_ooov, hb_ovvo, hb_ovov, hb_vvvo, hb_ovoo;
  DBuf hb_B, hb_Bp, hb_Lt, hb_C, hb_D, hb_X, hb_E, hb_G, hb_t1L, hb_Woo, hb_Q;
  // Lambda (cclambda.cc): l1/l2 and their previous values, Goo/Gvv, the raw
  // l1 residual and the half residual Z of l2, the update's partial sums
  DBuf lam_l1, lam_l1o, lam_l2, lam_l2o, lam_Goo, lam_Gvv, lam_R1, lam_Z;
  bool have_lambda = false;
  // CCSD densities (ccdensity.cc): the ten blocks, and the intermediates of
  // their terms with three or more amplitudes
  DBuf dn_oo, dn_vv, dn_ov, dn_vo, dn_oooo, dn_vvvv, dn_ooov, dn_vvvo, dn_ovov, dn_oovv;
  DBuf dx_P, dx_Q, dx_h, dx_k, dx_k2, dx_S, dx_B, dx_C, dx_Y, dx_S1, dx_C1;
  DBuf dx_R, dx_R2, dx_R3, dx_U, dx_V, dx_W, dx_Y2, dx_Ut2, dx_TU, dx_A1, dx_A2, dx_A3, dx_B1, dx_B2, dx_perm;
  // (T) gradient terms of CCSD(T) (ccwfn.cc:854-1098): the Lambda sources
  // s1, s2 and the (T) density parts; tg_valid while they belong to the
  // current amplitudes; lam_with_t: the last Lambda solve included them (and
  // the densities then do)
  DBuf tg_s1, tg_s2, tg_oo, tg_vv, tg_ov, tg_ooov, tg_vvvo, tg_oovv;
  // one occupied slice of t3 and l3, [j][k][a][b][c]; the accumulators of the
  // l2 sources by slice, the [i][a][b][c] accumulator of Gvvvo; t1s, t2s;
  // <zy|an> and L_xyan with the slice index first
  DBuf tg_T, tg_L, tg_Y2a, tg_Y2b, tg_Y2l, tg_X, tg_Gt, tg_t1s, tg_t2s, tg_ooovT, tg_LooovT;
  bool tg_valid = false, lam_with_t = false;
  // linear response: perturbation slots (ugx_pert_*), the scratch of their
  // residuals (R1, half Z, the Y source S1/S2, W), the HBAR diagonals as
  // denominators, and named intermediates of the term tables
  std::vector<std::unique_ptr<PertSlot>> perts;
  DBuf rp_R1, rp_S1, rp_Z, rp_S2, rp_W, rp_eps, rp_A1, rp_A2;
  std::map<std::string, std::unique_ptr<DBuf>> rp_tmp;
  std::map<std::string, long> rp_tmp_size;
  // residual of the last ugx_residual, for ugx_intermediates
  bool have_intermediates = false;
  double last_flops = 0.0;

  long ov() const { return (long)no * nv; }
  long oovv_n() const { return ov() * ov(); }
};

namespace {

// A named host-side range of an entry point or phase (roctx): rocprofv3
// --marker-trace records it beside the kernel trace; without a profiler
// attached the push and pop are no-ops
struct Phase {
  explicit Phase(const char *name) { roctxRangePushA(name); }
  ~Phase() { roctxRangePop(); }
  Phase(const Phase &) = delete;
  Phase &operator=(const Phase &) = delete;
};

void require_loaded(ugx_ctx *c) {
  if (!c->loaded) throw UgxError(UGX_NOT_READY, "no Hamiltonian loaded (call ugx_load_hamiltonian)");
}

void upload(DBuf &b, const double *h, size_t n, hipStream_t st) {
  b.alloc(n * 8);
  check(hipMemcpyAsync(b.p, h, n * 8, hipMemcpyHostToDevice, st), "host-to-device copy");
}

double host_sum(ugx_ctx *c, int n) {
  double *h = c->hpin.p;
  check(hipMemcpyAsync(h, c->part.p, n * 8, hipMemcpyDeviceToHost, c->st), "partials copy");
  check(hipStreamSynchronize(c->st), "stream sync");
  double s = 0.0;
  for (int i = 0; i < n; i++) s += h[i];
  return s;
}

int npart_max(ugx_ctx *c) {
  return std::max(update_nparts(c->no, c->nv), energy_nparts(c->no, c->nv)) * 8 + 16;
}

// -------------------------------------------------------------- energy
// ccwfn.cc:260-282 for the amplitudes in t1 and tau.
double energy_of(ugx_ctx *c, int form, const double *t1dev) {
  const int no = c->no, nv = c->nv;
  int np = energy_nparts(no, nv);
  launch_energy(no, nv, c->fov.d(), t1dev, c->tau.d(), c->Loovv.d(), c->part.d(), np, c->st);
  // the tau.L partials and t1 come back with one synchronisation; f_ov
  // is kept on the host since load
  double *h = c->hpin.p, *t1 = c->hpin.p + np;
  check(hipMemcpyAsync(h, c->part.p, np * 8, hipMemcpyDeviceToHost, c->st), "partials copy");
  check(hipMemcpyAsync(t1, t1dev, c->ov() * 8, hipMemcpyDeviceToHost, c->st), "t1 copy");
  check(hipStreamSynchronize(c->st), "stream sync");
  double two = 0.0;
  for (int i = 0; i < np; i++) two += h[i];
  const std::vector<double> &fov = c->h_fov;
  double one = 0.0;
  if (form == UGX_ENERGY_REFERENCE) {
    // the '=' at ccwfn.cc:272 keeps only the last (i,a) term
    if (c->ov() > 0) one = 2 * fov[c->ov() - 1] * t1[c->ov() - 1];
  } else {
    for (long k = 0; k < c->ov(); k++) one += 2 * fov[k] * t1[k];
  }
  return one + two;
}

// Energy of the current amplitudes (t1, tau).
double energy(ugx_ctx *c, int form) { return energy_of(c, form, c->t1.d()); }

double t1norm(ugx_ctx *c) {
  std::vector<double> t1(c->ov());
  check(hipMemcpy(t1.data(), c->t1.p, c->ov() * 8, hipMemcpyDeviceToHost), "t1 copy");
  double s = 0.0;
  for (double x : t1) s += x * x;
  return std::sqrt(s / (2 * c->no));
}

void build_tau(ugx_ctx *c) { launch_tau(c->no, c->nv, c->t1.d(), c->t2.d(), c->tau.d(), c->ttau.d(), c->st); }

// Residual evaluation modes.  RES_BASE and RES_DELTA are the refinement
// polish: R(T) = R(T_s) + [R(T) - R(T_s)] around a base point T_s.
enum { RES_PLAIN = 0, RES_BASE = 1, RES_DELTA = 2 };

void ensure_refine(ugx_ctx *c) {
  const long o = c->no, v = c->nv, ov = o * v, oovv = ov * ov;
  for (DBuf *b : {&c->rf_t2, &c->rf_tau, &c->rf_ttau, &c->rf_U, &c->rf_Y, &c->rf_WB, &c->rf_WJ, &c->rs_WJ,
                  &c->rs_WB, &c->rs_X, &c->rd_t2, &c->rd_tau, &c->rd_ttau, &c->rd_U, &c->rd_Y, &c->rd_WB, &c->rd_WJ})
    b->alloc(oovv * 8);
  for (DBuf *b : {&c->rf_FaeP, &c->rs_Fvv, &c->rd_FaeP}) b->alloc(v * v * 8);
  for (DBuf *b : {&c->rf_FmiP, &c->rs_Foo, &c->rd_FmiP}) b->alloc(o * o * 8);
  for (DBuf *b : {&c->rf_WO, &c->rs_WO, &c->rd_WO}) b->alloc(o * o * o * o * 8);
  c->rs_Z.alloc(o * v * o * o * 8);
  c->rs_R1.alloc(ov * 8);
  if (c->lad_packed) {
    for (DBuf *b : {&c->rf_S, &c->rd_S}) b->alloc(c->lad_S.bytes);
    for (DBuf *b : {&c->rf_A, &c->rd_A}) b->alloc(c->lad_A.bytes);
  }
}

// ------------------------------------------------------------ residual
// Residuals of the amplitudes (t1o, t2o) with tau/ttau built from them:
// R1 -> c->R1, the T2 half X -> c->X.  prec is the arithmetic of the
// contractions; "large" below are the o^2v^4 / o^3v^3 ones.
//   RES_PLAIN: every term at the current amplitudes, in prec.
//   RES_BASE:  the same in FP64 (prec must be FP64); it also keeps the base
//              point: the operands of the large contractions (rf_*) and,
//              per output, the sum of its large contractions (rs_*).
//   RES_DELTA: each output's large contractions as its base sum plus the
//              contraction of the differences from the base point,
//              A.B - A_s.B_s = A.(B - B_s) + (A - A_s).B_s, in prec; the
//              small terms are evaluated at the current amplitudes in FP64.
//              The error of prec then scales with |T - T_s|, not |T|.
void residual(ugx_ctx *c, int prec, int mode = RES_PLAIN) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const hipStream_t st = c->st;
  const int P64 = PREC_FP64;
  const bool base = mode == RES_BASE, delta = mode == RES_DELTA, plain = mode == RES_PLAIN;
  // the engine's sizing mode (prepare_solve): contractions are only planned,
  // and the streaming kernels and copies between them are skipped
  const bool dry = E.sizing();
  if (base && prec != P64) throw UgxError(UGX_BAD_INPUT, "residual: the base point is evaluated in FP64");
  if (!plain) ensure_refine(c);
  // the small (t1) terms: FP64 in the polish; in a plain evaluation prec,
  // except for small systems (o v < 1024), where the operand conversions of
  // the low precision cost more launches than the FP64 arithmetic saves
  const int SP = (plain && c->ov() >= 1024) ? prec : P64;
  const double *t1 = c->t1o.d(), *t2 = c->t2o.d(), *tau = c->tau.d(), *ttau = c->ttau.d();
  auto V = [](const double *p, std::vector<int> d) { return vref(p, std::move(d)); };
  auto Cn = [](const double *p, std::vector<int> d) { return cref(p, std::move(d)); };
  const auto T1 = V(t1, {o, v}), T2 = V(t2, {o, o, v, v}), TAU = V(tau, {o, o, v, v}), TTAU = V(ttau, {o, o, v, v});
  const auto OOOO = Cn(c->oooo.d(), {o, o, o, o}), OOOV = Cn(c->ooov.d(), {o, o, o, v});
  const auto OOVV = Cn(c->oovv.d(), {o, o, v, v}), OVOV = Cn(c->ovov.d(), {o, v, o, v});
  const auto OVVV = Cn(c->ovvv.d(), {o, v, v, v}), VVVV = Cn(c->vvvv.d(), {v, v, v, v});
  const auto LOOVV = Cn(c->Loovv.d(), {o, o, v, v}), LOOOV = Cn(c->Looov.d(), {o, o, o, v});
  const auto LOVVV = Cn(c->Lovvv.d(), {o, v, v, v});
  const auto FOV = Cn(c->fov.d(), {o, v});
  const long n4 = c->oovv_n();

  // Operands of the large contractions that several of them read (Y, the
  // pair matrices S and A, their differences) are cached like constants:
  // each is re-converted in place, for the element types of prec, as soon as
  // this evaluation has rebuilt it, and every later read is a cache hit.
  const int cdt_lo = prec == P64 ? 0 : (prec == PREC_FP32 ? 1 : 2);
  const int cdt_hi = prec == P64 ? 0 : (prec == PREC_FP32 ? 1 : 3);
  auto rebuilt = [&](const void *p) { E.refresh(p, cdt_lo, cdt_hi); };
  // base point: keep a copy of an operand; differences from it
  auto keep = [&](DBuf &dst, const void *src, size_t bytes) {
    if (!dry) check(hipMemcpyAsync(dst.p, src, bytes, hipMemcpyDeviceToDevice, st), "copy");
  };
  auto dif = [&](DBuf &dst, const double *x, const DBuf &xs, long n) {
    if (!dry) launch_diff(n, x, xs.d(), dst.d(), st);
  };
  // adds an output's base sum of large contractions (RES_BASE, RES_DELTA)
  auto add_sum = [&](const DBuf &sum, double *out, long n) {
    if (!plain && !dry) launch_axpby(n, 1.0, sum.d(), 1.0, out, st);
  };
  // one large contraction out += alpha A.B; rs is the output's base sum
  // (first: the first large term of that output).  dA / dB are the
  // differences of variable operands (null for constants), Bs the base
  // value of a variable B.
  auto big = [&](const char *spec, double alpha, const TRef &A, const TRef *dA, const TRef &B, const TRef *dB,
                 const TRef *Bs, const TRef &out, const TRef &rs, bool first, const char *label) {
    if (plain) {
      E.contract(spec, alpha, A, B, 1.0, out, prec, label);
    } else if (base) {
      E.contract(spec, alpha, A, B, first ? 0.0 : 1.0, rs, P64, label);
    } else {
      if (dB) E.contract(spec, alpha, A, *dB, 1.0, out, prec, label);
      if (dA) E.contract(spec, alpha, *dA, Bs ? *Bs : B, 1.0, out, prec, label);
    }
  };

  // U = 2 t2 - t2(ab swapped), the spin-adapted combination used by build_t1
  // (ccwfn.cc:455,461) and the Wmbej ring term (ccwfn.cc:518-519).
  if (!dry) launch_build_U(o, v, t2, c->U.d(), st);
  const auto UU = V(c->U.d(), {o, o, v, v});
  if (base) {
    keep(c->rf_t2, t2, n4 * 8); keep(c->rf_tau, tau, n4 * 8); keep(c->rf_ttau, ttau, n4 * 8);
    keep(c->rf_U, c->U.p, n4 * 8);
  } else if (delta) {
    dif(c->rd_t2, t2, c->rf_t2, n4); dif(c->rd_tau, tau, c->rf_tau, n4); dif(c->rd_ttau, ttau, c->rf_ttau, n4);
    dif(c->rd_U, c->U.d(), c->rf_U, n4);
  }
  const auto dT2 = V(c->rd_t2.d(), {o, o, v, v}), dTAU = V(c->rd_tau.d(), {o, o, v, v});
  const auto dTTAU = V(c->rd_ttau.d(), {o, o, v, v}), dU = V(c->rd_U.d(), {o, o, v, v});

  // ---- build_F, ccwfn.cc:311-356
  const auto FVV = V(c->Fvv.d(), {v, v}), FOO = V(c->Foo.d(), {o, o}), FOVI = V(c->Fov.d(), {o, v});
  if (!dry) check(hipMemcpyAsync(c->Fvv.p, c->fvv.p, (size_t)v * v * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("me,ma->ae", -0.5, FOV, T1, 1.0, FVV, SP, "F");
  E.contract("mf,mafe->ae", 1.0, T1, LOVVV, 1.0, FVV, SP, "F");
  big("mnaf,mnef->ae", -1.0, TTAU, &dTTAU, LOOVV, nullptr, nullptr, FVV, V(c->rs_Fvv.d(), {v, v}), true, "F");
  add_sum(c->rs_Fvv, c->Fvv.d(), (long)v * v);
  if (!dry) check(hipMemcpyAsync(c->Foo.p, c->foo.p, (size_t)o * o * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("ie,me->mi", 0.5, T1, FOV, 1.0, FOO, SP, "F");
  E.contract("ne,mnie->mi", 1.0, T1, LOOOV, 1.0, FOO, SP, "F");
  big("inef,mnef->mi", 1.0, TTAU, &dTTAU, LOOVV, nullptr, nullptr, FOO, V(c->rs_Foo.d(), {o, o}), true, "F");
  add_sum(c->rs_Foo, c->Foo.d(), (long)o * o);
  if (!dry) check(hipMemcpyAsync(c->Fov.p, c->fov.p, (size_t)o * v * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("nf,mnef->me", 1.0, T1, LOOVV, 1.0, FOVI, SP, "F");

  // ---- build_W, ccwfn.cc:358-430
  const long n_oooo = (long)o * o * o * o;
  const auto WO = V(c->Woooo.d(), {o, o, o, o});
  if (!dry) check(hipMemcpyAsync(c->Woooo.p, c->oooo.p, (size_t)n_oooo * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("je,mnie->mnij", 1.0, T1, OOOV, 1.0, WO, SP, "Woooo");
  E.contract("ie,nmje->mnij", 1.0, T1, OOOV, 1.0, WO, SP, "Woooo");  // <mn|ej> = <nm|je>
  // the pair matrices S, A of tau (launch_ladder_pack), shared by the tau
  // terms of Woooo, Zmbij and the ladder
  const int Q = o * (o + 1) / 2, Qs = o * (o - 1) / 2, P = v * (v + 1) / 2, Ps = v * (v - 1) / 2;
  const int ldp = (int)ladder_ld(P), ldm = (int)ladder_ld(Ps);  // k padded with zeros
  // rows x (e,f) pairs of a constant block times the pair matrices S / A,
  // unpacked into out: sum_ef in[r][e][f] tau_ij^ef on pairs i>=j / i>j and
  // e>=f / e>f
  auto pair_term = [&](int rows, DBuf &vp, DBuf &vm, DBuf &zp, DBuf &zm, const double *S, const double *A,
                       double beta, double *out, int pr, const char *label) {
    E.contract("ne,qe->nq", 1.0, Cn(vp.d(), {rows, ldp}), Cn(S, {Q, ldp}), 0.0, V(zp.d(), {rows, Q}), pr, label);
    if (Qs > 0) {
      if (Ps > 0)
        E.contract("ne,qe->nq", 1.0, Cn(vm.d(), {rows, ldm}), Cn(A, {Qs, ldm}), 0.0, V(zm.d(), {rows, Qs}), pr,
                   label);
      else
        if (!dry) launch_fill((long)rows * Qs, 0.0, zm.d(), st);
    }
    if (!dry) launch_pairs_unpack(rows, o, zp.d(), zm.d(), beta, out, st);
  };
  // the pair matrices this evaluation contracts: S, A (plain, base) or
  // their differences from the base point (delta)
  const double *pS = c->lad_S.d(), *pA = c->lad_A.d();
  if (c->lad_packed) {
    if (!dry) launch_ladder_pack(o, v, tau, c->lad_S.d(), c->lad_A.d(), st);
    if (base) {
      keep(c->rf_S, c->lad_S.p, c->lad_S.bytes); keep(c->rf_A, c->lad_A.p, c->lad_A.bytes);
    } else if (delta) {
      dif(c->rd_S, c->lad_S.d(), c->rf_S, (long)(c->lad_S.bytes / 8));
      dif(c->rd_A, c->lad_A.d(), c->rf_A, (long)(c->lad_A.bytes / 8));
      pS = c->rd_S.d(); pA = c->rd_A.d();
    }
    rebuilt(pS);
    rebuilt(pA);
    if (plain)
      pair_term(o * o, c->wo_vp, c->wo_vm, c->wo_P, c->wo_M, pS, pA, 1.0, c->Woooo.d(), prec, "Woooo");
    else if (base)
      pair_term(o * o, c->wo_vp, c->wo_vm, c->wo_P, c->wo_M, pS, pA, 0.0, c->rs_WO.d(), P64, "Woooo");
    else
      pair_term(o * o, c->wo_vp, c->wo_vm, c->wo_P, c->wo_M, pS, pA, 1.0, c->Woooo.d(), prec, "Woooo");
  } else {
    big("ijef,mnef->mnij", 1.0, TAU, &dTAU, OOVV, nullptr, nullptr, WO, V(c->rs_WO.d(), {o, o, o, o}), true,
        "Woooo");
  }
  add_sum(c->rs_WO, c->Woooo.d(), n_oooo);
  if (base) {
    keep(c->rf_WO, c->Woooo.p, n_oooo * 8);
    E.refresh(c->rf_WO.p);
  }
  else if (delta) dif(c->rd_WO, c->Woooo.d(), c->rf_WO, n_oooo);

  // Y[n][f][j][b] = 0.5 t2[j][n][f][b] + t1[j][f] t1[n][b]
  if (!dry) launch_build_Y(o, v, t1, t2, c->Y.d(), st);
  rebuilt(c->Y.p);
  const auto YY = Cn(c->Y.d(), {o, v, o, v});
  if (base) keep(c->rf_Y, c->Y.p, n4 * 8);
  else if (delta) dif(c->rd_Y, c->Y.d(), c->rf_Y, n4);
  if (delta) rebuilt(c->rd_Y.p);
  const auto dY = Cn(c->rd_Y.d(), {o, v, o, v});
  // Wmbje stored as WJ[m][e][j][b]
  const auto WJ = V(c->WJ.d(), {o, v, o, v});
  if (!dry) check(hipMemcpyAsync(c->WJ.p, c->WJ0.p, (size_t)n4 * 8, hipMemcpyDeviceToDevice, st), "copy");  // -<mb|je>
  E.contract("jf,mbfe->mejb", -1.0, T1, OVVV, 1.0, WJ, SP, "Wmbje");
  E.contract("nb,mnje->mejb", 1.0, T1, OOOV, 1.0, WJ, SP, "Wmbje");
  const auto RSJ = V(c->rs_WJ.d(), {o, v, o, v});
  big("mnfe,nfjb->mejb", 1.0, OOVV, nullptr, YY, &dY, nullptr, WJ, RSJ, true, "Wmbje");
  add_sum(c->rs_WJ, c->WJ.d(), n4);
  // Wmbej stored as WB[m][e][j][b];  <mb|ej> = <mj|eb>
  const auto WB = V(c->WB.d(), {o, v, o, v});
  if (!dry) check(hipMemcpyAsync(c->WB.p, c->WB0.p, (size_t)n4 * 8, hipMemcpyDeviceToDevice, st), "copy");  // <mj|eb>
  E.contract("jf,mbef->mejb", 1.0, T1, OVVV, 1.0, WB, SP, "Wmbej");
  E.contract("nb,nmje->mejb", -1.0, T1, OOOV, 1.0, WB, SP, "Wmbej");
  const auto RSB = V(c->rs_WB.d(), {o, v, o, v});
  big("mnef,nfjb->mejb", -1.0, OOVV, nullptr, YY, &dY, nullptr, WB, RSB, true, "Wmbej");
  big("mnef,njfb->mejb", 0.5, LOOVV, nullptr, T2, &dT2, nullptr, WB, RSB, false, "Wmbej");
  add_sum(c->rs_WB, c->WB.d(), n4);
  if (base) {
    keep(c->rf_WJ, c->WJ.p, n4 * 8); keep(c->rf_WB, c->WB.p, n4 * 8);
    E.refresh(c->rf_WJ.p); E.refresh(c->rf_WB.p);
  } else if (delta) {
    dif(c->rd_WJ, c->WJ.d(), c->rf_WJ, n4); dif(c->rd_WB, c->WB.d(), c->rf_WB, n4);
  }
  // Zmbij = sum_ef <mb|ef> tau_ijef
  const long n_ovoo = (long)o * v * o * o;
  const auto ZZ = V(c->Z.d(), {o, v, o, o});
  if (c->lad_packed) {
    // Zmbij on pairs: half the flops
    if (base) {
      pair_term(o * v, c->zm_vp, c->zm_vm, c->zm_P, c->zm_M, pS, pA, 0.0, c->rs_Z.d(), P64, "Zmbij");
      if (!dry) launch_axpby(n_ovoo, 1.0, c->rs_Z.d(), 0.0, c->Z.d(), st);
    } else {
      pair_term(o * v, c->zm_vp, c->zm_vm, c->zm_P, c->zm_M, pS, pA, 0.0, c->Z.d(), prec, "Zmbij");
      if (delta) add_sum(c->rs_Z, c->Z.d(), n_ovoo);
    }
  } else {
    if (plain) {
      E.contract("mbef,ijef->mbij", 1.0, OVVV, TAU, 0.0, ZZ, prec, "Zmbij");
    } else if (base) {
      E.contract("mbef,ijef->mbij", 1.0, OVVV, TAU, 0.0, V(c->rs_Z.d(), {o, v, o, o}), P64, "Zmbij");
      if (!dry) launch_axpby(n_ovoo, 1.0, c->rs_Z.d(), 0.0, c->Z.d(), st);
    } else {
      E.contract("mbef,ijef->mbij", 1.0, OVVV, dTAU, 0.0, ZZ, prec, "Zmbij");
      add_sum(c->rs_Z, c->Z.d(), n_ovoo);
    }
  }

  // ---- build_t1, ccwfn.cc:432-469
  const auto RR1 = V(c->R1.d(), {o, v});
  E.permute("ai->ia", 1.0, Cn(c->fvo.d(), {v, o}), 0.0, RR1);
  E.contract("ie,ae->ia", 1.0, T1, FVV, 1.0, RR1, SP, "t1");
  E.contract("ma,mi->ia", -1.0, T1, FOO, 1.0, RR1, SP, "t1");
  E.contract("imae,me->ia", 1.0, UU, FOVI, 1.0, RR1, SP, "t1");
  E.contract("nf,nifa->ia", 2.0, T1, OOVV, 1.0, RR1, SP, "t1");   // 2<na|fi> = 2<ni|fa>
  E.contract("nf,naif->ia", -1.0, T1, OVOV, 1.0, RR1, SP, "t1");  // -<na|if>
  // sum_mef U_mief <ma|ef>, read as U_imfe (U_mief = U_imfe, the pair
  // symmetry of t2): the stored order, so U needs no permute
  big("imfe,maef->ia", 1.0, UU, &dU, OVVV, nullptr, nullptr, RR1, V(c->rs_R1.d(), {o, v}), true, "t1");
  add_sum(c->rs_R1, c->R1.d(), (long)o * v);
  // -sum_mne t2_mnae L<nm|ei>, L<nm|ei> = 2<mn|ie> - <nm|ie> = Looov[m][n][i][e]
  E.contract("mnae,mnie->ia", -1.0, T2, LOOOV, 1.0, RR1, SP, "t1");

  // ---- build_t2 (half X), ccwfn.cc:471-539
  const auto XX = V(c->X.d(), {o, o, v, v});
  const auto RSX = V(c->rs_X.d(), {o, o, v, v});
  E.permute("ijab->ijab", 0.5, OOVV, 0.0, XX);
  // Fae - 0.5 t1.Fme and Fmi + 0.5 t1.Fme fold the t1*Fme terms (:489-491, :496-498)
  const auto FAEP = V(c->FaeP.d(), {v, v}), FMIP = V(c->FmiP.d(), {o, o});
  if (!dry) check(hipMemcpyAsync(c->FaeP.p, c->Fvv.p, (size_t)v * v * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("mb,me->be", -0.5, T1, FOVI, 1.0, FAEP, SP, "t2");
  if (!dry) check(hipMemcpyAsync(c->FmiP.p, c->Foo.p, (size_t)o * o * 8, hipMemcpyDeviceToDevice, st), "copy");
  E.contract("je,me->mj", 0.5, T1, FOVI, 1.0, FMIP, SP, "t2");
  if (base) {
    keep(c->rf_FaeP, c->FaeP.p, (size_t)v * v * 8); keep(c->rf_FmiP, c->FmiP.p, (size_t)o * o * 8);
    E.refresh(c->rf_FaeP.p); E.refresh(c->rf_FmiP.p);
  } else if (delta) {
    dif(c->rd_FaeP, c->FaeP.d(), c->rf_FaeP, (long)v * v); dif(c->rd_FmiP, c->FmiP.d(), c->rf_FmiP, (long)o * o);
  }
  // the base-point operands are constant during the polish: their
  // conversions are cached, and re-converted in place at each base point
  const auto dFAEP = V(c->rd_FaeP.d(), {v, v}), sFAEP = Cn(c->rf_FaeP.d(), {v, v});
  const auto dFMIP = V(c->rd_FmiP.d(), {o, o}), sFMIP = Cn(c->rf_FmiP.d(), {o, o});
  const auto dWO = V(c->rd_WO.d(), {o, o, o, o}), sWO = Cn(c->rf_WO.d(), {o, o, o, o});
  big("ijae,be->ijab", 1.0, T2, &dT2, FAEP, &dFAEP, &sFAEP, XX, RSX, true, "t2_F");
  big("imab,mj->ijab", -1.0, T2, &dT2, FMIP, &dFMIP, &sFMIP, XX, RSX, false, "t2_F");
  big("mnab,mnij->ijab", 0.5, TAU, &dTAU, WO, &dWO, &sWO, XX, RSX, false, "t2_Wmnij");
  if (c->lad_packed) {
    // sum_ef tau_ij^ef <ab|ef> = (L+ + L-)/2 over the pairs i>=j, a>=b, e>=f
    // (tau_ij^ef = tau_ji^fe, <ab|ef> = <ba|fe>): a quarter of the flops
    // S and A (or their differences) were packed from tau above
    const int pr = base ? P64 : prec;
    E.contract("qe,ae->qa", 1.0, Cn(pS, {Q, ldp}), Cn(c->lad_vp.d(), {P, ldp}), 0.0, V(c->lad_LP.d(), {Q, P}), pr,
               "ladder");
    if (Qs > 0 && Ps > 0)
      E.contract("qe,ae->qa", 1.0, Cn(pA, {Qs, ldm}), Cn(c->lad_vm.d(), {Ps, ldm}), 0.0,
                 V(c->lad_LM.d(), {Qs, Ps}), pr, "ladder");
    if (!dry) launch_ladder_unpack(o, v, c->lad_LP.d(), c->lad_LM.d(), 0.25, base ? c->rs_X.d() : c->X.d(), st);
  } else {
    big("ijef,abef->ijab", 0.5, TAU, &dTAU, VVVV, nullptr, nullptr, XX, RSX, false, "ladder");
  }
  E.contract("ie,jeba->ijab", 1.0, T1, OVVV, 1.0, XX, SP, "t2_t1");  // <ab|ej> = <je|ba>
  // Z' = Zmbij + sum_e t1_ie <mb|ej> + <mb|ij>  (the t1 t1 <mb|ej> and
  // t1 <mb|ij> terms of :512-515, :528-535 share Zmbij's t1_ma contraction)
  const auto ZX = V(c->Zx.d(), {o, v, o, o});
  E.permute("mbij->mbij", 1.0, ZZ, 0.0, ZX);
  E.contract("ie,mjeb->mbij", 1.0, T1, OOVV, 1.0, ZX, SP, "t2_t1");
  E.permute("ijmb->mbij", 1.0, OOOV, 1.0, ZX);
  E.contract("ma,mbij->ijab", -1.0, T1, ZX, 1.0, XX, SP, "t2_t1");
  // P2[m][a][j][i] = sum_e t1_ie <ma|je>
  const auto PP2 = V(c->P2.d(), {o, v, o, o});
  E.contract("ie,maje->maji", 1.0, T1, OVOV, 0.0, PP2, SP, "t2_t1");
  E.contract("mb,maji->ijab", -1.0, T1, PP2, 1.0, XX, SP, "t2_t1");
  // ring terms, ccwfn.cc:516-526
  const auto dWB = V(c->rd_WB.d(), {o, v, o, v}), sWB = Cn(c->rf_WB.d(), {o, v, o, v});
  const auto dWJ = V(c->rd_WJ.d(), {o, v, o, v}), sWJ = Cn(c->rf_WJ.d(), {o, v, o, v});
  big("imae,mejb->ijab", 1.0, UU, &dU, WB, &dWB, &sWB, XX, RSX, false, "ring");
  big("imae,mejb->ijab", 1.0, T2, &dT2, WJ, &dWJ, &sWJ, XX, RSX, false, "ring");
  big("mjae,meib->ijab", 1.0, T2, &dT2, WJ, &dWJ, &sWJ, XX, RSX, false, "ring");
  add_sum(c->rs_X, c->X.d(), n4);
  c->have_intermediates = true;
}

// t = t_old + R/D, returns ||R||_2 (ccwfn.cc:555-582)
void drop_graphs(ugx_ctx *c) {
  for (auto &g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
  for (int &x : c->graph_eager) x = 0;
}

// One residual evaluation of the iteration.  With graphs on, the first
// residual of a precision runs eagerly (it sizes the engine's workspaces and
// fills its caches); the next one of each amplitude-buffer parity is
// captured into a hipGraph, which later iterations replay.  The captured
// work is the same kernels with the same arguments, so results are
// bitwise identical to eager execution.
// The base point of the refinement polish always runs eagerly; graphs of the
// other modes are keyed by (amplitude buffer, precision, mode).
void residual_step(ugx_ctx *c, int prec, bool use_graph, int mode = RES_PLAIN) {
  const int key = prec + 4 * (mode == RES_DELTA);
  if (!use_graph || mode == RES_BASE || c->graph_eager[key] == 0) {
    residual(c, prec, mode);
    if (mode != RES_BASE) c->graph_eager[key]++;
    return;
  }
  for (auto &g : c->graphs)
    if (g.t1o == c->t1o.p && g.prec == key) {
      check(hipGraphLaunch(g.exec, c->st), "graph launch");
      c->eng->add_flops(g.flops);
      return;
    }
  Engine &E = *c->eng;
  const bool prof = E.profiling();
  E.set_profiling(false);
  const double f0 = E.flops_total();
  hipGraph_t graph = nullptr;
  check(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    residual(c, prec, mode);
  } catch (...) {
    (void)hipStreamEndCapture(c->st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    E.set_profiling(prof);
    throw;
  }
  check(hipStreamEndCapture(c->st, &graph), "end capture");
  E.set_profiling(prof);
  ugx_ctx::ResidualGraph g;
  g.t1o = c->t1o.p;
  g.prec = key;
  g.flops = E.flops_total() - f0;
  hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(e, "graph instantiate");
  c->graphs.push_back(g);
  // the capture recorded the work without running it
  check(hipGraphLaunch(g.exec, c->st), "graph launch");
}

double increment(ugx_ctx *c, double *R2out) {
  int np = update_nparts(c->no, c->nv);
  launch_update(c->no, c->nv, c->eps.d(), c->R1.d(), c->X.d(), c->t1o.d(), c->t2o.d(), c->t1.d(), c->t2.d(), R2out,
                c->part.d(), np, c->st);
  return std::sqrt(host_sum(c, np));
}

void amp_save(ugx_ctx *c) {
  std::swap(c->t1.p, c->t1o.p);
  std::swap(c->t1.bytes, c->t1o.bytes);
  std::swap(c->t2.p, c->t2o.p);
  std::swap(c->t2.bytes, c->t2o.bytes);
}

// ---------------------------------------------------------------- DIIS
// Pulay DIIS: error = t - t_old over (t1 | t2), B_pq = e_p . e_q with a -1
// border, scaled by max |B_pq|, B c = (0,..,0,-1); t = sum_p c_p t_p.
// Subspace handling follows the oracle (attic/diis.cc:13-161 semantics).
void diis_reset(ugx_ctx *c) {
  c->diis_count = 0;
  c->diis_next = 0;
}

int lin_solve(std::vector<double> &A, std::vector<double> &b, int n) {
  for (int col = 0; col < n; col++) {
    int piv = col;
    for (int r = col + 1; r < n; r++)
      if (std::fabs(A[r * n + col]) > std::fabs(A[piv * n + col])) piv = r;
    if (A[piv * n + col] == 0.0) return -1;
    if (piv != col) {
      for (int k = 0; k < n; k++) std::swap(A[col * n + k], A[piv * n + k]);
      std::swap(b[col], b[piv]);
    }
    for (int r = col + 1; r < n; r++) {
      double f = A[r * n + col] / A[col * n + col];
      for (int k = col; k < n; k++) A[r * n + k] -= f * A[col * n + k];
      b[r] -= f * b[col];
    }
  }
  for (int r = n - 1; r >= 0; r--) {
    double x = b[r];
    for (int k = r + 1; k < n; k++) x -= A[r * n + k] * b[k];
    b[r] = x / A[r * n + r];
  }
  return 0;
}

// Error and amplitude vectors of an nvec-vector subspace over (t1 | t2),
// allocated once per Hamiltonian and subspace size.
void diis_alloc(ugx_ctx *c, int nvec) {
  const long len = c->ov() + c->oovv_n();
  if (c->diis_nvec != nvec || (int)c->diis_err.size() != nvec) {
    c->diis_err.clear();
    c->diis_amp.clear();
    for (int p = 0; p < nvec; p++) {
      c->diis_err.emplace_back(new DBuf());
      c->diis_amp.emplace_back(new DBuf());
    }
    c->diis_nvec = nvec;
    c->diis_B.assign((size_t)nvec * nvec, 0.0);
    c->diis_norm.assign(nvec, 0.0);
  }
  for (int p = 0; p < nvec; p++) {
    c->diis_err[p]->alloc(len * 8);
    c->diis_amp[p]->alloc(len * 8);
  }
  diis_reset(c);
}

// One Pulay step over (x1 | x2) with their previous values (x1o | x2o): the
// CCSD amplitudes or the Lambda amplitudes, which share the subspace storage
// (the Lambda iterations run after the CCSD ones and reset it).
int diis_step_x(ugx_ctx *c, double *x1, const double *x1o, double *x2, const double *x2o, int nvec, int policy,
                int min_vecs) {
  const long n1 = c->ov(), n2 = c->oovv_n(), len = n1 + n2;
  if (c->diis_nvec != nvec || (int)c->diis_err.size() != nvec) diis_alloc(c, nvec);
  int slot;
  if (c->diis_count < nvec) slot = c->diis_count;
  else if (policy == 0) slot = c->diis_next;
  else {
    slot = 0;
    for (int p = 1; p < nvec; p++)
      if (c->diis_norm[p] > c->diis_norm[slot]) slot = p;
  }
  DBuf &e = *c->diis_err[slot];
  DBuf &x = *c->diis_amp[slot];
  launch_diis_store(n1, n2, x1, x1o, x2, x2o, e.d(), x.d(), c->st);
  if (c->diis_count < nvec) c->diis_count++;
  c->diis_next = (slot + 1) % nvec;
  const int n = c->diis_count;
  // dots of the new error with every stored error (slot order)
  std::vector<const double *> vecs(n);
  for (int p = 0; p < n; p++) vecs[p] = c->diis_err[p]->d();
  // about 64 elements per thread, between 64 and kDiisDotBlocks blocks
  const int np = (int)std::min<long>(kDiisDotBlocks, std::max<long>(64, (len + 256 * 64 - 1) / (256 * 64)));
  launch_diis_dots(len, e.d(), vecs.data(), n, c->part.d(), np, c->st);
  const double *h = c->hpin.p;
  check(hipMemcpyAsync(c->hpin.p, c->part.p, (size_t)n * np * 8, hipMemcpyDeviceToHost, c->st), "diis dots copy");
  check(hipStreamSynchronize(c->st), "stream sync");
  for (int p = 0; p < n; p++) {
    double s = 0.0;
    for (int b = 0; b < np; b++) s += h[(size_t)p * np + b];
    c->diis_B[(size_t)slot * nvec + p] = c->diis_B[(size_t)p * nvec + slot] = s;
  }
  c->diis_norm[slot] = c->diis_B[(size_t)slot * nvec + slot];
  if (n < min_vecs) return 0;
  const int m = n + 1;
  std::vector<double> B((size_t)m * m, 0.0), C(m, 0.0);
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) B[p * m + q] = c->diis_B[(size_t)p * nvec + q];
  for (int p = 0; p < n; p++) { B[p * m + n] = -1; B[n * m + p] = -1; }
  double mx = std::fabs(B[0]);
  for (int p = 0; p < n; p++)
    for (int q = 0; q < n; q++) mx = std::max(mx, std::fabs(B[p * m + q]));
  if (mx > 0)
    for (int p = 0; p < n; p++)
      for (int q = 0; q < n; q++) B[p * m + q] /= mx;
  C[n] = -1;
  if (lin_solve(B, C, m) != 0) return 0;
  std::vector<const double *> amps(n);
  for (int p = 0; p < n; p++) amps[p] = c->diis_amp[p]->d();
  launch_diis_combine(n1, n2, amps.data(), C.data(), n, x1, x2, c->st);
  return n;
}

int diis_step(ugx_ctx *c, int nvec, int policy, int min_vecs) {
  return diis_step_x(c, c->t1.d(), c->t1o.d(), c->t2.d(), c->t2o.d(), nvec, policy, min_vecs);
}

// -------------------------------------------------------------- setup
void alloc_work(ugx_ctx *c) {
  const long o = c->no, v = c->nv, ov = o * v, oovv = ov * ov;
  for (DBuf *b : {&c->t1, &c->t1o, &c->R1}) b->alloc(ov * 8);
  for (DBuf *b : {&c->t2, &c->t2o, &c->tau, &c->ttau, &c->U, &c->Y, &c->WJ, &c->WB, &c->X}) b->alloc(oovv * 8);
  c->Fvv.alloc(v * v * 8); c->FaeP.alloc(v * v * 8);
  c->Foo.alloc(o * o * 8); c->FmiP.alloc(o * o * 8);
  c->Fov.alloc(ov * 8);
  c->Woooo.alloc(o * o * o * o * 8);
  c->Z.alloc(o * v * o * o * 8); c->Zx.alloc(o * v * o * o * 8); c->P2.alloc(o * v * o * o * 8);
  // the DIIS dots write one partial per block for each stored vector
  const size_t npart = (size_t)std::max(npart_max(c), kDiisDotBlocks * kDiisMax);
  c->part.alloc(npart * 8);
  // readbacks: the partials, and the energy's partials followed by t1
  c->hpin.alloc(std::max<size_t>(npart, (size_t)energy_nparts(c->no, c->nv) + (size_t)ov));
}

// Everything a ugx_ccsd with options opt needs, allocated and converted
// before its iterations: the DIIS vectors, the refinement buffers, the
// learned guess's output, and the engine's workspaces, cached integral
// conversions and tuned tiles for each residual form the solve evaluates
// (one evaluation of each in the engine's sizing mode, which runs no
// contraction; the streaming kernels of those evaluations write work
// buffers that every real evaluation rewrites).  Done once per Hamiltonian
// and options signature.
void guess_t1_first_order(ugx_ctx *c);

void prepare_solve(ugx_ctx *c, const ugx_options &opt) {
  const int prec = opt.precision;
  const bool refine = prec != PREC_FP64 && opt.polish_mode == UGX_POLISH_REFINE;
  const std::string key = std::to_string(prec) + "/" + std::to_string(refine) + "/" +
                          std::to_string(opt.diis ? opt.diis_nvec : 0) + "/" + std::to_string(opt.guess) + "/" +
                          std::to_string(c->eng->autotune());
  if (c->prepared_for == key) return;
  if (opt.diis) diis_alloc(c, opt.diis_nvec);
  if (opt.guess == UGX_GUESS_MLP) c->G.alloc((size_t)c->oovv_n() * 4);
  if (refine) ensure_refine(c);
  Engine &E = *c->eng;
  E.set_sizing(true);
  try {
    if (opt.guess == UGX_GUESS_MLP) guess_t1_first_order(c);
    residual(c, prec, RES_PLAIN);
    if (prec != PREC_FP64) {
      if (refine) {
        residual(c, PREC_FP64, RES_BASE);
        residual(c, prec, RES_DELTA);
      } else {
        residual(c, PREC_FP64, RES_PLAIN);
      }
    }
  } catch (...) {
    E.set_sizing(false);
    throw;
  }
  E.set_sizing(false);
  check(hipStreamSynchronize(c->st), "stream sync");
  (void)E.take_allocations();
  c->prepared_for = key;
}

void guess_mp2(ugx_ctx *c) {
  launch_guess(c->no, c->nv, c->fock.d(), c->oovv.d(), c->t1.d(), c->t2.d(), c->st);
}

// First-order T1 at the current t2: one Jacobi step of the T1 equation
// from t1 = 0, t1 = R1(t1 = 0, t2) / D1.  With t1 = 0 every t1 term of
// build_t1 (ccwfn.cc:432-469) vanishes and Fme = f_me, which leaves
//   R1_ia = f_ai + sum_me U_imae f_me + sum_mef U_imfe <ma|ef>
//           - sum_mne t2_mnae L<mn|ie>
// (U = 2 t2 - t2 with a and b exchanged): three FP64 contractions,
// 2 o^2 v^3 flops at most.  The learned guess pairs its T2 with this T1.
void guess_t1_first_order(ugx_ctx *c) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const int P64 = PREC_FP64;
  const bool dry = E.sizing();
  if (!dry) launch_build_U(o, v, c->t2.d(), c->U.d(), c->st);
  const auto UU = vref(c->U.d(), {o, o, v, v}), T2 = vref(c->t2.d(), {o, o, v, v});
  const auto RR1 = vref(c->R1.d(), {o, v});
  E.permute("ai->ia", 1.0, cref(c->fvo.d(), {v, o}), 0.0, RR1);
  E.contract("imae,me->ia", 1.0, UU, cref(c->fov.d(), {o, v}), 1.0, RR1, P64, "guess_t1");
  E.contract("imfe,maef->ia", 1.0, UU, cref(c->ovvv.d(), {o, v, v, v}), 1.0, RR1, P64, "guess_t1");
  E.contract("mnae,mnie->ia", -1.0, T2, cref(c->Looov.d(), {o, o, o, v}), 1.0, RR1, P64, "guess_t1");
  if (!dry) launch_t1_divide(o, v, c->eps.d(), c->R1.d(), c->t1.d(), c->st);
}

// MP2 guess, then t2 += y_scale (f(x_ijab) + f(x_jiba)) / 2 with the network
// evaluated by the fused BF16 kernel on features formed from oovv and eps,
// and t1 the first-order T1 at that t2 (guess_t1_first_order)
void guess_mlp(ugx_ctx *c) {
  MlpDevice P;
  double ys = 0.0;
  if (!mlp_inference_view(c->mlp, c->device, &P, &ys))
    throw UgxError(UGX_BAD_INPUT, "UGX_GUESS_MLP: no network on this device (ugx_set_guess_network)");
  guess_mp2(c);
  c->G.alloc((size_t)c->oovv_n() * 4);
  c->eng->timed("guess_mlp", 6.784e4 * (double)c->oovv_n(), [&]() {
    launch_mlp_infer_ham(c->no, c->nv, c->eps.d(), c->oovv.d(), P, c->G.f(), c->st);
  });
  launch_apply_guess(c->no, c->nv, c->G.f(), ys, c->t2.d(), c->st);
  guess_t1_first_order(c);
}

// ----------------------------------------------------------------- HBAR
// HBAR::HBAR (hbar.cc:8-228) from the current amplitudes, in FP64, into the
// context's hb_* blocks (the reference's layouts).  Every term of the
// reference is one engine contraction on the Dirac blocks and L; a term with
// two amplitudes and one integral (t1 t2 <mn|ef>, t1 t1 t1 <mn|ef>) goes
// through one intermediate of at most o^2 v^2 elements:
//   B[m][n][e][i]  = sum_f <mn|ef> t1_if      Bp[m][n][i][f] = sum_e <mn|ef> t1_ie
//   Lt[m][n][i][f] = sum_e L<mn|ef> t1_ie     G[n][e] = sum_mf t1_mf L<mn|fe>
//   C[m][e][b][i]  = sum_nf <mn|ef> t2_in^fb  D[n][e][i][a] = sum_mf <mn|ef> t2_mi^af
//   X[n][e][i][b]  = sum_mf t2_mi^fb L<mn|fe> E[m][b][e][i] = sum_n t1_nb B[m][n][e][i]
//   t1L[m][e] = sum_nf t1_nf L<mn|ef>        Woo[m][n][i][j] = sum_ef <mn|ef> t2_ij^ef
//   Q[m][n][i][j] = sum_f Bp[m][n][i][f] t1_jf
// Hoooo and Hvvvv keep the reference's doubled single-t1 term (hbar.cc:80,
// :93).
void hbar(ugx_ctx *c) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const hipStream_t st = c->st;
  const int P = PREC_FP64;
  const long oo = (long)o * o, ov = (long)o * v, vv = (long)v * v;
  for (auto q : {std::make_pair(&c->hb_ov, ov), std::make_pair(&c->hb_oo, oo), std::make_pair(&c->hb_vv, vv),
                 std::make_pair(&c->hb_oooo, oo * oo), std::make_pair(&c->hb_vvvv, vv * vv),
                 std::make_pair(&c->hb_vovv, ov * vv), std::make_pair(&c->hb_ooov, oo * ov),
                 std::make_pair(&c->hb_ovvo, ov * ov), std::make_pair(&c->hb_ovov, ov * ov),
                 std::make_pair(&c->hb_vvvo, vv * ov), std::make_pair(&c->hb_ovoo, ov * oo),
                 std::make_pair(&c->hb_B, oo * ov), std::make_pair(&c->hb_Bp, oo * ov),
                 std::make_pair(&c->hb_Lt, oo * ov), std::make_pair(&c->hb_C, ov * ov),
                 std::make_pair(&c->hb_D, ov * ov), std::make_pair(&c->hb_X, ov * ov),
                 std::make_pair(&c->hb_E, ov * ov), std::make_pair(&c->hb_G, ov), std::make_pair(&c->hb_t1L, ov),
                 std::make_pair(&c->hb_Woo, oo * oo), std::make_pair(&c->hb_Q, oo * oo)})
    q.first->alloc((size_t)q.second * 8);
  launch_tau(o, v, c->t1.d(), c->t2.d(), c->tau.d(), c->ttau.d(), st);
  auto V = [](const DBuf &b, std::vector<int> d) { return vref(b.d(), std::move(d)); };
  auto Cn = [](const DBuf &b, std::vector<int> d) { return cref(b.d(), std::move(d)); };
  const auto T1 = V(c->t1, {o, v}), T2 = V(c->t2, {o, o, v, v}), TAU = V(c->tau, {o, o, v, v});
  const auto FOV = Cn(c->fov, {o, v}), FOO = Cn(c->foo, {o, o}), FVV = Cn(c->fvv, {v, v});
  const auto OOOO = Cn(c->oooo, {o, o, o, o}), OOOV = Cn(c->ooov, {o, o, o, v}), OOVV = Cn(c->oovv, {o, o, v, v});
  const auto OVOV = Cn(c->ovov, {o, v, o, v}), OVVV = Cn(c->ovvv, {o, v, v, v}), VVVV = Cn(c->vvvv, {v, v, v, v});
  const auto LOOVV = Cn(c->Loovv, {o, o, v, v}), LOOOV = Cn(c->Looov, {o, o, o, v});
  const auto LOVVV = Cn(c->Lovvv, {o, v, v, v});
  auto K = [&](const char *spec, double a, const TRef &A, const TRef &B, double b, const TRef &C) {
    E.contract(spec, a, A, B, b, C, P, "hbar");
  };
  // hbar.cc:38-45, Hov[m][e]
  const auto HOV = V(c->hb_ov, {o, v});
  E.permute("me->me", 1.0, FOV, 0.0, HOV);
  K("nf,mnef->me", 1.0, T1, LOOVV, 1.0, HOV);
  // :47-59, Hoo[m][i]; L<mn|ie> = Looov[m][n][i][e]
  const auto HOO = V(c->hb_oo, {o, o});
  E.permute("mi->mi", 1.0, FOO, 0.0, HOO);
  K("ie,me->mi", 1.0, T1, FOV, 1.0, HOO);
  K("ne,mnie->mi", 1.0, T1, LOOOV, 1.0, HOO);
  K("inef,mnef->mi", 1.0, TAU, LOOVV, 1.0, HOO);
  // :61-73, Hvv[a][e]; L<am|ef> = Lovvv[m][a][f][e]
  const auto HVV = V(c->hb_vv, {v, v});
  E.permute("ae->ae", 1.0, FVV, 0.0, HVV);
  K("ma,me->ae", -1.0, T1, FOV, 1.0, HVV);
  K("mf,mafe->ae", 1.0, T1, LOVVV, 1.0, HVV);
  K("mnfa,mnfe->ae", -1.0, TAU, LOOVV, 1.0, HVV);
  // :75-86, Hoooo[m][n][i][j]
  const auto HOOOO = V(c->hb_oooo, {o, o, o, o});
  E.permute("mnij->mnij", 1.0, OOOO, 0.0, HOOOO);
  K("je,mnie->mnij", 2.0, T1, OOOV, 1.0, HOOOO);
  K("ijef,mnef->mnij", 1.0, TAU, OOVV, 1.0, HOOOO);
  // :88-99, Hvvvv[a][b][e][f]; <am|ef> = ovvv[m][a][f][e]
  const auto HVVVV = V(c->hb_vvvv, {v, v, v, v});
  E.permute("abef->abef", 1.0, VVVV, 0.0, HVVVV);
  K("mb,mafe->abef", -2.0, T1, OVVV, 1.0, HVVVV);
  K("mnab,mnef->abef", 1.0, TAU, OOVV, 1.0, HVVVV);
  // :101-108, Hvovv[a][m][e][f]
  const auto HVOVV = V(c->hb_vovv, {v, o, v, v});
  E.permute("mafe->amef", 1.0, OVVV, 0.0, HVOVV);
  K("na,nmef->amef", -1.0, T1, OOVV, 1.0, HVOVV);
  // :110-117, Hooov[m][n][i][e]
  const auto HOOOV = V(c->hb_ooov, {o, o, o, v});
  E.permute("mnie->mnie", 1.0, OOOV, 0.0, HOOOV);
  K("if,nmef->mnie", 1.0, T1, OOVV, 1.0, HOOOV);
  // :119-134, Hovvo[m][b][e][j]; <mb|ej> = oovv[m][j][e][b], <mn|ej> = ooov[n][m][j][e]
  const auto HOVVO = V(c->hb_ovvo, {o, v, v, o});
  E.permute("mjeb->mbej", 1.0, OOVV, 0.0, HOVVO);
  K("jf,mbef->mbej", 1.0, T1, OVVV, 1.0, HOVVO);
  K("nb,nmje->mbej", -1.0, T1, OOOV, 1.0, HOVVO);
  K("mnef,njbf->mbej", -1.0, OOVV, TAU, 1.0, HOVVO);
  K("njfb,mnef->mbej", 1.0, T2, LOOVV, 1.0, HOVVO);
  // :136-149, Hovov[m][b][j][e]; <bm|ef> = ovvv[m][b][f][e]
  const auto HOVOV = V(c->hb_ovov, {o, v, o, v});
  E.permute("mbje->mbje", 1.0, OVOV, 0.0, HOVOV);
  K("jf,mbfe->mbje", 1.0, T1, OVVV, 1.0, HOVOV);
  K("nb,mnje->mbje", -1.0, T1, OOOV, 1.0, HOVOV);
  K("nmef,jnfb->mbje", -1.0, OOVV, TAU, 1.0, HOVOV);
  // intermediates of the t1 t2 and t1 t1 t1 terms
  const auto B = V(c->hb_B, {o, o, v, o}), Bp = V(c->hb_Bp, {o, o, o, v}), Lt = V(c->hb_Lt, {o, o, o, v});
  const auto CC = V(c->hb_C, {o, v, v, o}), DD = V(c->hb_D, {o, v, o, v}), XX = V(c->hb_X, {o, v, o, v});
  const auto EE = V(c->hb_E, {o, v, v, o}), G = V(c->hb_G, {o, v}), T1L = V(c->hb_t1L, {o, v});
  const auto WOO = V(c->hb_Woo, {o, o, o, o}), QQ = V(c->hb_Q, {o, o, o, o});
  K("mnef,if->mnei", 1.0, OOVV, T1, 0.0, B);
  K("mnef,ie->mnif", 1.0, OOVV, T1, 0.0, Bp);
  K("mnef,ie->mnif", 1.0, LOOVV, T1, 0.0, Lt);
  K("mnef,infb->mebi", 1.0, OOVV, T2, 0.0, CC);
  K("mnef,miaf->neia", 1.0, OOVV, T2, 0.0, DD);
  K("mf,mnfe->ne", 1.0, T1, LOOVV, 0.0, G);
  K("mifb,mnfe->neib", 1.0, T2, LOOVV, 0.0, XX);
  K("nb,mnei->mbei", 1.0, T1, B, 0.0, EE);
  K("nf,mnef->me", 1.0, T1, LOOVV, 0.0, T1L);
  K("mnef,ijef->mnij", 1.0, OOVV, T2, 0.0, WOO);
  K("mnif,jf->mnij", 1.0, Bp, T1, 0.0, QQ);
  // :151-187, Hvvvo[a][b][e][i]; <ab|ei> = ovvv[i][e][b][a], <am|ei> = ovov[m][a][i][e],
  // <bm|ie> = oovv[m][i][e][b], <mn|ei> = ooov[n][m][i][e], L<am|ef> = Lovvv[m][a][f][e]
  const auto HVVVO = V(c->hb_vvvo, {v, v, v, o});
  E.permute("ieba->abei", 1.0, OVVV, 0.0, HVVVO);
  K("if,abef->abei", 1.0, T1, VVVV, 1.0, HVVVO);
  K("mb,maie->abei", -1.0, T1, OVOV, 1.0, HVVVO);
  K("ma,mieb->abei", -1.0, T1, OOVV, 1.0, HVVVO);
  K("imfa,mbef->abei", -1.0, TAU, OVVV, 1.0, HVVVO);
  K("imfb,mafe->abei", -1.0, TAU, OVVV, 1.0, HVVVO);
  K("mnab,nmie->abei", 1.0, TAU, OOOV, 1.0, HVVVO);
  K("me,miab->abei", -1.0, FOV, T2, 1.0, HVVVO);
  K("mifb,mafe->abei", 1.0, T2, LOVVV, 1.0, HVVVO);
  K("mnab,mnei->abei", 1.0, T2, B, 1.0, HVVVO);    // <mn|ef> t1_if t2_mn^ab
  K("ma,mebi->abei", 1.0, T1, CC, 1.0, HVVVO);     // <mn|ef> t1_ma t2_in^fb
  K("nb,neia->abei", 1.0, T1, DD, 1.0, HVVVO);     // <mn|ef> t1_nb t2_mi^af
  K("niab,ne->abei", -1.0, T2, G, 1.0, HVVVO);     // t1_mf t2_ni^ab L<mn|fe>
  K("na,neib->abei", -1.0, T1, XX, 1.0, HVVVO);    // t1_na t2_mi^fb L<mn|fe>
  K("ma,mbei->abei", 1.0, T1, EE, 1.0, HVVVO);     // t1_if t1_ma t1_nb <mn|ef>
  // :189-223, Hovoo[m][b][i][j]; <mb|ij> = ooov[i][j][m][b], <mb|ie> = ovov[m][b][i][e],
  // <bm|je> = oovv[m][j][e][b], <nm|je> = ooov[n][m][j][e], L<mn|ie> = Looov[m][n][i][e]
  const auto HOVOO = V(c->hb_ovoo, {o, v, o, o});
  E.permute("ijmb->mbij", 1.0, OOOV, 0.0, HOVOO);
  K("je,mbie->mbij", 1.0, T1, OVOV, 1.0, HOVOO);
  K("ie,mjeb->mbij", 1.0, T1, OOVV, 1.0, HOVOO);
  K("nb,mnij->mbij", -1.0, T1, OOOO, 1.0, HOVOO);
  K("ineb,nmje->mbij", -1.0, TAU, OOOV, 1.0, HOVOO);
  K("jneb,mnie->mbij", -1.0, TAU, OOOV, 1.0, HOVOO);
  K("mbef,ijef->mbij", 1.0, OVVV, TAU, 1.0, HOVOO);
  K("me,ijeb->mbij", 1.0, FOV, T2, 1.0, HOVOO);
  K("mnie,njeb->mbij", 1.0, LOOOV, T2, 1.0, HOVOO);
  K("ineb,mnej->mbij", -1.0, T2, B, 1.0, HOVOO);   // <mn|ef> t1_jf t2_in^eb
  K("jnfb,mnif->mbij", -1.0, T2, Bp, 1.0, HOVOO);  // <mn|ef> t1_ie t2_jn^fb
  K("nb,mnij->mbij", -1.0, T1, WOO, 1.0, HOVOO);   // <mn|ef> t1_nb t2_ij^ef
  K("njfb,mnif->mbij", 1.0, T2, Lt, 1.0, HOVOO);   // t1_ie t2_nj^fb L<mn|ef>
  K("me,ijeb->mbij", 1.0, T1L, T2, 1.0, HOVOO);    // t1_nf t2_ij^eb L<mn|ef>
  K("nb,mnij->mbij", -1.0, T1, QQ, 1.0, HOVOO);    // t1_jf t1_ie t1_nb <mn|ef>
}

// -------------------------------------------------------------- Lambda
// CCLambda (cclambda.cc:13-289), the CCSD Lagrange multipliers over the HBAR
// of the context's amplitudes, in FP64.  build_G, build_l1 and build_l2 are
// engine contractions on the hb_* blocks; the update l = l_old + R/D, with
// R2 = Z + Z^T and the 2-norm of the raw residual, is the CCSD update kernel;
// DIIS is the CCSD Pulay step on (l1 | l2).  The CCSD_T / DERTYPE=FIRST
// terms s1/s2 (cclambda.cc:184, :235) belong to the (T) gradient, which
// this library does not build.
void lambda_alloc(ugx_ctx *c) {
  const long ov = c->ov(), oovv = c->oovv_n();
  for (DBuf *b : {&c->lam_l1, &c->lam_l1o, &c->lam_R1}) b->alloc(ov * 8);
  for (DBuf *b : {&c->lam_l2, &c->lam_l2o, &c->lam_Z}) b->alloc(oovv * 8);
  c->lam_Goo.alloc((size_t)c->no * c->no * 8);
  c->lam_Gvv.alloc((size_t)c->nv * c->nv * 8);
}

// build_G (cclambda.cc:137-156), build_l1 (:158-210) and build_l2 (:212-268)
// at (l1, l2): the raw l1 residual into R1 and the half residual Z of l2
// (R2_ijab = Z_ijab + Z_jiba), Goo/Gvv into lam_Goo/lam_Gvv.
// with_t: the CCSD(T) sources s1, s2/2 of tgrad() enter the residuals;
// inhom = false drops the inhomogeneous terms 2 Hov and L(oovv) (the left-hand
// perturbed equations apply the same operator to their amplitudes)
void lambda_residual(ugx_ctx *c, const double *l1p, const double *l2p, double *R1p, double *Zp, bool with_t,
                     bool inhom = true) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const int P = PREC_FP64;
  auto V = [](const double *p, std::vector<int> d) { return vref(const_cast<double *>(p), std::move(d)); };
  auto K = [&](const char *spec, double a, const TRef &A, const TRef &B, double b, const TRef &C) {
    E.contract(spec, a, A, B, b, C, P, "lambda");
  };
  const auto L1 = V(l1p, {o, v}), L2 = V(l2p, {o, o, v, v}), T2 = V(c->t2.d(), {o, o, v, v});
  const auto R1 = V(R1p, {o, v}), Z = V(Zp, {o, o, v, v});
  const auto GOO = V(c->lam_Goo.d(), {o, o}), GVV = V(c->lam_Gvv.d(), {v, v});
  const auto HOV = V(c->hb_ov.d(), {o, v}), HOO = V(c->hb_oo.d(), {o, o}), HVV = V(c->hb_vv.d(), {v, v});
  const auto HOOOO = V(c->hb_oooo.d(), {o, o, o, o}), HVVVV = V(c->hb_vvvv.d(), {v, v, v, v});
  const auto HVOVV = V(c->hb_vovv.d(), {v, o, v, v}), HOOOV = V(c->hb_ooov.d(), {o, o, o, v});
  const auto HOVVO = V(c->hb_ovvo.d(), {o, v, v, o}), HOVOV = V(c->hb_ovov.d(), {o, v, o, v});
  const auto HVVVO = V(c->hb_vvvo.d(), {v, v, v, o}), HOVOO = V(c->hb_ovoo.d(), {o, v, o, o});
  const auto LOOVV = cref(c->Loovv.d(), {o, o, v, v});
  // build_G
  K("mjab,ijab->mi", 1.0, T2, L2, 0.0, GOO);
  K("ijeb,ijab->ae", -1.0, T2, L2, 0.0, GVV);
  // build_l1
  E.permute("ia->ia", inhom ? 2.0 : 0.0, HOV, 0.0, R1);
  if (with_t && inhom) E.permute("ia->ia", 1.0, V(c->tg_s1.d(), {o, v}), 1.0, R1);  // cclambda.cc:181
  K("ie,ea->ia", 1.0, L1, HVV, 1.0, R1);
  K("ma,im->ia", -1.0, L1, HOO, 1.0, R1);
  K("me,ieam->ia", 2.0, L1, HOVVO, 1.0, R1);
  K("me,iema->ia", -1.0, L1, HOVOV, 1.0, R1);
  K("imef,efam->ia", 1.0, L2, HVVVO, 1.0, R1);
  K("mnae,iemn->ia", -1.0, L2, HOVOO, 1.0, R1);
  K("ef,eifa->ia", -2.0, GVV, HVOVV, 1.0, R1);
  K("ef,eiaf->ia", 1.0, GVV, HVOVV, 1.0, R1);
  K("mn,mina->ia", -2.0, GOO, HOOOV, 1.0, R1);
  K("mn,imna->ia", 1.0, GOO, HOOOV, 1.0, R1);
  // build_l2, Z before the pair symmetrization
  E.permute("ijab->ijab", inhom ? 1.0 : 0.0, LOOVV, 0.0, Z);
  if (with_t && inhom) E.permute("ijab->ijab", 0.5, V(c->tg_s2.d(), {o, o, v, v}), 1.0, Z);  // cclambda.cc:243
  if (!E.sizing()) launch_lambda_outer(o, v, l1p, c->hb_ov.d(), Zp, c->st);
  K("ijeb,ea->ijab", 1.0, L2, HVV, 1.0, Z);
  K("mjab,im->ijab", -1.0, L2, HOO, 1.0, Z);
  K("mnab,ijmn->ijab", 0.5, L2, HOOOO, 1.0, Z);
  K("ijef,efab->ijab", 0.5, L2, HVVVV, 1.0, Z);
  K("ie,ejab->ijab", 2.0, L1, HVOVV, 1.0, Z);
  K("ie,ejba->ijab", -1.0, L1, HVOVV, 1.0, Z);
  K("mb,jima->ijab", -2.0, L1, HOOOV, 1.0, Z);
  K("mb,ijma->ijab", 1.0, L1, HOOOV, 1.0, Z);
  K("ieam,mjeb->ijab", 2.0, HOVVO, L2, 1.0, Z);
  K("iema,mjeb->ijab", -1.0, HOVOV, L2, 1.0, Z);
  K("jema,mibe->ijab", -1.0, HOVOV, L2, 1.0, Z);
  K("jeam,mieb->ijab", -1.0, HOVVO, L2, 1.0, Z);
  K("ae,ijeb->ijab", 1.0, GVV, LOOVV, 1.0, Z);
  K("mi,mjab->ijab", -1.0, GOO, LOOVV, 1.0, Z);
}

// pseudoenergy (cclambda.cc:270-288): 1/2 sum <ij|ab> l2_ij^ab
double lambda_pseudo(ugx_ctx *c, const double *l2) {
  const int np = 64;
  launch_dot(c->oovv_n(), c->oovv.d(), l2, c->part.d(), np, c->st);
  return 0.5 * host_sum(c, np);
}

// CCLambda::CCLambda + compute_lambda (cclambda.cc:28-42, :93-135): HBAR of
// the current amplitudes, guess l1 = 2 t1, l2 = 2 (2 t2 - t2^T), Jacobi
// updates, break on rms < r_convergence before DIIS, DIIS extrapolating from
// diis_min_vecs stored vectors.  trace (maxiter + 1 values, or null): the
// pseudoenergy of the guess, after each iteration's DIIS, and of the final l2
// last.  Returns whether it converged.
bool lambda_solve(ugx_ctx *c, const ugx_options &opt, double *pseudo, int *niter_out, double *rms_out,
                  double *trace) {
  hbar(c);
  lambda_alloc(c);
  const long ov = c->ov(), oovv = c->oovv_n();
  const int o = c->no, v = c->nv;
  // guess (cclambda.cc:28-42): l1 = 2 t1, l2 = 4 t2 - 2 t2^T
  launch_axpby(ov, 2.0, c->t1.d(), 0.0, c->lam_l1.d(), c->st);
  c->eng->permute("ijba->ijab", -2.0, vref(c->t2.d(), {o, o, v, v}), 0.0, vref(c->lam_l2.d(), {o, o, v, v}));
  launch_axpby(oovv, 4.0, c->t2.d(), 1.0, c->lam_l2.d(), c->st);
  if (trace) trace[0] = lambda_pseudo(c, c->lam_l2.d());
  if (opt.diis) diis_alloc(c, opt.diis_nvec);
  double rms = 0.0;
  int niter = 0;
  for (int iter = 1; iter <= opt.maxiter; iter++) {
    std::swap(c->lam_l1.p, c->lam_l1o.p);
    std::swap(c->lam_l1.bytes, c->lam_l1o.bytes);
    std::swap(c->lam_l2.p, c->lam_l2o.p);
    std::swap(c->lam_l2.bytes, c->lam_l2o.bytes);
    lambda_residual(c, c->lam_l1o.d(), c->lam_l2o.d(), c->lam_R1.d(), c->lam_Z.d(), c->lam_with_t);
    const int np = update_nparts(o, v);
    launch_update(o, v, c->eps.d(), c->lam_R1.d(), c->lam_Z.d(), c->lam_l1o.d(), c->lam_l2o.d(), c->lam_l1.d(),
                  c->lam_l2.d(), nullptr, c->part.d(), np, c->st);
    rms = std::sqrt(host_sum(c, np));
    niter = iter;
    if (opt.print) std::fprintf(stderr, "ugx: Lambda iter %3d  rms %.3e\n", iter, rms);
    if (rms < opt.r_convergence) break;
    if (opt.diis)
      diis_step_x(c, c->lam_l1.d(), c->lam_l1o.d(), c->lam_l2.d(), c->lam_l2o.d(), opt.diis_nvec, opt.diis_policy,
                  opt.diis_min_vecs);
    if (trace) trace[iter] = lambda_pseudo(c, c->lam_l2.d());
  }
  const double pe = lambda_pseudo(c, c->lam_l2.d());
  if (trace) trace[niter] = pe;
  if (opt.diis) diis_reset(c);
  c->have_lambda = true;
  *pseudo = pe;
  *niter_out = niter;
  *rms_out = rms;
  return rms < opt.r_convergence;
}

// ----------------------------------------------------------- densities
// CCDensity::onepdm and twopdm (ccdensity.cc:89-339) from the amplitudes and
// the Lambda amplitudes of the last ugx_lambda, in FP64, into the dn_*
// blocks (the reference's layouts), and the energies they give.  Products of
// three or more amplitudes go through pair intermediates (at most o^2 v^2
// elements each):
//   P[i][k] = sum_mef t2_im^ef l2_km^ef      Q[c][a] = sum_mne l2_nm^ce t2_nm^ae
//   S[i][j][k][m] = sum_ef t2_ij^ef l2_km^ef B[k][m][i][f] = sum_e l2_km^ef t1_ie
//   C[k][m][i][j] = sum_f B[k][m][i][f] t1_jf
//   R[k][e][j][a]  = sum_mf l2_km^ef t2_jm^af   R2[k][e][j][a] = sum_mf l2_km^ef t2_mj^af
//   R3[k][f][i][a] = sum_me l2_km^ef t2_im^ea
//   U[n][c][i][b] = sum_me l2_nm^ce t2_im^be  V[m][c][i][a] = sum_ne l2_nm^ce t2_ni^ae
//   W[n][c][i][b] = sum_me l2_nm^ce t2_mi^be  Y[n][m][c][i] = sum_e l2_nm^ce t1_ie
//   Y2[n][b][c][i] = sum_m t1_mb Y[n][m][c][i]
//   h[j][b] = sum_me (2 t2_jm^be - t2_jm^eb) l1_me   k[i][m] = sum_e t1_ie l1_me
//   k2[e][a] = sum_m l1_me t1_ma
//   A1[i][a][n][f] = sum_me t2_mi^ae l2_mn^ef  A2[j][a][m][f] = sum_ne t2_nj^ae l2_mn^ef
//   A3[i][a][m][e] = sum_nf t2_in^af l2_mn^ef
//   S1[i][j][a][n] = sum_m S[i][j][m][n] t1_ma  B1[a][n][i][f] = sum_m t1_ma B[m][n][i][f]
//   B2[m][a][j][f] = sum_n t1_na B[m][n][j][f]  C1[a][n][i][j] = sum_m t1_ma C[m][n][i][j]
// The (T) parts that CCSD_T adds first (ccdensity.cc:43-74) are not built.
void density(ugx_ctx *c) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const int P = PREC_FP64;
  const long oo = (long)o * o, ov = (long)o * v, vv = (long)v * v;
  for (auto q : {std::make_pair(&c->dn_oo, oo), std::make_pair(&c->dn_vv, vv), std::make_pair(&c->dn_ov, ov),
                 std::make_pair(&c->dn_vo, ov), std::make_pair(&c->dn_oooo, oo * oo),
                 std::make_pair(&c->dn_vvvv, vv * vv), std::make_pair(&c->dn_ooov, oo * ov),
                 std::make_pair(&c->dn_vvvo, vv * ov), std::make_pair(&c->dn_ovov, ov * ov),
                 std::make_pair(&c->dn_oovv, ov * ov), std::make_pair(&c->dx_P, oo), std::make_pair(&c->dx_Q, vv),
                 std::make_pair(&c->dx_h, ov), std::make_pair(&c->dx_k, oo), std::make_pair(&c->dx_k2, vv),
                 std::make_pair(&c->dx_S, oo * oo), std::make_pair(&c->dx_B, oo * ov), std::make_pair(&c->dx_C, oo * oo),
                 std::make_pair(&c->dx_Y, oo * ov), std::make_pair(&c->dx_S1, oo * ov),
                 std::make_pair(&c->dx_C1, oo * ov), std::make_pair(&c->dx_R, ov * ov), std::make_pair(&c->dx_R2, ov * ov),
                 std::make_pair(&c->dx_R3, ov * ov), std::make_pair(&c->dx_U, ov * ov), std::make_pair(&c->dx_V, ov * ov),
                 std::make_pair(&c->dx_W, ov * ov), std::make_pair(&c->dx_Y2, ov * ov),
                 std::make_pair(&c->dx_Ut2, ov * ov), std::make_pair(&c->dx_TU, ov * ov),
                 std::make_pair(&c->dx_A1, ov * ov), std::make_pair(&c->dx_A2, ov * ov),
                 std::make_pair(&c->dx_A3, ov * ov), std::make_pair(&c->dx_B1, ov * ov),
                 std::make_pair(&c->dx_B2, ov * ov), std::make_pair(&c->dx_perm, vv * ov)})
    q.first->alloc((size_t)q.second * 8);
  launch_tau(o, v, c->t1.d(), c->t2.d(), c->tau.d(), c->ttau.d(), c->st);
  auto V = [](const DBuf &b, std::vector<int> d) { return vref(b.d(), std::move(d)); };
  auto K = [&](const char *spec, double a, const TRef &A, const TRef &B, double b, const TRef &C) {
    E.contract(spec, a, A, B, b, C, P, "density");
  };
  const auto T1 = V(c->t1, {o, v}), T2 = V(c->t2, {o, o, v, v}), TAU = V(c->tau, {o, o, v, v});
  const auto L1 = V(c->lam_l1, {o, v}), L2 = V(c->lam_l2, {o, o, v, v});
  const auto DOO = V(c->dn_oo, {o, o}), DVV = V(c->dn_vv, {v, v}), DOV = V(c->dn_ov, {o, v});
  const auto DVO = V(c->dn_vo, {v, o}), GOOOO = V(c->dn_oooo, {o, o, o, o}), GVVVV = V(c->dn_vvvv, {v, v, v, v});
  const auto GOOOV = V(c->dn_ooov, {o, o, o, v}), GVVVO = V(c->dn_vvvo, {v, v, v, o});
  const auto GOVOV = V(c->dn_ovov, {o, v, o, v}), GOOVV = V(c->dn_oovv, {o, o, v, v});
  const auto XP = V(c->dx_P, {o, o}), XQ = V(c->dx_Q, {v, v}), XH = V(c->dx_h, {o, v});
  const auto XK = V(c->dx_k, {o, o}), XK2 = V(c->dx_k2, {v, v}), XS = V(c->dx_S, {o, o, o, o});
  const auto XB = V(c->dx_B, {o, o, o, v}), XC = V(c->dx_C, {o, o, o, o}), XY = V(c->dx_Y, {o, o, v, o});
  const auto XS1 = V(c->dx_S1, {o, o, v, o}), XC1 = V(c->dx_C1, {v, o, o, o});
  const auto XR = V(c->dx_R, {o, v, o, v}), XR2 = V(c->dx_R2, {o, v, o, v}), XR3 = V(c->dx_R3, {o, v, o, v});
  const auto XU = V(c->dx_U, {o, v, o, v}), XV = V(c->dx_V, {o, v, o, v}), XW = V(c->dx_W, {o, v, o, v});
  const auto XY2 = V(c->dx_Y2, {o, v, v, o}), UT2 = V(c->dx_Ut2, {o, o, v, v}), TU = V(c->dx_TU, {o, o, v, v});
  const auto XA1 = V(c->dx_A1, {o, v, o, v}), XA2 = V(c->dx_A2, {o, v, o, v}), XA3 = V(c->dx_A3, {o, v, o, v});
  const auto XB1 = V(c->dx_B1, {v, o, o, v}), XB2 = V(c->dx_B2, {o, v, o, v});
  // intermediates
  K("imef,kmef->ik", 1.0, T2, L2, 0.0, XP);
  K("nmce,nmae->ca", 1.0, L2, T2, 0.0, XQ);
  K("ijef,kmef->ijkm", 1.0, T2, L2, 0.0, XS);
  K("kmef,ie->kmif", 1.0, L2, T1, 0.0, XB);
  K("kmif,jf->kmij", 1.0, XB, T1, 0.0, XC);
  K("kmef,jmaf->keja", 1.0, L2, T2, 0.0, XR);
  K("kmef,mjaf->keja", 1.0, L2, T2, 0.0, XR2);
  K("kmef,imea->kfia", 1.0, L2, T2, 0.0, XR3);
  K("nmce,imbe->ncib", 1.0, L2, T2, 0.0, XU);
  K("nmce,niae->mcia", 1.0, L2, T2, 0.0, XV);
  K("nmce,mibe->ncib", 1.0, L2, T2, 0.0, XW);
  K("nmce,ie->nmci", 1.0, L2, T1, 0.0, XY);
  K("mb,nmci->nbci", 1.0, T1, XY, 0.0, XY2);
  E.permute("jmbe->jmbe", 2.0, T2, 0.0, UT2);
  E.permute("jmeb->jmbe", -1.0, T2, 1.0, UT2);
  E.permute("jmba->jmba", 2.0, TAU, 0.0, TU);
  E.permute("jmab->jmba", -1.0, TAU, 1.0, TU);
  K("jmbe,me->jb", 1.0, UT2, L1, 0.0, XH);
  K("ie,me->im", 1.0, T1, L1, 0.0, XK);
  K("me,ma->ea", 1.0, L1, T1, 0.0, XK2);
  K("miae,mnef->ianf", 1.0, T2, L2, 0.0, XA1);
  K("njae,mnef->jamf", 1.0, T2, L2, 0.0, XA2);
  K("inaf,mnef->iame", 1.0, T2, L2, 0.0, XA3);
  K("ijmn,ma->ijan", 1.0, XS, T1, 0.0, XS1);
  K("ma,mnif->anif", 1.0, T1, XB, 0.0, XB1);
  K("na,mnjf->majf", 1.0, T1, XB, 0.0, XB2);
  K("ma,mnij->anij", 1.0, T1, XC, 0.0, XC1);
  // onepdm (ccdensity.cc:89-137)
  K("ie,je->ij", -1.0, T1, L1, 0.0, DOO);
  K("imef,jmef->ij", -1.0, T2, L2, 1.0, DOO);
  K("mb,ma->ab", 1.0, T1, L1, 0.0, DVV);
  K("mnbe,mnae->ab", 1.0, T2, L2, 1.0, DVV);
  E.permute("ia->ai", 1.0, L1, 0.0, DVO);
  E.permute("ia->ia", 2.0, T1, 0.0, DOV);
  K("me,imae->ia", 2.0, L1, T2, 1.0, DOV);
  K("me,miae->ia", -1.0, L1, TAU, 1.0, DOV);
  K("ma,im->ia", -1.0, T1, XP, 1.0, DOV);
  K("ie,ea->ia", -1.0, T1, XQ, 1.0, DOV);
  // twopdm (ccdensity.cc:149-339)
  K("ijef,klef->ijkl", 1.0, TAU, L2, 0.0, GOOOO);
  K("mnab,mncd->abcd", 1.0, TAU, L2, 0.0, GVVVV);
  // Gooov[i][j][k][a] (:181-200)
  K("ke,ijea->ijka", -2.0, L1, TAU, 0.0, GOOOV);
  K("ke,ijae->ijka", 1.0, L1, TAU, 1.0, GOOOV);
  K("ie,jkae->ijka", -1.0, T1, L2, 1.0, GOOOV);
  K("ja,ik->ijka", -2.0, T1, XP, 1.0, GOOOV);
  K("ie,keja->ijka", -2.0, T1, XR, 1.0, GOOOV);
  K("ia,jk->ijka", 1.0, T1, XP, 1.0, GOOOV);
  K("je,keia->ijka", 1.0, T1, XR, 1.0, GOOOV);
  K("ma,ijkm->ijka", 1.0, T1, XS, 1.0, GOOOV);
  K("ie,keja->ijka", 1.0, T1, XR2, 1.0, GOOOV);
  K("jf,kfia->ijka", 1.0, T1, XR3, 1.0, GOOOV);
  K("ma,kmij->ijka", 1.0, T1, XC, 1.0, GOOOV);
  // Gvvvo[a][b][c][i] (:210-229)
  K("mc,miab->abci", 2.0, L1, TAU, 0.0, GVVVO);
  K("mc,imab->abci", -1.0, L1, TAU, 1.0, GVVVO);
  K("ma,imbc->abci", 1.0, T1, L2, 1.0, GVVVO);
  K("na,ncib->abci", 2.0, T1, XU, 1.0, GVVVO);
  K("ib,ca->abci", 2.0, T1, XQ, 1.0, GVVVO);
  K("nb,ncia->abci", -1.0, T1, XU, 1.0, GVVVO);
  K("ia,cb->abci", -1.0, T1, XQ, 1.0, GVVVO);
  K("mb,mcia->abci", -1.0, T1, XV, 1.0, GVVVO);
  K("na,ncib->abci", -1.0, T1, XW, 1.0, GVVVO);
  K("nmab,nmci->abci", -1.0, T2, XY, 1.0, GVVVO);
  K("na,nbci->abci", -1.0, T1, XY2, 1.0, GVVVO);
  // Govov[i][a][j][b] (:239-246)
  K("ia,jb->iajb", -1.0, T1, L1, 0.0, GOVOV);
  K("mibe,jmea->iajb", -1.0, TAU, L2, 1.0, GOVOV);
  K("imbe,mjea->iajb", -1.0, T2, L2, 1.0, GOVOV);
  // Goovv[i][j][a][b] (:256-299)
  K("ia,jb->ijab", 4.0, T1, L1, 0.0, GOOVV);
  E.permute("ijab->ijab", 4.0, TAU, 1.0, GOOVV);
  E.permute("ijba->ijab", -2.0, TAU, 1.0, GOOVV);
  E.permute("ijab->ijab", 1.0, L2, 1.0, GOOVV);
  K("ia,jb->ijab", 4.0, T1, XH, 1.0, GOOVV);
  K("ib,ja->ijab", -2.0, T1, XH, 1.0, GOOVV);
  K("im,jmba->ijab", -2.0, XK, TU, 1.0, GOOVV);
  K("ea,ijeb->ijab", -2.0, XK2, UT2, 1.0, GOOVV);
  K("imae,mjeb->ijab", 4.0, T2, L2, 1.0, GOOVV);
  K("mjbe,imae->ijab", -2.0, TAU, L2, 1.0, GOOVV);
  K("mnab,ijmn->ijab", 1.0, T2, XS, 1.0, GOOVV);
  K("ianf,njbf->ijab", 1.0, XA1, T2, 1.0, GOOVV);
  K("jamf,imfb->ijab", 1.0, XA2, T2, 1.0, GOOVV);
  K("ijae,eb->ijab", -4.0, TAU, XQ, 1.0, GOOVV);
  K("imab,jm->ijab", -4.0, TAU, XP, 1.0, GOOVV);
  K("mjbe,iame->ijab", -4.0, TAU, XA3, 1.0, GOOVV);
  K("imae,jbme->ijab", 4.0, T2, XA3, 1.0, GOOVV);
  K("ijbe,ea->ijab", 2.0, TAU, XQ, 1.0, GOOVV);
  K("imba,jm->ijab", 2.0, TAU, XP, 1.0, GOOVV);
  K("mjae,ibme->ijab", 2.0, TAU, XA3, 1.0, GOOVV);
  K("imbe,jame->ijab", -2.0, T2, XA3, 1.0, GOOVV);
  K("ijan,nb->ijab", 1.0, XS1, T1, 1.0, GOOVV);
  K("mnab,mnij->ijab", 1.0, T2, XC, 1.0, GOOVV);
  K("anif,njbf->ijab", 1.0, XB1, T2, 1.0, GOOVV);
  K("miae,bmje->ijab", 1.0, T2, XB1, 1.0, GOOVV);
  K("imfb,majf->ijab", 1.0, T2, XB2, 1.0, GOOVV);
  K("njae,nbie->ijab", 1.0, T2, XB2, 1.0, GOOVV);
  K("nb,anij->ijab", 1.0, T1, XC1, 1.0, GOOVV);
  if (c->lam_with_t) {
    // the (T) parts CCSD_T copies in first (ccdensity.cc:43-74)
    E.permute("ij->ij", 1.0, V(c->tg_oo, {o, o}), 1.0, DOO);
    E.permute("ab->ab", 1.0, V(c->tg_vv, {v, v}), 1.0, DVV);
    E.permute("ia->ia", 1.0, V(c->tg_ov, {o, v}), 1.0, DOV);
    E.permute("ijka->ijka", 1.0, V(c->tg_ooov, {o, o, o, v}), 1.0, GOOOV);
    E.permute("abci->abci", 1.0, V(c->tg_vvvo, {v, v, v, o}), 1.0, GVVVO);
    E.permute("ijab->ijab", 1.0, V(c->tg_oovv, {o, o, v, v}), 1.0, GOOVV);
  }
}

// CCWfn::tgrad (ccwfn.cc:854-1098): the (T) gradient terms of CCSD(T) from
// the context's amplitudes, in FP64 on the device.  The reference holds t3
// and l3 whole (o^3 v^3 each); here they exist one occupied slice at a time,
// T[j][k][a][b][c] = t3[p][j][k][a][b][c] and the same for l3, and every
// term is rewritten with the slice index first through the symmetry of t3
// and l3 under simultaneous permutations of their (occupied, virtual) pairs,
// t3[x][y][z][a][b][c] = t3[y][x][z][b][a][c] = ...  Each slice is
//   T = W3 / D with W3 = the twelve terms of ccwfn.cc:664-689 (i = p),
//   L = (sum over the six pair permutations of the l3_tmp numerator,
//        ccwfn.cc:923-941) / D  (D is invariant under them),
// then adds its share of s1 (:874-881), of the T3 and L3 sources of s2
// (:884-915, :958-983) and of the (T) density parts (:985-1083).  Every
// term is one FP64 engine contraction; the peak extra memory is three
// o^2 v^3 arrays (T, L and the engine's permute workspace).
void tgrad(ugx_ctx *c) {
  Engine &E = *c->eng;
  const int o = c->no, v = c->nv;
  const int PR = PREC_FP64;
  const long ov = (long)o * v, oovv = ov * ov, v3 = (long)v * v * v, oov = (long)o * o * v;
  const long s5 = (long)o * o * v3;
  c->tg_s1.alloc(ov * 8); c->tg_s2.alloc(oovv * 8); c->tg_oo.alloc((size_t)o * o * 8);
  c->tg_vv.alloc((size_t)v * v * 8); c->tg_ov.alloc(ov * 8); c->tg_ooov.alloc((size_t)o * oov * 8);
  c->tg_vvvo.alloc((size_t)o * v3 * 8); c->tg_oovv.alloc(oovv * 8);
  c->tg_T.alloc(s5 * 8); c->tg_L.alloc(s5 * 8);
  for (DBuf *b : {&c->tg_Y2a, &c->tg_Y2b, &c->tg_Y2l, &c->tg_X, &c->tg_t2s}) b->alloc(oovv * 8);
  c->tg_Gt.alloc((size_t)o * v3 * 8); c->tg_t1s.alloc(ov * 8);
  c->tg_ooovT.alloc((size_t)o * oov * 8); c->tg_LooovT.alloc((size_t)o * oov * 8);
  auto V = [](const DBuf &b, std::vector<int> d) { return vref(b.d(), std::move(d)); };
  auto VP = [](const double *p, std::vector<int> d) { return vref(const_cast<double *>(p), std::move(d)); };
  auto K = [&](const std::string &spec, double a, const TRef &A, const TRef &B, double b, const TRef &C) {
    E.contract(spec, a, A, B, b, C, PR, "tgrad");
  };
  // t1s = 2 t1, t2s = 4 t2 - 2 t2^T (build_tstar, ccwfn.cc:628-645)
  const auto T1 = V(c->t1, {o, v}), T2 = V(c->t2, {o, o, v, v});
  const auto T1S = V(c->tg_t1s, {o, v}), T2S = V(c->tg_t2s, {o, o, v, v});
  E.permute("ia->ia", 2.0, T1, 0.0, T1S);
  E.permute("ijab->ijab", 4.0, T2, 0.0, T2S);
  E.permute("ijba->ijab", -2.0, T2, 1.0, T2S);
  // ooovT[z][y][n][a] = ooov[y][z][n][a] = <zy|an>; LooovT[x][y][n][a] =
  // Looov[y][x][n][a] = L_xyan: the slice index of these terms first
  E.permute("yzna->zyna", 1.0, cref(c->ooov.d(), {o, o, o, v}), 0.0, V(c->tg_ooovT, {o, o, o, v}));
  E.permute("yxna->xyna", 1.0, cref(c->Looov.d(), {o, o, o, v}), 0.0, V(c->tg_LooovT, {o, o, o, v}));
  for (DBuf *b : {&c->tg_s1, &c->tg_oo, &c->tg_vv, &c->tg_ov, &c->tg_ooov, &c->tg_oovv, &c->tg_Y2a, &c->tg_Y2b,
                  &c->tg_Y2l, &c->tg_Gt})
    if (!E.sizing()) launch_fill(b->bytes / 8, 0.0, b->d(), c->st);
  const auto OVVV = cref(c->ovvv.d(), {o, v, v, v}), OOOV = cref(c->ooov.d(), {o, o, o, v});
  const auto LOOVV = cref(c->Loovv.d(), {o, o, v, v}), LOVVV = cref(c->Lovvv.d(), {o, v, v, v});
  const auto LOOOV = cref(c->Looov.d(), {o, o, o, v}), FOV = cref(c->fov.d(), {o, v});
  const auto OOOVT = cref(c->tg_ooovT.d(), {o, o, o, v}), LOOOVT = cref(c->tg_LooovT.d(), {o, o, o, v});
  const auto TS = V(c->tg_T, {o, o, v, v, v}), LS = V(c->tg_L, {o, o, v, v, v});
  const auto S1 = V(c->tg_s1, {o, v}), Y2A = V(c->tg_Y2a, {o, o, v, v}), DOO = V(c->tg_oo, {o, o});
  const auto DVV = V(c->tg_vv, {v, v}), DOV = V(c->tg_ov, {o, v});
  struct Opd {
    TRef r;
    std::string l;
  };
  for (int p = 0; p < o; p++) {
    // slices of the operands at occupied index p
    const auto ovvv_p = VP(c->ovvv.d() + p * v3, {v, v, v}), lovvv_p = VP(c->Lovvv.d() + p * v3, {v, v, v});
    const auto ooov_p = VP(c->ooov.d() + p * oov, {o, o, v}), ooovT_p = VP(c->tg_ooovT.d() + p * oov, {o, o, v});
    const auto looov_p = VP(c->Looov.d() + p * oov, {o, o, v});
    const auto looovT_p = VP(c->tg_LooovT.d() + p * oov, {o, o, v});
    const auto t2_p = VP(c->t2.d() + p * ov * v, {o, v, v}), t2s_p = VP(c->tg_t2s.d() + p * ov * v, {o, v, v});
    const auto loovv_p = VP(c->Loovv.d() + p * ov * v, {o, v, v});
    const auto fov_p = VP(c->fov.d() + (long)p * v, {v}), t1s_p = VP(c->tg_t1s.d() + (long)p * v, {v});
    // ---- T = W3[p][j][k][a][b][c] / D (ccwfn.cc:664-689 with i = p)
    K("eab,kjce->jkabc", 1.0, ovvv_p, T2, 0.0, TS);
    K("eac,jkbe->jkabc", 1.0, ovvv_p, T2, 1.0, TS);
    K("keca,jeb->jkabc", 1.0, OVVV, t2_p, 1.0, TS);
    K("kecb,jae->jkabc", 1.0, OVVV, t2_p, 1.0, TS);
    K("jebc,kae->jkabc", 1.0, OVVV, t2_p, 1.0, TS);
    K("jeba,kec->jkabc", 1.0, OVVV, t2_p, 1.0, TS);
    K("jkmc,mab->jkabc", -1.0, OOOV, t2_p, 1.0, TS);
    K("kjmb,mac->jkabc", -1.0, OOOV, t2_p, 1.0, TS);
    K("jmb,kmca->jkabc", -1.0, ooov_p, T2, 1.0, TS);
    K("jma,kmcb->jkabc", -1.0, ooovT_p, T2, 1.0, TS);
    K("kma,jmbc->jkabc", -1.0, ooovT_p, T2, 1.0, TS);
    K("kmc,jmba->jkabc", -1.0, ooov_p, T2, 1.0, TS);
    if (!E.sizing()) launch_div_d3(o, v, p, c->eps.d(), c->tg_T.d(), c->st);
    // ---- L = l3[p][j][k][a][b][c]: the l3_tmp numerator N[x][y][z][al][be][ga]
    // (ccwfn.cc:923-941) summed over the six pair permutations (:947-952),
    // each operand taken with the slice index P first
    auto str = [](std::initializer_list<char> cs) { return std::string(cs); };
    auto pairsym = [&](const TRef &full, const TRef &slice, char y, char z, char b1, char b2) -> Opd {
      if (y == 'P') return {slice, str({z, b1, b2})};
      if (z == 'P') return {slice, str({y, b2, b1})};
      return {full, str({y, z, b1, b2})};
    };
    auto row = [&](const TRef &full, const TRef &slice, char x, char a1) -> Opd {
      if (x == 'P') return {slice, str({a1})};
      return {full, str({x, a1})};
    };
    auto t2s_ = [&](char y, char z, char b1, char b2) { return pairsym(T2S, t2s_p, y, z, b1, b2); };
    auto loovv_ = [&](char x, char y, char a1, char b1) { return pairsym(LOOVV, loovv_p, x, y, a1, b1); };
    auto fov_ = [&](char x, char a1) { return row(FOV, fov_p, x, a1); };
    auto t1s_ = [&](char z, char g) { return row(T1S, t1s_p, z, g); };
    // L_fyab = Lovvv[y][f][b][a];  <zf|ab> = ovvv[z][f][a][b]
    auto lvovv_ = [&](char f, char y, char a1, char b1) -> Opd {
      if (y == 'P') return {lovvv_p, str({f, b1, a1})};
      return {LOVVV, str({y, f, b1, a1})};
    };
    auto ovvv_ = [&](char z, char f, char a1, char b1) -> Opd {
      if (z == 'P') return {ovvv_p, str({f, a1, b1})};
      return {OVVV, str({z, f, a1, b1})};
    };
    // L_xyan = Looov[y][x][n][a];  <zy|an> = ooov[y][z][n][a]
    auto loovo_ = [&](char x, char y, char a1, char n) -> Opd {
      if (y == 'P') return {looov_p, str({x, n, a1})};
      if (x == 'P') return {looovT_p, str({y, n, a1})};
      return {LOOOV, str({y, x, n, a1})};
    };
    auto oovo_ = [&](char z, char y, char a1, char n) -> Opd {
      if (y == 'P') return {ooov_p, str({z, n, a1})};
      if (z == 'P') return {ooovT_p, str({y, n, a1})};
      return {OOOV, str({y, z, n, a1})};
    };
    bool first = true;
    auto term = [&](double coef, const Opd &A, const Opd &B) {
      K(A.l + "," + B.l + "->jkabc", coef, A.r, B.r, first ? 0.0 : 1.0, LS);
      first = false;
    };
    const char sig[6][6] = {{'P', 'j', 'k', 'a', 'b', 'c'}, {'P', 'k', 'j', 'a', 'c', 'b'},
                            {'j', 'P', 'k', 'b', 'a', 'c'}, {'j', 'k', 'P', 'b', 'c', 'a'},
                            {'k', 'P', 'j', 'c', 'a', 'b'}, {'k', 'j', 'P', 'c', 'b', 'a'}};
    for (const auto &g : sig) {
      const char x = g[0], y = g[1], z = g[2], al = g[3], be = g[4], ga = g[5];
      term(1.0, fov_(x, al), t2s_(y, z, be, ga));
      term(-1.0, fov_(y, al), t2s_(x, z, be, ga));
      term(1.0, loovv_(x, y, al, be), t1s_(z, ga));
      term(-1.0, loovv_(x, y, al, ga), t1s_(z, be));
      term(1.0, lvovv_('f', y, al, be), t2s_(x, z, 'f', ga));
      term(-1.0, ovvv_(z, 'f', al, be), t2s_(x, y, ga, 'f'));
      term(-1.0, loovo_(x, y, al, 'n'), t2s_('n', z, be, ga));
      term(1.0, oovo_(z, y, al, 'n'), t2s_('n', x, be, ga));
    }
    if (!E.sizing()) launch_div_d3(o, v, p, c->eps.d(), c->tg_L.d(), c->st);
    // ---- s1 (ccwfn.cc:874-881): i = p, then k = p
    K("klacd,klcd->a", 2.0, TS, LOOVV, 1.0, VP(c->tg_s1.d() + (long)p * v, {v}));
    K("ilacd,lcd->ia", -2.0, TS, loovv_p, 1.0, S1);
    // ---- Y2 of the T3 sources (ccwfn.cc:884-904): terms with i = p into
    // Y2a[i][j][a][b], with k = p into Y2a, with j = p into Y2b[j][i][a][b]
    const auto y2a_p = VP(c->tg_Y2a.d() + p * ov * v, {o, v, v}), y2b_p = VP(c->tg_Y2b.d() + p * ov * v, {o, v, v});
    K("jkabc,kc->jab", 1.0, TS, FOV, 1.0, y2a_p);
    K("jiabc,c->ijab", -1.0, TS, fov_p, 1.0, Y2A);
    K("klbac,klic->iab", -2.0, TS, OOOV, 1.0, y2b_p);
    K("lkbac,klic->iab", 1.0, TS, OOOV, 1.0, y2b_p);
    K("lkcba,klic->iab", 1.0, TS, OOOV, 1.0, y2b_p);
    K("ikbcd,kadc->iab", 2.0, TS, OVVV, 1.0, y2b_p);
    K("kibcd,kadc->iab", -1.0, TS, OVVV, 1.0, y2b_p);
    K("kidbc,kadc->iab", -1.0, TS, OVVV, 1.0, y2b_p);
    // ---- Y2 of the L3 sources (ccwfn.cc:958-977), i = p
    const auto y2l_p = VP(c->tg_Y2l.d() + p * ov * v, {o, v, v});
    K("klacb,lkjc->jab", -1.0, LS, OOOV, 1.0, y2l_p);
    K("kjacd,kbcd->jab", 1.0, LS, OVVV, 1.0, y2l_p);
    // ---- (T) density parts (ccwfn.cc:985-1083)
    K("imedf,jmedf->ij", -0.5, TS, LS, 1.0, DOO);
    K("mnbde,mnade->ab", 0.5, TS, LS, 1.0, DVV);
    K("niefa,nef->ia", 1.0, TS, t2s_p, 1.0, DOV);
    K("inefa,nef->ia", -1.0, TS, t2s_p, 1.0, DOV);
    const auto g3_p = VP(c->tg_ooov.d() + p * oov, {o, o, v});
    K("jmeaf,kmef->jka", -2.0, TS, T2S, 1.0, g3_p);
    K("jmaef,kmef->jka", 1.0, TS, T2S, 1.0, g3_p);
    K("mjeaf,kmef->jka", 1.0, TS, T2S, 1.0, g3_p);
    K("mjefa,kmef->jka", -1.0, LS, T2, 1.0, g3_p);
    const auto gt_p = VP(c->tg_Gt.d() + p * v3, {v, v, v});
    K("nmbae,mnec->abc", 2.0, TS, T2S, 1.0, gt_p);
    K("nmabe,mnec->abc", -1.0, TS, T2S, 1.0, gt_p);
    K("nmeab,mnec->abc", -1.0, TS, T2S, 1.0, gt_p);
    K("nmbae,mnec->abc", 1.0, LS, T2, 1.0, gt_p);
    const auto goovv_p = VP(c->tg_oovv.d() + p * ov * v, {o, v, v});
    K("jmabe,me->jab", 8.0, TS, T1, 1.0, goovv_p);
    K("jmaeb,me->jab", -8.0, TS, T1, 1.0, goovv_p);
    K("jmbae,me->jab", -4.0, TS, T1, 1.0, goovv_p);
    K("jmbea,me->jab", 4.0, TS, T1, 1.0, goovv_p);
  }
  // Y2 = Y2a + Y2b^T(ij); X2 = Y2 + Y2^T(pair); s2 = 4 X2 - 2 X2(ab swapped)
  // (ccwfn.cc:906-915), then + Y2l + Y2l^T(pair) (:979-983)
  const auto Y2B = V(c->tg_Y2b, {o, o, v, v}), Y2L = V(c->tg_Y2l, {o, o, v, v}), X2 = V(c->tg_X, {o, o, v, v});
  const auto S2 = V(c->tg_s2, {o, o, v, v});
  E.permute("jiab->ijab", 1.0, Y2B, 1.0, Y2A);
  E.permute("ijab->ijab", 1.0, Y2A, 0.0, X2);
  E.permute("jiba->ijab", 1.0, Y2A, 1.0, X2);
  E.permute("ijab->ijab", 4.0, X2, 0.0, S2);
  E.permute("ijba->ijab", -2.0, X2, 1.0, S2);
  E.permute("ijab->ijab", 1.0, Y2L, 1.0, S2);
  E.permute("jiba->ijab", 1.0, Y2L, 1.0, S2);
  // Gvvvo[a][b][c][i] from the [i][a][b][c] accumulator
  E.permute("iabc->abci", 1.0, V(c->tg_Gt, {o, v, v, v}), 0.0, V(c->tg_vvvo, {v, v, v, o}));
}

// sum x.y on the device, summed on the host
double device_dot(ugx_ctx *c, long n, const double *x, const double *y) {
  const int np = 64;
  launch_dot(n, x, y, c->part.d(), np, c->st);
  return host_sum(c, np);
}

// The energies of the density blocks (ccdensity.cc:138-146 and the six
// parts of twopdm): E[0] = E_one, E[1] = E_two, E[2..7] = the OOOO, VVVV,
// OOOV, VVVO, OVOV, OOVV parts.  <ab|ci> = <ia|bc> = ovvv[i][a][b][c].
void density_energies(ugx_ctx *c, double *E) {
  const int o = c->no, v = c->nv;
  const long oo = (long)o * o, ov = (long)o * v, vv = (long)v * v;
  E[0] = device_dot(c, oo, c->foo.d(), c->dn_oo.d()) + device_dot(c, vv, c->fvv.d(), c->dn_vv.d());
  E[2] = 0.5 * device_dot(c, oo * oo, c->oooo.d(), c->dn_oooo.d());
  E[3] = 0.5 * device_dot(c, vv * vv, c->vvvv.d(), c->dn_vvvv.d());
  E[4] = device_dot(c, oo * ov, c->ooov.d(), c->dn_ooov.d());
  c->eng->permute("abci->iabc", 1.0, vref(c->dn_vvvo.d(), {v, v, v, o}), 0.0, vref(c->dx_perm.d(), {o, v, v, v}));
  E[5] = device_dot(c, vv * ov, c->ovvv.d(), c->dx_perm.d());
  E[6] = device_dot(c, ov * ov, c->ovov.d(), c->dn_ovov.d());
  E[7] = 0.5 * device_dot(c, ov * ov, c->oovv.d(), c->dn_oovv.d());
  E[1] = E[2] + E[3] + E[4] + E[5] + E[6] + E[7];
}

// ------------------------------------------------------ linear response
// CCPert (ccpert.cc) and CCResp::linresp (ccresp.cc:27-124) in FP64 over the
// HBAR and Lambda amplitudes of the last ugx_lambda.  Each perturbation slot
// holds the MO perturbation's blocks, its similarity transform (pertbar,
// ccpert.cc:112-176), the right-hand (X) and left-hand (Y) perturbed
// amplitudes and their previous values.  The terms of pertbar, of the X
// residual (build_X1/X2, :382-470) and of the part of the Y residual that
// does not depend on Y (the terms of build_Y1/Y2 in the right-hand
// amplitudes, the Lambda amplitudes and the perturbation, the reference's (i)
// .. (x), :516-882) are tables of einsum contractions; the Y-dependent part
// is the Lambda operator (build_G/l1/l2 without their inhomogeneous terms),
// so each Y iteration costs one Lambda residual.  Index letters fix the
// extents: i j k l m n o occupied, a b c d e f g virtual.
struct PTerm {
  const char *spec;
  double alpha;
  const char *a, *b;  // b null: a permute of a into c
  double beta;
  const char *c;
};

// pertbar, ccpert.cc:112-176 (Aov is the perturbation's ov block itself)
const PTerm kPertBar[] = {
    {"mi->mi", 1.0, "Poo", nullptr, 0.0, "Aoo"},
    {"ie,me->mi", 1.0, "t1", "Pov", 1.0, "Aoo"},
    {"ae->ae", 1.0, "Pvv", nullptr, 0.0, "Avv"},
    {"ma,me->ae", -1.0, "t1", "Pov", 1.0, "Avv"},
    {"ai->ai", 1.0, "Pvo", nullptr, 0.0, "Avo"},
    {"ie,ae->ai", 1.0, "t1", "Pvv", 1.0, "Avo"},
    {"ma,mi->ai", -1.0, "t1", "Poo", 1.0, "Avo"},
    {"miea,me->ai", 2.0, "t2", "Pov", 1.0, "Avo"},
    {"imea,me->ai", -1.0, "t2", "Pov", 1.0, "Avo"},
    {"ma,me->ae", 1.0, "t1", "Pov", 0.0, "Qae"},
    {"ie,ae->ai", -1.0, "t1", "Qae", 1.0, "Avo"},
    {"ijeb,me->mbij", 1.0, "t2", "Pov", 0.0, "Aovoo"},
    {"miab,me->abei", -1.0, "t2", "Pov", 0.0, "Avvvo"},
    {"ijeb,ae->abij", 1.0, "t2", "Avv", 0.0, "Avvoo"},
    {"mjab,mi->abij", -1.0, "t2", "Aoo", 1.0, "Avvoo"},
};

// build_X1 / build_X2 (ccpert.cc:382-470) at (X1o, X2o) without the omega
// terms: R1 and the half residual Z (R2 = Z + Z^T)
const PTerm kPertX[] = {
    {"ai->ia", 1.0, "Avo", nullptr, 0.0, "R1"},
    {"ie,ae->ia", 1.0, "X1o", "Hvv", 1.0, "R1"},
    {"ma,mi->ia", -1.0, "X1o", "Hoo", 1.0, "R1"},
    {"me,maei->ia", 2.0, "X1o", "Hovvo", 1.0, "R1"},
    {"me,maie->ia", -1.0, "X1o", "Hovov", 1.0, "R1"},
    {"me,miea->ia", 2.0, "Hov", "X2o", 1.0, "R1"},
    {"me,imea->ia", -1.0, "Hov", "X2o", 1.0, "R1"},
    {"imef,amef->ia", 2.0, "X2o", "Hvovv", 1.0, "R1"},
    {"imef,amfe->ia", -1.0, "X2o", "Hvovv", 1.0, "R1"},
    {"mnae,mnie->ia", -2.0, "X2o", "Hooov", 1.0, "R1"},
    {"mnae,nmie->ia", 1.0, "X2o", "Hooov", 1.0, "R1"},
    {"amef,mf->ae", 2.0, "Hvovv", "X1o", 0.0, "Zvv"},
    {"amfe,mf->ae", -1.0, "Hvovv", "X1o", 1.0, "Zvv"},
    {"mnef,mnaf->ae", -1.0, "Loovv", "X2o", 1.0, "Zvv"},
    {"mnie,ne->mi", -2.0, "Hooov", "X1o", 0.0, "Zoo"},
    {"nmie,ne->mi", 1.0, "Hooov", "X1o", 1.0, "Zoo"},
    {"mnef,inef->mi", -1.0, "Loovv", "X2o", 1.0, "Zoo"},
    {"abij->ijab", 1.0, "Avvoo", nullptr, 0.0, "Z"},
    {"ie,abej->ijab", 1.0, "X1o", "Hvvvo", 1.0, "Z"},
    {"ma,mbij->ijab", -1.0, "X1o", "Hovoo", 1.0, "Z"},
    {"mi,mjab->ijab", 1.0, "Zoo", "t2", 1.0, "Z"},
    {"ae,ijeb->ijab", 1.0, "Zvv", "t2", 1.0, "Z"},
    {"ijeb,ae->ijab", 1.0, "X2o", "Hvv", 1.0, "Z"},
    {"mjab,mi->ijab", -1.0, "X2o", "Hoo", 1.0, "Z"},
    {"mnab,mnij->ijab", 0.5, "X2o", "Hoooo", 1.0, "Z"},
    {"ijef,abef->ijab", 0.5, "X2o", "Hvvvv", 1.0, "Z"},
    {"imeb,maje->ijab", -1.0, "X2o", "Hovov", 1.0, "Z"},
    {"imea,mbej->ijab", -1.0, "X2o", "Hovvo", 1.0, "Z"},
    {"miea,mbej->ijab", 2.0, "X2o", "Hovvo", 1.0, "Z"},
    {"miea,mbje->ijab", -1.0, "X2o", "Hovov", 1.0, "Z"},
};

// The Y-independent terms of build_Y1 (into S1) and build_Y2 (into the half
// S2), ccpert.cc:516-882, from the converged X1/X2, l1/l2, the perturbation
// and HBAR; products of three or four factors go through the intermediates
// Q*, P*, T*, U*, X* (at most o v^3 elements)
const PTerm kPertYSource[] = {
    {"ia->ia", 2.0, "Pov", nullptr, 0.0, "S1"},
    // (i)
    {"me,imae->ia", 2.0, "X1", "Loovv", 1.0, "S1"},
    // (m), (n)
    {"ma,im->ia", -1.0, "l1", "Aoo", 1.0, "S1"},
    {"ie,ea->ia", 1.0, "l1", "Avv", 1.0, "S1"},
    {"imfe,feam->ia", 1.0, "l2", "Avvvo", 1.0, "S1"},
    {"mnea,ienm->ia", -0.5, "l2", "Aovoo", 1.0, "S1"},
    {"mnae,iemn->ia", -0.5, "l2", "Aovoo", 1.0, "S1"},
    // (q)
    {"me,ma->ea", 1.0, "X1", "Hov", 0.0, "Q1"},
    {"ie,ea->ia", -1.0, "l1", "Q1", 1.0, "S1"},
    {"me,ie->mi", 1.0, "X1", "Hov", 0.0, "Q2"},
    {"ma,mi->ia", -1.0, "l1", "Q2", 1.0, "S1"},
    {"me,ne->mn", 1.0, "X1", "l1", 0.0, "Q3"},
    {"mn,mina->ia", -2.0, "Q3", "Hooov", 1.0, "S1"},
    {"mn,imna->ia", 1.0, "Q3", "Hooov", 1.0, "S1"},
    {"me,imne->in", 2.0, "X1", "Hooov", 0.0, "Q4"},
    {"me,mine->in", -1.0, "X1", "Hooov", 1.0, "Q4"},
    {"in,na->ia", -1.0, "Q4", "l1", 1.0, "S1"},
    {"me,fmae->fa", 2.0, "X1", "Hvovv", 0.0, "Q5"},
    {"me,fmea->fa", -1.0, "X1", "Hvovv", 1.0, "Q5"},
    {"if,fa->ia", 1.0, "l1", "Q5", 1.0, "S1"},
    {"me,mf->ef", 1.0, "X1", "l1", 0.0, "Q6"},
    {"ef,fiea->ia", 2.0, "Q6", "Hvovv", 1.0, "S1"},
    {"ef,fiae->ia", -1.0, "Q6", "Hvovv", 1.0, "S1"},
    // (r)
    {"mnef,nf->me", 1.0, "X2", "l1", 0.0, "Q7"},
    {"me,imae->ia", 2.0, "Q7", "Loovv", 1.0, "S1"},
    {"mnef,ne->mf", 1.0, "X2", "l1", 0.0, "Q8"},
    {"mf,imaf->ia", -1.0, "Q8", "Loovv", 1.0, "S1"},
    {"mnef,mief->ni", 1.0, "X2", "Loovv", 0.0, "Q9"},
    {"ni,na->ia", -1.0, "Q9", "l1", 1.0, "S1"},
    {"mnef,nmfa->ea", 1.0, "X2", "Loovv", 0.0, "Q10"},
    {"ie,ea->ia", -1.0, "l1", "Q10", 1.0, "S1"},
    // (s)
    {"me,mfna->efna", 1.0, "X1", "Hovov", 0.0, "Tvvov"},
    {"efna,nief->ia", -1.0, "Tvvov", "l2", 1.0, "S1"},
    {"me,ifne->mifn", 1.0, "X1", "Hovov", 0.0, "Tooo1"},
    {"mifn,nmaf->ia", -1.0, "Tooo1", "l2", 1.0, "S1"},
    {"me,mfan->efan", 1.0, "X1", "Hovvo", 0.0, "Tvvvo"},
    {"efan,nife->ia", -1.0, "Tvvvo", "l2", 1.0, "S1"},
    {"me,ifen->mifn", 1.0, "X1", "Hovvo", 0.0, "Tooo1"},
    {"mifn,nmfa->ia", -1.0, "Tooo1", "l2", 1.0, "S1"},
    {"imfg,fgae->imae", 1.0, "l2", "Hvvvv", 0.0, "T2"},
    {"imae,me->ia", 0.5, "T2", "X1", 1.0, "S1"},
    {"mifg,fgea->miea", 1.0, "l2", "Hvvvv", 0.0, "T2"},
    {"miea,me->ia", 0.5, "T2", "X1", 1.0, "S1"},
    {"imno,noae->imae", 1.0, "Hoooo", "l2", 0.0, "T2"},
    {"imae,me->ia", 0.5, "T2", "X1", 1.0, "S1"},
    {"mino,noea->miea", 1.0, "Hoooo", "l2", 0.0, "T2"},
    {"miea,me->ia", 0.5, "T2", "X1", 1.0, "S1"},
    // the four-factor terms, ccpert.cc:629-640
    {"jmfe,jmbe->fb", 1.0, "t2", "l2", 0.0, "P1"},
    {"nb,fb->nf", 1.0, "X1", "P1", 0.0, "P2"},
    {"nf,inaf->ia", -1.0, "P2", "Loovv", 1.0, "S1"},
    {"jnfb,jnab->fa", 1.0, "t2", "l2", 0.0, "P3"},
    {"me,mief->if", 1.0, "X1", "Loovv", 0.0, "P4"},
    {"if,fa->ia", -1.0, "P4", "P3", 1.0, "S1"},
    {"jnfb,jifb->ni", 1.0, "t2", "l2", 0.0, "P5"},
    {"me,mnea->na", 1.0, "X1", "Loovv", 0.0, "P6"},
    {"ni,na->ia", -1.0, "P5", "P6", 1.0, "S1"},
    {"nmeb,jmeb->nj", 1.0, "t2", "l2", 0.0, "P7"},
    {"nj,jf->nf", 1.0, "P7", "X1", 0.0, "P8"},
    {"nf,inaf->ia", -1.0, "P8", "Loovv", 1.0, "S1"},
    // (t)
    {"mnef,nife->mi", 1.0, "X2", "l2", 0.0, "Q14"},
    {"mi,ma->ia", -1.0, "Q14", "Hov", 1.0, "S1"},
    {"mnef,nmfa->ea", 1.0, "X2", "l2", 0.0, "Q15"},
    {"ie,ea->ia", -1.0, "Hov", "Q15", 1.0, "S1"},
    {"mnef,imfg->ineg", 1.0, "X2", "l2", 0.0, "T2"},
    {"ineg,gnea->ia", -1.0, "T2", "Hvovv", 1.0, "S1"},
    {"mnef,mifg->ineg", 1.0, "X2", "l2", 0.0, "T2"},
    {"ineg,gnae->ia", -1.0, "T2", "Hvovv", 1.0, "S1"},
    {"mnef,gief->mngi", 1.0, "X2", "Hvovv", 0.0, "Tooo1"},
    {"mngi,nmag->ia", -1.0, "Tooo1", "l2", 1.0, "S1"},
    {"mnef,nifg->meig", 1.0, "X2", "l2", 0.0, "T2"},
    {"meig,gmae->ia", 2.0, "T2", "Hvovv", 1.0, "S1"},
    {"meig,gmea->ia", -1.0, "T2", "Hvovv", 1.0, "S1"},
    {"mnef,nmfg->eg", 1.0, "X2", "l2", 0.0, "Q20"},
    {"eg,giea->ia", 2.0, "Q20", "Hvovv", 1.0, "S1"},
    {"eg,giae->ia", -1.0, "Q20", "Hvovv", 1.0, "S1"},
    {"mnef,oief->mnoi", 1.0, "X2", "l2", 0.0, "Toooo"},
    {"mnoi,mnoa->ia", 1.0, "Toooo", "Hooov", 1.0, "S1"},
    {"mnef,omaf->neoa", 1.0, "X2", "l2", 0.0, "T2"},
    {"neoa,inoe->ia", 1.0, "T2", "Hooov", 1.0, "S1"},
    {"mnef,onea->mfoa", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfoa,miof->ia", 1.0, "T2", "Hooov", 1.0, "S1"},
    {"mnef,nofe->mo", 1.0, "X2", "l2", 0.0, "Q24"},
    {"mo,mioa->ia", -2.0, "Q24", "Hooov", 1.0, "S1"},
    {"mo,imoa->ia", 1.0, "Q24", "Hooov", 1.0, "S1"},
    {"mnef,nofa->meoa", 1.0, "X2", "l2", 0.0, "T2"},
    {"meoa,imoe->ia", -2.0, "T2", "Hooov", 1.0, "S1"},
    {"meoa,mioe->ia", 1.0, "T2", "Hooov", 1.0, "S1"},
    // build_Y2: (o), (p)
    {"jb,ia->ijab", 2.0, "l1", "Pov", 0.0, "S2"},
    {"ib,ja->ijab", -1.0, "l1", "Pov", 1.0, "S2"},
    {"ijeb,ea->ijab", 1.0, "l2", "Avv", 1.0, "S2"},
    {"mjab,im->ijab", -1.0, "l2", "Aoo", 1.0, "S2"},
    // (u)
    {"me,mieb->ib", 1.0, "X1", "Loovv", 0.0, "U1"},
    {"ja,ib->ijab", -1.0, "l1", "U1", 1.0, "S2"},
    {"ijae,eb->ijab", -1.0, "Loovv", "Q6", 1.0, "S2"},
    {"me,ie->im", 1.0, "X1", "l1", 0.0, "U3"},
    {"im,jmba->ijab", -1.0, "U3", "Loovv", 1.0, "S2"},
    {"me,imae->ia", 1.0, "X1", "Loovv", 0.0, "U4"},
    {"jb,ia->ijab", 2.0, "l1", "U4", 1.0, "S2"},
    // (w)
    {"jibe,ea->ijab", -1.0, "l2", "Q1", 1.0, "S2"},
    {"mi,jmba->ijab", -1.0, "Q2", "l2", 1.0, "S2"},
    {"jife,me->jifm", 1.0, "l2", "X1", 0.0, "Tooo1"},
    {"jifm,fmba->ijab", -1.0, "Tooo1", "Hvovv", 1.0, "S2"},
    {"me,mifb->eifb", 1.0, "X1", "l2", 0.0, "Tvovv"},
    {"eifb,fjea->ijab", -1.0, "Tvovv", "Hvovv", 1.0, "S2"},
    {"me,jmfa->jefa", 1.0, "X1", "l2", 0.0, "Tvovv"},
    {"jefa,fibe->ijab", -1.0, "Tvovv", "Hvovv", 1.0, "S2"},
    {"jibf,fa->ijab", 1.0, "l2", "Q5", 1.0, "S2"},
    {"me,jmbf->jebf", 1.0, "X1", "l2", 0.0, "Tvovv"},
    {"jebf,fiea->ijab", 2.0, "Tvovv", "Hvovv", 1.0, "S2"},
    {"jebf,fiae->ijab", -1.0, "Tvovv", "Hvovv", 1.0, "S2"},
    {"nibe,me->nibm", 1.0, "l2", "X1", 0.0, "Tooo1"},
    {"nibm,jmna->ijab", 1.0, "Tooo1", "Hooov", 1.0, "S2"},
    {"nieb,me->nibm", 1.0, "l2", "X1", 0.0, "Tooo1"},
    {"nibm,mjna->ijab", 1.0, "Tooo1", "Hooov", 1.0, "S2"},
    {"jine,me->jinm", 1.0, "Hooov", "X1", 0.0, "Toooo"},
    {"jinm,nmba->ijab", 1.0, "Toooo", "l2", 1.0, "S2"},
    {"jnbe,me->jnbm", 1.0, "l2", "X1", 0.0, "Tooo1"},
    {"jnbm,mina->ijab", -2.0, "Tooo1", "Hooov", 1.0, "S2"},
    {"jnbm,imna->ijab", 1.0, "Tooo1", "Hooov", 1.0, "S2"},
    {"in,jnba->ijab", -1.0, "Q4", "l2", 1.0, "S2"},
    // (x)
    {"mnef,jief->mnji", 1.0, "X2", "l2", 0.0, "Toooo"},
    {"mnji,nmab->ijab", 0.25, "Toooo", "oovv", 1.0, "S2"},
    {"mnef,jife->mnji", 1.0, "X2", "l2", 0.0, "Toooo"},
    {"mnji,mnab->ijab", 0.25, "Toooo", "oovv", 1.0, "S2"},
    {"mnef,mifb->neib", 1.0, "X2", "l2", 0.0, "T2"},
    {"neib,jnae->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,nieb->mfib", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfib,jmaf->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,mibf->neib", 1.0, "X2", "l2", 0.0, "T2"},
    {"neib,njae->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,nibe->mfib", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfib,mjaf->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,imfb->neib", 1.0, "X2", "l2", 0.0, "T2"},
    {"neib,jnea->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,ineb->mfib", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfib,jmfa->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,imbf->neib", 1.0, "X2", "l2", 0.0, "T2"},
    {"neib,njea->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"mnef,inbe->mfib", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfib,mjfa->ijab", 0.25, "T2", "oovv", 1.0, "S2"},
    {"jief,mnef->jimn", 1.0, "oovv", "X2", 0.0, "Toooo"},
    {"jimn,nmab->ijab", 0.25, "Toooo", "l2", 1.0, "S2"},
    {"jife,mnef->jimn", 1.0, "oovv", "X2", 0.0, "Toooo"},
    {"jimn,mnab->ijab", 0.25, "Toooo", "l2", 1.0, "S2"},
    {"mnef,jnbe->mfjb", 1.0, "X2", "l2", 0.0, "T2"},
    {"mfjb,imaf->ijab", -0.5, "T2", "Loovv", 1.0, "S2"},
    {"mnef,jmbf->nejb", 1.0, "X2", "l2", 0.0, "T2"},
    {"nejb,inae->ijab", -0.5, "T2", "Loovv", 1.0, "S2"},
    {"mief,mnef->in", 1.0, "Loovv", "X2", 0.0, "Xo"},
    {"in,jnba->ijab", -0.5, "Xo", "l2", 1.0, "S2"},
    {"mnea,mnef->fa", 1.0, "Loovv", "X2", 0.0, "Xp"},
    {"jibf,fa->ijab", -0.5, "l2", "Xp", 1.0, "S2"},
    {"jibe,ea->ijab", -0.5, "l2", "Q10", 1.0, "S2"},
    {"nife,mnef->im", 1.0, "Loovv", "X2", 0.0, "Xr"},
    {"im,jmba->ijab", -0.5, "Xr", "l2", 1.0, "S2"},
    {"ijae,eb->ijab", -1.0, "Loovv", "Q15", 1.0, "S2"},
    {"imab,mj->ijab", -1.0, "Loovv", "Q14", 1.0, "S2"},
    {"mnef,nifb->meib", 1.0, "X2", "l2", 0.0, "T2"},
    {"meib,mjea->ijab", -1.0, "T2", "Loovv", 1.0, "S2"},
    {"mejb,imae->ijab", 2.0, "T2", "Loovv", 1.0, "S2"},
};

// extent of an index letter
inline int letter_extent(const ugx_ctx *c, char ch) {
  if (std::strchr("ijklmno", ch)) return c->no;
  if (std::strchr("abcdefgh", ch)) return c->nv;
  throw std::runtime_error(std::string("ugx: index letter with no extent: ") + ch);
}

std::vector<int> letter_dims(const ugx_ctx *c, const std::string &l) {
  std::vector<int> d;
  for (char ch : l) d.push_back(letter_extent(c, ch));
  return d;
}

// The operands of a slot's tables that are not scratch: fixed names, then
// the slot's aliases and buffers; null for a scratch intermediate
double *pert_operand(ugx_ctx *c, PertSlot &P, const std::string &n) {
  static const char *const fixed[] = {"t1",    "t2",    "l1",    "l2",    "Loovv", "oovv",  "Hov",
                                      "Hoo",   "Hvv",   "Hoooo", "Hvvvv", "Hvovv", "Hooov", "Hovvo",
                                      "Hovov", "Hvvvo", "Hovoo", "R1",    "Z",     "S1",    "S2"};
  DBuf *const bufs[] = {&c->t1,      &c->t2,      &c->lam_l1,  &c->lam_l2,  &c->Loovv,   &c->oovv,    &c->hb_ov,
                        &c->hb_oo,   &c->hb_vv,   &c->hb_oooo, &c->hb_vvvv, &c->hb_vovv, &c->hb_ooov, &c->hb_ovvo,
                        &c->hb_ovov, &c->hb_vvvo, &c->hb_ovoo, &c->rp_R1,   &c->rp_Z,    &c->rp_S1,   &c->rp_S2};
  static_assert(sizeof(fixed) / sizeof(fixed[0]) == sizeof(bufs) / sizeof(bufs[0]), "one buffer per name");
  for (size_t k = 0; k < sizeof(fixed) / sizeof(fixed[0]); k++)
    if (n == fixed[k]) return bufs[k]->d();
  auto at = P.aliases.find(n);
  if (at != P.aliases.end()) return const_cast<double *>(at->second);
  auto it = P.bufs.find(n);
  if (it != P.bufs.end()) return it->second->d();
  return nullptr;
}

// Every operand of a slot's tables: pert_operand, then the context's scratch
// intermediates
double *pert_name(ugx_ctx *c, PertSlot &P, const std::string &n) {
  if (double *p = pert_operand(c, P, n)) return p;
  auto jt = c->rp_tmp.find(n);
  if (jt != c->rp_tmp.end()) return jt->second->d();
  throw std::runtime_error("ugx: response operand without a buffer: " + n);
}

// Parses "A,B->C" / "A->C" into the letters of each operand.
void split_spec(const std::string &s, std::string &a, std::string &b, std::string &cc) {
  const size_t arrow = s.find("->"), comma = s.find(',');
  cc = s.substr(arrow + 2);
  if (comma == std::string::npos || comma > arrow) {
    a = s.substr(0, arrow);
    b.clear();
  } else {
    a = s.substr(0, comma);
    b = s.substr(comma + 1, arrow - comma - 1);
  }
}

// Allocates every scratch intermediate a table names (its largest use), so a
// table's run allocates nothing.
void size_terms(ugx_ctx *c, PertSlot &P, const PTerm *t, int n) {
  auto need = [&](const char *name, const std::string &letters) {
    // a scratch buffer from an earlier, smaller problem is resized here too
    if (!name || pert_operand(c, P, name)) return;
    long sz = 1;
    for (int d : letter_dims(c, letters)) sz *= d;
    auto &slot = c->rp_tmp_size[name];
    slot = std::max(slot, sz);
  };
  for (int k = 0; k < n; k++) {
    std::string a, b, cc;
    split_spec(t[k].spec, a, b, cc);
    need(t[k].c, cc);
  }
  for (auto &kv : c->rp_tmp_size) {
    auto &buf = c->rp_tmp[kv.first];
    if (!buf) buf.reset(new DBuf());
    buf->alloc((size_t)kv.second * 8);
  }
}

void run_terms(ugx_ctx *c, PertSlot &P, const PTerm *t, int n) {
  Engine &E = *c->eng;
  size_terms(c, P, t, n);
  for (int k = 0; k < n; k++) {
    std::string a, b, cc;
    split_spec(t[k].spec, a, b, cc);
    const TRef A = vref(pert_name(c, P, t[k].a), letter_dims(c, a));
    const TRef C = vref(pert_name(c, P, t[k].c), letter_dims(c, cc));
    if (!t[k].b) {
      E.permute(a + "->" + cc, t[k].alpha, A, t[k].beta, C);
    } else {
      const TRef B = vref(pert_name(c, P, t[k].b), letter_dims(c, b));
      E.contract(t[k].spec, t[k].alpha, A, B, t[k].beta, C, PREC_FP64, "response");
    }
  }
}

void pert_alloc_scratch(ugx_ctx *c) {
  const long ov = c->ov(), oovv = c->oovv_n();
  for (DBuf *b : {&c->rp_R1, &c->rp_S1}) b->alloc(ov * 8);
  for (DBuf *b : {&c->rp_Z, &c->rp_S2, &c->rp_W}) b->alloc(oovv * 8);
  c->rp_eps.alloc((size_t)(c->no + c->nv) * 8);
}

// The HBAR diagonals as "orbital energies": the perturbed denominators are
// Hoo_ii - Hvv_aa (+ ...) (ccpert.cc:44-56)
void pert_denominators(ugx_ctx *c) {
  const int o = c->no, v = c->nv;
  check(hipMemcpy2DAsync(c->rp_eps.p, 8, c->hb_oo.p, (size_t)(o + 1) * 8, 8, o, hipMemcpyDeviceToDevice, c->st),
        "Hoo diagonal");
  check(hipMemcpy2DAsync(c->rp_eps.d() + o, 8, c->hb_vv.p, (size_t)(v + 1) * 8, 8, v, hipMemcpyDeviceToDevice,
                         c->st),
        "Hvv diagonal");
}

// build_X1/X2 (hand 0) or build_G/Y1/Y2 (hand 1) at (A1, A2) into rp_R1 and
// the half residual rp_Z; the Y source (rp_S1, rp_S2) must be current
void pert_residual(ugx_ctx *c, PertSlot &P, int hand, const double *A1, const double *A2) {
  const long ov = c->ov(), oovv = c->oovv_n();
  if (hand == 0) {
    P.alias("X1o", A1);
    P.alias("X2o", A2);
    run_terms(c, P, kPertX, (int)(sizeof(kPertX) / sizeof(kPertX[0])));
    launch_axpby(ov, -P.omega, A1, 1.0, c->rp_R1.d(), c->st);
    launch_axpby(oovv, -0.5 * P.omega, A2, 1.0, c->rp_Z.d(), c->st);
  } else {
    lambda_residual(c, A1, A2, c->rp_R1.d(), c->rp_Z.d(), false, false);
    launch_axpby(ov, P.omega, A1, 1.0, c->rp_R1.d(), c->st);
    launch_axpby(ov, 1.0, c->rp_S1.d(), 1.0, c->rp_R1.d(), c->st);
    launch_axpby(oovv, 0.5 * P.omega, A2, 1.0, c->rp_Z.d(), c->st);
    launch_axpby(oovv, 1.0, c->rp_S2.d(), 1.0, c->rp_Z.d(), c->st);
  }
}

void pert_y_source(ugx_ctx *c, PertSlot &P) {
  run_terms(c, P, kPertYSource, (int)(sizeof(kPertYSource) / sizeof(kPertYSource[0])));
}

// CCPert::pseudoresponse, ccpert.cc:229-249
double pert_pseudo(ugx_ctx *c, PertSlot &P, int hand) {
  const int o = c->no, v = c->nv;
  Engine &E = *c->eng;
  const double *X1 = hand == 0 ? P.buf("X1") : P.buf("Y1"), *X2 = hand == 0 ? P.buf("X2") : P.buf("Y2");
  const auto W = vref(c->rp_W.d(), {o, o, v, v});
  E.permute("abij->ijab", 1.0, vref(P.buf("Avvoo"), {v, v, o, o}), 0.0, W);
  E.permute("ai->ia", 1.0, vref(P.buf("Avo"), {v, o}), 0.0, vref(c->rp_R1.d(), {o, v}));
  // 2 (2 X2 - X2^T(ab)) . Avvoo = 4 X2.W - 2 X2.W^T(ab)
  E.permute("ijba->ijab", 1.0, vref(const_cast<double *>(X2), {o, o, v, v}), 0.0, vref(c->rp_Z.d(), {o, o, v, v}));
  const double p1 = 2.0 * device_dot(c, c->ov(), X1, c->rp_R1.d());
  const double p2 = 4.0 * device_dot(c, c->oovv_n(), X2, c->rp_W.d()) - 2.0 * device_dot(c, c->oovv_n(), c->rp_Z.d(), c->rp_W.d());
  return -2.0 * (p1 + p2);
}

// CCPert::solve, ccpert.cc:178-225, with the library's DIIS on (x1 | x2)
bool pert_solve(ugx_ctx *c, PertSlot &P, int hand, const ugx_options &opt, double *pseudo, int *niter_out,
                double *rms_out) {
  const int o = c->no, v = c->nv;
  pert_alloc_scratch(c);
  pert_denominators(c);
  if (hand == 1) pert_y_source(c, P);
  const char *n1 = hand == 0 ? "X1" : "Y1", *n2 = hand == 0 ? "X2" : "Y2";
  const char *n1o = hand == 0 ? "X1old" : "Y1old", *n2o = hand == 0 ? "X2old" : "Y2old";
  if (opt.diis) diis_alloc(c, opt.diis_nvec);
  double rms = 0.0;
  int niter = 0;
  for (int iter = 1; iter <= opt.maxiter; iter++) {
    P.swap(n1, n1o);
    P.swap(n2, n2o);
    pert_residual(c, P, hand, P.buf(n1o), P.buf(n2o));
    const int np = update_nparts(o, v);
    launch_update_shift(o, v, c->rp_eps.d(), P.omega, c->rp_R1.d(), c->rp_Z.d(), P.buf(n1o), P.buf(n2o), P.buf(n1),
                        P.buf(n2), c->part.d(), np, c->st);
    rms = std::sqrt(host_sum(c, np));
    niter = iter;
    if (opt.print) std::fprintf(stderr, "ugx: perturbed (%s) iter %3d  rms %.3e\n", hand ? "left" : "right", iter, rms);
    if (rms < opt.r_convergence) break;
    if (opt.diis)
      diis_step_x(c, P.buf(n1), P.buf(n1o), P.buf(n2), P.buf(n2o), opt.diis_nvec, opt.diis_policy, opt.diis_min_vecs);
  }
  if (opt.diis) diis_reset(c);
  *pseudo = pert_pseudo(c, P, hand);
  *niter_out = niter;
  *rms_out = rms;
  return rms < opt.r_convergence;
}

// CCResp::linresp, ccresp.cc:27-124: A from slot Px, amplitudes from Py
double pert_linresp(ugx_ctx *c, PertSlot &Px, PertSlot &Py) {
  const int o = c->no, v = c->nv;
  const long ov = c->ov(), oovv = c->oovv_n();
  Engine &E = *c->eng;
  auto V = [](const double *p, std::vector<int> d) { return vref(const_cast<double *>(p), std::move(d)); };
  auto K = [&](const char *spec, double a, const TRef &A, const TRef &B, double b, const TRef &C) {
    E.contract(spec, a, A, B, b, C, PREC_FP64, "response");
  };
  pert_alloc_scratch(c);
  const auto X1 = V(Py.buf("X1"), {o, v}), X2 = V(Py.buf("X2"), {o, o, v, v});
  const auto L2 = V(c->lam_l2.d(), {o, o, v, v});
  const auto R1 = V(c->rp_R1.d(), {o, v}), W = V(c->rp_W.d(), {o, o, v, v}), Z = V(c->rp_Z.d(), {o, o, v, v});
  const auto AVO = V(Px.buf("Avo"), {v, o}), AVVOO = V(Px.buf("Avvoo"), {v, v, o, o});
  const auto AOO = V(Px.buf("Aoo"), {o, o}), AVV = V(Px.buf("Avv"), {v, v}), POV = V(Px.buf("Pov"), {o, v});
  const auto AVVVO = V(Px.buf("Avvvo"), {v, v, v, o}), AOVOO = V(Px.buf("Aovoo"), {o, v, o, o});
  // polar1 = Avo . Y1 + 1/2 (Avvoo + Avvoo^T) . Y2
  E.permute("ai->ia", 1.0, AVO, 0.0, R1);
  double polar1 = device_dot(c, ov, c->rp_R1.d(), Py.buf("Y1"));
  E.permute("abij->ijab", 0.5, AVVOO, 0.0, W);
  E.permute("baji->ijab", 0.5, AVVOO, 1.0, W);
  polar1 += device_dot(c, oovv, c->rp_W.d(), Py.buf("Y2"));
  // polar2, the l1 terms: l1 . [ (2 X2 - X2^T) . Pov + Avv.X1 - Aoo.X1 ]
  E.permute("ijab->ijab", 2.0, X2, 0.0, W);
  E.permute("ijba->ijab", -1.0, X2, 1.0, W);
  K("ijab,jb->ia", 1.0, W, POV, 0.0, R1);
  K("ac,ic->ia", 1.0, AVV, X1, 1.0, R1);
  K("ki,ka->ia", -1.0, AOO, X1, 1.0, R1);
  double polar2 = device_dot(c, ov, c->rp_R1.d(), c->lam_l1.d());
  // X1 . (l2 . Avvvo)
  K("ijbc,bcaj->ia", 1.0, L2, AVVVO, 0.0, R1);
  polar2 += device_dot(c, ov, c->rp_R1.d(), Py.buf("X1"));
  // the l2 terms: l2 . Z
  K("ka,kbij->ijab", -0.5, X1, AOVOO, 0.0, Z);
  K("kb,kaji->ijab", -0.5, X1, AOVOO, 1.0, Z);
  K("kjab,ki->ijab", -0.5, X2, AOO, 1.0, Z);
  K("kiba,kj->ijab", -0.5, X2, AOO, 1.0, Z);
  K("ijac,bc->ijab", 0.5, X2, AVV, 1.0, Z);
  K("jibc,ac->ijab", 0.5, X2, AVV, 1.0, Z);
  polar2 += device_dot(c, oovv, c->rp_Z.d(), c->lam_l2.d());
  polar2 += 2.0 * device_dot(c, ov, Px.buf("Pov"), Py.buf("X1"));
  return -1.0 * (polar1 + polar2);
}

// ------------------------------------------------------------------ (T)
// W3 of a triplet (ccwfn.cc:1396-1413) is the sum over the six permutations
// of the (ia, jb, kc) pairs of
//   X(p,q,r)[x][y][z] = sum_e <pe|xy> t2_rq^ze - sum_m t2_pm^xy <qr|mz>
//   W = X(ijk)[abc] + X(ikj)[acb] + X(kij)[cab] + X(kji)[cba] + X(jki)[bca]
//       + X(jik)[bac],
// each an MFMA GEMM with K = v (particle) + o (hole).
void triples_prepare(ugx_ctx *c, int prec) {
  const int o = c->no, v = c->nv;
  const long v3 = (long)v * v * v;
  if (!c->tables_ready) {
    // P[p][x][y][e] = <pe|xy> = ovvv[p][e][x][y]
    c->tP.alloc((size_t)o * v3 * 8);
    c->eng->permute("pexy->pxye", 1.0, cref(c->ovvv.d(), {o, v, v, v}), 0.0, vref(c->tP.d(), {o, v, v, v}));
    // Q[p][y][x][e] = <pe|xy>: P with the row pair swapped
    c->tQ.alloc((size_t)o * v3 * 8);
    c->eng->permute("pexy->pyxe", 1.0, cref(c->ovvv.d(), {o, v, v, v}), 0.0, vref(c->tQ.d(), {o, v, v, v}));
    c->trip_host.clear();
    for (int i = 0; i < o; i++)
      for (int j = 0; j <= i; j++)
        for (int k = 0; k <= j; k++) {
          if (i == j && j == k) continue;  // the bracket vanishes for i = j = k
          c->trip_host.push_back(i);
          c->trip_host.push_back(j);
          c->trip_host.push_back(k);
        }
    c->tables_ready = true;
  }
  if (prec == PREC_BF16X3) {
    Engine &E = *c->eng;
    auto pair_of = [&](DBuf &dst, const std::string &spec, const double *src, std::vector<int> dims,
                       std::vector<int> odims) {
      long n = 1;
      for (int x : dims) n *= x;
      dst.alloc((size_t)n * 4);
      TRef out = vref(nullptr, odims);
      out.p = dst.p;
      out.dtype = 2;
      E.permute(spec, 1.0, vref(src, dims), 0.0, out);
    };
    if (!c->b_P.p) {
      // P and Q, the streamed operands of the particle terms, as MFMA panels
      // (rows (x,y), k = e): each 16-row, 32-deep slice one contiguous block
      const long v2 = (long)v * v, np = panel_size(o, v2, v);
      c->b_P.alloc((size_t)np * 4);
      launch_pack_panel(c->tP.d(), v3, v, 1, o, v2, v, (unsigned short *)c->b_P.p, c->st);
      c->b_Q.alloc((size_t)np * 4);
      launch_pack_panel(c->tQ.d(), v3, v, 1, o, v2, v, (unsigned short *)c->b_Q.p, c->st);
      pair_of(c->b_OVt, "qrmz->qrzm", c->ooov.d(), {o, o, o, v}, {o, o, v, o});  // OVt[q][r][z][m]
    }
    if (c->t_dma > 1) {
      // every operand as a panel: OV[q][r] (rows z, k = m) from ooov once,
      // the t2 slices (rows x, k = y) and T2t / T2tq (rows (x,y) / (y,x),
      // k = m) every call
      const long v2 = (long)v * v;
      if (!c->p_OV.p) {
        c->p_OV.alloc((size_t)panel_size(o * o, v, o) * 4);
        launch_pack_panel(c->ooov.d(), (long)o * v, 1, v, o * o, v, o, (unsigned short *)c->p_OV.p, c->st);
      }
      c->p_t2.alloc((size_t)panel_size(o * o, v, v) * 4);
      launch_pack_panel(c->t2.d(), v2, v, 1, o * o, v, v, (unsigned short *)c->p_t2.p, c->st);
      c->p_T2t.alloc((size_t)panel_size(o, v2, o) * 4);
      launch_pack_panel(c->t2.d(), (long)o * v2, 1, v2, o, v2, o, (unsigned short *)c->p_T2t.p, c->st);
      c->p_T2tq.alloc((size_t)panel_size(o, v2, o) * 4);
      launch_pack_panel2(c->t2.d(), (long)o * v2, 1, v, v, v2, o, v2, o, (unsigned short *)c->p_T2tq.p, c->st);
      if (c->t_merge) {
        // padded rows (x, y) -> x VP + y; rows y >= v are zeros
        const long VP = (v + 15) / 16 * 16, R = (long)v * VP;
        if (!c->q_P.p) {
          c->q_P.alloc((size_t)panel_size(o, R, v) * 4);
          launch_pack_panel_rows(c->tP.d(), v3, v2, v, v, VP, 1, o, R, v, (unsigned short *)c->q_P.p, c->st);
          c->q_Q.alloc((size_t)panel_size(o, R, v) * 4);
          launch_pack_panel_rows(c->tQ.d(), v3, v2, v, v, VP, 1, o, R, v, (unsigned short *)c->q_Q.p, c->st);
        }
        // T2t[p][x][y][m] = t2[p][m][x][y];  T2tq[p][y][x][m] = t2[p][m][x][y]
        c->q_T2t.alloc((size_t)panel_size(o, R, o) * 4);
        launch_pack_panel_rows(c->t2.d(), (long)o * v2, v, 1, v, VP, v2, o, R, o, (unsigned short *)c->q_T2t.p,
                               c->st);
        c->q_T2tq.alloc((size_t)panel_size(o, R, o) * 4);
        launch_pack_panel_rows(c->t2.d(), (long)o * v2, 1, v, v, VP, v2, o, R, o, (unsigned short *)c->q_T2tq.p,
                               c->st);
      }
      return;
    }
    // amplitude-dependent pairs, rebuilt every call
    pair_of(c->b_t2, "ijab->ijab", c->t2.d(), {o, o, v, v}, {o, o, v, v});
    pair_of(c->b_T2t, "pmxy->pxym", c->t2.d(), {o, o, v, v}, {o, v, v, o});   // T2t[p][x][y][m] = t2[p][m][x][y]
    pair_of(c->b_T2tq, "pmxy->pyxm", c->t2.d(), {o, o, v, v}, {o, v, v, o});  // T2tq[p][y][x][m] = t2[p][m][x][y]
  }
  if (prec == PREC_FP32 && !c->f_tP.p) {
    auto conv = [&](DBuf &dst, const double *src, long n) {
      dst.alloc(n * 4);
      PermDesc pd{};
      pd.rank = 1; pd.od[0] = (int)n; pd.istr[0] = 1; pd.in = src; pd.out = dst.p; pd.in_t = 0; pd.out_t = 1;
      pd.alpha = 1.0; pd.beta = 0.0;
      launch_permute(pd, c->st);
    };
    conv(c->f_tP, c->tP.d(), (long)o * v3);
    conv(c->f_tQ, c->tQ.d(), (long)o * v3);
    conv(c->f_ooov, c->ooov.d(), (long)o * o * o * v);
  }
}

// One W3 to build: triplet (i, j, k) of the problem whose operands live in
// context src (the (T) of one context uses src = itself; a batch groups the
// triplets of several equally sized problems).
struct TEnt {
  const ugx_ctx *src;
  int i, j, k;
};

// Output offset tables of the three W launches into the blocked layout of
// block edge T:  z = c: rows (a,b), cols c;  z = b: rows b, cols (a,c);
// z = a: rows a, cols (b,c).
void ensure_w_tables(ugx_ctx *c, int T) {
  const int v = c->nv;
  if (c->tOff_blk == T && c->tOff_v == v && c->tOff.p) return;
  const long vv = (long)v * v, nbk = tw_nb(v, T);
  std::vector<long> off((size_t)3 * (vv + v));
  long *g0m = off.data(), *g0n = g0m + vv, *g1m = g0n + v, *g1n = g1m + v, *g2m = g1n + vv, *g2n = g2m + v;
  for (long x = 0; x < v; x++) {
    g0n[x] = tw_off_c(x, T);
    g1m[x] = tw_off_b(x, nbk, T);
    g2m[x] = tw_off_a(x, nbk, T);
    for (long y = 0; y < v; y++) {
      g0m[x * v + y] = tw_off_a(x, nbk, T) + tw_off_b(y, nbk, T);
      g1n[x * v + y] = tw_off_a(x, nbk, T) + tw_off_c(y, T);
      g2n[x * v + y] = tw_off_b(x, nbk, T) + tw_off_c(y, T);
    }
  }
  // rows r..r+3 (r % 4 == 0) of table rt contiguous and 4-aligned, and
  // every entry of table ct a multiple of 4
  auto runs4 = [](const long *rt, long nr, const long *ct, long nc) {
    if (nr % 4) return false;
    for (long r = 0; r < nr; r += 4)
      if (rt[r] % 4 || rt[r + 1] != rt[r] + 1 || rt[r + 2] != rt[r] + 2 || rt[r + 3] != rt[r] + 3) return false;
    for (long x = 0; x < nc; x++)
      if (ct[x] % 4) return false;
    return true;
  };
  c->tOff_row4[0] = runs4(g0n, v, g0m, vv);
  c->tOff_row4[1] = runs4(g1n, vv, g1m, v);
  c->tOff_row4[2] = runs4(g2n, vv, g2m, v);
  c->tOff_merge4 = runs4(g0n, v, g1m, v);
  for (long x = 0; x < v; x++)
    if (g2m[x] % 4) c->tOff_merge4 = false;
  check(hipStreamSynchronize(c->st), "stream sync");  // the old tables may still be in use
  c->tOff.alloc(off.size() * 8);
  check(hipMemcpy(c->tOff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice), "W offset tables");
  c->tOff_blk = T;
  c->tOff_v = v;
}

// Output tables of launch g (0: z = c, 1: z = b, 2: z = a) into the blocked
// W; transposed launches swap the row and column tables.
void set_w_tables(ugx_ctx *c, GemmDesc &d, int g, bool transposed = false) {
  ensure_w_tables(c, tw_blk(c->tW_f32));
  const long v = c->nv, vv = v * v;
  const long *t = (const long *)c->tOff.p;
  const long *m0 = t, *n0 = m0 + vv, *m1 = n0 + v, *n1 = m1 + v, *m2 = n1 + vv, *n2 = m2 + v;
  d.offm = g == 0 ? m0 : g == 1 ? m1 : m2;
  d.offn = g == 0 ? n0 : g == 1 ? n1 : n2;
  if (transposed) std::swap(d.offm, d.offn);
}

// Builds W for triplets [t0, t0 + nb) of the table into c->tWbuf.
//
// The six permutations pair up by their third output index z, and the two
// members of a pair share one accumulator (one grouped GEMM with four K
// segments: two particle terms, K = v, and two hole terms, K = o):
//   z = c: X(ijk)[abc] + X(jik)[bac]   rows (a,b), columns c
//   z = b: X(ikj)[acb] + X(kij)[cab]   rows b, columns (a,c)
//   z = a: X(jki)[bca] + X(kji)[cba]   rows a, columns (b,c)
// The column index of each launch is the one with unit stride in W, so the
// 16 lanes of an MFMA output row write consecutive addresses.  The second
// member of a pair reads its integrals from Q[p][y][x][e] = <pe|xy> and its
// hole amplitudes as t2_pm^xy = t2_mp^yx, which puts its rows/columns in
// the first member's order.  W is written once and accumulated twice.
void triples_build_w(ugx_ctx *c, int prec, const TEnt *ents, int nb) {
  const int o = c->no, v = c->nv;
  const long v2 = (long)v * v, v3 = v2 * v, ov = (long)o * v, ov2 = (long)o * v2;
  const bool f32 = (prec != PREC_FP64);
  const size_t es = f32 ? 4 : 8;
  auto at = [&](const char *base, long off) { return (const void *)(base + (size_t)off * es); };
  std::vector<GemmPtrs> tab((size_t)3 * nb);
  for (int b = 0; b < nb; b++) {
    const ugx_ctx *src = ents[b].src;  // the entry's own operands (same no, nv)
    const char *P = (const char *)(f32 ? src->f_tP.p : src->tP.p);
    const char *Q = (const char *)(f32 ? src->f_tQ.p : src->tQ.p);
    const char *T2 = (const char *)(f32 ? src->f_t2.p : src->t2.p);
    const char *OV = (const char *)(f32 ? src->f_ooov.p : src->ooov.p);
    const long i = ents[b].i, j = ents[b].j, k = ents[b].k;
    double *W = (double *)((char *)c->tWbuf.p + (size_t)b * tw_stride(v, tw_blk(c->tW_f32)) * (c->tW_f32 ? 4 : 8));
    GemmPtrs &A = tab[b], &B = tab[(size_t)nb + b], &C = tab[(size_t)2 * nb + b];
    // z = c
    A.A[0] = at(P, i * v3);              A.B[0] = at(T2, (k * o + j) * v2);
    A.A[1] = at(T2, i * ov2);            A.B[1] = at(OV, (j * o + k) * ov);
    A.A[2] = at(Q, j * v3);              A.B[2] = at(T2, (k * o + i) * v2);
    A.A[3] = at(T2, j * v2);             A.B[3] = at(OV, (i * o + k) * ov);
    A.C = W;
    // z = b
    B.A[0] = at(T2, (j * o + k) * v2);   B.B[0] = at(P, i * v3);
    B.A[1] = at(OV, (k * o + j) * ov);   B.B[1] = at(T2, i * ov2);
    B.A[2] = at(T2, (j * o + i) * v2);   B.B[2] = at(Q, k * v3);
    B.A[3] = at(OV, (i * o + j) * ov);   B.B[3] = at(T2, k * v2);
    B.C = W;
    // z = a
    C.A[0] = at(T2, (i * o + k) * v2);   C.B[0] = at(P, j * v3);
    C.A[1] = at(OV, (k * o + i) * ov);   C.B[1] = at(T2, j * ov2);
    C.A[2] = at(T2, (i * o + j) * v2);   C.B[2] = at(Q, k * v3);
    C.A[3] = at(OV, (j * o + i) * ov);   C.B[3] = at(T2, k * v2);
    C.C = W;
  }
  c->tptrs.alloc(tab.size() * sizeof(GemmPtrs));
  check(hipMemcpyAsync(c->tptrs.p, tab.data(), tab.size() * sizeof(GemmPtrs), hipMemcpyHostToDevice, c->st),
        "pointer table copy");
  auto tile_rows_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 3 : n <= 80 ? 8 : n <= 128 ? 0 : n <= 160 ? 5 : 0; };
  auto tile_cols_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 4 : n <= 80 ? 10 : n <= 128 ? 0 : n <= 160 ? 9 : 0; };
  const double fl = 4.0 * (double)nb * (double)v3 * (double)(v + o);
  for (int g = 0; g < 3; g++) {
    GemmDesc d{};
    d.nseg = 4;
    d.batch = nb;
    for (int s = 0; s < 4; s++) {
      d.seg[s].K = (s % 2 == 0) ? v : o;
      d.seg[s].sign = (s % 2 == 0) ? 1.0 : -1.0;
    }
    int tile;
    if (g == 0) {
      d.M = (int)v2; d.N = v;
      d.seg[0].sam = v; d.seg[0].sak = 1; d.seg[0].sbk = 1; d.seg[0].sbn = v;
      d.seg[1].sam = 1; d.seg[1].sak = v2; d.seg[1].sbk = v; d.seg[1].sbn = 1;
      d.seg[2].sam = v; d.seg[2].sak = 1; d.seg[2].sbk = 1; d.seg[2].sbn = v;
      d.seg[3].sam = 1; d.seg[3].sak = ov2; d.seg[3].sbk = v; d.seg[3].sbn = 1;
      d.scm = v; d.scn = 1;
      tile = tile_rows_v(v);
    } else {
      d.M = v; d.N = (int)v2;
      d.seg[0].sam = v; d.seg[0].sak = 1; d.seg[0].sbk = 1; d.seg[0].sbn = v;
      d.seg[1].sam = 1; d.seg[1].sak = v; d.seg[1].sbk = v2; d.seg[1].sbn = 1;
      d.seg[2].sam = v; d.seg[2].sak = 1; d.seg[2].sbk = 1; d.seg[2].sbn = v;
      d.seg[3].sam = 1; d.seg[3].sak = v; d.seg[3].sbk = ov2; d.seg[3].sbn = 1;
      if (g == 1) { d.scm = v; d.ndiv = v; d.scn1 = v2; d.scn = 1; }
      else { d.scm = v2; d.scn = 1; }
      tile = tile_cols_v(v);
    }
    d.alpha = 1.0;
    d.beta = (g == 0) ? 0.0 : 1.0;
    d.c_f32 = c->tW_f32;
    d.tag = 1;
    set_w_tables(c, d, g);
    tile = c->eng->tile_for(d, tab.data() + (size_t)g * nb, prec, tile, tw_stride(v, tw_blk(c->tW_f32)));
    c->eng->timed(prec == PREC_FP32 ? "triples_w:fp32" : "triples_w", fl, [&]() {
      launch_gemm(d, (const GemmPtrs *)c->tptrs.p + (size_t)g * nb, prec, tile, c->st);
    });
  }
}

// BF16x3 form of triples_build_w: the same three grouped GEMMs on BF16
// hi/lo pair operands that all have k contiguous -- the hole terms read
// T2t[p][x][y][m] = t2_pm^xy, T2tq[p][y][x][m] = t2_pm^xy and
// OVt[q][r][z][m] = <qr|mz> instead of the row-contiguous t2 and ooov, and
// P and Q are MFMA panels.  Each GEMM is the transpose of its
// triples_build_w counterpart, so that its rows carry W's contiguous index c:
//   z = c: rows c, columns (a,b);  z = b: rows (a,c), columns b;
//   z = a: rows (b,c), columns a,
// and the four rows an MFMA lane accumulates are four consecutive floats of
// the blocked FP32 W (one 16-byte read-modify-write, GemmDesc::row4).
void triples_build_w_pair(ugx_ctx *c, const TEnt *ents, int nb) {
  const int o = c->no, v = c->nv;
  const long v2 = (long)v * v, v3 = v2 * v, ov2o = (long)o * v2, vo = (long)v * o;
  // P and Q are panels (launch_pack_panel): pP halves per p, nP per array
  const long pP = panel_size(1, v2, v), nP = panel_size(o, v2, v), pblk = 16L * ((v + 31) / 32 * 32);
  const long nT = (long)o * o * v2, nOV = (long)o * o * o * v;
  std::vector<GemmPtrs> tab((size_t)3 * nb);
  for (int b = 0; b < nb; b++) {
    const ugx_ctx *src = ents[b].src;  // the entry's own operands (same no, nv)
    const unsigned short *P = (const unsigned short *)src->b_P.p, *Q = (const unsigned short *)src->b_Q.p;
    const unsigned short *T2 = (const unsigned short *)src->b_t2.p, *T2t = (const unsigned short *)src->b_T2t.p;
    const unsigned short *T2q = (const unsigned short *)src->b_T2tq.p, *OVt = (const unsigned short *)src->b_OVt.p;
    const long i = ents[b].i, j = ents[b].j, k = ents[b].k;
    double *W = (double *)((char *)c->tWbuf.p + (size_t)b * tw_stride(v, tw_blk(c->tW_f32)) * (c->tW_f32 ? 4 : 8));
    GemmPtrs &A = tab[b], &B = tab[(size_t)nb + b], &C = tab[(size_t)2 * nb + b];
    A.A[0] = T2 + (k * o + j) * v2;   A.B[0] = P + i * pP;
    A.A[1] = OVt + (j * o + k) * vo;  A.B[1] = T2t + i * ov2o;
    A.A[2] = T2 + (k * o + i) * v2;   A.B[2] = Q + j * pP;
    A.A[3] = OVt + (i * o + k) * vo;  A.B[3] = T2q + j * ov2o;
    A.C = W;
    B.A[0] = P + i * pP;              B.B[0] = T2 + (j * o + k) * v2;
    B.A[1] = T2t + i * ov2o;          B.B[1] = OVt + (k * o + j) * vo;
    B.A[2] = Q + k * pP;              B.B[2] = T2 + (j * o + i) * v2;
    B.A[3] = T2q + k * ov2o;          B.B[3] = OVt + (i * o + j) * vo;
    B.C = W;
    C.A[0] = P + j * pP;              C.B[0] = T2 + (i * o + k) * v2;
    C.A[1] = T2t + j * ov2o;          C.B[1] = OVt + (k * o + i) * vo;
    C.A[2] = Q + k * pP;              C.B[2] = T2 + (i * o + j) * v2;
    C.A[3] = T2q + k * ov2o;          C.B[3] = OVt + (j * o + i) * vo;
    C.C = W;
  }
  c->tptrs.alloc(tab.size() * sizeof(GemmPtrs));
  check(hipMemcpyAsync(c->tptrs.p, tab.data(), tab.size() * sizeof(GemmPtrs), hipMemcpyHostToDevice, c->st),
        "pointer table copy");
  auto tile_rows_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 3 : n <= 80 ? 8 : n <= 128 ? 0 : n <= 160 ? 5 : 0; };
  auto tile_cols_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 4 : n <= 80 ? 10 : n <= 128 ? 0 : n <= 160 ? 9 : 0; };
  const double fl = 4.0 * (double)nb * (double)v3 * (double)(v + o);
  for (int g = 0; g < 3; g++) {
    GemmDesc d{};
    d.nseg = 4;
    d.batch = nb;
    d.vec_ok = (v % 8 == 0 && o % 8 == 0) ? 1 : 0;
    // g = 0: A = the t2 / OVt rows of c, B = the panel / T2t columns (a,b);
    // g > 0: A = the panel / T2t rows (x,c), B = the t2 / OVt columns
    for (int s = 0; s < 4; s++) {
      const bool part = (s % 2 == 0);
      GemmSeg &q = d.seg[s];
      q.K = part ? v : o;
      q.sign = part ? 1.0 : -1.0;
      q.sak = 1; q.sbk = 1;
      const long lo_small = part ? nT : nOV, lo_big = part ? nP : nT;
      const long ld_small = part ? v : o;
      if (g == 0) {
        q.sam = ld_small; q.loA = lo_small;
        q.loB = lo_big;
        if (part) { q.pblkB = pblk; q.sbn = BK_PANEL; } else q.sbn = o;
      } else {
        q.sbn = ld_small; q.loB = lo_small;
        q.loA = lo_big;
        if (part) { q.pblkA = pblk; q.sam = BK_PANEL; } else q.sam = o;
      }
    }
    int tile;
    if (g == 0) { d.M = v; d.N = (int)v2; tile = tile_cols_v(v); }
    else { d.M = (int)v2; d.N = v; tile = tile_rows_v(v); }
    d.alpha = 1.0;
    d.beta = (g == 0) ? 0.0 : 1.0;
    d.c_f32 = c->tW_f32;
    d.tag = 1;
    set_w_tables(c, d, g, true);
    d.row4 = (c->tW_f32 && c->tOff_row4[g]) ? 1 : 0;
    // consecutive triplets share i (and often j): run them tile by tile so
    // their common <pe|xy> panel tiles come from L2
    d.zgroup = c->t_zgroup;
    tile = c->eng->tile_for(d, tab.data() + (size_t)g * nb, PREC_BF16X3_PRESPLIT, tile, tw_stride(v, tw_blk(c->tW_f32)));
    c->eng->timed("triples_w:bf16x3", fl, [&]() {
      launch_gemm(d, (const GemmPtrs *)c->tptrs.p + (size_t)g * nb, PREC_BF16X3_PRESPLIT, tile, c->st);
    });
  }
}

// The z = c and z = b permutation groups as one launch.  For a fixed a both
// cover the W slab [a][b][c]: group c as rows c (t2 / <qr|mz> slices) times
// columns (a,b) (rows of the big panels P_i, T2t_i, Q_j, T2tq_j), group b as
// rows (a,c) (P_i, T2t_i, Q_k, T2tq_k) times columns b (slices).  Taking
// the rows (a, .) of the big panels as one operand block (the padded q_*
// panels start every a-slab on a panel block), both become a GEMM with rows
// c and columns b, eight K segments (2 (v + o) + 2 (v + o) = 768 at
// (H2O)8) summed in one accumulator, and one write of the slab.  Each
// (triplet, a) is one sub-entry (GemmDesc::zsub = v) of the triplet's table
// entry; W's blocked offsets split into c (rows), b (columns) and a (offz).
// Entries are in the order of the z = c launch (its pair of big panels),
// groups of zgroup triplets running a by a on one XCD.
void triples_w_merged(ugx_ctx *c, const TEnt *ents, int nb, const std::vector<int> &pos, double beta) {
  const int o = c->no, v = c->nv;
  const long v2 = (long)v * v;
  const long VP = (v + 15) / 16 * 16, R = (long)v * VP;
  const long pT = panel_size(1, v, v), nT = panel_size(o * o, v, v);     // t2 slices
  const long pO = panel_size(1, v, o), nO = panel_size(o * o, v, o);     // OV slices
  const long qP = panel_size(1, R, v), nP = panel_size(o, R, v);         // padded P, Q
  const long qT = panel_size(1, R, o), nTT = panel_size(o, R, o);        // padded T2t, T2tq
  const long blk_v = 16L * ((v + 31) / 32 * 32), blk_o = 16L * ((o + 31) / 32 * 32);
  std::vector<GemmPtrs> tab((size_t)nb);
  for (int b = 0; b < nb; b++) {
    const ugx_ctx *src = ents[b].src;
    const unsigned short *P = (const unsigned short *)src->q_P.p, *Q = (const unsigned short *)src->q_Q.p;
    const unsigned short *T2 = (const unsigned short *)src->p_t2.p, *OV = (const unsigned short *)src->p_OV.p;
    const unsigned short *T2t = (const unsigned short *)src->q_T2t.p, *T2q = (const unsigned short *)src->q_T2tq.p;
    const long i = ents[b].i, j = ents[b].j, k = ents[b].k;
    double *W = (double *)((char *)c->tWbuf.p + (size_t)b * tw_stride(v, tw_blk(c->tW_f32)) * (c->tW_f32 ? 4 : 8));
    GemmPtrs &A = tab[pos[b]];
    // group c: rows c from the slices, columns b from the a-slab of the big panels
    A.A[0] = T2 + (k * o + j) * pT;   A.B[0] = P + i * qP;
    A.A[1] = OV + (j * o + k) * pO;   A.B[1] = T2t + i * qT;
    A.A[2] = T2 + (k * o + i) * pT;   A.B[2] = Q + j * qP;
    A.A[3] = OV + (i * o + k) * pO;   A.B[3] = T2q + j * qT;
    // group b: rows c from the a-slab of the big panels, columns b from the slices
    A.A[4] = P + i * qP;              A.B[4] = T2 + (j * o + k) * pT;
    A.A[5] = T2t + i * qT;            A.B[5] = OV + (k * o + j) * pO;
    A.A[6] = Q + k * qP;              A.B[6] = T2 + (j * o + i) * pT;
    A.A[7] = T2q + k * qT;            A.B[7] = OV + (i * o + j) * pO;
    A.C = W;
  }
  c->tptrs2.alloc(tab.size() * sizeof(GemmPtrs));
  check(hipMemcpyAsync(c->tptrs2.p, tab.data(), tab.size() * sizeof(GemmPtrs), hipMemcpyHostToDevice, c->st),
        "pointer table copy");
  ensure_w_tables(c, tw_blk(c->tW_f32));
  const long *t = (const long *)c->tOff.p;
  // the v-entry tables: tw_off_c (rows c), tw_off_b (columns b), tw_off_a (a)
  const long *offc = t + v2, *offb = offc + v, *offa = offb + v + v2;
  GemmDesc d{};
  d.nseg = 8;
  d.batch = nb;
  d.vec_ok = 1;
  for (int s = 0; s < 8; s++) {
    const bool part = (s % 2 == 0), big_b = (s < 4);
    GemmSeg &q = d.seg[s];
    q.K = part ? v : o;
    q.sign = part ? 1.0 : -1.0;
    q.sak = 1; q.sbk = 1;
    q.sam = BK_PANEL; q.sbn = BK_PANEL;
    const long lo_small = part ? nT : nO, lo_big = part ? nP : nTT, blk = part ? blk_v : blk_o;
    q.pblkA = blk; q.pblkB = blk;
    const long slab = VP / 16 * blk;  // one a-slab of a big panel, elements
    if (big_b) { q.loA = lo_small; q.loB = lo_big; q.subB = slab; }
    else { q.loA = lo_big; q.loB = lo_small; q.subA = slab; }
  }
  d.M = v;
  d.N = v;
  d.alpha = 1.0;
  d.beta = beta;
  d.c_f32 = c->tW_f32;
  d.tag = 1;
  d.offm = offc;
  d.offn = offb;
  d.offz = offa;
  d.zsub = v;
  d.row4 = (c->tW_f32 && c->tOff_merge4) ? 1 : 0;
  d.zgroup = c->t_zgroup_m;
  d.dma = c->t_dma;
  const int tile = v <= 32 ? 2 : v <= 64 ? 1 : v <= 128 ? 0 : 11;  // one tile covers the slab
  const double fl = 8.0 * (double)nb * (double)v2 * v * (double)(v + o);
  c->eng->timed("triples_w:bf16x3", fl, [&]() {
    launch_gemm(d, (const GemmPtrs *)c->tptrs2.p, PREC_BF16X3_PRESPLIT, tile, c->st);
  });
}

// triples_build_w_pair with every operand an MFMA panel, for the kernel
// that copies panel blocks straight into LDS (GemmDesc::dma): the same
// three transposed grouped GEMMs, pointer choices and K segments.
void triples_build_w_dma(ugx_ctx *c, const TEnt *ents, int nb) {
  const int o = c->no, v = c->nv;
  const long v2 = (long)v * v, v3 = v2 * v;
  const long pP = panel_size(1, v2, v), nP = panel_size(o, v2, v);          // P, Q
  const long pT = panel_size(1, v, v), nT = panel_size(o * o, v, v);         // t2 slices
  const long pO = panel_size(1, v, o), nO = panel_size(o * o, v, o);         // OV slices
  const long pTT = panel_size(1, v2, o), nTT = panel_size(o, v2, o);         // T2t, T2tq
  const long blk_v = 16L * ((v + 31) / 32 * 32), blk_o = 16L * ((o + 31) / 32 * 32);
  std::vector<GemmPtrs> tab((size_t)3 * nb);
  // Table position of entry b in each launch.  The big panel operands of
  // launch z = c are P_i and Q_j, of z = b P_i and Q_k, of z = a P_j and Q_k:
  // ordering each launch's entries by its own pair makes the zgroup
  // neighbours (consecutive positions, run tile by tile on one XCD) share
  // both panels through L2.  The order does not change any entry's result.
  std::vector<int> pos[3];
  for (int g = 0; g < 3; g++) {
    std::vector<int> ord(nb);
    for (int b = 0; b < nb; b++) ord[b] = b;
    if (c->t_sort && g > 0) {
      std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        const TEnt &X = ents[x], &Y = ents[y];
        if (X.src != Y.src) return X.src < Y.src;
        const int x0 = g == 1 ? X.i : X.j, y0 = g == 1 ? Y.i : Y.j;
        if (x0 != y0) return x0 < y0;
        return X.k < Y.k;
      });
    }
    pos[g].assign(nb, 0);
    for (int q = 0; q < nb; q++) pos[g][ord[q]] = q;
  }
  for (int b = 0; b < nb; b++) {
    const ugx_ctx *src = ents[b].src;  // the entry's own operands (same no, nv)
    const unsigned short *P = (const unsigned short *)src->b_P.p, *Q = (const unsigned short *)src->b_Q.p;
    const unsigned short *T2 = (const unsigned short *)src->p_t2.p, *OV = (const unsigned short *)src->p_OV.p;
    const unsigned short *T2t = (const unsigned short *)src->p_T2t.p, *T2q = (const unsigned short *)src->p_T2tq.p;
    const long i = ents[b].i, j = ents[b].j, k = ents[b].k;
    double *W = (double *)((char *)c->tWbuf.p + (size_t)b * tw_stride(v, tw_blk(c->tW_f32)) * (c->tW_f32 ? 4 : 8));
    GemmPtrs &A = tab[pos[0][b]], &B = tab[(size_t)nb + pos[1][b]], &C = tab[(size_t)2 * nb + pos[2][b]];
    A.A[0] = T2 + (k * o + j) * pT;   A.B[0] = P + i * pP;
    A.A[1] = OV + (j * o + k) * pO;   A.B[1] = T2t + i * pTT;
    A.A[2] = T2 + (k * o + i) * pT;   A.B[2] = Q + j * pP;
    A.A[3] = OV + (i * o + k) * pO;   A.B[3] = T2q + j * pTT;
    A.C = W;
    B.A[0] = P + i * pP;              B.B[0] = T2 + (j * o + k) * pT;
    B.A[1] = T2t + i * pTT;           B.B[1] = OV + (k * o + j) * pO;
    B.A[2] = Q + k * pP;              B.B[2] = T2 + (j * o + i) * pT;
    B.A[3] = T2q + k * pTT;           B.B[3] = OV + (i * o + j) * pO;
    B.C = W;
    C.A[0] = P + j * pP;              C.B[0] = T2 + (i * o + k) * pT;
    C.A[1] = T2t + j * pTT;           C.B[1] = OV + (k * o + i) * pO;
    C.A[2] = Q + k * pP;              C.B[2] = T2 + (i * o + j) * pT;
    C.A[3] = T2q + k * pTT;           C.B[3] = OV + (j * o + i) * pO;
    C.C = W;
  }
  c->tptrs.alloc(tab.size() * sizeof(GemmPtrs));
  check(hipMemcpyAsync(c->tptrs.p, tab.data(), tab.size() * sizeof(GemmPtrs), hipMemcpyHostToDevice, c->st),
        "pointer table copy");
  auto tile_rows_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 3 : n <= 80 ? 8 : n <= 128 ? 0 : n <= 160 ? 5 : 0; };
  auto tile_cols_v = [&](int n) { return n <= 32 ? 2 : n <= 64 ? 4 : n <= 80 ? 10 : n <= 128 ? 0 : n <= 160 ? 9 : 0; };
  const double fl = 4.0 * (double)nb * (double)v3 * (double)(v + o);
  // the z = c and z = b groups as one launch, with the z = a launch before
  // (writing W) or after it (accumulating)
  const bool merged = c->t_merge && v <= 160, za_first = merged && c->t_za_first;
  if (merged && !za_first) triples_w_merged(c, ents, nb, pos[0], 0.0);
  for (int g = merged ? 2 : 0; g < 3; g++) {
    GemmDesc d{};
    d.nseg = 4;
    d.batch = nb;
    d.vec_ok = 1;  // panel bases are whole kilobytes apart
    for (int s = 0; s < 4; s++) {
      const bool part = (s % 2 == 0);
      GemmSeg &q = d.seg[s];
      q.K = part ? v : o;
      q.sign = part ? 1.0 : -1.0;
      q.sak = 1; q.sbk = 1;
      q.sam = BK_PANEL; q.sbn = BK_PANEL;
      const long lo_small = part ? nT : nO, lo_big = part ? nP : nTT, blk = part ? blk_v : blk_o;
      q.pblkA = blk; q.pblkB = blk;
      if (g == 0) { q.loA = lo_small; q.loB = lo_big; }
      else { q.loA = lo_big; q.loB = lo_small; }
    }
    int tile;
    if (g == 0) { d.M = v; d.N = (int)v2; tile = tile_cols_v(v); }
    else { d.M = (int)v2; d.N = v; tile = tile_rows_v(v); }
    d.alpha = 1.0;
    d.beta = (g == 0 || za_first) ? 0.0 : 1.0;
    d.c_f32 = c->tW_f32;
    d.tag = 1;
    set_w_tables(c, d, g, true);
    d.row4 = (c->tW_f32 && c->tOff_row4[g]) ? 1 : 0;
    d.zgroup = c->t_zgroup;
    d.dma = c->t_dma;
    tile = c->eng->tile_for(d, tab.data() + (size_t)g * nb, PREC_BF16X3_PRESPLIT, tile, tw_stride(v, tw_blk(c->tW_f32)));
    c->eng->timed("triples_w:bf16x3", fl, [&]() {
      launch_gemm(d, (const GemmPtrs *)c->tptrs.p + (size_t)g * nb, PREC_BF16X3_PRESPLIT, tile, c->st);
    });
  }
  if (za_first) triples_w_merged(c, ents, nb, pos[0], 1.0);
}

// The (T) operands of a context at a precision: the integral tables and
// the amplitude-dependent copies (FP32 t2, or the BF16 hi/lo pairs).
void triples_operands(ugx_ctx *c, int prec) {
  triples_prepare(c, prec);
  if (prec == PREC_FP32) {
    const long n = c->oovv_n();
    c->f_t2.alloc(n * 4);
    PermDesc pd{};
    pd.rank = 1; pd.od[0] = (int)n; pd.istr[0] = 1; pd.in = c->t2.p; pd.out = c->f_t2.p; pd.in_t = 0; pd.out_t = 1;
    pd.alpha = 1.0; pd.beta = 0.0;
    launch_permute(pd, c->st);
  }
}

// (T) energies of a list of triplet entries, computed on the lead context's
// stream and buffers.  Entry e is triplet (i, j, k) of the problem in
// context ents[e].src (all of one size, operands prepared by
// triples_operands and complete on the device); prob[e] indexes probs, the
// matching (eps, oovv, t1) of each problem.  Batches of entries get one W
// buffer of up to t_wbytes (48 GiB by default): three grouped W3 GEMM launches whose pointer
// tables span every entry of the batch, then one persistent energy kernel
// over all (entry, block pair) items.
std::vector<double> triples_entries(ugx_ctx *c, int prec, const std::vector<TEnt> &ents,
                                    const std::vector<int> &prob, const std::vector<TripProb> &probs) {
  const int o = c->no, v = c->nv;
  const long nt = (long)ents.size();
  std::vector<double> e(nt);
  if (nt == 0) return e;
  const int wf = (prec != PREC_FP64) ? 1 : 0;  // W is FP32 in the low-precision forms
  const size_t wes = wf ? 4 : 8;
  const long ws = tw_stride(v, tw_blk(wf));
  // fewer, larger batches: fewer launches and energy-kernel tails (48 GiB
  // holds ~3400 (H2O)8 triplets: 4 batches of the 11480)
  size_t mem_free = 0, mem_total = 0;
  check(hipMemGetInfo(&mem_free, &mem_total), "memory info");
  const long budget = std::min<long>(c->t_wbytes, (long)(mem_free / 4));
  long nb = std::max<long>(1, std::min<long>(nt, budget / (ws * (long)wes)));
  nb = std::min<long>(nb, c->t_batch > 0 ? c->t_batch : 4096);
  c->tW_f32 = wf;
  c->tWbuf.alloc((size_t)nb * ws * wes);
  c->tE.alloc((size_t)nt * 8);
  const int npairs = triples_energy_npairs(v, wf);
  c->tPart.alloc((size_t)nb * npairs * 8);
  c->tProb.alloc(probs.size() * sizeof(TripProb));
  check(hipMemcpyAsync(c->tProb.p, probs.data(), probs.size() * sizeof(TripProb), hipMemcpyHostToDevice, c->st),
        "problem table");
  std::vector<int> t4((size_t)nt * 4);
  for (long t = 0; t < nt; t++) {
    t4[4 * t] = ents[t].i; t4[4 * t + 1] = ents[t].j; t4[4 * t + 2] = ents[t].k; t4[4 * t + 3] = prob[t];
  }
  c->ttrip.alloc(t4.size() * 4);
  check(hipMemcpyAsync(c->ttrip.p, t4.data(), t4.size() * 4, hipMemcpyHostToDevice, c->st), "entry table");
  c->tCounter.alloc(64);
  // the f_ov t3 terms of the reference's tcorr (ccwfn.cc:706-709) whenever
  // any entry's problem has a non-zero occupied-virtual Fock element
  int fov = 0;
  for (const TEnt &en : ents)
    if (en.src->has_fov) { fov = 1; break; }
  // the pair arrays of the f_ov kernel, per problem, from its final t2 / t1
  std::vector<TripProb> pv = probs;
  if (fov) {
    for (size_t q = 0; q < probs.size(); q++) {
      ugx_ctx *s = nullptr;
      for (const TEnt &en : ents)
        if (en.src->eps.p == probs[q].eps) { s = const_cast<ugx_ctx *>(en.src); break; }
      if (!s) continue;
      const size_t pe = wf ? 8 : 16;
      s->tOT2.alloc((size_t)s->oovv_n() * pe);
      s->tT1F.alloc((size_t)s->ov() * pe);
      launch_interleave(s->oovv_n(), s->oovv.d(), s->t2.d(), s->tOT2.p, wf, s->st);
      launch_interleave(s->ov(), s->t1.d(), s->fov.d(), s->tT1F.p, wf, s->st);
      if (s != c) check(hipStreamSynchronize(s->st), "stream sync");
      pv[q].ot2 = s->tOT2.p;
      pv[q].t1f = s->tT1F.p;
    }
    check(hipMemcpyAsync(c->tProb.p, pv.data(), pv.size() * sizeof(TripProb), hipMemcpyHostToDevice, c->st),
          "problem table");
  }
  // batch ranges of at most nb entries, ending at a block boundary (all
  // triplets of one i of one problem) when one is within the range: the
  // triplets a launch orders by (i, k) or (j, k) then share their panels
  std::vector<std::pair<long, long>> ranges;
  for (long t0 = 0; t0 < nt;) {
    long t1 = std::min<long>(nt, t0 + nb);
    if (c->t_align && t1 < nt) {
      long t = t1;
      while (t > t0 && ents[t - 1].src == ents[t].src && ents[t - 1].i == ents[t].i) t--;
      if (t > t0) t1 = t;
    }
    ranges.emplace_back(t0, t1);
    t0 = t1;
  }
  for (auto &rg : ranges) {
    const long t0 = rg.first;
    int b = (int)(rg.second - rg.first);
    if (prec == PREC_BF16X3 && c->t_dma > 1) triples_build_w_dma(c, ents.data() + t0, b);
    else if (prec == PREC_BF16X3) triples_build_w_pair(c, ents.data() + t0, b);
    else triples_build_w(c, prec, ents.data() + t0, b);
    c->eng->timed("triples_e", 0.0, [&]() {
      launch_triples_energy(o, v, b, (const int *)c->ttrip.p + 4 * t0, (const TripProb *)c->tProb.p, c->tWbuf.p, wf,
                            ws, fov, c->tE.d() + t0, c->tPart.d(), (int *)c->tCounter.p, c->st);
    });
  }
  check(hipMemcpyAsync(e.data(), c->tE.p, nt * 8, hipMemcpyDeviceToHost, c->st), "triples copy");
  check(hipStreamSynchronize(c->st), "stream sync");
  return e;
}

// (T) over all triplets; per-triplet energies (table order, i = j = k
// entries reported as 0) into etrip_out when non-null.
double triples(ugx_ctx *c, int prec, std::vector<double> *etrip_out) {
  const int o = c->no;
  triples_operands(c, prec);
  const long nt = (long)c->trip_host.size() / 3;
  std::vector<TEnt> ents(nt);
  for (long t = 0; t < nt; t++) ents[t] = TEnt{c, c->trip_host[3 * t], c->trip_host[3 * t + 1], c->trip_host[3 * t + 2]};
  std::vector<int> prob(nt, 0);
  std::vector<double> e = triples_entries(c, prec, ents, prob, {TripProb{c->eps.d(), c->oovv.d(), c->t1.d(), nullptr, nullptr}});
  double et = 0.0;
  for (double x : e) et += x;
  if (etrip_out) {
    etrip_out->clear();
    long k = 0;
    for (int i = 0; i < o; i++)
      for (int j = 0; j <= i; j++)
        for (int kk = 0; kk <= j; kk++) {
          if (i == j && j == kk) etrip_out->push_back(0.0);
          else etrip_out->push_back(e[k++]);
        }
  }
  return et;
}

double triples_flops(int o, int v) {
  double ntrip = (double)o * (o + 1) * (o + 2) / 6.0 - o;
  return 12.0 * ntrip * ((double)v * v * v * v + (double)o * v * v * v);
}

void validate_options(ugx_ctx *c, const ugx_options &opt, const char *who) {
  const std::string w(who);
  if (opt.maxiter < 0 || opt.diis_nvec < 1 || opt.diis_nvec > kDiisMax || opt.diis_min_vecs < 1 ||
      opt.precision < 0 || opt.precision > 2 || !(opt.r_convergence > 0) ||
      (opt.polish_mode != UGX_POLISH_PLAIN && opt.polish_mode != UGX_POLISH_REFINE))
    throw UgxError(UGX_BAD_INPUT, w + ": invalid options");
  if (opt.guess < UGX_GUESS_MP2 || opt.guess > UGX_GUESS_MLP) throw UgxError(UGX_BAD_INPUT, w + ": unknown guess");
  if (opt.guess == UGX_GUESS_USER && !c->have_amps)
    throw UgxError(UGX_BAD_INPUT, w + ": UGX_GUESS_USER without ugx_set_amplitudes");
  if (opt.guess == UGX_GUESS_MLP && !c->mlp)
    throw UgxError(UGX_BAD_INPUT, w + ": UGX_GUESS_MLP without ugx_set_guess_network");
}

// BF16x3 GEMM of FP32 operands in any layout (d's segment 0 as built by
// ugx_gemm): the operands are split into BF16 hi/lo pairs A[m][k] at pa
// and B[n][k] at pb (M K and N K pairs, 4 bytes each), then multiplied by
// the pre-split kernel.  Nothing is allocated here.
void gemm_bf16x3_split(ugx_ctx *c, GemmDesc &d, void *pa, void *pb, int tile) {
  GemmSeg &g = d.seg[0];
  const int M = d.M, N = d.N, K = g.K;
  PermDesc pd{};
  pd.rank = 2; pd.in_t = 1; pd.out_t = 2; pd.alpha = 1.0; pd.beta = 0.0;
  pd.od[0] = M; pd.od[1] = K; pd.istr[0] = g.sam; pd.istr[1] = g.sak; pd.in = g.A; pd.out = pa;
  if ((long)M * K > 0) launch_permute(pd, c->st);
  pd.od[0] = N; pd.od[1] = K; pd.istr[0] = g.sbn; pd.istr[1] = g.sbk; pd.in = g.B; pd.out = pb;
  if ((long)N * K > 0) launch_permute(pd, c->st);
  g.A = pa; g.B = pb; g.sam = K; g.sak = 1; g.sbn = K; g.sbk = 1;
  g.loA = (long)M * K; g.loB = (long)N * K;
  launch_gemm(d, nullptr, PREC_BF16X3_PRESPLIT, tile, c->st);
}

// ------------------------------------------------------- static scratch
// Staging of ugx_gemm_bf16x3_static (the BF16 hi/lo pairs of its operands):
// a file-scope device array of kStaticScratchBytes, one copy per GPU, held
// until the process ends.  In host code the array's name is not a device
// address: ugx_create fetches each device's address of it with
// hipGetSymbolAddress (with that device current) into g_scratch, and only
// those addresses are used.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kStaticScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch[kStaticScratchDevices] = {};
bool g_scratch_tried[kStaticScratchDevices] = {};

// Once per device, with that device current.
void static_scratch_setup(int device) {
  if (device < 0 || device >= kStaticScratchDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_tried[device]) return;
  g_scratch_tried[device] = true;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
  }
  g_scratch[device] = p;
}

void *static_scratch(int device) {
  if (device < 0 || device >= kStaticScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

template <typename F>
int guarded(ugx_ctx *c, F &&f) {
  if (!c) return UGX_BAD_INPUT;
  try {
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) throw UgxError(UGX_HIP_ERROR, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return f();
  } catch (const UgxError &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    c->err = "host allocation failed";
    return UGX_OOM;
  } catch (const std::exception &e) {
    c->err = e.what();
    return UGX_HIP_ERROR;
  }
}

}  // namespace

// ====================================================================== ABI
extern "C" {

void ugx_default_options(ugx_options *opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->wfn_ccsd_t = 0;
  opt->maxiter = 100;
  opt->diis = 1;
  opt->r_convergence = 1e-7;
  opt->ooc = 0;
  opt->diis_nvec = 8;
  opt->diis_min_vecs = 3;
  opt->diis_policy = 0;
  opt->precision = UGX_PREC_FP64;
  opt->polish_threshold = 1e-5;
  opt->guess = UGX_GUESS_MP2;
  opt->energy_form = UGX_ENERGY_REFERENCE;
  opt->print = 0;
  opt->polish_keep_diis = 1;
  opt->graph = 0;
  opt->polish_mode = UGX_POLISH_PLAIN;
}

int ugx_create(ugx_ctx **out, int device) {
  if (!out) return UGX_BAD_INPUT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return UGX_NO_DEVICE;
  }
  if (device < 0 || device >= n) return UGX_NO_DEVICE;
  ugx_ctx *c = new (std::nothrow) ugx_ctx();
  if (!c) return UGX_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return UGX_HIP_ERROR;
  }
  c->eng.reset(new Engine(c->st));
  static_scratch_setup(device);  // the device's address of the static scratch (device current)
  // tuning knobs: LDS stages of the BF16x3 kernel (1 or 2); triplets per
  // (T) batch; triplets per output tile of a W3 launch; the W3 kernel (0:
  // register-staged operands, 2 or 3: panels copied into LDS, with that many
  // stages); the (T) W buffer budget in GiB; UGX_LADDER=full keeps the
  // full ladder contraction
  if (const char *e = std::getenv("UGX_B3_STAGES")) gemm_set_b3_stages(std::atoi(e));
  if (const char *e = std::getenv("UGX_T_BATCH")) c->t_batch = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_ZGROUP")) c->t_zgroup = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_DMA")) c->t_dma = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_SORT")) c->t_sort = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_ALIGN")) c->t_align = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_WGIB")) c->t_wbytes = std::max(1L, std::atol(e)) << 30;
  if (const char *e = std::getenv("UGX_T_MERGE")) c->t_merge = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_ZAFIRST")) c->t_za_first = std::atoi(e);
  if (const char *e = std::getenv("UGX_T_ZGROUP_M")) c->t_zgroup_m = std::atoi(e);
  if (const char *e = std::getenv("UGX_LADDER")) c->lad_full = std::string(e) == "full";
  if (c->t_dma != 0 && c->t_dma != 2 && c->t_dma != 3) c->t_dma = 0;
  // a tuned tile table to start from (once per process)
  static std::once_flag tiles_once;
  std::call_once(tiles_once, []() {
    if (const char *e = std::getenv("UGX_TILE_FILE")) (void)tiles_load(e);
  });
  *out = c;
  return UGX_OK;
}

void ugx_destroy(ugx_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  drop_graphs(c);
  c->eng.reset();
  hipStream_t st = c->st;
  delete c;
  if (st) (void)hipStreamDestroy(st);
}

const char *ugx_last_error(const ugx_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ugx_load_hamiltonian(ugx_ctx *c, int no, int nv, const double *fock, const double *oooo, const double *ooov,
                         const double *oovv, const double *ovov, const double *ovvv, const double *vvvv) {
  return guarded(c, [&]() {
    Phase ph("ugx_load_hamiltonian");
    if (no <= 0 || nv <= 0 || !fock || !oooo || !ooov || !oovv || !ovov || !ovvv || !vvvv)
      throw UgxError(UGX_BAD_INPUT, "ugx_load_hamiltonian: sizes must be positive and all blocks given");
    check(hipStreamSynchronize(c->st), "stream sync");
    drop_graphs(c);
    c->eng->drop_cache();
    c->loaded = false;
    c->tables_ready = false;
    c->have_amps = false;
    c->tg_valid = false;
    c->perts.clear();
    c->have_intermediates = false;
    c->f_tP.release(); c->f_tQ.release(); c->f_ooov.release(); c->f_t2.release();
  c->b_P.release(); c->b_Q.release(); c->b_OVt.release();
  c->p_OV.release(); c->p_t2.release(); c->p_T2t.release(); c->p_T2tq.release();
  c->q_P.release(); c->q_Q.release(); c->q_T2t.release(); c->q_T2tq.release();
    c->diis_err.clear(); c->diis_amp.clear(); c->diis_nvec = 0;
    c->prepared_for.clear();
    c->no = no; c->nv = nv;
    const long n = no + nv, o = no, v = nv;
    for (long k = 0; k < n * n; k++)
      if (!std::isfinite(fock[k])) throw UgxError(UGX_NONFINITE, "non-finite Fock matrix element");
    std::vector<double> h(n * n);
    upload(c->fock, fock, n * n, c->st);
    std::vector<double> eps(n), fov(o * v), fvo(v * o), foo(o * o), fvv(v * v);
    for (long p = 0; p < n; p++) eps[p] = fock[p * n + p];
    for (long i = 0; i < o; i++)
      for (long a = 0; a < v; a++) { fov[i * v + a] = fock[i * n + o + a]; fvo[a * o + i] = fock[(o + a) * n + i]; }
    for (long i = 0; i < o; i++)
      for (long j = 0; j < o; j++) foo[i * o + j] = fock[i * n + j];
    for (long a = 0; a < v; a++)
      for (long b = 0; b < v; b++) fvv[a * v + b] = fock[(o + a) * n + o + b];
    upload(c->eps, eps.data(), n, c->st);
    upload(c->fov, fov.data(), o * v, c->st);
    c->h_fov = fov;
    c->has_fov = false;
    for (double f : fov)
      if (f != 0.0) c->has_fov = true;
    upload(c->fvo, fvo.data(), o * v, c->st);
    upload(c->foo, foo.data(), o * o, c->st);
    upload(c->fvv, fvv.data(), v * v, c->st);
    upload(c->oooo, oooo, o * o * o * o, c->st);
    upload(c->ooov, ooov, o * o * o * v, c->st);
    upload(c->oovv, oovv, o * o * v * v, c->st);
    upload(c->ovov, ovov, o * v * o * v, c->st);
    upload(c->ovvv, ovvv, o * v * v * v, c->st);
    upload(c->vvvv, vvvv, v * v * v * v, c->st);
    // non-finite integrals would only surface as a non-finite residual
    // norm iterations later: one scan of the blocks on the device
    {
      DBuf flag;
      flag.alloc(8);
      launch_fill(1, 0.0, flag.d(), c->st);
      const std::pair<const DBuf *, long> blk[] = {{&c->oooo, o * o * o * o}, {&c->ooov, o * o * o * v},
                                                   {&c->oovv, o * o * v * v}, {&c->ovov, o * v * o * v},
                                                   {&c->ovvv, o * v * v * v}, {&c->vvvv, v * v * v * v}};
      for (auto &b : blk) launch_nonfinite(b.second, b.first->d(), flag.d(), c->st);
      double h = 0.0;
      check(hipMemcpyAsync(&h, flag.p, 8, hipMemcpyDeviceToHost, c->st), "flag copy");
      check(hipStreamSynchronize(c->st), "stream sync");
      if (h != 0.0) throw UgxError(UGX_NONFINITE, "non-finite two-electron integral");
    }
    Engine &E = *c->eng;
    // The packed ladder needs <ab|ef> = <ba|fe> (real orbitals), checked on
    // a deterministic sample of the block, and grids within limits;
    // otherwise the full contraction runs.
    {
      bool sym = true;
      double vmax = 0.0;
      unsigned long long x = 0x9E3779B97F4A7C15ull;
      auto rnd = [&](long m) { x = x * 6364136223846793005ull + 1442695040888963407ull; return (long)((x >> 33) % (unsigned long long)m); };
      std::vector<std::pair<double, double>> smp;
      for (int q = 0; q < 4096; q++) {
        const long a = rnd(v), b = rnd(v), e = rnd(v), f = rnd(v);
        const double p1 = vvvv[((a * v + b) * v + e) * v + f], p2 = vvvv[((b * v + a) * v + f) * v + e];
        vmax = std::max(vmax, std::fabs(p1));
        smp.emplace_back(p1, p2);
      }
      for (auto &pr : smp)
        if (std::fabs(pr.first - pr.second) > 1e-12 * vmax + 1e-300) sym = false;
      const long P = v * (v + 1) / 2, Ps = v * (v - 1) / 2, Q = o * (o + 1) / 2, Qs = o * (o - 1) / 2;
      c->lad_packed = sym && !c->lad_full && P <= 65535 && Q <= 65535;
      if (c->lad_packed) {
        c->lad_vp.alloc((size_t)P * ladder_ld(P) * 8);
        c->lad_vm.alloc((size_t)std::max<long>(1, Ps * ladder_ld(Ps)) * 8);
        c->lad_S.alloc((size_t)Q * ladder_ld(P) * 8);
        c->lad_A.alloc((size_t)std::max<long>(1, Qs * ladder_ld(Ps)) * 8);
        c->lad_LP.alloc((size_t)Q * P * 8);
        c->lad_LM.alloc((size_t)std::max<long>(1, Qs * Ps) * 8);
        launch_ladder_vpm(nv, c->vvvv.d(), c->lad_vp.d(), c->lad_vm.d(), c->st);
        c->zm_vp.alloc((size_t)o * v * ladder_ld(P) * 8);
        c->zm_vm.alloc((size_t)std::max<long>(1, o * v * ladder_ld(Ps)) * 8);
        c->zm_P.alloc((size_t)o * v * Q * 8);
        c->zm_M.alloc((size_t)std::max<long>(1, o * v * Qs) * 8);
        launch_rows_vpm(o * v, nv, c->ovvv.d(), c->zm_vp.d(), c->zm_vm.d(), c->st);
        c->wo_vp.alloc((size_t)o * o * ladder_ld(P) * 8);
        c->wo_vm.alloc((size_t)std::max<long>(1, o * o * ladder_ld(Ps)) * 8);
        c->wo_P.alloc((size_t)o * o * Q * 8);
        c->wo_M.alloc((size_t)std::max<long>(1, o * o * Qs) * 8);
        launch_rows_vpm(o * o, nv, c->oovv.d(), c->wo_vp.d(), c->wo_vm.d(), c->st);
      }
    }
    // L = 2<pq|rs> - <pq|sr> for the blocks the equations read (hamiltonian.cc:96-102)
    c->Loovv.alloc(o * o * v * v * 8);
    E.permute("mnef->mnef", 2.0, cref(c->oovv.d(), {no, no, nv, nv}), 0.0, vref(c->Loovv.d(), {no, no, nv, nv}));
    E.permute("mnfe->mnef", -1.0, cref(c->oovv.d(), {no, no, nv, nv}), 1.0, vref(c->Loovv.d(), {no, no, nv, nv}));
    c->WJ0.alloc(o * o * v * v * 8);
    c->WB0.alloc(o * o * v * v * 8);
    E.permute("mbje->mejb", -1.0, cref(c->ovov.d(), {no, nv, no, nv}), 0.0, vref(c->WJ0.d(), {no, nv, no, nv}));
    E.permute("mjeb->mejb", 1.0, cref(c->oovv.d(), {no, no, nv, nv}), 0.0, vref(c->WB0.d(), {no, nv, no, nv}));
    // L<mn|ie> = 2<mn|ie> - <mn|ei> = 2 ooov[m][n][i][e] - ooov[n][m][i][e]
    c->Looov.alloc(o * o * o * v * 8);
    E.permute("mnie->mnie", 2.0, cref(c->ooov.d(), {no, no, no, nv}), 0.0, vref(c->Looov.d(), {no, no, no, nv}));
    E.permute("nmie->mnie", -1.0, cref(c->ooov.d(), {no, no, no, nv}), 1.0, vref(c->Looov.d(), {no, no, no, nv}));
    // L<ma|fe> = 2<ma|fe> - <ma|ef>
    c->Lovvv.alloc(o * v * v * v * 8);
    E.permute("mafe->mafe", 2.0, cref(c->ovvv.d(), {no, nv, nv, nv}), 0.0, vref(c->Lovvv.d(), {no, nv, nv, nv}));
    E.permute("maef->mafe", -1.0, cref(c->ovvv.d(), {no, nv, nv, nv}), 1.0, vref(c->Lovvv.d(), {no, nv, nv, nv}));
    alloc_work(c);
    check(hipStreamSynchronize(c->st), "stream sync");
    c->loaded = true;
    c->have_lambda = false;
    return UGX_OK;
  });
}

int ugx_set_amplitudes(ugx_ctx *c, const double *t1, const double *t2) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!t1 || !t2) throw UgxError(UGX_BAD_INPUT, "ugx_set_amplitudes: null amplitudes");
    check(hipMemcpyAsync(c->t1.p, t1, c->ov() * 8, hipMemcpyHostToDevice, c->st), "t1 copy");
    check(hipMemcpyAsync(c->t2.p, t2, c->oovv_n() * 8, hipMemcpyHostToDevice, c->st), "t2 copy");
    check(hipStreamSynchronize(c->st), "stream sync");
    c->have_amps = true;
    c->tg_valid = false;
    return UGX_OK;
  });
}

int ugx_get_amplitudes(ugx_ctx *c, double *t1, double *t2) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (t1) check(hipMemcpyAsync(t1, c->t1.p, c->ov() * 8, hipMemcpyDeviceToHost, c->st), "t1 copy");
    if (t2) check(hipMemcpyAsync(t2, c->t2.p, c->oovv_n() * 8, hipMemcpyDeviceToHost, c->st), "t2 copy");
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_energy(ugx_ctx *c, int form, double *e) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!e) throw UgxError(UGX_BAD_INPUT, "null output");
    build_tau(c);
    *e = energy(c, form);
    return UGX_OK;
  });
}

int ugx_prepare(ugx_ctx *c, const ugx_options *opt_in) {
  return guarded(c, [&]() {
    require_loaded(c);
    ugx_options opt;
    if (opt_in) opt = *opt_in; else ugx_default_options(&opt);
    validate_options(c, opt, "ugx_prepare");
    prepare_solve(c, opt);
    return UGX_OK;
  });
}

int ugx_ccsd(ugx_ctx *c, const ugx_options *opt_in, ugx_result *res) {
  return guarded(c, [&]() {
    Phase ph("ugx_ccsd");
    require_loaded(c);
    ugx_options opt;
    if (opt_in) opt = *opt_in; else ugx_default_options(&opt);
    validate_options(c, opt, "ugx_ccsd");
    ugx_result r;
    std::memset(&r, 0, sizeof(r));
    Engine &E = *c->eng;
    prepare_solve(c, opt);
    E.reset_counters();
    auto t0 = std::chrono::steady_clock::now();
    if (opt.guess == UGX_GUESS_MP2) guess_mp2(c);
    else if (opt.guess == UGX_GUESS_MLP) guess_mlp(c);
    build_tau(c);
    double emp2 = energy(c, opt.energy_form);
    r.e_mp2 = emp2;
    if (opt.print)
      std::printf("\t Iter   Correlation Energy   T1 Norm      RMS   \n\t  %3d  %20.15f\n", 0, emp2);
    double eccsd = emp2, rms = 0.0;
    int prec = opt.precision;
    diis_reset(c);
    int iter, niter = 0, nlow = 0, low_stall = 0;
    double low_best = 1e300;
    // mode of the next residual; polished: the residuals are FP64-accurate
    // (FP64, or the refinement polish's difference form)
    int mode = RES_PLAIN;
    bool polished = prec == PREC_FP64;
    for (iter = 1; iter <= opt.maxiter; iter++) {
      amp_save(c);
      residual_step(c, mode == RES_BASE ? (int)PREC_FP64 : prec, opt.graph != 0, mode);
      if (mode == RES_BASE) mode = RES_DELTA;
      rms = increment(c, nullptr);
      niter = iter;
      if (!polished) nlow++;
      if (!std::isfinite(rms)) throw UgxError(UGX_NONFINITE, "non-finite residual norm");
      if (rms < opt.r_convergence && polished) {
        // the reported energy is that of the amplitudes this iteration
        // started from (ccwfn.cc:216, 223-224): t1_old and the tau built
        // from them, both still in place
        if (!opt.print) eccsd = energy_of(c, opt.energy_form, c->t1o.d());
        break;
      }
      // low precision has a residual floor (BF16x3: ~1e-6): switch to the
      // FP64 polish when the threshold is reached or the residual stalls
      if (!polished) {
        if (rms < 0.8 * low_best) { low_best = rms; low_stall = 0; }
        else low_stall++;
      }
      if (!polished && (rms < opt.polish_threshold || low_stall >= 3)) {
        // polish: the remaining residuals are FP64-accurate, evaluated in
        // FP64 or (refinement) as the FP64 residual of the next amplitudes
        // plus low-precision differences from them; the subspace built from
        // low-precision residuals is kept (its error vectors are accurate to
        // the low precision) unless asked otherwise
        polished = true;
        if (opt.polish_mode == UGX_POLISH_REFINE) mode = RES_BASE;
        else prec = PREC_FP64;
        if (!opt.polish_keep_diis) diis_reset(c);
      }
      if (opt.diis) diis_step(c, opt.diis_nvec, opt.diis_policy, opt.diis_min_vecs);
      build_tau(c);
      // only printed per iteration; otherwise evaluated once, for the
      // iteration that ends the loop
      if (opt.print) {
        eccsd = energy(c, opt.energy_form);
        std::printf("\t  %3d  %20.15f  %8.6f  %8.6e\n", iter, eccsd, t1norm(c), rms);
      } else if (iter == opt.maxiter) {
        eccsd = energy(c, opt.energy_form);
      }
    }
    build_tau(c);
    check(hipStreamSynchronize(c->st), "stream sync");
    auto t1c = std::chrono::steady_clock::now();
    r.e_ccsd = eccsd;
    r.e_ccsd_final = energy(c, opt.energy_form);
    r.rms = rms;
    r.niter = niter;
    r.niter_low = nlow;
    r.seconds_ccsd = std::chrono::duration<double>(t1c - t0).count();
    r.flops_ccsd = E.flops_total();
    r.device_allocs = (int)std::min<long>(E.take_allocations(), 1L << 30);
    E.collect_stats();
    const bool conv = (rms < opt.r_convergence);
    if (conv && opt.wfn_ccsd_t) {
      auto t2c = std::chrono::steady_clock::now();
      {
        Phase pt("ugx (T)");
        r.e_t = triples(c, opt.precision, nullptr);
      }
      r.seconds_triples = std::chrono::duration<double>(std::chrono::steady_clock::now() - t2c).count();
      E.collect_stats();
    }
    if (res) *res = r;
    c->have_amps = true;
    c->tg_valid = false;
    if (!conv) {
      c->err = "Computation has not converged.";
      return UGX_NOT_CONVERGED;
    }
    return UGX_OK;
  });
}

int ugx_triples(ugx_ctx *c, int precision, double *e_t) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!e_t || precision < 0 || precision > 2) throw UgxError(UGX_BAD_INPUT, "ugx_triples: bad arguments");
    *e_t = triples(c, precision, nullptr);
    return UGX_OK;
  });
}

int ugx_triples_per_triplet(ugx_ctx *c, int precision, double *e) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!e || precision < 0 || precision > 2) throw UgxError(UGX_BAD_INPUT, "bad arguments");
    std::vector<double> per;
    triples(c, precision, &per);
    std::memcpy(e, per.data(), per.size() * 8);
    return UGX_OK;
  });
}

int ugx_w3(ugx_ctx *c, int i, int j, int k, int precision, double *W3) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!W3 || i < 0 || j < 0 || k < 0 || i >= c->no || j >= c->no || k >= c->no || precision < 0 ||
        precision > 2)
      throw UgxError(UGX_BAD_INPUT, "ugx_w3: bad arguments");
    triples_operands(c, precision);
    c->tW_f32 = (precision != PREC_FP64) ? 1 : 0;
    const int T = tw_blk(c->tW_f32);
    const long v = c->nv, ws = tw_stride(c->nv, T), nbk = tw_nb(c->nv, T);
    c->tWbuf.alloc((size_t)ws * 8);
    const TEnt one{c, i, j, k};
    if (precision == PREC_BF16X3 && c->t_dma > 1) triples_build_w_dma(c, &one, 1);
    else if (precision == PREC_BF16X3) triples_build_w_pair(c, &one, 1);
    else triples_build_w(c, precision, &one, 1);
    // blocked layout -> [a][b][c]
    std::vector<double> wb(ws);
    if (c->tW_f32) {
      std::vector<float> wf(ws);
      check(hipMemcpyAsync(wf.data(), c->tWbuf.p, ws * 4, hipMemcpyDeviceToHost, c->st), "W3 copy");
      check(hipStreamSynchronize(c->st), "stream sync");
      for (long x = 0; x < ws; x++) wb[x] = wf[x];
    } else {
      check(hipMemcpyAsync(wb.data(), c->tWbuf.p, ws * 8, hipMemcpyDeviceToHost, c->st), "W3 copy");
      check(hipStreamSynchronize(c->st), "stream sync");
    }
    for (long a = 0; a < v; a++)
      for (long b = 0; b < v; b++)
        for (long cc = 0; cc < v; cc++)
          W3[(a * v + b) * v + cc] = wb[tw_off_a(a, nbk, T) + tw_off_b(b, nbk, T) + tw_off_c(cc, T)];
    return UGX_OK;
  });
}

int ugx_residual(ugx_ctx *c, const double *t1, const double *t2, int precision, double *R1, double *R2) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!t1 || !t2 || precision < 0 || precision > 2) throw UgxError(UGX_BAD_INPUT, "ugx_residual: bad arguments");
    check(hipMemcpyAsync(c->t1o.p, t1, c->ov() * 8, hipMemcpyHostToDevice, c->st), "t1 copy");
    check(hipMemcpyAsync(c->t2o.p, t2, c->oovv_n() * 8, hipMemcpyHostToDevice, c->st), "t2 copy");
    launch_tau(c->no, c->nv, c->t1o.d(), c->t2o.d(), c->tau.d(), c->ttau.d(), c->st);
    residual(c, precision);
    // R2 = X + X^T via the update kernel (its t outputs go to scratch: U and R1 copy)
    DBuf r2;
    r2.alloc(c->oovv_n() * 8);
    DBuf t1s, t2s;
    t1s.alloc(c->ov() * 8);
    t2s.alloc(c->oovv_n() * 8);
    int np = update_nparts(c->no, c->nv);
    launch_update(c->no, c->nv, c->eps.d(), c->R1.d(), c->X.d(), c->t1o.d(), c->t2o.d(), t1s.d(), t2s.d(), r2.d(),
                  c->part.d(), np, c->st);
    if (R1) check(hipMemcpyAsync(R1, c->R1.p, c->ov() * 8, hipMemcpyDeviceToHost, c->st), "R1 copy");
    if (R2) check(hipMemcpyAsync(R2, r2.p, c->oovv_n() * 8, hipMemcpyDeviceToHost, c->st), "R2 copy");
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_intermediates(ugx_ctx *c, double *Fvv, double *Foo, double *Fov, double *Woooo, double *Wovov,
                      double *Wovvo, double *Wovoo) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!c->have_intermediates) throw UgxError(UGX_NOT_READY, "no residual evaluated yet");
    const int o = c->no, v = c->nv;
    const long ovov = (long)o * v * o * v;
    auto get = [&](double *dst, const void *src, long n) {
      if (dst) check(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, c->st), "copy");
    };
    get(Fvv, c->Fvv.p, (long)v * v);
    get(Foo, c->Foo.p, (long)o * o);
    get(Fov, c->Fov.p, (long)o * v);
    get(Woooo, c->Woooo.p, (long)o * o * o * o);
    get(Wovoo, c->Z.p, (long)o * v * o * o);
    DBuf tmp;
    tmp.alloc(ovov * 8);
    if (Wovov) {  // WJ[m][e][j][b] -> Wovov[m][b][j][e]
      c->eng->permute("mejb->mbje", 1.0, vref(c->WJ.d(), {o, v, o, v}), 0.0, vref(tmp.d(), {o, v, o, v}));
      get(Wovov, tmp.p, ovov);
      check(hipStreamSynchronize(c->st), "stream sync");
    }
    if (Wovvo) {  // WB[m][e][j][b] -> Wovvo[m][b][e][j]
      c->eng->permute("mejb->mbej", 1.0, vref(c->WB.d(), {o, v, o, v}), 0.0, vref(tmp.d(), {o, v, v, o}));
      get(Wovvo, tmp.p, ovov);
    }
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_hbar(ugx_ctx *c, double *Hov, double *Hoo, double *Hvv, double *Hoooo, double *Hvvvv, double *Hvovv,
             double *Hooov, double *Hovvo, double *Hovov, double *Hvvvo, double *Hovoo) {
  return guarded(c, [&]() {
    Phase ph("ugx_hbar");
    require_loaded(c);
    hbar(c);
    const long o = c->no, v = c->nv;
    auto get = [&](double *dst, const DBuf &src, long n) {
      if (dst) check(hipMemcpyAsync(dst, src.p, n * 8, hipMemcpyDeviceToHost, c->st), "HBAR copy");
    };
    get(Hov, c->hb_ov, o * v);
    get(Hoo, c->hb_oo, o * o);
    get(Hvv, c->hb_vv, v * v);
    get(Hoooo, c->hb_oooo, o * o * o * o);
    get(Hvvvv, c->hb_vvvv, v * v * v * v);
    get(Hvovv, c->hb_vovv, v * o * v * v);
    get(Hooov, c->hb_ooov, o * o * o * v);
    get(Hovvo, c->hb_ovvo, o * v * v * o);
    get(Hovov, c->hb_ovov, o * v * o * v);
    get(Hvvvo, c->hb_vvvo, v * v * v * o);
    get(Hovoo, c->hb_ovoo, o * v * o * o);
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_lambda(ugx_ctx *c, const ugx_options *opt_in, double *pseudoenergy, int *niter, double *rms,
               double *trace) {
  return guarded(c, [&]() {
    Phase ph("ugx_lambda");
    require_loaded(c);
    if (!c->have_amps) throw UgxError(UGX_NOT_READY, "ugx_lambda: no amplitudes (run ugx_ccsd first)");
    ugx_options opt;
    if (opt_in) opt = *opt_in;
    else ugx_default_options(&opt);
    validate_options(c, opt, "ugx_lambda");
    // WFN=CCSD_T with DERTYPE=FIRST adds the (T) sources (cclambda.cc:181,
    // :243): here, when the caller has run ugx_tgrad on these amplitudes
    c->lam_with_t = opt.wfn_ccsd_t && c->tg_valid;
    double pe = 0.0, r = 0.0;
    int n = 0;
    const bool conv = lambda_solve(c, opt, &pe, &n, &r, trace);
    if (pseudoenergy) *pseudoenergy = pe;
    if (niter) *niter = n;
    if (rms) *rms = r;
    if (!conv) {
      c->err = "Computation has not converged.";
      return UGX_NOT_CONVERGED;
    }
    return UGX_OK;
  });
}

int ugx_get_lambda(ugx_ctx *c, double *l1, double *l2) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!c->have_lambda) throw UgxError(UGX_NOT_READY, "ugx_get_lambda: no Lambda amplitudes (run ugx_lambda)");
    if (l1) check(hipMemcpyAsync(l1, c->lam_l1.p, c->ov() * 8, hipMemcpyDeviceToHost, c->st), "l1 copy");
    if (l2) check(hipMemcpyAsync(l2, c->lam_l2.p, c->oovv_n() * 8, hipMemcpyDeviceToHost, c->st), "l2 copy");
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_lambda_residual(ugx_ctx *c, const double *l1, const double *l2, double *R1, double *R2, double *Goo,
                        double *Gvv) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!c->have_amps) throw UgxError(UGX_NOT_READY, "ugx_lambda_residual: no amplitudes");
    if (!l1 || !l2) throw UgxError(UGX_BAD_INPUT, "ugx_lambda_residual: null l1/l2");
    hbar(c);
    lambda_alloc(c);
    const long ov = c->ov(), oovv = c->oovv_n();
    // the arguments go into the "old" buffers: l1/l2 of a finished solve stay
    check(hipMemcpyAsync(c->lam_l1o.p, l1, ov * 8, hipMemcpyHostToDevice, c->st), "l1 copy");
    check(hipMemcpyAsync(c->lam_l2o.p, l2, oovv * 8, hipMemcpyHostToDevice, c->st), "l2 copy");
    lambda_residual(c, c->lam_l1o.d(), c->lam_l2o.d(), c->lam_R1.d(), c->lam_Z.d(), c->tg_valid);
    if (R1) check(hipMemcpyAsync(R1, c->lam_R1.p, ov * 8, hipMemcpyDeviceToHost, c->st), "R1 copy");
    if (R2) {
      // R2 = Z + Z^T(ij,ab), formed in lam_l2o (its input copy is consumed)
      const int o = c->no, v = c->nv;
      c->eng->permute("jiba->ijab", 1.0, vref(c->lam_Z.d(), {o, o, v, v}), 0.0, vref(c->lam_l2o.d(), {o, o, v, v}));
      launch_axpby(oovv, 1.0, c->lam_Z.d(), 1.0, c->lam_l2o.d(), c->st);
      check(hipMemcpyAsync(R2, c->lam_l2o.p, oovv * 8, hipMemcpyDeviceToHost, c->st), "R2 copy");
    }
    if (Goo) check(hipMemcpyAsync(Goo, c->lam_Goo.p, (size_t)c->no * c->no * 8, hipMemcpyDeviceToHost, c->st), "Goo");
    if (Gvv) check(hipMemcpyAsync(Gvv, c->lam_Gvv.p, (size_t)c->nv * c->nv * 8, hipMemcpyDeviceToHost, c->st), "Gvv");
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

namespace {
// ugx_tgrad (ooc = 0) and ugx_tgrad_ooc (ooc = 1): tgrad_ooc
// (ccwfn.cc:1150-1316) gives tgrad's terms except that the (T) Doo and Dvv
// keep their diagonals only, for canonical Fock matrices; with f_ov != 0 its
// s2 takes another disconnected term than tgrad's l3, a difference linear in
// f_ov: f_ov beyond a converged SCF's residual (1e-8) is refused.
int tgrad_entry(ugx_ctx *c, int ooc, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
                double *Gvvvo, double *Goovv) {
  return guarded(c, [&]() {
    Phase ph("ugx_tgrad");
    require_loaded(c);
    if (!c->have_amps) throw UgxError(UGX_NOT_READY, "ugx_tgrad: no amplitudes (run ugx_ccsd first)");
    if (ooc) {
      double m = 0.0;
      for (double x : c->h_fov) m = std::max(m, std::fabs(x));
      if (m > 1e-8)
        throw UgxError(UGX_BAD_INPUT, "ugx_tgrad_ooc: the OOC=true (T) gradient terms are reproduced for canonical "
                                      "orbitals (|f_ov| <= 1e-8); largest |f_ov| is " + std::to_string(m));
    }
    c->tg_valid = false;
    tgrad(c);
    const long o = c->no, v = c->nv;
    if (ooc) {
      // ccwfn.cc:1197 and :1242 accumulate Dvv[a][a] and Doo[i][i] only
      for (auto ov : {std::make_pair(&c->tg_oo, o), std::make_pair(&c->tg_vv, v)}) {
        std::vector<double> h((size_t)ov.second * ov.second);
        check(hipMemcpyAsync(h.data(), ov.first->p, h.size() * 8, hipMemcpyDeviceToHost, c->st), "tgrad copy");
        check(hipStreamSynchronize(c->st), "stream sync");
        for (long p = 0; p < ov.second; p++)
          for (long q = 0; q < ov.second; q++)
            if (p != q) h[(size_t)p * ov.second + q] = 0.0;
        check(hipMemcpyAsync(ov.first->p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->st), "tgrad copy");
        check(hipStreamSynchronize(c->st), "stream sync");
      }
    }
    c->tg_valid = true;
    auto get = [&](double *dst, const DBuf &src, long n) {
      if (dst) check(hipMemcpyAsync(dst, src.p, n * 8, hipMemcpyDeviceToHost, c->st), "tgrad copy");
    };
    get(s1, c->tg_s1, o * v);
    get(s2, c->tg_s2, o * o * v * v);
    get(Doo, c->tg_oo, o * o);
    get(Dvv, c->tg_vv, v * v);
    get(Dov, c->tg_ov, o * v);
    get(Gooov, c->tg_ooov, o * o * o * v);
    get(Gvvvo, c->tg_vvvo, v * v * v * o);
    get(Goovv, c->tg_oovv, o * o * v * v);
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}
}  // namespace

int ugx_tgrad(ugx_ctx *c, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
              double *Gvvvo, double *Goovv) {
  return tgrad_entry(c, 0, s1, s2, Doo, Dvv, Dov, Gooov, Gvvvo, Goovv);
}

int ugx_tgrad_ooc(ugx_ctx *c, double *s1, double *s2, double *Doo, double *Dvv, double *Dov, double *Gooov,
                  double *Gvvvo, double *Goovv) {
  return tgrad_entry(c, 1, s1, s2, Doo, Dvv, Dov, Gooov, Gvvvo, Goovv);
}

namespace {
PertSlot &pert_slot(ugx_ctx *c, int h) {
  if (h < 0 || h >= (int)c->perts.size() || !c->perts[h])
    throw UgxError(UGX_BAD_INPUT, "no perturbation with handle " + std::to_string(h));
  return *c->perts[h];
}
}  // namespace

int ugx_pert_create(ugx_ctx *c, const double *pert, double omega, int *handle) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!c->have_lambda) throw UgxError(UGX_NOT_READY, "ugx_pert_create: no Lambda amplitudes (run ugx_lambda)");
    if (!pert || !handle) throw UgxError(UGX_BAD_INPUT, "ugx_pert_create: null argument");
    const int o = c->no, v = c->nv, n = o + v;
    const long ov = c->ov(), oovv = c->oovv_n();
    std::unique_ptr<PertSlot> S(new PertSlot());
    PertSlot &P = *S;
    P.omega = omega;
    P.make("P", (size_t)n * n);
    P.make("Poo", (size_t)o * o); P.make("Pov", ov); P.make("Pvo", ov); P.make("Pvv", (size_t)v * v);
    P.make("Aoo", (size_t)o * o); P.make("Avv", (size_t)v * v); P.make("Avo", ov);
    P.make("Aovoo", (size_t)o * ov * o); P.make("Avvvo", (size_t)v * v * ov); P.make("Avvoo", oovv);
    for (const char *x : {"X1", "X1old", "Y1", "Y1old"}) P.make(x, ov);
    for (const char *x : {"X2", "X2old", "Y2", "Y2old"}) P.make(x, oovv);
    check(hipMemcpyAsync(P.buf("P"), pert, (size_t)n * n * 8, hipMemcpyHostToDevice, c->st), "perturbation copy");
    auto block = [&](const char *dst, int r0, int c0, int nr, int nc) {
      check(hipMemcpy2DAsync(P.buf(dst), (size_t)nc * 8, P.buf("P") + (size_t)r0 * n + c0, (size_t)n * 8,
                             (size_t)nc * 8, nr, hipMemcpyDeviceToDevice, c->st),
            "perturbation block");
    };
    block("Poo", 0, 0, o, o);
    block("Pov", 0, o, o, v);
    block("Pvo", o, 0, v, o);
    block("Pvv", o, o, v, v);
    run_terms(c, P, kPertBar, (int)(sizeof(kPertBar) / sizeof(kPertBar[0])));
    // guesses X = Y = Abar / (D + omega) (ccpert.cc:58-70): one update from zero
    pert_alloc_scratch(c);
    pert_denominators(c);
    c->eng->permute("ai->ia", 1.0, vref(P.buf("Avo"), {v, o}), 0.0, vref(c->rp_R1.d(), {o, v}));
    c->eng->permute("abij->ijab", 1.0, vref(P.buf("Avvoo"), {v, v, o, o}), 0.0, vref(c->rp_Z.d(), {o, o, v, v}));
    launch_fill(ov, 0.0, P.buf("X1old"), c->st);
    launch_fill(oovv, 0.0, P.buf("X2old"), c->st);
    const int np = update_nparts(o, v);
    launch_update_shift(o, v, c->rp_eps.d(), omega, c->rp_R1.d(), c->rp_Z.d(), P.buf("X1old"), P.buf("X2old"),
                        P.buf("X1"), P.buf("X2"), c->part.d(), np, c->st);
    check(hipMemcpyAsync(P.buf("Y1"), P.buf("X1"), ov * 8, hipMemcpyDeviceToDevice, c->st), "Y1 guess");
    check(hipMemcpyAsync(P.buf("Y2"), P.buf("X2"), oovv * 8, hipMemcpyDeviceToDevice, c->st), "Y2 guess");
    check(hipStreamSynchronize(c->st), "stream sync");
    int h = -1;
    for (int k = 0; k < (int)c->perts.size(); k++)
      if (!c->perts[k]) { h = k; break; }
    if (h < 0) { h = (int)c->perts.size(); c->perts.emplace_back(); }
    c->perts[h] = std::move(S);
    *handle = h;
    return UGX_OK;
  });
}

int ugx_pert_solve(ugx_ctx *c, int handle, int hand, const ugx_options *opt_in, double *pseudoresponse, int *niter,
                   double *rms) {
  return guarded(c, [&]() {
    Phase ph("ugx_pert_solve");
    require_loaded(c);
    PertSlot &P = pert_slot(c, handle);
    if (hand != 0 && hand != 1) throw UgxError(UGX_BAD_INPUT, "ugx_pert_solve: hand is 0 (right) or 1 (left)");
    if (hand == 1 && !P.x_solved)
      throw UgxError(UGX_NOT_READY, "ugx_pert_solve: the left-hand equations need the right-hand amplitudes");
    ugx_options opt;
    if (opt_in) opt = *opt_in;
    else ugx_default_options(&opt);
    validate_options(c, opt, "ugx_pert_solve");
    double pr = 0.0, r = 0.0;
    int n = 0;
    const bool conv = pert_solve(c, P, hand, opt, &pr, &n, &r);
    if (hand == 0) P.x_solved = conv;
    if (pseudoresponse) *pseudoresponse = pr;
    if (niter) *niter = n;
    if (rms) *rms = r;
    if (!conv) {
      c->err = "Computation has not converged.";
      return UGX_NOT_CONVERGED;
    }
    return UGX_OK;
  });
}

int ugx_pert_amplitudes(ugx_ctx *c, int handle, double *X1, double *X2, double *Y1, double *Y2) {
  return guarded(c, [&]() {
    require_loaded(c);
    PertSlot &P = pert_slot(c, handle);
    const long ov = c->ov(), oovv = c->oovv_n();
    auto get = [&](double *dst, const char *n, long cnt) {
      if (dst) check(hipMemcpyAsync(dst, P.buf(n), cnt * 8, hipMemcpyDeviceToHost, c->st), "perturbed amplitudes");
    };
    get(X1, "X1", ov); get(X2, "X2", oovv); get(Y1, "Y1", ov); get(Y2, "Y2", oovv);
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_pert_residual(ugx_ctx *c, int handle, int hand, const double *A1, const double *A2, double *R1, double *R2) {
  return guarded(c, [&]() {
    require_loaded(c);
    PertSlot &P = pert_slot(c, handle);
    if (!A1 || !A2) throw UgxError(UGX_BAD_INPUT, "ugx_pert_residual: null amplitudes");
    if (hand != 0 && hand != 1) throw UgxError(UGX_BAD_INPUT, "ugx_pert_residual: hand is 0 (right) or 1 (left)");
    const int o = c->no, v = c->nv;
    const long ov = c->ov(), oovv = c->oovv_n();
    pert_alloc_scratch(c);
    c->rp_A1.alloc(ov * 8);
    c->rp_A2.alloc(oovv * 8);
    check(hipMemcpyAsync(c->rp_A1.p, A1, ov * 8, hipMemcpyHostToDevice, c->st), "amplitude copy");
    check(hipMemcpyAsync(c->rp_A2.p, A2, oovv * 8, hipMemcpyHostToDevice, c->st), "amplitude copy");
    if (hand == 1) pert_y_source(c, P);
    pert_residual(c, P, hand, c->rp_A1.d(), c->rp_A2.d());
    if (R1) check(hipMemcpyAsync(R1, c->rp_R1.p, ov * 8, hipMemcpyDeviceToHost, c->st), "R1 copy");
    if (R2) {
      c->eng->permute("jiba->ijab", 1.0, vref(c->rp_Z.d(), {o, o, v, v}), 0.0, vref(c->rp_W.d(), {o, o, v, v}));
      launch_axpby(oovv, 1.0, c->rp_Z.d(), 1.0, c->rp_W.d(), c->st);
      check(hipMemcpyAsync(R2, c->rp_W.p, oovv * 8, hipMemcpyDeviceToHost, c->st), "R2 copy");
    }
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_linresp(ugx_ctx *c, int hx, int hy, double *value) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (!value) throw UgxError(UGX_BAD_INPUT, "ugx_linresp: null value");
    PertSlot &Px = pert_slot(c, hx), &Py = pert_slot(c, hy);
    *value = pert_linresp(c, Px, Py);
    return UGX_OK;
  });
}

int ugx_pert_release(ugx_ctx *c, int handle) {
  return guarded(c, [&]() {
    (void)pert_slot(c, handle);
    c->perts[handle].reset();
    return UGX_OK;
  });
}

int ugx_density(ugx_ctx *c, double *energies, double *Doo, double *Dvv, double *Dov, double *Dvo, double *Goooo,
                double *Gvvvv, double *Gooov, double *Gvvvo, double *Govov, double *Goovv) {
  return guarded(c, [&]() {
    Phase ph("ugx_density");
    require_loaded(c);
    if (!c->have_lambda) throw UgxError(UGX_NOT_READY, "ugx_density: no Lambda amplitudes (run ugx_lambda)");
    density(c);
    double E[8];
    density_energies(c, E);
    if (energies)
      for (int k = 0; k < 8; k++) energies[k] = E[k];
    const long o = c->no, v = c->nv;
    auto get = [&](double *dst, const DBuf &src, long n) {
      if (dst) check(hipMemcpyAsync(dst, src.p, n * 8, hipMemcpyDeviceToHost, c->st), "density copy");
    };
    get(Doo, c->dn_oo, o * o);
    get(Dvv, c->dn_vv, v * v);
    get(Dov, c->dn_ov, o * v);
    get(Dvo, c->dn_vo, v * o);
    get(Goooo, c->dn_oooo, o * o * o * o);
    get(Gvvvv, c->dn_vvvv, v * v * v * v);
    get(Gooov, c->dn_ooov, o * o * o * v);
    get(Gvvvo, c->dn_vvvo, v * v * v * o);
    get(Govov, c->dn_ovov, o * v * o * v);
    get(Goovv, c->dn_oovv, o * o * v * v);
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

void ugx_set_profiling(ugx_ctx *c, int on) {
  if (c && c->eng) c->eng->set_profiling(on != 0);
}

size_t ugx_kernel_stats(ugx_ctx *c, char *buf, size_t buflen) {
  if (!c || !c->eng) return 0;
  std::string s;
  try {
    c->eng->collect_stats();
    char line[256];
    for (auto &kv : c->eng->stats()) {
      std::snprintf(line, sizeof(line), "%s %ld %.6f %.6f\n", kv.first.c_str(), kv.second.launches, kv.second.ms,
                    kv.second.flops * 1e-9);
      s += line;
    }
  } catch (...) {
    return 0;
  }
  if (buf && buflen) {
    size_t n = std::min(buflen - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return s.size() + 1;
}

int ugx_gemm(ugx_ctx *c, char opa, char opb, int M, int N, int K, double alpha, const void *A, long lda,
             const void *B, long ldb, double beta, double *C, long ldc, int precision, int tile) {
  return guarded(c, [&]() {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C || precision < 0 || precision > UGX_PREC_BF16X3_PANEL ||
        precision == UGX_PREC_BF16 || tile >= gemm_num_tiles())
      throw UgxError(UGX_BAD_INPUT, "ugx_gemm: bad arguments");
    if (precision == UGX_PREC_BF16X3_PAIR && !((opa == 'N' || opa == 'n') && (opb == 'T' || opb == 't')))
      throw UgxError(UGX_BAD_INPUT, "ugx_gemm: BF16 pair operands must be A[m][k] and B[n][k] (opa N, opb T)");
    GemmDesc d{};
    d.M = M; d.N = N; d.nseg = 1; d.batch = 1;
    GemmSeg &g = d.seg[0];
    g.A = A; g.B = B; g.K = K; g.sign = 1.0;
    const bool an = (opa == 'N' || opa == 'n'), bn = (opb == 'N' || opb == 'n');
    if (an) { g.sam = lda; g.sak = 1; } else { g.sam = 1; g.sak = lda; }
    if (bn) { g.sbk = ldb; g.sbn = 1; } else { g.sbk = 1; g.sbn = ldb; }
    d.C = C; d.scm = ldc; d.scn = 1; d.alpha = alpha; d.beta = beta;
    if (precision == UGX_PREC_BF16X3_PAIR) {
      g.loA = (long)M * lda; g.loB = (long)N * ldb;
      launch_gemm(d, nullptr, PREC_BF16X3_PRESPLIT, tile, c->st);
    } else if (precision == UGX_PREC_BF16X3_PANEL) {
      // FP64 operands in any layout packed into BF16 hi/lo MFMA panels
      // (A as [m][k], B as [n][k]), then the LDS-copy kernel of the (T) W3
      // build (tiles whose block copies do not split over the four waves
      // take the register-staged kernel on the same panels)
      if (M == 0 || N == 0) return UGX_OK;
      const long Kp = (K + 31) / 32 * 32;
      DBuf pa, pb;
      pa.alloc((size_t)panel_size(1, M, K) * 4);
      pb.alloc((size_t)panel_size(1, N, K) * 4);
      launch_pack_panel((const double *)A, 0, an ? lda : 1, an ? 1 : lda, 1, M, K, (unsigned short *)pa.p, c->st);
      launch_pack_panel((const double *)B, 0, bn ? 1 : ldb, bn ? ldb : 1, 1, N, K, (unsigned short *)pb.p, c->st);
      g.A = pa.p; g.B = pb.p;
      g.sam = BK_PANEL; g.sak = 1; g.sbn = BK_PANEL; g.sbk = 1;
      g.loA = panel_size(1, M, K); g.loB = panel_size(1, N, K);
      g.pblkA = 16 * Kp; g.pblkB = 16 * Kp;
      d.dma = 2;
      d.vec_ok = 1;
      launch_gemm(d, nullptr, PREC_BF16X3_PRESPLIT, tile, c->st);
      check(hipStreamSynchronize(c->st), "stream sync");
    } else if (precision == UGX_PREC_BF16X3) {
      // FP32 operands in any layout: split into BF16 hi/lo pairs A[m][k],
      // B[n][k] in two staging buffers first
      DBuf pa, pb;
      pa.alloc(std::max<size_t>(1, (size_t)M * K) * 4);
      pb.alloc(std::max<size_t>(1, (size_t)N * K) * 4);
      gemm_bf16x3_split(c, d, pa.p, pb.p, tile);
      check(hipStreamSynchronize(c->st), "stream sync");
    } else {
      launch_gemm(d, nullptr, precision, tile, c->st);
    }
    check(hipGetLastError(), "gemm launch");
    return UGX_OK;
  });
}

int ugx_gemm_bf16x3_static(ugx_ctx *c, char opa, char opb, int M, int N, int K, double alpha, const float *A,
                           long lda, const float *B, long ldb, double beta, double *C, long ldc, int tile) {
  return guarded(c, [&]() {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C || tile >= gemm_num_tiles())
      throw UgxError(UGX_BAD_INPUT, "ugx_gemm_bf16x3_static: bad arguments");
    // pairs of A from offset 0, pairs of B from the next 256-byte boundary
    const size_t bytes_a = ((size_t)M * K * 4 + 255) / 256 * 256, bytes_b = (size_t)N * K * 4;
    if (bytes_a + bytes_b > kStaticScratchBytes)
      throw UgxError(UGX_BAD_INPUT, "ugx_gemm_bf16x3_static: operands need " + std::to_string(bytes_a + bytes_b) +
                                        " bytes of staging; the static scratch holds " +
                                        std::to_string(kStaticScratchBytes) + " (64 MiB): (M + N) K <= 16 Mi");
    unsigned char *base = (unsigned char *)static_scratch(c->device);
    if (!base) throw UgxError(UGX_HIP_ERROR, "ugx_gemm_bf16x3_static: no static scratch on this device");
    GemmDesc d{};
    d.M = M; d.N = N; d.nseg = 1; d.batch = 1;
    GemmSeg &g = d.seg[0];
    g.A = A; g.B = B; g.K = K; g.sign = 1.0;
    const bool an = (opa == 'N' || opa == 'n'), bn = (opb == 'N' || opb == 'n');
    if (an) { g.sam = lda; g.sak = 1; } else { g.sam = 1; g.sak = lda; }
    if (bn) { g.sbk = ldb; g.sbn = 1; } else { g.sbk = 1; g.sbn = ldb; }
    d.C = C; d.scm = ldc; d.scn = 1; d.alpha = alpha; d.beta = beta;
    gemm_bf16x3_split(c, d, base, base + bytes_a, tile);
    check(hipGetLastError(), "gemm launch");
    return UGX_OK;
  });
}

int ugx_synchronize(ugx_ctx *c) {
  return guarded(c, [&]() {
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_batch_solve(ugx_ctx *c, int nprob, const ugx_problem *probs, const ugx_options *opt_in, int nworkers,
                    ugx_result *results, double *wall_seconds) {
  return guarded(c, [&]() {
    Phase ph("ugx_batch_solve");
    if (nprob <= 0 || !probs || !results || nworkers < 1 || nworkers > 64)
      throw UgxError(UGX_BAD_INPUT, "ugx_batch_solve: bad arguments");
    for (int p = 0; p < nprob; p++) {
      const ugx_problem &q = probs[p];
      if (q.no <= 0 || q.nv <= 0 || !q.fock || !q.oooo || !q.ooov || !q.oovv || !q.ovov || !q.ovvv || !q.vvvv)
        throw UgxError(UGX_BAD_INPUT, "ugx_batch_solve: problem " + std::to_string(p) + " is incomplete");
    }
    ugx_options opt;
    if (opt_in) opt = *opt_in; else ugx_default_options(&opt);
    if (opt.guess == UGX_GUESS_USER) throw UgxError(UGX_BAD_INPUT, "ugx_batch_solve: UGX_GUESS_USER is per context");
    const bool want_t = opt.wfn_ccsd_t != 0;
    opt.wfn_ccsd_t = 0;  // (T) runs batched per size class below
    auto t_start = std::chrono::steady_clock::now();
    // one context (stream, engine, buffers) per problem
    std::vector<std::unique_ptr<ugx_ctx, void (*)(ugx_ctx *)>> sub;
    sub.reserve(nprob);
    for (int p = 0; p < nprob; p++) {
      ugx_ctx *s = nullptr;
      int rc = ugx_create(&s, c->device);
      if (rc != UGX_OK) throw UgxError(rc, "ugx_batch_solve: context creation failed");
      sub.emplace_back(s, ugx_destroy);
      s->eng->set_autotune(c->eng->autotune());
      s->eng->set_profiling(c->eng->profiling());
      s->mlp = c->mlp;
      s->t_batch = c->t_batch;
      s->t_zgroup = c->t_zgroup;
      s->t_dma = c->t_dma;
      s->t_sort = c->t_sort;
      s->t_align = c->t_align;
      s->t_wbytes = c->t_wbytes;
      s->lad_full = c->lad_full;
    }
    // CCSD: workers take problems largest first (o^2 v^4 ladder cost)
    std::vector<int> order(nprob);
    for (int p = 0; p < nprob; p++) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      auto cost = [&](int p) { return (double)probs[p].no * probs[p].no * std::pow((double)probs[p].nv, 4); };
      return cost(a) > cost(b);
    });
    std::vector<int> rc(nprob, UGX_OK);
    std::vector<double> t_ccsd(nprob, 0.0), t_load(nprob, 0.0);
    std::atomic<int> next(0);
    // Phase 1 loads and prepares every problem (all device allocation of
    // the CCSD happens here); phase 2 iterates and allocates nothing, so no
    // worker allocates or frees device memory while another captures or
    // replays its graphs.
    auto loader = [&]() {
      for (;;) {
        const int q = next.fetch_add(1);
        if (q >= nprob) break;
        const int p = order[q];
        const ugx_problem &pr = probs[p];
        auto t0 = std::chrono::steady_clock::now();
        int r = ugx_load_hamiltonian(sub[p].get(), pr.no, pr.nv, pr.fock, pr.oooo, pr.ooov, pr.oovv, pr.ovov,
                                     pr.ovvv, pr.vvvv);
        if (r == UGX_OK) r = ugx_prepare(sub[p].get(), &opt);
        t_load[p] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rc[p] = r;
      }
    };
    auto solver = [&]() {
      for (;;) {
        const int q = next.fetch_add(1);
        if (q >= nprob) break;
        const int p = order[q];
        if (rc[p] != UGX_OK) continue;
        auto t0 = std::chrono::steady_clock::now();
        rc[p] = ugx_ccsd(sub[p].get(), &opt, &results[p]);
        t_ccsd[p] = t_load[p] + std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
    };
    const int nthr = std::min(nworkers, nprob);
    auto run_pool = [&](const std::function<void()> &fn) {
      std::vector<std::thread> pool;
      next.store(0);
      try {
        for (int w = 0; w < nthr; w++) pool.emplace_back(fn);
      } catch (...) {
        next.store(nprob);  // the started workers stop after their current problem
        for (auto &th : pool) th.join();
        throw;
      }
      for (auto &th : pool) th.join();
    };
    run_pool(loader);
    run_pool(solver);
    int status = UGX_OK;
    for (int p = 0; p < nprob; p++) {
      if (rc[p] == UGX_NOT_CONVERGED) status = UGX_NOT_CONVERGED;
      else if (rc[p] != UGX_OK)
        throw UgxError(rc[p], "ugx_batch_solve: problem " + std::to_string(p) + ": " + sub[p]->err);
      results[p].seconds_ccsd = t_ccsd[p];
      results[p].e_t = 0.0;
      results[p].seconds_triples = 0.0;
    }
    // (T): one batched computation per (no, nv) class over every triplet of
    // every converged problem of the class, on the class lead's stream
    if (want_t) {
      std::map<std::pair<int, int>, std::vector<int>> classes;
      for (int p = 0; p < nprob; p++)
        if (rc[p] == UGX_OK) classes[{probs[p].no, probs[p].nv}].push_back(p);
      for (auto &kv : classes) {
        const std::vector<int> &mem = kv.second;
        auto t0 = std::chrono::steady_clock::now();
        for (int p : mem) triples_operands(sub[p].get(), opt.precision);
        for (int p : mem) check(hipStreamSynchronize(sub[p]->st), "stream sync");
        ugx_ctx *lead = sub[mem[0]].get();
        std::vector<TEnt> ents;
        std::vector<int> prob;
        std::vector<TripProb> tp;
        for (size_t q = 0; q < mem.size(); q++) {
          ugx_ctx *s = sub[mem[q]].get();
          tp.push_back(TripProb{s->eps.d(), s->oovv.d(), s->t1.d(), nullptr, nullptr});
          const std::vector<int> &tr = s->trip_host;
          for (size_t t = 0; t + 2 < tr.size(); t += 3) {
            ents.push_back(TEnt{s, tr[t], tr[t + 1], tr[t + 2]});
            prob.push_back((int)q);
          }
        }
        std::vector<double> e = triples_entries(lead, opt.precision, ents, prob, tp);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (size_t t = 0; t < e.size(); t++) results[mem[prob[t]]].e_t += e[t];
        for (int p : mem) results[p].seconds_triples = dt / (double)mem.size();
      }
    }
    if (wall_seconds)
      *wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    // profiling on the calling context: its statistics gain every
    // problem's device time per label (ugx_kernel_stats)
    if (c->eng->profiling()) {
      c->eng->reset_counters();
      for (auto &s : sub) {
        s->eng->collect_stats();
        c->eng->merge_stats(s->eng->stats());
      }
    }
    if (status == UGX_NOT_CONVERGED) c->err = "Computation has not converged.";
    return status;
  });
}

int ugx_set_autotune(ugx_ctx *c, int on) {
  return guarded(c, [&]() {
    c->eng->set_autotune(on != 0);
    return UGX_OK;
  });
}

int ugx_save_tile_table(ugx_ctx *c, const char *path) {
  return guarded(c, [&]() {
    if (!path) throw UgxError(UGX_BAD_INPUT, "ugx_save_tile_table: null path");
    if (!tiles_save(path)) throw UgxError(UGX_BAD_INPUT, std::string("cannot write the tile table to ") + path);
    return UGX_OK;
  });
}

int ugx_load_tile_table(ugx_ctx *c, const char *path) {
  return guarded(c, [&]() {
    if (!path) throw UgxError(UGX_BAD_INPUT, "ugx_load_tile_table: null path");
    if (tiles_load(path) < 0) throw UgxError(UGX_BAD_INPUT, std::string("cannot read the tile table ") + path);
    return UGX_OK;
  });
}

size_t ugx_tile_table(ugx_ctx *c, char *buf, size_t buflen) {
  if (!c) return 0;
  std::string s = c->eng->tile_table();
  if (buf && buflen) {
    size_t n = std::min(buflen - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return s.size() + 1;
}

int ugx_features(ugx_ctx *c, long n, const long *idx, float *X) {
  return guarded(c, [&]() {
    require_loaded(c);
    if (n <= 0 || !X) throw UgxError(UGX_BAD_INPUT, "ugx_features: bad arguments");
    const long total = c->oovv_n();
    DBuf di, dx;
    if (idx) {
      for (long r = 0; r < n; r++)
        if (idx[r] < 0 || idx[r] >= total) throw UgxError(UGX_BAD_INPUT, "ugx_features: index out of range");
      di.alloc((size_t)n * 8);
      check(hipMemcpyAsync(di.p, idx, (size_t)n * 8, hipMemcpyHostToDevice, c->st), "index copy");
    } else if (n > total) {
      throw UgxError(UGX_BAD_INPUT, "ugx_features: more rows than amplitudes");
    }
    dx.alloc((size_t)n * MLP_NFEAT * 4);
    launch_features(c->no, c->nv, c->eps.d(), c->oovv.d(), n, idx ? (const long *)di.p : nullptr, dx.f(), c->st);
    check(hipMemcpyAsync(X, dx.p, (size_t)n * MLP_NFEAT * 4, hipMemcpyDeviceToHost, c->st), "features copy");
    check(hipStreamSynchronize(c->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_set_guess_network(ugx_ctx *c, const ugx_mlp *net) {
  return guarded(c, [&]() {
    MlpDevice P;
    double ys;
    if (net && !mlp_inference_view(net, c->device, &P, &ys))
      throw UgxError(UGX_BAD_INPUT, "ugx_set_guess_network: the network lives on another device");
    c->mlp = net;
    return UGX_OK;
  });
}

}  // extern "C"
