// Internal declarations shared by the gfx950 kernels, the contraction engine
// and the C-ABI layer.  Nothing here crosses the library boundary.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace ugx {

// Arithmetic of a dense contraction.
enum Prec : int {
  PREC_FP64 = 0,    // FP64 operands, v_mfma_f64_16x16x4_f64, FP64 accumulate
  PREC_FP32 = 1,    // FP32 operands, v_mfma_f32_16x16x4_f32, FP32 accumulate
  PREC_BF16X3 = 2,  // FP32 operands split hi+lo into BF16, three
                    // v_mfma_f32_16x16x32_bf16 products (hi*hi + hi*lo +
                    // lo*hi), FP32 accumulate
  // Same arithmetic on operands that are already split into BF16 hi/lo
  // arrays with k contiguous (A[m][k], B[n][k]); the form the engine and the
  // (T) build use.  Internal launch mode, not an ABI precision.
  PREC_BF16X3_PRESPLIT = 3,
};

constexpr int GEMM_MAXSEG = 8;
constexpr int BK_PANEL = 32;  // k depth of one MFMA panel slice (GemmSeg::pblkA/B)

// One K segment of a contraction C(m,n) += sign * sum_k A(m,k) B(k,n).
//   A(m,k) = A[m*sam + k*sak]       B(k,n) = B[k*sbk + n*sbn]
// Element type is double for PREC_FP64 and float otherwise.
struct GemmSeg {
  const void *A;
  const void *B;
  long sam, sak, sbk, sbn;
  long sAz, sBz;  // per-batch offsets (elements) when no pointer table
  int K;
  double sign;
  // PREC_BF16X3_PRESPLIT: element offsets from the hi halves to the lo
  // halves of A and B (operands stored as two BF16 arrays, k contiguous)
  long loA, loB;
  // PREC_BF16X3_PRESPLIT: > 0 marks an operand packed as MFMA panels (see
  // launch_pack_panel): element (r, k) at (r/16)*pblk + (k/32)*512 +
  // (r%16)*32 + 8*panel_chunk(r%16, (k%32)/8) + k%8, pblk = 16 * K rounded
  // up to 32; its sam / sbn are then ignored.  0 = row-major with stride
  // sam / sbn.
  long pblkA, pblkB;
  // GemmDesc::zsub > 0: element offsets of A and B per sub-entry index
  long subA, subB;
};

// The 8-k chunk c of panel row rr (0..15) is stored at chunk position
// c ^ ((rr >> 1) & 3) (an involution): a panel block copied verbatim into
// LDS (64-byte rows) is then read conflict-free by MFMA fragment loads,
// whose 8 lanes per cycle take rows 0..7 at one logical chunk.
__host__ __device__ inline int panel_chunk(int rr, int c) { return c ^ ((rr >> 1) & 3); }

// C(m,n) lives at C[(m / cdiv) * scm1 + (m % cdiv) * scm
//                   + (n / ndiv) * scn1 + (n % ndiv) * scn];
// cdiv / ndiv <= 0 (or >= M / N) mean no split: C[m*scm + n*scn].  The
// splits let a (T) permutation write W[a][b][c] in any index order.  C is
// always double: the
// FP32/BF16 forms widen their FP32 accumulators in the epilogue, so every
// result tensor stays FP64.
struct GemmDesc {
  int M, N;
  int nseg;
  GemmSeg seg[GEMM_MAXSEG];
  double *C;
  long scm, scn, scm1, scn1, sCz;
  int cdiv, ndiv;
  double alpha, beta;
  int batch;
  // split-K: when > 0, batch entry z covers k in [z*K, min((z+1)*K, ktotal))
  // of segment 0 (its sAz/sBz advance by K*sak / K*sbk)
  long ktotal;
  // pointer-table launches of pre-split BF16x3 operands: the caller
  // asserts every table pointer is 16-byte aligned (enables 16-byte loads)
  int vec_ok;
  // C is float (its pointers hold float data despite their type)
  int c_f32;
  // when offm is non-null, C(m,n) lives at C[offm[m] + offn[n]] (device
  // tables of M and N entries; any separable layout, e.g. the blocked (T) W)
  const long *offm, *offn;
  // 1: launch the TAG = 1 instantiation (same code; the (T) W3 launches use
  // it so that profiles attribute them apart)
  int tag;
  // PREC_BF16X3_PRESPLIT with c_f32 and offm: the caller asserts M % 4 == 0
  // and, for every m % 4 == 0, offm[m + r] = offm[m] + r (r < 4) with
  // offm[m] + offn[n] a multiple of 4 and C 16-byte aligned, so the epilogue
  // moves four rows of C as one 16-byte access
  int row4;
  // PREC_BF16X3_PRESPLIT pointer-table launches: > 1 runs groups of zgroup
  // consecutive batch entries of one output tile back to back on one XCD
  // (entries that share operand tiles then reuse them from L2)
  int zgroup;
  // PREC_BF16X3_PRESPLIT with every operand of every segment a panel: > 1
  // selects the kernel that copies panel blocks straight into LDS (global
  // loads to LDS, no register staging) with that many LDS stages (2 or 3)
  int dma;
  // LDS-copy kernel with a pointer table: > 0 expands every table entry z
  // into zsub sub-entries u = 0..zsub-1, sub-entry u reading segment s's A
  // at ptrs[z].A[s] + u*subA and B at ptrs[z].B[s] + u*subB and writing C
  // at ptrs[z].C + offz[u]; the grid is (tiles * zsub, batch).  Groups of
  // zgroup table entries run sub-entry by sub-entry on one XCD, so the
  // sub-entry-dependent operand blocks they share come from L2.
  int zsub;
  const long *offz;
};

// Per-batch pointers for grouped launches (used by the (T) build where each
// batch entry is one triplet): A/B of every segment and C.
struct GemmPtrs {
  const void *A[GEMM_MAXSEG];
  const void *B[GEMM_MAXSEG];
  double *C;
};

// Launches C = alpha * sum_seg sign * A B + beta * C over desc.batch
// entries.  ptrs (device memory, desc.batch entries) overrides the strided
// batch addressing when non-null.  tile = -1 picks a tile shape by size.
void launch_gemm(const GemmDesc &d, const GemmPtrs *ptrs, int prec, int tile, hipStream_t st);
int gemm_num_tiles();
const char *gemm_tile_name(int tile);
int gemm_pick_tile(const GemmDesc &d);
void gemm_set_b3_stages(int stages);  // 1 or 2 LDS stages in the BF16x3 kernel
void gemm_tile_dims(int tile, int &bm, int &bn);
// C[m*scm + n*scn] = alpha * sum_z W[z][m][n] + beta * C  (split-K reduce)
void launch_splitk_reduce(int S, int M, int N, const double *W, double alpha, double beta, double *C, long scm,
                          long scn, hipStream_t st);

// ------------------------------------------------------------ streaming
// out[idx] = alpha * in[perm(idx)] + beta * out[idx] over a row-major output
// of rank <= 6 with dims od[]; istr[] gives, per output dim, the stride of
// the input element.  Element types: 0 = double, 1 = float (in or out),
// 2 = split BF16 pair (out only, beta = 0): n halves hi = bf16(x), then n
// halves lo = bf16(x - hi).
// out_t: 0 FP64, 1 FP32, 2 BF16 hi/lo pair arrays, 3 BF16 hi/lo MFMA panels
// of the output viewed as [n / pk][pk] (rows x k; row blocks pblk apart, lo
// halves at plo; see launch_pack_panel; k padding is not written)
struct PermDesc {
  int rank;
  int od[6];
  long istr[6];
  const void *in;
  void *out;
  int in_t, out_t;
  double alpha, beta;
  long pk, pblk, plo;
};
void launch_permute(const PermDesc &p, hipStream_t st);
// Packs np FP64 matrices in[p*sp + r*sr + k*sk] (r < R, k < K) as split
// BF16 MFMA panels: hi halves first, then lo halves (at offset
// panel_size(np, R, K)), each [p][r/16][k/32][r%16][k%32] (8-k chunks
// swizzled by panel_chunk) with R rounded up to 16 and K to 32, zeros in
// the padding.
long panel_size(int np, long R, int K);
void launch_pack_panel(const double *in, long sp, long sr, long sk, int np, long R, int K, unsigned short *out,
                       hipStream_t st);
// The same with a two-level row index: row r reads in[p*sp + (r / R2)*sr1
// + (r % R2)*sr2 + k*sk].
// launch_pack_panel2 with the rows' inner index r % R2 valid only below
// R2v (rows at or past it pack as zeros): a row set padded per outer index
void launch_pack_panel_rows(const double *in, long sp, long sr1, long sr2, long R2v, long R2, long sk, int np,
                            long R, int K, unsigned short *out, hipStream_t st);
void launch_pack_panel2(const double *in, long sp, long sr1, long sr2, long R2, long sk, int np, long R, int K,
                        unsigned short *out, hipStream_t st);

void launch_fill(long n, double v, double *y, hipStream_t st);
// t1_ia = R1_ia / (eps_i - eps_a) (the update's T1 denominator, ccwfn.cc:62-65)
// Z_ijab += 2 l1_ia Hov_jb - l1_ja Hov_ib (the l1 x Hov terms of the l2 residual)
void launch_lambda_outer(int no, int nv, const double *l1, const double *hov, double *Z, hipStream_t st);
void launch_t1_divide(int no, int nv, const double *eps, const double *R1, double *t1, hipStream_t st);
// flag[0] = 1 when any of x[0..n) is NaN or infinite (flag untouched otherwise)
void launch_nonfinite(long n, const double *x, double *flag, hipStream_t st);
// Packed symmetric / antisymmetric ladder.  Pairs p(a,b) = a(a+1)/2 + b for
// a >= b (P = v(v+1)/2 of them) and s(a,b) = a(a-1)/2 + b for a > b
// (P' = v(v-1)/2); likewise over (i,j).  The (e,f) index -- the GEMMs'
// k -- is stored with leading dimension ladder_ld(pairs), a multiple of 32
// with zeros in the padding (aligned rows for the BF16x3 kernels).
long ladder_ld(long pairs);
//   V+[p(ab)][p(ef)] = <ab|ef> + <ab|fe>, V-[s(ab)][s(ef)] = <ab|ef> - <ab|fe>
void launch_ladder_vpm(int nv, const double *vvvv, double *vp, double *vm, hipStream_t st);
//   S[p(ij)][p(ef)] = (tau_ij^ef + tau_ij^fe)(1/2 if e = f), A[s(ij)][s(ef)] = tau_ij^ef - tau_ij^fe
void launch_ladder_pack(int no, int nv, const double *tau, double *S, double *A, hipStream_t st);
// The same pair split of the (e,f) index of nrows slabs in[r][e][f]:
//   Vp[r][p(ef)] = in[r][e][f] + in[r][f][e], Vm[r][s(ef)] = in[r][e][f] - in[r][f][e]
void launch_rows_vpm(long nrows, int nv, const double *in, double *vp, double *vm, hipStream_t st);
//   Z[r][i][j] = beta Z[r][i][j] + (Z+[r][p(ij)] + s_ij Z-[r][s(ij)]) / 2,
//   s_ij = +1 for i > j, -1 below; beta 0 or 1
void launch_pairs_unpack(long nrows, int no, const double *ZP, const double *ZM, double beta, double *Z,
                         hipStream_t st);
//   X_ijab += scale (L+[p(ij)][p(ab)] +- L-[s(ij)][s(ab)]), the sign s_ij s_ab
void launch_ladder_unpack(int no, int nv, const double *LP, const double *LM, double scale, double *X,
                          hipStream_t st);
void launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t st);  // y = a x + b y
// x[j][k][a][b][c] /= eps[p] + eps[j] + eps[k] - eps[no+a] - eps[no+b] - eps[no+c]
void launch_div_d3(int no, int nv, int p, const double *eps, double *x, hipStream_t st);
void launch_diff(long n, const double *x, const double *y, double *z, hipStream_t st);       // z = x - y
// tau = t2 + t1 (x) t1, ttau = t2 + 0.5 t1 (x) t1 (ttau may be null)
void launch_tau(int no, int nv, const double *t1, const double *t2, double *tau, double *ttau, hipStream_t st);
// Y[n][f][j][b] = 0.5 t2[j][n][f][b] + t1[j][f] t1[n][b]
// U[i][j][a][b] = 2 t2[i][j][a][b] - t2[i][j][b][a]
void launch_build_U(int no, int nv, const double *t2, double *U, hipStream_t st);
void launch_build_Y(int no, int nv, const double *t1, const double *t2, double *Y, hipStream_t st);
// R2 = X + X^T(ij,ab);  rsq partials of |R1|^2 + |R2|^2;  t = t_old + R/D
// with D from the Fock diagonal eps (nact).  R1/R2 are written out when the
// pointers are non-null.
int update_nparts(int no, int nv);
void launch_update(int no, int nv, const double *eps, const double *R1, const double *X, const double *t1old,
                   const double *t2old, double *t1, double *t2, double *R2out, double *partials, int npart,
                   hipStream_t st);
// the same update with every denominator shifted by `shift` (x = x_old +
// R / (D + shift), the perturbed-amplitude equations at a frequency)
void launch_update_shift(int no, int nv, const double *eps, double shift, const double *R1, const double *X,
                         const double *t1old, const double *t2old, double *t1, double *t2, double *partials, int npart,
                         hipStream_t st);
// MP2 guess t1 = f_ia/D1, t2 = <ij|ab>/D2 from the full Fock matrix
void launch_guess(int no, int nv, const double *fock, const double *oovv, double *t1, double *t2, hipStream_t st);
// partial sums of x.y
void launch_dot(long n, const double *x, const double *y, double *partials, int npart, hipStream_t st);
void launch_reduce(const double *partials, int n, double *out, hipStream_t st);
// DIIS: err = t - t_old, amp = t (over concatenated t1|t2)
void launch_diis_store(long n1, long n2, const double *t1, const double *t1old, const double *t2,
                       const double *t2old, double *err, double *amp, hipStream_t st);
// partials[p*npart + blk] = sum err_new . err_p over nvec stored vectors
constexpr int kDiisMax = 16;  // largest DIIS subspace
constexpr int kDiisDotBlocks = 2048;  // blocks (and partials per vector) of the DIIS dots
void launch_diis_dots(long len, const double *e, const double *const *vecs, int nvec, double *partials,
                      int npart, hipStream_t st);
void launch_diis_combine(long n1, long n2, const double *const *amps, const double *coef, int nvec, double *t1,
                         double *t2, hipStream_t st);
// partials of sum tau_ijab L_ijab (the one-electron term 2 f_ia t1_ia is
// added on the host)
void launch_energy(int no, int nv, const double *fov, const double *t1, const double *tau, const double *Loovv,
                   double *partials, int npart, hipStream_t st);
int energy_nparts(int no, int nv);

// ------------------------------------------------------------ (T)
// W3 of a triplet is stored blocked: T^3 sub-cubes (T = tw_blk), each
// contiguous in [a][b][c] order, the sub-cubes in [A][B][C] order over
// nb = ceil(nv / T) blocks per index, so W[a][b][c] lives at
//   tw_off_a(a) + tw_off_b(b) + tw_off_c(c)
// with tw_off_a(x) = (x / T) nb^2 T^3 + (x % T) T^2,
//      tw_off_b(x) = (x / T) nb T^3 + (x % T) T,  tw_off_c(x) = (x / T) T^3 + x % T,
// and one triplet occupies (nb T)^3 elements (positions past nv unused).
// T = 12 for both element types: the energy kernel's LDS holds six padded
// 12^3 sub-cubes plus the <pq|xy> blocks with room for two workgroups per
// CU at FP32 (measured faster than T = 16 with one).
inline int tw_blk(int w_f32) { (void)w_f32; return 12; }
inline long tw_nb(int nv, int T) { return (nv + T - 1) / T; }
inline long tw_stride(int nv, int T) { const long e = tw_nb(nv, T) * T; return e * e * e; }
inline long tw_off_a(long x, long nb, int T) { return (x / T) * nb * nb * T * T * T + (x % T) * T * T; }
inline long tw_off_b(long x, long nb, int T) { return (x / T) * nb * T * T * T + (x % T) * T; }
inline long tw_off_c(long x, int T) { return (x / T) * T * T * T + x % T; }
// Energy epilogue over a batch of triplet entries: entry t is (i, j, k,
// problem) = trip4[4t .. 4t+3]; its W (blocked layout above) starts at
// W + t * wstride; probs[problem] gives that problem's Fock diagonal, oovv
// block and t1 (all problems share no and nv).  Sums the (T) bracket over
// a>=b>=c into etrip[t].  W is FP64, or FP32 when w_f32.  A persistent
// kernel: as many workgroups as fit on the device drain a queue of
// (entry, block pair) items; counter is one device int (reset by the launch
// with a kernel), part holds ntrip * triples_energy_npairs(nv, w_f32)
// doubles.
// Per-problem operands of the (T) energy kernel: orbital energies, the oovv
// block and t1; for the f_ov variant instead of oovv and t1 the pair arrays
// ot2[i][j][a][b] = (<ij|ab>, t2_ij^ab) and t1f[i][a] = (t1_ia, f_ia) in the
// kernel's pair type (float2 with an FP32 W, double2 with an FP64 W; see
// launch_interleave).
struct TripProb {
  const double *eps, *oovv, *t1;
  const void *ot2, *t1f;
};
// out[k] = (x[k], y[k]) for k < n, as float2 (f32) or double2
void launch_interleave(long n, const double *x, const double *y, void *out, int f32, hipStream_t st);
int triples_energy_npairs(int nv, int w_f32);
// fov != 0: includes the f_ov t3 terms of the reference's tcorr (any
// problem of the batch with a non-zero occupied-virtual Fock block)
void launch_triples_energy(int no, int nv, int ntrip, const int *trip4, const TripProb *probs, const void *W,
                           int w_f32, long wstride, int fov, double *etrip, double *part, int *counter,
                           hipStream_t st);

// ------------------------------------------------------------ learned guess
constexpr int MLP_NFEAT = 8;
constexpr int MLP_WIDTH = 128;
// Inference view of a network: the first layer folded with the feature
// standardisation (FP32), the two hidden 128x128 layers as BF16 MFMA A
// fragments in the kernel's permuted K order (8 x 4 x 64 lanes x 16 bytes
// each), biases and the output layer in FP32.
struct MlpDevice {
  const float *w1f, *b1f;
  const void *w2p;
  const float *b2;
  const void *w3p;
  const float *b3, *w4;
  float b4;
};
// X[r][8] = features of amplitude idx[r] (flattened ijab; idx null: r)
void launch_features(int no, int nv, const double *eps, const double *oovv, long n, const long *idx, float *X,
                     hipStream_t st);
void launch_mlp_infer_x(long N, const float *X, const MlpDevice &P, float *out, hipStream_t st);
void launch_mlp_infer_ham(int no, int nv, const double *eps, const double *oovv, const MlpDevice &P, float *out,
                          hipStream_t st);
void launch_apply_guess(int no, int nv, const float *G, double ys, double *t2, hipStream_t st);
void launch_mlp_normalize(long n, const float *X, const float *shift, const float *inv, float *Xn, hipStream_t st);
void launch_mlp_gather(int B, const float *X, const float *y, const int *perm, float *Xb, float *yb, hipStream_t st);
void launch_mlp_bias_gelu(long rows, const float *b, float *Z, float *H, hipStream_t st);
void launch_mlp_head(int B, const float *H, const float *w4, const float *b4, const float *target, float *y, float *dy,
                     double *sq, hipStream_t st);
void launch_mlp_head_back(int B, const float *dy, const float *w4, const float *Z, float *dZ, hipStream_t st);
void launch_mlp_gelu_back(long n, const float *dH, const float *Z, float *dZ, hipStream_t st);
int mlp_colsum_chunks(int B);
// out[f] = sum_r w[r] M[r][f] (w null: 1); P holds mlp_colsum_chunks(B) * width floats
void launch_mlp_colsum(int B, int width, const float *M, const float *w, float *P, float *out, hipStream_t st);
void launch_mlp_sum_parts(int S, long n, const float *P, float *out, hipStream_t st);
void launch_mlp_loss(int n, const double *sq, double *out, hipStream_t st);
void launch_mlp_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t st);

}  // namespace ugx
