// Dense contractions on gfx950 matrix cores.
//
// One kernel family serves every contraction of the CCSD residual and the
// (T) W3 build:
//   C(m,n) = alpha * sum_seg sign_s * sum_k A_s(m,k) B_s(k,n) + beta * C(m,n)
// with up to four K segments that have their own operands and strides (the
// (T) particle terms, K = v, and hole terms, K = o, of two output
// permutations share one accumulator that way), a batch dimension (strided
// or a device table of pointers), and row/column splits of the output
// address so that a (T) permutation can write W[a][b][c] in any index order.
//
// Work decomposition: 256 threads = 4 wave64s arranged WGM x WGN; each wave
// owns TM x TN blocks of 16x16 MFMA outputs, so a workgroup tile is
// BM x BN = (16*WGM*TM) x (16*WGN*TN).  Operands are staged through LDS in
// BK = 16 slices stored k-major ([k][m], [k][n]) so the MFMA operand reads of
// a 16-lane group are consecutive; the next slice is prefetched into
// registers while the MFMAs work on the current one.  Global loads follow
// whichever of the two operand strides is 1, so both row- and column-major
// operands are read coalesced.
//
// Launch: a 1D grid over output tiles whose block ids are remapped so that
// the tiles dispatched to one XCD (block ids are dealt round-robin to the
// eight XCDs) form a contiguous band of the output, which keeps a band's
// shared A rows in that XCD's L2.
//
// MFMA forms:
//   FP64: v_mfma_f64_16x16x4_f64   (A: lane l -> A[l%16][l/16],
//                                     B: lane l -> B[l/16][l%16],
//                                     D: lane l, item r -> D[l/16 + 4r][l%16],
//                                     measured on gfx950)
//   FP32: v_mfma_f32_16x16x4_f32   (A, B as FP64;
//                                     D: lane l, item r -> D[4(l/16)+r][l%16])
//   BF16X3: v_mfma_f32_16x16x32_bf16 on hi/lo BF16 splits of FP32 operands
//           (A: lane l -> A[l%16][8(l/16) .. 8(l/16)+7], B likewise along k)
#include "ugx_internal.h"

#include <cstdint>
#include <stdexcept>

namespace ugx {
namespace {

constexpr int NT = 256;
constexpr int BK = 16;
constexpr int MAXSEG = GEMM_MAXSEG;

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef dbl4 acc_t;
  __device__ static inline acc_t zero() { return dbl4{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of accumulator item r held by lane group kq = lane / 16
  __device__ static inline int row(int kq, int r) { return kq + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef flt4 acc_t;
  __device__ static inline acc_t zero() { return flt4{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int kq, int r) { return 4 * kq + r; }
};

// Per-thread loader of one operand tile (R rows x KB k-values per slice).
// Element e = tid + q*NT (q < R*KB/NT) of the slice maps to (row, k) along
// the operand's unit-stride side: k-contiguous operands are read as KB
// consecutive k per row (16 lanes cover one row segment), row-contiguous ones
// as R consecutive rows per k.  The mapping is a bijection for any R that is
// a multiple of 16.  The slice pointer is set up once per segment and
// advanced by KB k per slice.
template <typename T, int R, int KB>
struct OpLoader {
  static constexpr int N_ = R * KB / NT;  // elements per thread per slice
  static_assert(R % 16 == 0 && (R * KB) % NT == 0, "tile rows must be a multiple of 16");
  const T *ptr;   // element (row0, k0) of the current slice
  long srow, sk;
  int rows_left;  // rows of the operand from row0 on
  int K;
  bool kc;

  __device__ static inline void rk(bool kc, int e, int &r, int &kk) {
    if (kc) { r = e / KB; kk = e % KB; } else { r = e % R; kk = e / R; }
  }
  __device__ inline void setup(const T *base, long srow_, long sk_, int row0, int rows, int K_) {
    srow = srow_; sk = sk_; K = K_;
    kc = (sk == 1);
    ptr = base + (long)row0 * srow;
    rows_left = rows - row0;
  }
  __device__ inline void load(T (&v)[N_], int k0, T scale, int tid) const {
#pragma unroll
    for (int q = 0; q < N_; q++) {
      int r, kk;
      rk(kc, tid + q * NT, r, kk);
      T x = T(0);
      if (r < rows_left && k0 + kk < K) x = ptr[(long)r * srow + (long)kk * sk];
      v[q] = x * scale;
    }
  }
  __device__ inline void advance() { ptr += (long)KB * sk; }
  // k-major LDS tile S[k][row]
  template <int LD>
  __device__ inline void store_kmajor(const T (&v)[N_], T (*S)[LD], int tid) const {
#pragma unroll
    for (int q = 0; q < N_; q++) {
      int r, kk;
      rk(kc, tid + q * NT, r, kk);
      S[kk][r] = v[q];
    }
  }
};

// Segment bookkeeping shared by the kernels: slice counts per segment (the
// split-K tail applied to segment 0).  The operand bases, K and sign of a
// segment are read from the kernel argument and the pointer table when the
// segment starts: a runtime index into a local array would place it in
// scratch memory, so only nt[] is kept, and only constant indices touch it.
template <typename T, int KB>
struct Segs {
  int nt[MAXSEG];
  int n, ntiles;
  __device__ inline void init(const GemmDesc &d, long z) {
    n = d.nseg;
    ntiles = 0;
#pragma unroll
    for (int s = 0; s < MAXSEG; s++) {
      nt[s] = s < n ? (kof(d, z, s) + KB - 1) / KB : 0;
      ntiles += nt[s];
    }
  }
  __device__ static inline int kof(const GemmDesc &d, long z, int s) {
    int k = d.seg[s].K;
    if (s == 0 && d.ktotal > 0) {
      long rem = d.ktotal - z * (long)k;
      if (rem < k) k = rem > 0 ? (int)rem : 0;
    }
    return k;
  }
  __device__ static inline const T *a(const GemmDesc &d, const GemmPtrs *ptrs, long z, int s) {
    return ptrs ? (const T *)ptrs[z].A[s] : (const T *)d.seg[s].A + z * d.seg[s].sAz;
  }
  __device__ static inline const T *b(const GemmDesc &d, const GemmPtrs *ptrs, long z, int s) {
    return ptrs ? (const T *)ptrs[z].B[s] : (const T *)d.seg[s].B + z * d.seg[s].sBz;
  }
  // +-1 as a float whatever T is (T may be an unsigned BF16 carrier)
  __device__ static inline float sg(const GemmDesc &d, int s) { return d.seg[s].sign < 0 ? -1.0f : 1.0f; }
  // slice t -> segment and its k0
  __device__ inline int locate(int t, int &k0) const {
    int s = 0;
#pragma unroll
    for (int q = 0; q < MAXSEG - 1; q++)
      if (q < n - 1 && s == q && t >= nt[q]) { t -= nt[q]; s = q + 1; }
    k0 = t * KB;
    return s;
  }
};

// Maps the linear block id to an output tile (tm, tn).  Block ids are dealt
// round-robin to the 8 XCDs; the remap gives XCD x the contiguous id range
// [x*per, (x+1)*per), and tiles are walked n-fastest inside bands of GROUP
// tile rows, so the workgroups resident on one XCD share A rows and B
// columns in its L2.
__device__ inline void tile_of(int bid, int nblk, int tiles_m, int tiles_n, int &tm, int &tn) {
  constexpr int NXCD = 8;
  int id = bid;
  if (nblk % NXCD == 0) {
    int per = nblk / NXCD;
    id = (bid % NXCD) * per + bid / NXCD;
  }
  constexpr int GROUP = 8;
  int band = id / (GROUP * tiles_n);
  int first = band * GROUP;
  int rows = tiles_m - first < GROUP ? tiles_m - first : GROUP;
  int in = id - band * GROUP * tiles_n;
  tm = first + in % rows;
  tn = in / rows;
}

// Batched launches with d.zgroup = G > 1: block (x, y) of the (tiles x
// batch) grid -> (tile, batch entry).  The hardware deals linear block ids
// round-robin to the 8 XCDs; XCD x is given the contiguous range of work
// items [start_x, start_x + count_x) in the order (group of G entries, tile,
// entry), so G consecutive entries of one output tile run back to back on
// one XCD and share, through its L2, the operand tiles they have in common
// (the (T) entries of a batch share i and often j).  Any grid size: XCD x
// holds count_x = q + (x < rem) ids, start_x = x q + min(x, rem).
__device__ inline void zgroup_of(int G, int &bx, long &z) {
  constexpr long NXCD = 8;
  const long tiles = gridDim.x, nb = gridDim.y, total = tiles * nb;
  const long L = (long)blockIdx.y * tiles + blockIdx.x;
  const long q = total / NXCD, rem = total % NXCD, x = L % NXCD;
  const long w = x * q + (x < rem ? x : rem) + L / NXCD;  // work item
  const long g = w / (G * tiles), r = w % (G * tiles);
  const long gz = (nb - g * G < G) ? nb - g * G : G;      // entries in group g
  bx = (int)(r / gz);
  z = g * G + r % gz;
}

// Launches with d.zsub > 0 (grid (tiles * zsub, batch)): block -> (tile,
// table entry z, sub-entry u).  As in zgroup_of, XCD x takes a contiguous
// range of work items, here in the order (group of G table entries,
// sub-entry, tile, entry): the G entries of a group run sub-entry by
// sub-entry on one XCD, so the operand blocks that depend on the sub-entry
// (shared by entries with equal table pointers) and the entries' own blocks
// (used by every sub-entry) are reused through its L2.
__device__ inline void zsub_of(int G, int zsub, int tiles, int &bx, long &z, int &u) {
  constexpr long NXCD = 8;
  const long nb = gridDim.y, total = (long)gridDim.x * nb;
  const long L = (long)blockIdx.y * gridDim.x + blockIdx.x;
  const long q = total / NXCD, rem = total % NXCD, x = L % NXCD;
  const long w = x * q + (x < rem ? x : rem) + L / NXCD;  // work item
  const long per = (long)G * zsub * tiles;
  const long g = w / per, r = w % per;
  const long gz = (nb - g * G < G) ? nb - g * G : G;      // entries in group g
  const long ut = r / gz;                                 // (sub-entry, tile)
  u = (int)(ut / tiles);
  bx = (int)(ut % tiles);
  z = g * G + r % gz;
}

// Address parts of C(m, n): C[c_row_off(m) + c_col_off(n)] (an offset table,
// a two-level split, or a plain stride per dimension).
__device__ inline long c_row_off(const GemmDesc &d, int m) {
  if (d.offm) return d.offm[m];
  return d.cdiv > 0 ? (long)(m / d.cdiv) * d.scm1 + (long)(m % d.cdiv) * d.scm : (long)m * d.scm;
}
__device__ inline long c_col_off(const GemmDesc &d, int n) {
  if (d.offn) return d.offn[n];
  return d.ndiv > 0 ? (long)(n / d.ndiv) * d.scn1 + (long)(n % d.ndiv) * d.scn : (long)n * d.scn;
}

// Epilogue: C = alpha * acc + beta * C over this thread's accumulators,
// rows ro[i][r] (item r of block i) and columns co[j] (-1 = outside C).
// With beta != 0 the old values are read first and combined afterwards,
// each masked element reading C's first element instead (always in bounds,
// the value is discarded): a load behind a per-element branch would wait for
// its own round trip before the next one is issued.  One row block (TN*4
// values) is read at a time, which bounds the registers the epilogue holds.
// The arithmetic is the same on every path.
template <int TM, int TN, typename ACC>
__device__ inline void store_tile(const GemmDesc &d, void *Cv, const long (&ro)[TM][4], const long (&co)[TN],
                                  const ACC (&acc)[TM][TN]) {
  const double alpha = d.alpha, beta = d.beta;
  if (d.c_f32 && d.row4) {
    // a lane's four rows are four consecutive, 16-byte aligned floats
    float *C = (float *)Cv;
#pragma unroll
    for (int i = 0; i < TM; i++) {
      float4 old[TN];
      if (beta != 0.0) {
#pragma unroll
        for (int j = 0; j < TN; j++) {
          const bool ok = ro[i][0] >= 0 && co[j] >= 0;
          old[j] = *reinterpret_cast<const float4 *>(C + (ok ? ro[i][0] + co[j] : 0));
        }
      }
#pragma unroll
      for (int j = 0; j < TN; j++)
        if (ro[i][0] >= 0 && co[j] >= 0) {
          double v0 = alpha * (double)acc[i][j][0], v1 = alpha * (double)acc[i][j][1];
          double v2 = alpha * (double)acc[i][j][2], v3 = alpha * (double)acc[i][j][3];
          if (beta != 0.0) {
            v0 += beta * (double)old[j].x; v1 += beta * (double)old[j].y;
            v2 += beta * (double)old[j].z; v3 += beta * (double)old[j].w;
          }
          float4 w;
          w.x = (float)v0; w.y = (float)v1; w.z = (float)v2; w.w = (float)v3;
          *reinterpret_cast<float4 *>(C + ro[i][0] + co[j]) = w;
        }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < TM; i++) {
    if (d.c_f32) {
      float *C = (float *)Cv;
      float old[TN][4];
      if (beta != 0.0) {
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) old[j][r] = C[(ro[i][r] >= 0 && co[j] >= 0) ? ro[i][r] + co[j] : 0];
      }
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (ro[i][r] >= 0 && co[j] >= 0) {
            double v = alpha * (double)acc[i][j][r];
            if (beta != 0.0) v += beta * (double)old[j][r];
            C[ro[i][r] + co[j]] = (float)v;
          }
    } else {
      double *C = (double *)Cv;
      double old[TN][4];
      if (beta != 0.0) {
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) old[j][r] = C[(ro[i][r] >= 0 && co[j] >= 0) ? ro[i][r] + co[j] : 0];
      }
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (ro[i][r] >= 0 && co[j] >= 0) {
            double v = alpha * (double)acc[i][j][r];
            if (beta != 0.0) v += beta * old[j][r];
            C[ro[i][r] + co[j]] = v;
          }
    }
  }
}

// FP64 / FP32 main kernel.  Two LDS stages: while the MFMAs consume slice t
// from one stage, slice t+1 (loaded into registers one iteration earlier) is
// written to the other stage and the global loads of slice t+2 are issued,
// so each global load has a full slice of MFMA work to land and there is one
// barrier per slice.
// TAG only distinguishes instantiations (the (T) W3 launches use TAG = 1,
// so profiles report them apart from the CCSD contractions).
// Tiles of up to 4 x 4 blocks per wave are held at two waves per SIMD
// (the 128 x 128 FP64 tile fits without spills; larger ones would spill).
template <typename T, int WGM, int WGN, int TM, int TN, int TAG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(TM * TN <= 16 ? 2 : 1))) void k_gemm(
    GemmDesc d, const GemmPtrs *__restrict__ ptrs) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  constexpr int LDA = BM + (sizeof(T) == 8 ? 1 : 2), LDB = BN + (sizeof(T) == 8 ? 1 : 2);
  static_assert(BM * BK / NT >= 1 && BN * BK / NT >= 1, "tile too small for 256 threads");
  __shared__ T As[2][BK][LDA];
  __shared__ T Bs[2][BK][LDB];

  const int M = d.M, N = d.N;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const long z = blockIdx.y;
  const int tid = threadIdx.x;

  Segs<T, BK> S;
  S.init(d, z);
  const int ntiles = S.ntiles;

  OpLoader<T, BM, BK> la;
  OpLoader<T, BN, BK> lb;
  int cur_seg = -1;
  auto seg_setup = [&](int s) {
    const GemmSeg &g = d.seg[s];
    la.setup(S.a(d, ptrs, z, s), g.sam, g.sak, m0, M, S.kof(d, z, s));
    lb.setup(S.b(d, ptrs, z, s), g.sbn, g.sbk, n0, N, S.kof(d, z, s));
    cur_seg = s;
  };

  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int r16 = lane & 15, kq = lane >> 4;

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = MF::zero();

  T ra[OpLoader<T, BM, BK>::N_], rb[OpLoader<T, BN, BK>::N_];
  // slice t -> (segment, k0); the loaders walk the slices in order
  auto fetch = [&](int t) {
    int k0;
    int s = S.locate(t, k0);
    if (s != cur_seg) seg_setup(s);
    la.load(ra, k0, T(S.sg(d, s)), tid);
    lb.load(rb, k0, T(1), tid);
    la.advance();
    lb.advance();
  };

  if (ntiles > 0) {
    fetch(0);
    la.template store_kmajor<LDA>(ra, As[0], tid);
    lb.template store_kmajor<LDB>(rb, Bs[0], tid);
    if (ntiles > 1) fetch(1);
  }
  __syncthreads();

  for (int t = 0; t < ntiles; t++) {
    const int st = t & 1;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = As[st][kk + kq][(wm * TM + i) * 16 + r16];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bs[st][kk + kq][(wn * TN + j) * 16 + r16];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = MF::op(a[i], b[j], acc[i][j]);
    }
    if (t + 1 < ntiles) {
      // the other stage was last read in iteration t-1, which every wave
      // finished before the barrier that ended it
      la.template store_kmajor<LDA>(ra, As[st ^ 1], tid);
      lb.template store_kmajor<LDB>(rb, Bs[st ^ 1], tid);
      if (t + 2 < ntiles) fetch(t + 2);
    }
    __syncthreads();
  }

  // C is FP64, or FP32 when d.c_f32 (the (T) W buffer of the low-precision
  // forms); per-batch offsets count elements of C's type
  void *Cv = ptrs ? (void *)ptrs[z].C
                  : (d.c_f32 ? (void *)((float *)d.C + z * d.sCz) : (void *)(d.C + z * d.sCz));
  // this thread's TM*4 output rows and TN columns, addressed once
  long ro[TM][4], co[TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = m0 + (wm * TM + i) * 16 + MF::row(kq, r);
      ro[i][r] = m < M ? c_row_off(d, m) : -1;
    }
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int n = n0 + (wn * TN + j) * 16 + r16;
    co[j] = n < N ? c_col_off(d, n) : -1;
  }
  store_tile<TM, TN>(d, Cv, ro, co, acc);
}

// ------------------------------------------------------------ BF16x3
// Each FP32 operand x is carried as hi = bf16(x) and lo = bf16(x - hi)
// (round to nearest even); the product is hi*hi + hi*lo + lo*hi with FP32
// accumulation (the dropped lo*lo term is ~2^-16 relative).
constexpr int BK3 = 32;
static_assert(BK3 == BK_PANEL, "MFMA panels are packed one BF16x3 slice deep");
constexpr int LD3 = BK3 + 8;  // halves per LDS row (80 bytes: 16-byte aligned rows)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline unsigned short bf16_bits_rne(float x) {
  unsigned int u = __float_as_uint(x);
  unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(r >> 16);
}
__device__ inline float bf16_to_float(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// ------------------------------------------------- BF16x3, pre-split operands
// Operands arrive as BF16 hi/lo pairs with k contiguous (A[m][k], B[n][k]):
// no conversion in the kernel, 16-byte global loads of eight k values when
// every stride and K are multiples of 8 (VEC), element loads otherwise.
// Element (r, k) of a slice lies at hi + (r/16)*sblk + (r%16)*srow + k%32:
// row-major operands have sblk = 16*srow and step BK3 halves per slice;
// MFMA panels (GemmSeg::pblkA/B) have srow = 32, sblk = pblk and step 512,
// so the 16 rows x 32 k of a slice are one contiguous kilobyte, and their
// 8-k chunks are swizzled (panel_chunk).
template <int R>
struct PairLoader {
  static_assert(R % 16 == 0, "tile rows must be a multiple of 16");
  static constexpr int CH = R * (BK3 / 8);             // 16-byte chunks per array per slice
  static constexpr int NQ = (CH + NT - 1) / NT;        // chunks per thread
  const unsigned short *hi;                            // row0, k0 of the current slice
  long lo, srow, sblk, kstep;
  int rows_left, K;
  bool swz;
  __device__ inline void setup(const unsigned short *base, long lo_, long srow_, long pblk, int row0, int rows,
                               int K_) {
    if (pblk > 0) { srow = BK3; sblk = pblk; kstep = 16 * BK3; swz = true; }
    else { srow = srow_; sblk = 16 * srow_; kstep = BK3; swz = false; }
    hi = base + (long)(row0 >> 4) * sblk;  // row0 is a multiple of 16
    lo = lo_; rows_left = rows - row0; K = K_;
  }
  template <bool VEC>
  __device__ inline void load(uint4 (&h)[NQ], uint4 (&l)[NQ], int k0, int tid) const {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      int c = tid + q * NT;
      int r, kc;
      chunk(c, r, kc);
      uint4 vh = {0u, 0u, 0u, 0u}, vl = {0u, 0u, 0u, 0u};
      if (c < CH && r < rows_left) {
        const int kp = swz ? panel_chunk(r & 15, kc >> 3) << 3 : kc;  // stored chunk of a panel row
        const unsigned short *ph = hi + (long)(r >> 4) * sblk + (long)(r & 15) * srow + kp;
        if (VEC) {
          if (k0 + kc < K) {
            vh = *reinterpret_cast<const uint4 *>(ph);
            vl = *reinterpret_cast<const uint4 *>(ph + lo);
          }
        } else {
          unsigned short th[8], tl[8];
#pragma unroll
          for (int e = 0; e < 8; e++) {
            bool ok = (k0 + kc + e < K);
            th[e] = ok ? ph[e] : (unsigned short)0;
            tl[e] = ok ? ph[e + lo] : (unsigned short)0;
          }
          vh.x = th[0] | ((unsigned)th[1] << 16); vh.y = th[2] | ((unsigned)th[3] << 16);
          vh.z = th[4] | ((unsigned)th[5] << 16); vh.w = th[6] | ((unsigned)th[7] << 16);
          vl.x = tl[0] | ((unsigned)tl[1] << 16); vl.y = tl[2] | ((unsigned)tl[3] << 16);
          vl.z = tl[4] | ((unsigned)tl[5] << 16); vl.w = tl[6] | ((unsigned)tl[7] << 16);
        }
      }
      h[q] = vh;
      l[q] = vl;
    }
  }
  // 16-byte chunk c of a slice -> (row, k offset): each run of 64 chunks
  // (one wave) covers 16 rows x 4 chunks with the row fastest, so the 8
  // lanes an LDS write serves per cycle hit 8 rows at one k (rows are 80
  // bytes apart: 8 disjoint 4-bank groups) -- the same pattern as the MFMA
  // fragment reads; with the chunk fastest, two rows share one cycle and
  // collide on a bank group
  __device__ static inline void chunk(int c, int &r, int &kc) {
    r = ((c >> 6) << 4) + (c & 15);
    kc = ((c >> 4) & 3) * 8;
  }
  __device__ inline void advance() { hi += kstep; }
  __device__ inline void store(const uint4 (&h)[NQ], const uint4 (&l)[NQ], unsigned short (*H)[LD3],
                               unsigned short (*L)[LD3], int tid) const {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      int c = tid + q * NT;
      if (c < CH) {
        int r, kc;
        chunk(c, r, kc);
        *reinterpret_cast<uint4 *>(&H[r][kc]) = h[q];
        *reinterpret_cast<uint4 *>(&L[r][kc]) = l[q];
      }
    }
  }
};

template <int WGM, int WGN, int TM, int TN, bool VEC, int STAGES, int TAG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_b3p(GemmDesc d,
                                                                                const GemmPtrs *__restrict__ ptrs) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  // STAGES = 2: ping-pong LDS, one barrier per slice.  STAGES = 1: half the
  // LDS (two workgroups per CU, so one's epilogue overlaps the other's
  // MFMAs), two barriers per slice.
  __shared__ __attribute__((aligned(16))) unsigned short Ah[STAGES][BM][LD3];
  __shared__ __attribute__((aligned(16))) unsigned short Al[STAGES][BM][LD3];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[STAGES][BN][LD3];
  __shared__ __attribute__((aligned(16))) unsigned short Bl[STAGES][BN][LD3];

  const int M = d.M, N = d.N;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn, bx = blockIdx.x;
  long z = blockIdx.y;
  if (d.zgroup > 1) {
    zgroup_of(d.zgroup, bx, z);
    tm = bx % tiles_m;  // the XCD mapping is zgroup_of's
    tn = bx / tiles_m;
  } else {
    tile_of(bx, gridDim.x, tiles_m, tiles_n, tm, tn);
  }
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x;

  Segs<unsigned short, BK3> S;
  S.init(d, z);
  const int ntiles = S.ntiles;
  PairLoader<BM> la;
  PairLoader<BN> lb;
  int cur_seg = -1;
  auto seg_setup = [&](int s) {
    const GemmSeg &g = d.seg[s];
    la.setup(S.a(d, ptrs, z, s), g.loA, g.sam, g.pblkA, m0, M, S.kof(d, z, s));
    lb.setup(S.b(d, ptrs, z, s), g.loB, g.sbn, g.pblkB, n0, N, S.kof(d, z, s));
    cur_seg = s;
  };

  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int r16 = lane & 15, kq = lane >> 4;

  flt4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};

  uint4 rah[PairLoader<BM>::NQ], ral[PairLoader<BM>::NQ], rbh[PairLoader<BN>::NQ], rbl[PairLoader<BN>::NQ];
  // a negative segment sign flips the sign bits of its B halves (exact)
  auto fetch = [&](int t) {
    int k0;
    int s = S.locate(t, k0);
    if (s != cur_seg) seg_setup(s);
    la.template load<VEC>(rah, ral, k0, tid);
    lb.template load<VEC>(rbh, rbl, k0, tid);
    if (S.sg(d, s) < 0) {
#pragma unroll
      for (int q = 0; q < PairLoader<BN>::NQ; q++) {
        rbh[q].x ^= 0x80008000u; rbh[q].y ^= 0x80008000u; rbh[q].z ^= 0x80008000u; rbh[q].w ^= 0x80008000u;
        rbl[q].x ^= 0x80008000u; rbl[q].y ^= 0x80008000u; rbl[q].z ^= 0x80008000u; rbl[q].w ^= 0x80008000u;
      }
    }
    la.advance();
    lb.advance();
  };
  if (ntiles > 0) {
    fetch(0);
    la.store(rah, ral, Ah[0], Al[0], tid);
    lb.store(rbh, rbl, Bh[0], Bl[0], tid);
    if (ntiles > 1) fetch(1);
  }
  __syncthreads();

  for (int t = 0; t < ntiles; t++) {
    const int st = (STAGES == 2) ? (t & 1) : 0;
    bf16x8 ah[TM], al[TM];
#pragma unroll
    for (int i = 0; i < TM; i++) {
      int row = (wm * TM + i) * 16 + r16;
      ah[i] = *reinterpret_cast<const bf16x8 *>(&Ah[st][row][8 * kq]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&Al[st][row][8 * kq]);
    }
#pragma unroll
    for (int j = 0; j < TN; j++) {
      int col = (wn * TN + j) * 16 + r16;
      bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&Bh[st][col][8 * kq]);
      bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&Bl[st][col][8 * kq]);
#pragma unroll
      for (int i = 0; i < TM; i++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
      }
    }
    if (STAGES == 2) {
      if (t + 1 < ntiles) {
        la.store(rah, ral, Ah[st ^ 1], Al[st ^ 1], tid);
        lb.store(rbh, rbl, Bh[st ^ 1], Bl[st ^ 1], tid);
        if (t + 2 < ntiles) fetch(t + 2);
      }
      __syncthreads();
    } else {
      __syncthreads();
      if (t + 1 < ntiles) {
        la.store(rah, ral, Ah[0], Al[0], tid);
        lb.store(rbh, rbl, Bh[0], Bl[0], tid);
        if (t + 2 < ntiles) fetch(t + 2);
      }
      __syncthreads();
    }
  }

  // C is FP64, or FP32 when d.c_f32 (the (T) W buffer of the low-precision
  // forms); per-batch offsets count elements of C's type
  void *Cv = ptrs ? (void *)ptrs[z].C
                  : (d.c_f32 ? (void *)((float *)d.C + z * d.sCz) : (void *)(d.C + z * d.sCz));
  // this thread's TM*4 output rows and TN columns, addressed once
  long ro[TM][4], co[TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = m0 + (wm * TM + i) * 16 + 4 * kq + r;
      ro[i][r] = m < M ? c_row_off(d, m) : -1;
    }
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int n = n0 + (wn * TN + j) * 16 + r16;
    co[j] = n < N ? c_col_off(d, n) : -1;
  }
  store_tile<TM, TN>(d, Cv, ro, co, acc);
}

// ------------------------------------------ BF16x3, panels copied into LDS
// Every operand of every segment is an MFMA panel (GemmSeg::pblkA/B > 0).
// A slice of the tile is (BM + BN) / 16 row blocks of each operand, hi and
// lo: OPS one-kilobyte panel blocks, copied verbatim into LDS by
// global_load_lds_dwordx4 (one block per wave instruction, lane l moving
// bytes 16l..16l+15), so no registers stage the tile and no LDS store
// instructions run.  The panel swizzle (panel_chunk) makes the fragment
// reads of the copied 64-byte rows conflict-free.  STAGES LDS stages; the
// blocks of slice t + STAGES - 1 are issued right after the barrier of
// slice t, so each copy has STAGES - 1 slices of MFMA work to land; one
// barrier per slice.  Every wave issues the same OPS / 4 copies per slice
// (blocks past the operand's last row block repeat that block, and the
// rows they feed are never stored), so "slice t has landed" is an exact
// vmcnt wait.  A negative segment sign flips the sign bits of the B
// fragments in registers.
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt is six bits");
  // gfx9 s_waitcnt: vmcnt[3:0], expcnt[6:4] (7: no wait), lgkmcnt[11:8]
  // (15: no wait), vmcnt[5:4] in bits [15:14]
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <int WGM, int WGN, int TM, int TN, int STAGES, int TAG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(STAGES == 2 ? 2 : 1))) void k_gemm_b3d(
    GemmDesc d, const GemmPtrs *__restrict__ ptrs) {
  static_assert(STAGES == 2 || STAGES == 3, "two or three LDS stages");
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  constexpr int RA = BM / 16, RB = BN / 16, OPS = 2 * (RA + RB), OPW = OPS / 4;
  static_assert(OPS % 4 == 0, "every wave issues the same number of block copies");
  // One array per stage, each named statically (the loop is unrolled by
  // STAGES): the compiler can then tell that fragment reads of one stage do
  // not alias copies in flight into another, and does not drain them.
  // Slots per stage: [A hi: RA][A lo: RA][B hi: RB][B lo: RB], 512 halves.
  __shared__ __attribute__((aligned(1024))) unsigned short L0[OPS][512];
  __shared__ __attribute__((aligned(1024))) unsigned short L1[OPS][512];
  __shared__ __attribute__((aligned(1024))) unsigned short L2[STAGES == 3 ? OPS : 1][512];

  const int M = d.M, N = d.N;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn, bx = blockIdx.x;
  long z = blockIdx.y;
  int u = 0;  // sub-entry (d.zsub > 0)
  if (d.zsub > 0) {
    zsub_of(d.zgroup > 1 ? d.zgroup : 1, d.zsub, tiles_m * tiles_n, bx, z, u);
    tm = bx % tiles_m;
    tn = bx / tiles_m;
  } else if (d.zgroup > 1) {
    zgroup_of(d.zgroup, bx, z);
    tm = bx % tiles_m;
    tn = bx / tiles_m;
  } else {
    tile_of(bx, gridDim.x, tiles_m, tiles_n, tm, tn);
  }
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / WGN, wn = wave % WGN;
  const int r16 = lane & 15, kq = lane >> 4;
  // fragment offset in a block: row r16, logical chunk kq
  const int foff = r16 * BK3 + panel_chunk(r16, kq) * 8;

  Segs<unsigned short, BK3> S;
  S.init(d, z);
  const int ntiles = S.ntiles;
  const int lastA = (M + 15) / 16 - 1, lastB = (N + 15) / 16 - 1;

  int cur_seg = -1;
  const unsigned short *bA = nullptr, *bB = nullptr;
  long loA = 0, loB = 0, pbA = 0, pbB = 0;
  auto issue = [&](int t, unsigned short (*Ls)[512]) {
    int k0;
    const int s = S.locate(t, k0);
    if (s != cur_seg) {
      bA = S.a(d, ptrs, z, s) + (long)u * d.seg[s].subA;
      bB = S.b(d, ptrs, z, s) + (long)u * d.seg[s].subB;
      loA = d.seg[s].loA; loB = d.seg[s].loB; pbA = d.seg[s].pblkA; pbB = d.seg[s].pblkB;
      cur_seg = s;
    }
    const long ks = k0 / BK3;
#pragma unroll
    for (int q = 0; q < OPW; q++) {
      const int op = wave + 4 * q;
      const unsigned short *src;
      if (op < 2 * RA) {
        const int h = op / RA, rb = min(m0 / 16 + op % RA, lastA);
        src = bA + (h ? loA : 0) + rb * pbA + ks * 512;
      } else {
        const int o2 = op - 2 * RA, h = o2 / RB, rb = min(n0 / 16 + o2 % RB, lastB);
        src = bB + (h ? loB : 0) + rb * pbB + ks * 512;
      }
      __builtin_amdgcn_global_load_lds((const void *)(src + lane * 8),
                                       (__attribute__((address_space(3))) void *)&Ls[op][0], 16, 0, 0);
    }
  };

  flt4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};

  // One slice: wait until this wave's copies of slice t have landed (at most
  // the copies of the slices after it outstanding), barrier (every wave's
  // copies are in; the stage read in slice t - 1 is free), read slice t's
  // fragments, refill the free stage with slice t + STAGES - 1, then the
  // MFMAs.  The barrier is a plain s_barrier: the explicit vmcnt wait is the
  // only ordering the copies need (a fenced barrier would drain the copies
  // just issued).  The copies are issued after the fragment reads because
  // the compiler waits for every copy in flight before an LDS read.
  constexpr int PRE = STAGES - 1;
  auto step = [&](int t, const unsigned short (*Lr)[512], unsigned short (*Li)[512]) {
    if (STAGES == 3 && t + 1 < ntiles) wait_vmcnt<OPW>();
    else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    int k0;
    const int s = S.locate(t, k0);
    const bool neg = S.sg(d, s) < 0;
    bf16x8 ah[TM], al[TM];
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const int blk = wm * TM + i;
      ah[i] = *reinterpret_cast<const bf16x8 *>(&Lr[blk][foff]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&Lr[RA + blk][foff]);
    }
    if (t + PRE < ntiles) issue(t + PRE, Li);
    // B columns software-pipelined: column j + 1's fragments are read
    // before column j's MFMAs, so each read overlaps MFMAs
    auto rdB = [&](int j, uint4 &uh, uint4 &ul) {
      const int blk = wn * TN + j;
      uh = *reinterpret_cast<const uint4 *>(&Lr[2 * RA + blk][foff]);
      ul = *reinterpret_cast<const uint4 *>(&Lr[2 * RA + RB + blk][foff]);
    };
    uint4 ch, cl;
    rdB(0, ch, cl);
#pragma unroll
    for (int j = 0; j < TN; j++) {
      uint4 nh = ch, nl = cl;
      if (j + 1 < TN) rdB(j + 1, nh, nl);
      if (neg) {
        ch.x ^= 0x80008000u; ch.y ^= 0x80008000u; ch.z ^= 0x80008000u; ch.w ^= 0x80008000u;
        cl.x ^= 0x80008000u; cl.y ^= 0x80008000u; cl.z ^= 0x80008000u; cl.w ^= 0x80008000u;
      }
      const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&ch), bl = *reinterpret_cast<const bf16x8 *>(&cl);
#pragma unroll
      for (int i = 0; i < TM; i++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
      }
      ch = nh;
      cl = nl;
    }
  };
  issue(0, L0);
  if (STAGES == 3 && ntiles > 1) issue(1, L1);
  for (int t = 0; t < ntiles; t += STAGES) {
    if (STAGES == 2) {
      step(t, L0, L1);
      if (t + 1 < ntiles) step(t + 1, L1, L0);
    } else {
      step(t, L0, L2);
      if (t + 1 < ntiles) step(t + 1, L1, L0);
      if (t + 2 < ntiles) step(t + 2, L2, L1);
    }
  }

  void *Cv = ptrs ? (void *)ptrs[z].C
                  : (d.c_f32 ? (void *)((float *)d.C + z * d.sCz) : (void *)(d.C + z * d.sCz));
  if (d.zsub > 0) Cv = d.c_f32 ? (void *)((float *)Cv + d.offz[u]) : (void *)((double *)Cv + d.offz[u]);
  long ro[TM][4], co[TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = m0 + (wm * TM + i) * 16 + 4 * kq + r;
      ro[i][r] = m < M ? c_row_off(d, m) : -1;
    }
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int n = n0 + (wn * TN + j) * 16 + r16;
    co[j] = n < N ? c_col_off(d, n) : -1;
  }
  store_tile<TM, TN>(d, Cv, ro, co, acc);
}

// ------------------------------------------------------------ dispatch
struct TileShape {
  int wgm, wgn, tm, tn;
  const char *name;
};
// Tile table.  Index order is the public tile id used by the autotuner.
const TileShape kTiles[] = {
    {2, 2, 4, 4, "128x128"}, {2, 2, 2, 2, "64x64"},  {2, 2, 1, 1, "32x32"},  {4, 1, 2, 4, "128x64"},
    {1, 4, 4, 2, "64x128"},  {2, 2, 4, 5, "128x160"}, {4, 1, 1, 4, "64x64w"}, {2, 2, 2, 5, "64x160"},
    {4, 1, 2, 5, "128x80"},  {2, 2, 5, 4, "160x128"}, {1, 4, 5, 2, "80x128"},  {2, 2, 5, 5, "160x160"},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

// Pre-split BF16x3 operands allow 16-byte loads when every segment's row
// strides, K and hi/lo offsets are multiples of 8 and the bases are 16-byte
// aligned (checked on the pointer table's entries by the caller via
// d.flags bit 0).
// LDS stages of the pre-split BF16x3 kernel (tunable, see gemm_set_b3_stages)
int g_b3_stages = 1;

bool b3_vectorizable(const GemmDesc &d) {
  for (int s = 0; s < d.nseg; s++) {
    const GemmSeg &g = d.seg[s];
    if (g.sak != 1 || g.sbk != 1) throw std::runtime_error("ugx: pre-split BF16x3 needs k-contiguous operands");
    if (g.sam % 8 || g.sbn % 8 || g.K % 8 || g.loA % 8 || g.loB % 8 || g.sAz % 8 || g.sBz % 8) return false;
    if (g.pblkA % 512 || g.pblkB % 512) return false;
    if (g.A && (((uintptr_t)g.A) & 15)) return false;
    if (g.B && (((uintptr_t)g.B) & 15)) return false;
  }
  return true;
}

template <typename T, int WGM, int WGN, int TM, int TN>
void launch_t(const GemmDesc &d, const GemmPtrs *ptrs, int prec, hipStream_t st) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  long tiles = (long)((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
  if (d.zsub > 0) tiles *= d.zsub;
  if (tiles > 0x7fffffffL || d.batch > 65535) throw std::runtime_error("ugx: GEMM grid too large");
  dim3 grid((unsigned)tiles, (unsigned)d.batch, 1);
  if (d.zsub > 0 && !(prec == PREC_BF16X3_PRESPLIT && d.dma > 1 && ptrs && d.offz && (2 * (BM + BN) / 16) % 4 == 0))
    throw std::runtime_error("ugx: sub-entries need the LDS-copy BF16x3 kernel, a pointer table and offz");
  if (prec == PREC_BF16X3_PRESPLIT) {
    // tiles whose block copies split evenly over the 4 waves can take the
    // LDS-copy kernel; the others fall back to the register-staged one
    if constexpr ((2 * (BM + BN) / 16) % 4 == 0) {
      if (d.dma > 1) {
        for (int q = 0; q < d.nseg; q++)
          if (d.seg[q].pblkA <= 0 || d.seg[q].pblkB <= 0)
            throw std::runtime_error("ugx: the LDS-copy BF16x3 kernel needs panel operands");
        if (d.dma == 2) {
          if (d.tag) hipLaunchKernelGGL((k_gemm_b3d<WGM, WGN, TM, TN, 2, 1>), grid, dim3(NT), 0, st, d, ptrs);
          else hipLaunchKernelGGL((k_gemm_b3d<WGM, WGN, TM, TN, 2, 0>), grid, dim3(NT), 0, st, d, ptrs);
        } else {
          if (d.tag) hipLaunchKernelGGL((k_gemm_b3d<WGM, WGN, TM, TN, 3, 1>), grid, dim3(NT), 0, st, d, ptrs);
          else hipLaunchKernelGGL((k_gemm_b3d<WGM, WGN, TM, TN, 3, 0>), grid, dim3(NT), 0, st, d, ptrs);
        }
        return;
      }
    }
    bool vec = b3_vectorizable(d) && (ptrs == nullptr || d.vec_ok);
    if (g_b3_stages == 1) {
      if (d.tag) {
        if (vec) hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, true, 1, 1>), grid, dim3(NT), 0, st, d, ptrs);
        else hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, false, 1, 1>), grid, dim3(NT), 0, st, d, ptrs);
      } else {
        if (vec) hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, true, 1, 0>), grid, dim3(NT), 0, st, d, ptrs);
        else hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, false, 1, 0>), grid, dim3(NT), 0, st, d, ptrs);
      }
    } else {
      if (vec) hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, true, 2, 0>), grid, dim3(NT), 0, st, d, ptrs);
      else hipLaunchKernelGGL((k_gemm_b3p<WGM, WGN, TM, TN, false, 2, 0>), grid, dim3(NT), 0, st, d, ptrs);
    }
  } else {
    if (d.tag) hipLaunchKernelGGL((k_gemm<T, WGM, WGN, TM, TN, 1>), grid, dim3(NT), 0, st, d, ptrs);
    else hipLaunchKernelGGL((k_gemm<T, WGM, WGN, TM, TN, 0>), grid, dim3(NT), 0, st, d, ptrs);
  }
}

template <typename T>
void launch_tile(int tile, const GemmDesc &d, const GemmPtrs *ptrs, int prec, hipStream_t st) {
  switch (tile) {
    case 0: launch_t<T, 2, 2, 4, 4>(d, ptrs, prec, st); break;
    case 1: launch_t<T, 2, 2, 2, 2>(d, ptrs, prec, st); break;
    case 2: launch_t<T, 2, 2, 1, 1>(d, ptrs, prec, st); break;
    case 3: launch_t<T, 4, 1, 2, 4>(d, ptrs, prec, st); break;
    case 4: launch_t<T, 1, 4, 4, 2>(d, ptrs, prec, st); break;
    case 5: launch_t<T, 2, 2, 4, 5>(d, ptrs, prec, st); break;
    case 6: launch_t<T, 4, 1, 1, 4>(d, ptrs, prec, st); break;
    case 7: launch_t<T, 2, 2, 2, 5>(d, ptrs, prec, st); break;
    case 8: launch_t<T, 4, 1, 2, 5>(d, ptrs, prec, st); break;
    case 9: launch_t<T, 2, 2, 5, 4>(d, ptrs, prec, st); break;
    case 10: launch_t<T, 1, 4, 5, 2>(d, ptrs, prec, st); break;
    case 11: launch_t<T, 2, 2, 5, 5>(d, ptrs, prec, st); break;
    default: throw std::runtime_error("ugx: bad GEMM tile id");
  }
}

// Default tile choice: the largest tile whose grid still puts at least two
// workgroups on each of the 256 CUs; skinny shapes get the matching
// rectangular tiles.
int pick_tile_impl(const GemmDesc &d) {
  auto tiles = [&](int bm, int bn) { return (long)((d.M + bm - 1) / bm) * ((d.N + bn - 1) / bn) * d.batch; };
  const long want = 512;
  if (d.M >= 128 && d.N >= 128 && tiles(128, 128) >= want) return 0;
  if (d.N <= 64 && d.M >= 128 && tiles(128, 64) >= want / 2) return 3;
  if (d.M <= 64 && d.N >= 128 && tiles(64, 128) >= want / 2) return 4;
  if (d.M >= 64 && d.N >= 64 && tiles(64, 64) >= want / 4) return 1;
  if (d.M <= 32 || d.N <= 32) return 2;
  return 1;
}

}  // namespace

int gemm_num_tiles() { return kNumTiles; }
void gemm_set_b3_stages(int s) { g_b3_stages = (s == 1) ? 1 : 2; }
int gemm_pick_tile(const GemmDesc &d) { return pick_tile_impl(d); }
void gemm_tile_dims(int tile, int &bm, int &bn) {
  const TileShape &t = kTiles[tile];
  bm = 16 * t.wgm * t.tm;
  bn = 16 * t.wgn * t.tn;
}
const char *gemm_tile_name(int tile) { return (tile >= 0 && tile < kNumTiles) ? kTiles[tile].name : "?"; }

void launch_gemm(const GemmDesc &d, const GemmPtrs *ptrs, int prec, int tile, hipStream_t st) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  if (d.nseg < 1 || d.nseg > MAXSEG) throw std::runtime_error("ugx: GEMM needs one to eight K segments");
  if (tile < 0) tile = pick_tile_impl(d);
  if (prec == PREC_FP64) launch_tile<double>(tile, d, ptrs, prec, st);
  else if (prec == PREC_FP32 || prec == PREC_BF16X3_PRESPLIT) launch_tile<float>(tile, d, ptrs, prec, st);
  else if (prec == PREC_BF16X3)
    throw std::runtime_error("ugx: BF16x3 contractions take pre-split operands (PREC_BF16X3_PRESPLIT)");
  else throw std::runtime_error("ugx: unknown contraction precision");
}

}  // namespace ugx
