// Dense contractions on gfx950 matrix cores.
//
// One kernel family serves every contraction of the CCSD residual and the
// (T) W3 build:
//   C(m,n) = alpha * sum_seg sign_s * sum_k A_s(m,k) B_s(k,n) + beta * C(m,n)
// with up to two K segments that have their own operands and strides (the
// (T) particle term, K = v, and hole term, K = o, of one output permutation
// share an accumulator that way), a batch dimension (strided or a device
// table of pointers), and an output row split so that a (T) permutation can
// write W[a][b][c] with any of the six index orders.
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

#include <stdexcept>

namespace ugx {
namespace {

constexpr int NT = 256;
constexpr int BK = 16;

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

struct SegView {
  const void *A;
  const void *B;
  long sam, sak, sbk, sbn;
  int K;
  double sign;
};

__device__ inline SegView seg_view(const GemmDesc &d, const GemmPtrs *ptrs, int s, long z) {
  const GemmSeg &g = d.seg[s];
  SegView v;
  v.sam = g.sam; v.sak = g.sak; v.sbk = g.sbk; v.sbn = g.sbn; v.K = g.K; v.sign = g.sign;
  if (ptrs) {
    v.A = ptrs[z].A[s];
    v.B = ptrs[z].B[s];
  } else {
    v.A = nullptr; v.B = nullptr;  // resolved by the caller with the element size
  }
  return v;
}

// Loads one BK-slice of A (BM rows) into registers, scaled by the segment
// sign.  Follows the unit stride: k-contiguous (sak == 1) reads BK
// consecutive k per row; otherwise m-contiguous.
template <typename T, int BM>
__device__ inline void load_a(T (&r)[BM * BK / NT], const T *A, const SegView &S, int m0, int k0, int M, int tid) {
  const bool kc = (S.sak == 1);
#pragma unroll
  for (int q = 0; q < BM * BK / NT; q++) {
    int e = tid + q * NT;
    int row, kk;
    if (kc) { row = e / BK; kk = e % BK; } else { kk = e / BM; row = e % BM; }
    int m = m0 + row, k = k0 + kk;
    T v = T(0);
    if (m < M && k < S.K) v = A[(long)m * S.sam + (long)k * S.sak];
    r[q] = v * T(S.sign);
  }
}

template <typename T, int BM, int LD>
__device__ inline void store_a(const T (&r)[BM * BK / NT], T (*As)[LD], bool kc, int tid) {
#pragma unroll
  for (int q = 0; q < BM * BK / NT; q++) {
    int e = tid + q * NT;
    int row, kk;
    if (kc) { row = e / BK; kk = e % BK; } else { kk = e / BM; row = e % BM; }
    As[kk][row] = r[q];
  }
}

template <typename T, int BN>
__device__ inline void load_b(T (&r)[BN * BK / NT], const T *B, const SegView &S, int n0, int k0, int N, int tid) {
  const bool kc = (S.sbk == 1);
#pragma unroll
  for (int q = 0; q < BN * BK / NT; q++) {
    int e = tid + q * NT;
    int col, kk;
    if (kc) { col = e / BK; kk = e % BK; } else { kk = e / BN; col = e % BN; }
    int n = n0 + col, k = k0 + kk;
    T v = T(0);
    if (n < N && k < S.K) v = B[(long)k * S.sbk + (long)n * S.sbn];
    r[q] = v;
  }
}

template <typename T, int BN, int LD>
__device__ inline void store_b(const T (&r)[BN * BK / NT], T (*Bs)[LD], bool kc, int tid) {
#pragma unroll
  for (int q = 0; q < BN * BK / NT; q++) {
    int e = tid + q * NT;
    int col, kk;
    if (kc) { col = e / BK; kk = e % BK; } else { kk = e / BN; col = e % BN; }
    Bs[kk][col] = r[q];
  }
}

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

template <typename T, int WGM, int WGN, int TM, int TN>
__global__ __launch_bounds__(NT) void k_gemm(GemmDesc d, const GemmPtrs *__restrict__ ptrs) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  constexpr int LDA = BM + (sizeof(T) == 8 ? 1 : 2), LDB = BN + (sizeof(T) == 8 ? 1 : 2);
  constexpr int RA = BM * BK / NT, RB = BN * BK / NT;
  static_assert(RA >= 1 && RB >= 1, "tile too small for 256 threads");
  __shared__ T As[BK][LDA];
  __shared__ T Bs[BK][LDB];

  const int M = d.M, N = d.N;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const long z = blockIdx.y;

  SegView S[2];
  const T *Ap[2];
  const T *Bp[2];
  int nt[2] = {0, 0};
  for (int s = 0; s < 2; s++) {
    if (s < d.nseg) {
      S[s] = seg_view(d, ptrs, s, z);
      if (ptrs) { Ap[s] = (const T *)S[s].A; Bp[s] = (const T *)S[s].B; }
      else {
        Ap[s] = (const T *)d.seg[s].A + z * d.seg[s].sAz;
        Bp[s] = (const T *)d.seg[s].B + z * d.seg[s].sBz;
      }
      nt[s] = (S[s].K + BK - 1) / BK;
    } else {
      S[s] = S[0]; Ap[s] = nullptr; Bp[s] = nullptr; nt[s] = 0;
    }
  }
  const int ntiles = nt[0] + nt[1];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int r16 = lane & 15, kq = lane >> 4;

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = MF::zero();

  T ra[RA], rb[RB];
  auto fetch = [&](int t) {
    int s = (t < nt[0]) ? 0 : 1;
    int k0 = (s == 0 ? t : t - nt[0]) * BK;
    load_a<T, BM>(ra, Ap[s], S[s], m0, k0, M, tid);
    load_b<T, BN>(rb, Bp[s], S[s], n0, k0, N, tid);
    return s;
  };
  if (ntiles > 0) {
    int s = fetch(0);
    store_a<T, BM, LDA>(ra, As, S[s].sak == 1, tid);
    store_b<T, BN, LDB>(rb, Bs, S[s].sbk == 1, tid);
  }
  __syncthreads();

  for (int t = 0; t < ntiles; t++) {
    const bool more = (t + 1 < ntiles);
    int sn = 0;
    if (more) sn = fetch(t + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = As[kk + kq][(wm * TM + i) * 16 + r16];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bs[kk + kq][(wn * TN + j) * 16 + r16];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = MF::op(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
    if (more) {
      store_a<T, BM, LDA>(ra, As, S[sn].sak == 1, tid);
      store_b<T, BN, LDB>(rb, Bs, S[sn].sbk == 1, tid);
    }
    __syncthreads();
  }

  double *C = ptrs ? ptrs[z].C : d.C + z * d.sCz;
  const double alpha = d.alpha, beta = d.beta;
  const int cdiv = d.cdiv;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int m = m0 + (wm * TM + i) * 16 + MF::row(kq, r);
        int n = n0 + (wn * TN + j) * 16 + r16;
        if (m < M && n < N) {
          long addr = (long)(m / cdiv) * d.scm1 + (long)(m % cdiv) * d.scm + (long)n * d.scn;
          double v = alpha * (double)acc[i][j][r];
          if (beta != 0.0) v += beta * C[addr];
          C[addr] = v;
        }
      }
}

// ------------------------------------------------------------ BF16x3
// FP32 operands are split on the LDS store into hi = bf16(x) and
// lo = bf16(x - hi); the product is hi*hi + hi*lo + lo*hi with FP32
// accumulation (the lo*lo term, ~2^-16 relative, is dropped).  LDS holds the
// splits m-major ([m][k], k fastest) so each lane's eight consecutive k
// values are one 16-byte read.
constexpr int BK3 = 32;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline unsigned short bf16_bits_rne(float x) {
  unsigned int u = __float_as_uint(x);
  unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(r >> 16);
}
__device__ inline float bf16_to_float(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <int BM>
__device__ inline void load3_a(float (&r)[BM * BK3 / NT], const float *A, const SegView &S, int m0, int k0, int M,
                               int tid) {
  const bool kc = (S.sak == 1);
#pragma unroll
  for (int q = 0; q < BM * BK3 / NT; q++) {
    int e = tid + q * NT;
    int row, kk;
    if (kc) { row = e / BK3; kk = e % BK3; } else { kk = e / BM; row = e % BM; }
    int m = m0 + row, k = k0 + kk;
    float v = 0.f;
    if (m < M && k < S.K) v = A[(long)m * S.sam + (long)k * S.sak];
    r[q] = v * (float)S.sign;
  }
}

template <int BN>
__device__ inline void load3_b(float (&r)[BN * BK3 / NT], const float *B, const SegView &S, int n0, int k0, int N,
                               int tid) {
  const bool kc = (S.sbk == 1);
#pragma unroll
  for (int q = 0; q < BN * BK3 / NT; q++) {
    int e = tid + q * NT;
    int col, kk;
    if (kc) { col = e / BK3; kk = e % BK3; } else { kk = e / BN; col = e % BN; }
    int n = n0 + col, k = k0 + kk;
    float v = 0.f;
    if (n < N && k < S.K) v = B[(long)k * S.sbk + (long)n * S.sbn];
    r[q] = v;
  }
}

// Stores split values: Hs/Ls[row][k] (row-major, LD = BK3 + 8 halves).
template <int BR, int LD>
__device__ inline void store3(const float (&r)[BR * BK3 / NT], unsigned short (*Hs)[LD], unsigned short (*Ls)[LD],
                              bool kc, int tid) {
#pragma unroll
  for (int q = 0; q < BR * BK3 / NT; q++) {
    int e = tid + q * NT;
    int row, kk;
    if (kc) { row = e / BK3; kk = e % BK3; } else { kk = e / BR; row = e % BR; }
    float x = r[q];
    unsigned short h = bf16_bits_rne(x);
    unsigned short l = bf16_bits_rne(x - bf16_to_float(h));
    Hs[row][kk] = h;
    Ls[row][kk] = l;
  }
}

template <int WGM, int WGN, int TM, int TN>
__global__ __launch_bounds__(NT) void k_gemm_bf16x3(GemmDesc d, const GemmPtrs *__restrict__ ptrs) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  constexpr int LD = BK3 + 8;
  constexpr int RA = BM * BK3 / NT, RB = BN * BK3 / NT;
  __shared__ __attribute__((aligned(16))) unsigned short Ah[BM][LD];
  __shared__ __attribute__((aligned(16))) unsigned short Al[BM][LD];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[BN][LD];
  __shared__ __attribute__((aligned(16))) unsigned short Bl[BN][LD];

  const int M = d.M, N = d.N;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int m0 = tm * BM, n0 = tn * BN;
  const long z = blockIdx.y;

  SegView S[2];
  const float *Ap[2];
  const float *Bp[2];
  int nt[2] = {0, 0};
  for (int s = 0; s < 2; s++) {
    if (s < d.nseg) {
      S[s] = seg_view(d, ptrs, s, z);
      if (ptrs) { Ap[s] = (const float *)S[s].A; Bp[s] = (const float *)S[s].B; }
      else {
        Ap[s] = (const float *)d.seg[s].A + z * d.seg[s].sAz;
        Bp[s] = (const float *)d.seg[s].B + z * d.seg[s].sBz;
      }
      nt[s] = (S[s].K + BK3 - 1) / BK3;
    } else {
      S[s] = S[0]; Ap[s] = nullptr; Bp[s] = nullptr; nt[s] = 0;
    }
  }
  const int ntiles = nt[0] + nt[1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int r16 = lane & 15, kq = lane >> 4;

  flt4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = flt4{0.f, 0.f, 0.f, 0.f};

  float ra[RA], rb[RB];
  auto fetch = [&](int t) {
    int s = (t < nt[0]) ? 0 : 1;
    int k0 = (s == 0 ? t : t - nt[0]) * BK3;
    load3_a<BM>(ra, Ap[s], S[s], m0, k0, M, tid);
    load3_b<BN>(rb, Bp[s], S[s], n0, k0, N, tid);
    return s;
  };
  if (ntiles > 0) {
    int s = fetch(0);
    store3<BM, LD>(ra, Ah, Al, S[s].sak == 1, tid);
    store3<BN, LD>(rb, Bh, Bl, S[s].sbk == 1, tid);
  }
  __syncthreads();

  for (int t = 0; t < ntiles; t++) {
    const bool more = (t + 1 < ntiles);
    int sn = 0;
    if (more) sn = fetch(t + 1);
    bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; i++) {
      int row = (wm * TM + i) * 16 + r16;
      ah[i] = *reinterpret_cast<const bf16x8 *>(&Ah[row][8 * kq]);
      al[i] = *reinterpret_cast<const bf16x8 *>(&Al[row][8 * kq]);
    }
#pragma unroll
    for (int j = 0; j < TN; j++) {
      int col = (wn * TN + j) * 16 + r16;
      bh[j] = *reinterpret_cast<const bf16x8 *>(&Bh[col][8 * kq]);
      bl[j] = *reinterpret_cast<const bf16x8 *>(&Bl[col][8 * kq]);
    }
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
    if (more) {
      store3<BM, LD>(ra, Ah, Al, S[sn].sak == 1, tid);
      store3<BN, LD>(rb, Bh, Bl, S[sn].sbk == 1, tid);
    }
    __syncthreads();
  }

  double *C = ptrs ? ptrs[z].C : d.C + z * d.sCz;
  const double alpha = d.alpha, beta = d.beta;
  const int cdiv = d.cdiv;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int m = m0 + (wm * TM + i) * 16 + 4 * kq + r;
        int n = n0 + (wn * TN + j) * 16 + r16;
        if (m < M && n < N) {
          long addr = (long)(m / cdiv) * d.scm1 + (long)(m % cdiv) * d.scm + (long)n * d.scn;
          double v = alpha * (double)acc[i][j][r];
          if (beta != 0.0) v += beta * C[addr];
          C[addr] = v;
        }
      }
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
    {4, 1, 2, 5, "128x80"},
};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

template <typename T, int WGM, int WGN, int TM, int TN>
void launch_t(const GemmDesc &d, const GemmPtrs *ptrs, int prec, hipStream_t st) {
  constexpr int BM = 16 * WGM * TM, BN = 16 * WGN * TN;
  long tiles = (long)((d.M + BM - 1) / BM) * ((d.N + BN - 1) / BN);
  if (tiles > 0x7fffffffL || d.batch > 65535) throw std::runtime_error("ugx: GEMM grid too large");
  dim3 grid((unsigned)tiles, (unsigned)d.batch, 1);
  if (prec == PREC_BF16X3) {
    if constexpr (BM * BK3 / NT >= 1 && BN * BK3 / NT >= 1)
      hipLaunchKernelGGL((k_gemm_bf16x3<WGM, WGN, TM, TN>), grid, dim3(NT), 0, st, d, ptrs);
  } else {
    hipLaunchKernelGGL((k_gemm<T, WGM, WGN, TM, TN>), grid, dim3(NT), 0, st, d, ptrs);
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
    default: throw std::runtime_error("ugx: bad GEMM tile id");
  }
}

// Default tile choice: the largest tile whose grid still puts at least two
// workgroups on each of the 256 CUs; skinny shapes get the matching
// rectangular tiles.
int pick_tile(const GemmDesc &d) {
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
const char *gemm_tile_name(int tile) { return (tile >= 0 && tile < kNumTiles) ? kTiles[tile].name : "?"; }

void launch_gemm(const GemmDesc &d, const GemmPtrs *ptrs, int prec, int tile, hipStream_t st) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  if (d.nseg < 1 || d.nseg > 2) throw std::runtime_error("ugx: GEMM needs one or two K segments");
  if (d.cdiv <= 0) throw std::runtime_error("ugx: GEMM row split must be positive");
  if (tile < 0) tile = pick_tile(d);
  if (prec == PREC_FP64) launch_tile<double>(tile, d, ptrs, prec, st);
  else if (prec == PREC_FP32 || prec == PREC_BF16X3) launch_tile<float>(tile, d, ptrs, prec, st);
  else throw std::runtime_error("ugx: unknown contraction precision");
}

}  // namespace ugx
