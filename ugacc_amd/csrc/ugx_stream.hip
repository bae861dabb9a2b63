// Streaming kernels: tensor permutation (with type conversion and
// accumulation), tau, the Jacobi update with the residual norm, dot products,
// the DIIS vectors and the correlation energy.  All are HBM-bound.  Each
// writes its output coalesced; the permutation kernel stages a 64x64 tile
// through LDS whenever the output's fastest index is not the input's, so
// both sides stay coalesced.  Reductions are deterministic: per-block
// partial sums (wave64 butterflies, then across the block's waves) followed
// by one single-block pass, all in FP64.
#include "ugx_internal.h"

#include <algorithm>
#include <cstdint>
#include <stdexcept>

namespace ugx {
namespace {

constexpr int NT = 256;

inline int nblocks(long n, int per = NT, int cap = 1 << 16) {
  long b = (n + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

__device__ inline double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[NT / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / 64; k++) s += red[k];
  return s;  // valid in thread 0
}

template <typename T>
__device__ inline double ld(const void *p, long i) { return (double)((const T *)p)[i]; }

// ---------------------------------------------------------------- permute
struct PermArgs {
  int rank;
  int od[6];
  long istr[6];
  const void *in;
  void *out;
  double alpha, beta;
  long n;  // total elements (offset of the lo half of a BF16 pair output)
  long pk, pblk, plo;  // MFMA panel output: k extent, row-block stride, lo offset
  // n < 2^31: output indices are split into coordinates in 32 bits, by
  // multiply-high with the magic numbers of od[] and pk (div31)
  int i32;
  unsigned dm[6], pkm;
  int ds[6], pks;
  // eight consecutive outputs per thread: k_perm_linear8 (vec8), the write
  // side of k_perm_tiled (vec8t)
  int vec8, vec8t;
};

// floor(x / d) for 0 <= x < 2^31 without a division (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994): with l = ceil(log2 d) >= 1 and m = floor(2^(31+l) / d) + 1, which
// is below 2^32, x m / 2^(31+l) = x / d + e with 0 < e < 2^-l <= 1 / d, so
// floor(x / d) = mulhi32(x, m) >> (l - 1).  m = 0 stands for d = 1.
__device__ inline unsigned div31(unsigned x, unsigned m, int s) { return m ? (__umulhi(x, m) >> s) : x; }

void magic31(long d, unsigned &m, int &s) {
  if (d <= 1) {
    m = 0;
    s = 0;
    return;
  }
  int l = 0;
  while ((1L << l) < d) l++;
  m = (unsigned)((1ULL << (31 + l)) / (unsigned long long)d + 1);
  s = l - 1;
}

// Output element policies: FP64, FP32, and the split BF16 pair (hi at
// out[idx], lo at out[n + idx], hi = bf16(x), lo = bf16(x - hi), both round
// to nearest even; accumulation is not offered for the pair).
struct OutD {
  typedef double tile_t;
  __device__ static double get(const PermArgs &p, long i) { return ((const double *)p.out)[i]; }
  __device__ static void put(const PermArgs &p, long i, double v) { ((double *)p.out)[i] = v; }
  __device__ static void put8(const PermArgs &p, long i, const double *v) {
    for (int u = 0; u < 8; u++) put(p, i + u, v[u]);
  }
};
__device__ inline unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
struct OutF {
  typedef float tile_t;
  __device__ static double get(const PermArgs &p, long i) { return ((const float *)p.out)[i]; }
  __device__ static void put(const PermArgs &p, long i, double v) { ((float *)p.out)[i] = (float)v; }
  __device__ static void put8(const PermArgs &p, long i, const double *v) {
    for (int u = 0; u < 8; u++) put(p, i + u, v[u]);
  }
};
// eight values as BF16 hi and lo halves, packed two per word
__device__ inline void split8(const double *v, uint4 &hi, uint4 &lo) {
  unsigned int h[8], l[8];
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const float x = (float)v[u];
    h[u] = bf16_rne(x);
    l[u] = bf16_rne(x - __uint_as_float(h[u] << 16));
  }
  hi = uint4{h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
  lo = uint4{l[0] | (l[1] << 16), l[2] | (l[3] << 16), l[4] | (l[5] << 16), l[6] | (l[7] << 16)};
}
struct OutB3 {
  typedef float tile_t;
  __device__ static double get(const PermArgs &, long) { return 0.0; }
  __device__ static void put(const PermArgs &p, long i, double v) {
    float x = (float)v;
    unsigned short h = bf16_rne(x);
    float hf = __uint_as_float(((unsigned int)h) << 16);
    ((unsigned short *)p.out)[i] = h;
    ((unsigned short *)p.out)[p.n + i] = bf16_rne(x - hf);
  }
  // elements i .. i + 7, i a multiple of 8 (16-byte aligned halves)
  __device__ static void put8(const PermArgs &p, long i, const double *v) {
    uint4 hi, lo;
    split8(v, hi, lo);
    *reinterpret_cast<uint4 *>((unsigned short *)p.out + i) = hi;
    *reinterpret_cast<uint4 *>((unsigned short *)p.out + p.n + i) = lo;
  }
};

// Split BF16 pair into MFMA panels: element i of the [n / pk][pk] output
// (row r, k) goes to panel block r / 16, k slice k / 32, row r % 16, stored
// 8-k chunk panel_chunk(r % 16, k % 32 / 8) (the layout of launch_pack_panel)
struct OutP {
  typedef float tile_t;
  __device__ static double get(const PermArgs &, long) { return 0.0; }
  __device__ static void put(const PermArgs &p, long i, double v) {
    long r, k;
    if (p.i32) {
      const unsigned q = div31((unsigned)i, p.pkm, p.pks);
      r = q;
      k = (unsigned)i - q * (unsigned)p.pk;
    } else {
      r = i / p.pk;
      k = i - r * p.pk;
    }
    const int rr = (int)(r & 15), kk = (int)(k & 31);
    const long a = (r >> 4) * p.pblk + (k >> 5) * 512 + rr * 32 + 8 * panel_chunk(rr, kk >> 3) + (kk & 7);
    float x = (float)v;
    unsigned short h = bf16_rne(x);
    float hf = __uint_as_float(((unsigned int)h) << 16);
    ((unsigned short *)p.out)[a] = h;
    ((unsigned short *)p.out)[p.plo + a] = bf16_rne(x - hf);
  }
  // elements i .. i + 7 of one row, i a multiple of 8 and pk too: one stored
  // 8-k chunk (n < 2^31)
  __device__ static void put8(const PermArgs &p, long i, const double *v) {
    const unsigned q = div31((unsigned)i, p.pkm, p.pks);
    const long r = q, k = (unsigned)i - q * (unsigned)p.pk;
    const int rr = (int)(r & 15), kk = (int)(k & 31);
    const long a = (r >> 4) * p.pblk + (k >> 5) * 512 + rr * 32 + 8 * panel_chunk(rr, kk >> 3);
    uint4 hi, lo;
    split8(v, hi, lo);
    *reinterpret_cast<uint4 *>((unsigned short *)p.out + a) = hi;
    *reinterpret_cast<uint4 *>((unsigned short *)p.out + p.plo + a) = lo;
  }
};

// Linear form: used when the output's fastest index is also the input's
// unit-stride index (or the tensor is small).
template <typename TI, class OW>
__global__ void k_perm_linear(PermArgs p, long n) {
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < n; idx += (long)gridDim.x * NT) {
    long off = 0;
    if (p.i32) {
      // all six dims, unrolled: the ones past the rank have extent 1 and
      // stride 0
      unsigned r = (unsigned)idx;
#pragma unroll
      for (int d = 5; d >= 0; d--) {
        const unsigned q = div31(r, p.dm[d], p.ds[d]);
        off += (long)(r - q * (unsigned)p.od[d]) * p.istr[d];
        r = q;
      }
    } else {
      long r = idx;
      for (int d = p.rank - 1; d >= 0; d--) {
        int c = (int)(r % p.od[d]);
        r /= p.od[d];
        off += c * p.istr[d];
      }
    }
    double v = p.alpha * (double)((const TI *)p.in)[off];
    if (p.beta != 0.0) v += p.beta * OW::get(p, idx);
    OW::put(p, idx, v);
  }
}

// Eight consecutive outputs per thread (vec8: BF16 pair outputs, n < 2^31,
// no accumulation, the innermost output dim of unit input stride and a
// multiple of 8 extent, so outputs idx .. idx + 7 read inputs off .. off + 7):
// one index split, eight loads, two 16-byte stores.
template <typename TI, class OW>
__global__ void k_perm_linear8(PermArgs p, long n8) {
  for (long t = blockIdx.x * (long)NT + threadIdx.x; t < n8; t += (long)gridDim.x * NT) {
    const unsigned idx = (unsigned)(t * 8);
    long off = 0;
    unsigned r = idx;
#pragma unroll
    for (int d = 5; d >= 0; d--) {
      const unsigned q = div31(r, p.dm[d], p.ds[d]);
      off += (long)(r - q * (unsigned)p.od[d]) * p.istr[d];
      r = q;
    }
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = p.alpha * (double)((const TI *)p.in)[off + u];
    OW::put8(p, idx, v);
  }
}

// Tiled form: q = the output dim with input stride 1, l = the last output
// dim.  A 64x64 (q,l) tile is read along q and written along l.
template <typename TI, class OW>
__global__ void k_perm_tiled(PermArgs p, int q, long nrest, int tq, int tl) {
  typedef typename OW::tile_t TT;
  __shared__ TT tile[64][65];
  const int L = p.rank - 1;
  const int Dq = p.od[q], Dl = p.od[L];
  long ostr[6];
  {
    long s = 1;
    for (int d = p.rank - 1; d >= 0; d--) { ostr[d] = s; s *= p.od[d]; }
  }
  long nb = (long)tq * tl * nrest;
  for (long b = blockIdx.x; b < nb; b += gridDim.x) {
    long r = b;
    int bl = (int)(r % tl); r /= tl;
    int bq = (int)(r % tq); r /= tq;
    // remaining dims (all except q and L), row-major over output order
    long ibase = 0, obase = 0;
    for (int d = p.rank - 1; d >= 0; d--) {
      if (d == q || d == L) continue;
      int c = (int)(r % p.od[d]);
      r /= p.od[d];
      ibase += c * p.istr[d];
      obase += c * ostr[d];
    }
    const int q0 = bq * 64, l0 = bl * 64;
    // read: threads along q (unit input stride)
    for (int e = threadIdx.x; e < 64 * 64; e += NT) {
      int lq = e & 63, ll = e >> 6;
      int gq = q0 + lq, gl = l0 + ll;
      TT v = 0;
      if (gq < Dq && gl < Dl) v = (TT)(p.alpha * (double)((const TI *)p.in)[ibase + gq + gl * p.istr[L]]);
      tile[lq][ll] = v;
    }
    __syncthreads();
    if (p.vec8t) {
      // eight consecutive l per thread (Dl and every output stride a
      // multiple of 8: one stored chunk each)
      for (int e = threadIdx.x; e < 64 * 8; e += NT) {
        const int ll = (e & 7) * 8, lq = e >> 3;
        const int gq = q0 + lq, gl = l0 + ll;
        if (gq < Dq && gl < Dl) {
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) v[u] = (double)tile[lq][ll + u];
          OW::put8(p, obase + gq * ostr[q] + gl, v);
        }
      }
    } else {
      for (int e = threadIdx.x; e < 64 * 64; e += NT) {
        int ll = e & 63, lq = e >> 6;
        int gq = q0 + lq, gl = l0 + ll;
        if (gq < Dq && gl < Dl) {
          long o = obase + gq * ostr[q] + gl;
          double v = (double)tile[lq][ll];
          if (p.beta != 0.0) v += p.beta * OW::get(p, o);
          OW::put(p, o, v);
        }
      }
    }
    __syncthreads();
  }
}

template <typename TI, class OW>
void perm_launch(const PermArgs &a, hipStream_t st) {
  long n = a.n;
  if (n <= 0) return;
  const int L = a.rank - 1;
  int q = -1;
  if (a.istr[L] != 1)
    for (int d = 0; d < L; d++)
      if (a.istr[d] == 1 && a.od[d] > 1) q = d;
  if (a.vec8) {
    hipLaunchKernelGGL((k_perm_linear8<TI, OW>), dim3(nblocks(n / 8)), dim3(NT), 0, st, a, n / 8);
  } else if (q < 0 || n < 4096) {
    hipLaunchKernelGGL((k_perm_linear<TI, OW>), dim3(nblocks(n)), dim3(NT), 0, st, a, n);
  } else {
    int tq = (a.od[q] + 63) / 64, tl = (a.od[L] + 63) / 64;
    long nrest = n / ((long)a.od[q] * a.od[L]);
    long nb = (long)tq * tl * nrest;
    hipLaunchKernelGGL((k_perm_tiled<TI, OW>), dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(NT), 0, st, a, q,
                       nrest, tq, tl);
  }
}

// ---------------------------------------------------------------- misc
__global__ void k_axpby(long n, double a, const double *x, double b, double *y) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT)
    y[i] = a * x[i] + (b != 0.0 ? b * y[i] : 0.0);
}

// x[j][k][a][b][c] /= eps_p + eps_j + eps_k - eps_a - eps_b - eps_c (virtual
// energies at no + a, ...): one occupied slice p of a six-index triples
// amplitude over its denominators
__global__ void k_div_d3(int no, int nv, int p, const double *__restrict__ eps, double *x) {
  const long n = (long)no * no * nv * nv * nv;
  const double ep = eps[p];
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    long r = i;
    const int c = (int)(r % nv); r /= nv;
    const int b = (int)(r % nv); r /= nv;
    const int a = (int)(r % nv); r /= nv;
    const int k = (int)(r % no);
    const int j = (int)(r / no);
    x[i] /= ep + eps[j] + eps[k] - eps[no + a] - eps[no + b] - eps[no + c];
  }
}

__global__ void k_diff(long n, const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ z) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) z[i] = x[i] - y[i];
}

__global__ void k_fill(long n, double v, double *y) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) y[i] = v;
}

template <typename P>
__global__ void k_interleave(long n, const double *__restrict__ x, const double *__restrict__ y, P *__restrict__ out) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    P p;
    p.x = x[i];
    p.y = y[i];
    out[i] = p;
  }
}

// t1_ia = R1_ia / (f_ii - f_aa)
__global__ void k_t1_divide(int no, int nv, const double *__restrict__ eps, const double *__restrict__ R1,
                            double *__restrict__ t1) {
  const long n = (long)no * nv;
  for (long q = blockIdx.x * (long)NT + threadIdx.x; q < n; q += (long)gridDim.x * NT) {
    const int i = (int)(q / nv), a = (int)(q % nv);
    t1[q] = R1[q] / (eps[i] - eps[no + a]);
  }
}

// Z_ijab += 2 l1_ia Hov_jb - l1_ja Hov_ib (cclambda.cc:238), b fastest
__global__ void k_lambda_outer(int no, int nv, const double *__restrict__ l1, const double *__restrict__ hov,
                               double *__restrict__ Z) {
  const long n = (long)no * no * nv * nv;
  for (long q = blockIdx.x * (long)NT + threadIdx.x; q < n; q += (long)gridDim.x * NT) {
    const int b = (int)(q % nv);
    const long r = q / nv;
    const int a = (int)(r % nv);
    const long ij = r / nv;
    const int j = (int)(ij % no), i = (int)(ij / no);
    Z[q] += 2.0 * l1[(long)i * nv + a] * hov[(long)j * nv + b] - l1[(long)j * nv + a] * hov[(long)i * nv + b];
  }
}

// flag[0] |= 1 when any of the n values is NaN or infinite
__global__ void k_nonfinite(long n, const double *__restrict__ x, double *flag) {
  bool bad = false;
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) bad |= !isfinite(x[i]);
  if (bad) flag[0] = 1.0;  // every writer stores the same value
}

// U_ijab = 2 t2_ijab - t2_ijba (ccwfn.cc:455, 518-519), one 32 x 32 (a, b)
// tile of one (i, j) plane per block: the tile and its mirror are read
// along b into LDS, so both global reads and the write are coalesced.
__global__ void k_build_U(int nv, long nplane, const double *__restrict__ t2, double *__restrict__ U) {
  __shared__ double T[32][33], Tm[32][33];
  const int nt = (nv + 31) / 32;
  const long ntile = nplane * nt * nt;
  const long v2 = (long)nv * nv;
  for (long blk = blockIdx.x; blk < ntile; blk += gridDim.x) {
    const long ij = blk / (nt * nt);
    const int ta = (int)((blk / nt) % nt), tb = (int)(blk % nt);
    const double *P = t2 + ij * v2;
    for (int e = threadIdx.x; e < 32 * 32; e += NT) {
      const int r = e >> 5, q = e & 31;
      const int a = ta * 32 + r, b = tb * 32 + q;     // tile (a, b)
      const int am = tb * 32 + r, bm = ta * 32 + q;   // mirror tile (b, a), rows b
      T[r][q] = (a < nv && b < nv) ? P[(long)a * nv + b] : 0.0;
      Tm[r][q] = (am < nv && bm < nv) ? P[(long)am * nv + bm] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += NT) {
      const int r = e >> 5, q = e & 31;
      const int a = ta * 32 + r, b = tb * 32 + q;
      if (a < nv && b < nv) U[ij * v2 + (long)a * nv + b] = 2.0 * T[r][q] - Tm[q][r];
    }
    __syncthreads();
  }
}

// ccwfn.cc:284-298
__global__ void k_tau(int no, int nv, const double *t1, const double *t2, double *tau, double *ttau) {
  long n = (long)no * no * nv * nv;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < n; idx += (long)gridDim.x * NT) {
    long r = idx;
    int b = r % nv; r /= nv;
    int a = r % nv; r /= nv;
    int j = r % no; r /= no;
    int i = (int)r;
    double x = t1[i * nv + a] * t1[j * nv + b];
    double t = t2[idx];
    tau[idx] = t + x;
    if (ttau) ttau[idx] = t + 0.5 * x;
  }
}

// Y[n][f][j][b] = 0.5 t2[j][n][f][b] + t1[j][f] t1[n][b] (ccwfn.cc:393,410),
// written in the (n f)(j b) order the Wmbej/Wmbje GEMMs read.
__global__ void k_build_Y(int no, int nv, const double *t1, const double *t2, double *Y) {
  long tot = (long)no * no * nv * nv;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < tot; idx += (long)gridDim.x * NT) {
    long r = idx;
    int b = r % nv; r /= nv;
    int j = r % no; r /= no;
    int f = r % nv; r /= nv;
    int n = (int)r;
    Y[idx] = 0.5 * t2[(((long)j * no + n) * nv + f) * nv + b] + t1[j * nv + f] * t1[n * nv + b];
  }
}

// Jacobi update, ccwfn.cc:555-582, with the t2 residual assembled from its
// half X (R2_ijab = X_ijab + X_jiba): ||R||^2 partials, t = t_old + R/D,
// D_ijab = (f_ii + f_jj) - (f_aa + f_bb) so that D_ijab == D_jiba exactly.
__global__ __launch_bounds__(NT) void k_update(int no, int nv, const double *eps, const double *R1, const double *X,
                                               const double *t1old, const double *t2old, double *t1, double *t2,
                                               double *R2out, double *partials, double shift) {
  // R_ijab = X_ijab + X_jiba, one 32 x 32 (a, b) tile of one (i, j) pair at
  // a time: X_jiba is read along a into LDS and taken back transposed, so
  // both reads of X are coalesced (a per-element read of X_jiba strides by
  // nv along b and spreads each cache line over the XCDs' caches)
  __shared__ double sT[32][33];
  const int tb = (nv + 31) / 32;
  const long ntiles = (long)no * no * tb * tb;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  constexpr int TY = NT / 32;
  const long v2 = (long)nv * nv;
  double s = 0.0;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int tbb = (int)(tile % tb);
    long r = tile / tb;
    const int tba = (int)(r % tb);
    r /= tb;
    const int j = (int)(r % no), i = (int)(r / no);
    const int a0 = tba * 32, b0 = tbb * 32;
    const double *Xji = X + ((long)j * no + i) * v2;
    __syncthreads();  // the previous tile's reads of sT are done
    for (int y = ty; y < 32; y += TY) {  // sT[b - b0][a - a0] = X_jiba
      const int b = b0 + y, a = a0 + tx;
      sT[y][tx] = (b < nv && a < nv) ? Xji[(long)b * nv + a] : 0.0;
    }
    __syncthreads();
    const long base = ((long)i * no + j) * v2;
    for (int y = ty; y < 32; y += TY) {
      const int a = a0 + y, b = b0 + tx;
      if (a < nv && b < nv) {
        const long idx = base + (long)a * nv + b;
        const double R = X[idx] + sT[tx][y];
        const double D = (eps[i] + eps[j]) - (eps[no + a] + eps[no + b]) + shift;
        s += R * R;
        if (R2out) R2out[idx] = R;
        t2[idx] = t2old[idx] + R / D;
      }
    }
  }
  const long n1 = (long)no * nv;
  for (long q = blockIdx.x * (long)NT + threadIdx.x; q < n1; q += (long)gridDim.x * NT) {
    const int i = (int)(q / nv), a = (int)(q % nv);
    const double R = R1[q];
    s += R * R;
    t1[q] = t1old[q] + R / (eps[i] - eps[no + a] + shift);
  }
  double tot = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// MP2 guess, ccwfn.cc:75-87: t1 = f_ia / D1, t2 = <ij|ab> / D2.
__global__ void k_guess(int no, int nv, const double *fock, const double *oovv, double *t1, double *t2) {
  const int n = no + nv;
  long n1 = (long)no * nv, n2 = n1 * n1;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < n2 + n1; idx += (long)gridDim.x * NT) {
    if (idx < n2) {
      long r = idx;
      int b = r % nv; r /= nv;
      int a = r % nv; r /= nv;
      int j = r % no; r /= no;
      int i = (int)r;
      double D = (fock[(long)i * n + i] + fock[(long)j * n + j]) -
                 (fock[(long)(no + a) * n + no + a] + fock[(long)(no + b) * n + no + b]);
      t2[idx] = oovv[idx] / D;
    } else {
      long q = idx - n2;
      int i = q / nv, a = q % nv;
      t1[q] = fock[(long)i * n + no + a] / (fock[(long)i * n + i] - fock[(long)(no + a) * n + no + a]);
    }
  }
}

__global__ void k_splitk_reduce(int S, int M, int N, const double *W, double alpha, double beta, double *C,
                                long scm, long scn) {
  long mn = (long)M * N;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < mn; idx += (long)gridDim.x * NT) {
    double s = 0.0;
    for (int z = 0; z < S; z++) s += W[(long)z * mn + idx];
    long m = idx / N, n = idx % N;
    long o = m * scm + n * scn;
    double v = alpha * s;
    if (beta != 0.0) v += beta * C[o];
    C[o] = v;
  }
}

__global__ void k_dot(long n, const double *x, const double *y, double *partials) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) s += x[i] * y[i];
  double tot = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ void k_reduce(const double *partials, int n, double *out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) s += partials[i];
  double tot = block_sum(s);
  if (threadIdx.x == 0) out[0] = tot;
}

// ccwfn.cc:260-282.  The one-electron part is written into partials[npart]
// separately so the host can apply either the reference's form (only the
// last (i,a) term survives the '=' at ccwfn.cc:272) or the full sum.
__global__ void k_energy(int no, int nv, const double *fov, const double *t1, const double *tau,
                         const double *Loovv, double *partials) {
  long n1 = (long)no * nv, n2 = n1 * n1;
  double s = 0.0;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < n2; idx += (long)gridDim.x * NT)
    s += tau[idx] * Loovv[idx];
  double tot = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ void k_diis_store(long n1, long n2, const double *t1, const double *t1old, const double *t2,
                             const double *t2old, double *err, double *amp) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n1 + n2; i += (long)gridDim.x * NT) {
    double x, xo;
    if (i < n1) { x = t1[i]; xo = t1old[i]; } else { x = t2[i - n1]; xo = t2old[i - n1]; }
    err[i] = x - xo;
    amp[i] = x;
  }
}

struct VecList {
  const double *v[kDiisMax];
  double c[kDiisMax];
};

__global__ void k_diis_dots(long len, const double *e, VecList vl, int nvec, double *partials, int npart) {
  double s[kDiisMax];
#pragma unroll
  for (int p = 0; p < kDiisMax; p++) s[p] = 0.0;
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < len; i += (long)gridDim.x * NT) {
    double x = e[i];
#pragma unroll
    for (int p = 0; p < kDiisMax; p++)
      if (p < nvec) s[p] += x * vl.v[p][i];
  }
  for (int p = 0; p < nvec; p++) {
    double tot = block_sum(s[p]);
    if (threadIdx.x == 0) partials[p * npart + blockIdx.x] = tot;
  }
}

__global__ void k_diis_combine(long n1, long n2, VecList vl, int nvec, double *t1, double *t2) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n1 + n2; i += (long)gridDim.x * NT) {
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < kDiisMax; p++)
      if (p < nvec) s += vl.c[p] * vl.v[p][i];
    if (i < n1) t1[i] = s; else t2[i - n1] = s;
  }
}

// One thread per panel element: i -> (p, row block, k slice, row, stored
// 8-k chunk, k in chunk); the stored chunk of row rr holds logical chunk
// chunk ^ ((rr >> 1) & 3) (the swizzle of panel_chunk).
__global__ void k_pack_panel(const double *in, long sp, long sr1, long sr2, long R2, long R2v, long sk, long R, int K,
                             int nrb, int nks, long n, unsigned short *out) {
  for (long i = blockIdx.x * (long)NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) {
    const int kk = (int)(i & 31), rr = (int)((i >> 5) & 15);
    const long rest = i >> 9;
    const int ks = (int)(rest % nks);
    const long rb = (rest / nks) % nrb, p = rest / ((long)nks * nrb);
    const long r = rb * 16 + rr;
    const int k = ks * 32 + panel_chunk(rr, kk >> 3) * 8 + (kk & 7);
    float x = 0.0f;
    if (r < R && k < K && r % R2 < R2v) x = (float)in[p * sp + (r / R2) * sr1 + (r % R2) * sr2 + k * sk];
    const unsigned short h = bf16_rne(x);
    out[i] = h;
    out[n + i] = bf16_rne(x - __uint_as_float(((unsigned int)h) << 16));
  }
}

// ---- packed symmetric / antisymmetric ladder
// Pair indices: p(a,b) = a(a+1)/2 + b for a >= b, s(a,b) = a(a-1)/2 + b for
// a > b.  pair_of_index inverts them.
__device__ inline void pair_of_index(long p, bool strict, int &a, int &b) {
  // x = floor((sqrt(8p+1)-1)/2), corrected for rounding; a strict pair
  // s(a,b) is the pair p(a-1, b)
  int x = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while ((long)(x + 1) * (x + 2) / 2 <= p) x++;
  while ((long)x * (x + 1) / 2 > p) x--;
  if (strict) { a = x + 1; b = (int)(p - (long)x * (x + 1) / 2); }
  else { a = x; b = (int)(p - (long)x * (x + 1) / 2); }
}

// V+[p(ab)][p(ef)] = <ab|ef> + <ab|fe>,  V-[s(ab)][s(ef)] = <ab|ef> - <ab|fe>,
// rows ldk elements apart with zeros past the last pair
__global__ void k_ladder_vpm(int nv, const double *__restrict__ vvvv, double *__restrict__ vp,
                             double *__restrict__ vm, long ldp, long ldm) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2, v = nv;
  const long row = blockIdx.y;
  int a, b;
  if (row < P) {
    pair_of_index(row, false, a, b);
    const double *Vab = vvvv + ((long)a * v + b) * v * v;
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldp; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < P ? col : 0, false, e, f);
      vp[row * ldp + col] = col < P ? Vab[(long)e * v + f] + Vab[(long)f * v + e] : 0.0;
    }
  }
  if (row < Ps) {
    pair_of_index(row, true, a, b);
    const double *Vab = vvvv + ((long)a * v + b) * v * v;
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldm; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < Ps ? col : 0, true, e, f);
      vm[row * ldm + col] = col < Ps ? Vab[(long)e * v + f] - Vab[(long)f * v + e] : 0.0;
    }
  }
}

// S[p(ij)][p(ef)] = (tau_ij^ef + tau_ij^fe) (1/2 if e = f),
// A[s(ij)][s(ef)] = tau_ij^ef - tau_ij^fe, rows ldp / ldm elements apart with
// zeros past the last pair
__global__ void k_ladder_pack(int no, int nv, const double *__restrict__ tau, double *__restrict__ S,
                              double *__restrict__ A, long ldp, long ldm) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2, v = nv;
  const long Q = (long)no * (no + 1) / 2, Qs = (long)no * (no - 1) / 2;
  const long row = blockIdx.y;
  int i, j;
  if (row < Q) {
    pair_of_index(row, false, i, j);
    const double *T = tau + ((long)i * no + j) * v * v;
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldp; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < P ? col : 0, false, e, f);
      const double x = T[(long)e * v + f] + T[(long)f * v + e];
      S[row * ldp + col] = col < P ? (e == f ? 0.5 * x : x) : 0.0;
    }
  }
  if (row < Qs) {
    pair_of_index(row, true, i, j);
    const double *T = tau + ((long)i * no + j) * v * v;
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldm; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < Ps ? col : 0, true, e, f);
      A[row * ldm + col] = col < Ps ? T[(long)e * v + f] - T[(long)f * v + e] : 0.0;
    }
  }
}

// Row r of in (a v x v slab at in + r v^2, e.g. <mb|..> of ovvv):
// Vp[r][p(ef)] = in_r[e][f] + in_r[f][e], Vm[r][s(ef)] = in_r[e][f] - in_r[f][e],
// rows ldp / ldm elements apart with zeros past the last pair
__global__ void k_rows_vpm(long nrows, int nv, const double *__restrict__ in, double *__restrict__ vp,
                           double *__restrict__ vm, long ldp, long ldm) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2, v = nv;
  for (long row = blockIdx.y; row < nrows; row += gridDim.y) {
    const double *R = in + row * v * v;
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldp; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < P ? col : 0, false, e, f);
      vp[row * ldp + col] = col < P ? R[(long)e * v + f] + R[(long)f * v + e] : 0.0;
    }
    for (long col = blockIdx.x * (long)NT + threadIdx.x; col < ldm; col += (long)gridDim.x * NT) {
      int e, f;
      pair_of_index(col < Ps ? col : 0, true, e, f);
      vm[row * ldm + col] = col < Ps ? R[(long)e * v + f] - R[(long)f * v + e] : 0.0;
    }
  }
}

// Z[r][i][j] = beta Z[r][i][j] + (Z+[r][p(ij)] + s_ij Z-[r][s(ij)]) / 2 over
// nrows rows r (no Z- term for i = j; beta 0 or 1)
__global__ void k_pairs_unpack(long nrows, int no, const double *__restrict__ ZP, const double *__restrict__ ZM,
                               double beta, double *__restrict__ Z) {
  const long Q = (long)no * (no + 1) / 2, Qs = (long)no * (no - 1) / 2, oo = (long)no * no;
  const long n = nrows * oo;
  for (long idx = blockIdx.x * (long)NT + threadIdx.x; idx < n; idx += (long)gridDim.x * NT) {
    const long r = idx / oo;
    const int j = (int)(idx % no), i = (int)((idx / no) % no);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    double x = ZP[r * Q + (long)hi * (hi + 1) / 2 + lo];
    if (i != j) {
      const double m = ZM[r * Qs + (long)hi * (hi - 1) / 2 + lo];
      x += i > j ? m : -m;
    }
    Z[idx] = beta != 0.0 ? Z[idx] + 0.5 * x : 0.5 * x;
  }
}

// X[i][j][a][b] += scale (L+[p(ij)][p(ab)] + s_ij s_ab L-[s(ij)][s(ab)]),
// indices sorted into pairs, s = +1 for i > j (a > b), -1 below, and no
// L- term on a diagonal.  One 32 x 32 (a, b) tile of one (i, j) pair at a
// time, with the packed rows read along their pair index: along b in a tile
// on or below the diagonal (p(ab) = a(a+1)/2 + b for a >= b), along a in a
// tile above it (p(ba) = b(b+1)/2 + a), staged in LDS and written to X
// along b.
__global__ __launch_bounds__(NT) void k_ladder_unpack(int no, int nv, const double *__restrict__ LP,
                                                      const double *__restrict__ LM, double scale,
                                                      double *__restrict__ X) {
  __shared__ double sL[32][33];
  const long v = nv, P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2;
  const int tb = (nv + 31) / 32;
  const long ntiles = (long)no * no * tb * tb;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  constexpr int TY = NT / 32;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int tbb = (int)(tile % tb);
    long r = tile / tb;
    const int tba = (int)(r % tb);
    r /= tb;
    const int j = (int)(r % no), i = (int)(r / no);
    const int a0 = tba * 32, b0 = tbb * 32;
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const double *lp = LP + ((long)hi * (hi + 1) / 2 + lo) * P;
    const double *lm = i != j ? LM + ((long)hi * (hi - 1) / 2 + lo) * Ps : LM;  // read only for i != j
    auto val = [&](int a, int b) {
      const int ha = a > b ? a : b, la = a > b ? b : a;
      double x = lp[(long)ha * (ha + 1) / 2 + la];
      if (i != j && a != b) {
        const double m = lm[(long)ha * (ha - 1) / 2 + la];
        x += ((i > j) == (a > b)) ? m : -m;
      }
      return x;
    };
    double *Xij = X + ((long)i * no + j) * v * v;
    if (tba >= tbb) {
      for (int y = ty; y < 32; y += TY) {
        const int a = a0 + y, b = b0 + tx;
        if (a < nv && b < nv) Xij[(long)a * v + b] += scale * val(a, b);
      }
    } else {
      __syncthreads();  // the previous tile's reads of sL are done
      for (int y = ty; y < 32; y += TY) {  // sL[b - b0][a - a0]
        const int b = b0 + y, a = a0 + tx;
        sL[y][tx] = (a < nv && b < nv) ? val(a, b) : 0.0;
      }
      __syncthreads();
      for (int y = ty; y < 32; y += TY) {
        const int a = a0 + y, b = b0 + tx;
        if (a < nv && b < nv) Xij[(long)a * v + b] += scale * sL[tx][y];
      }
    }
  }
}

}  // namespace

long panel_size(int np, long R, int K) { return (long)np * ((R + 15) / 16 * 16) * ((K + 31) / 32 * 32); }

long ladder_ld(long pairs) { return (pairs + 31) / 32 * 32; }

void launch_ladder_vpm(int nv, const double *vvvv, double *vp, double *vm, hipStream_t st) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2;
  if (P > 65535) throw std::runtime_error("ugx: packed ladder needs v(v+1)/2 <= 65535");
  const int gx = (int)std::min<long>((ladder_ld(P) + NT - 1) / NT, 64);
  hipLaunchKernelGGL(k_ladder_vpm, dim3(gx, (unsigned)P), dim3(NT), 0, st, nv, vvvv, vp, vm, ladder_ld(P),
                     ladder_ld(Ps));
}

void launch_ladder_pack(int no, int nv, const double *tau, double *S, double *A, hipStream_t st) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2, Q = (long)no * (no + 1) / 2;
  if (Q > 65535) throw std::runtime_error("ugx: packed ladder needs o(o+1)/2 <= 65535");
  const int gx = (int)std::min<long>((ladder_ld(P) + NT - 1) / NT, 64);
  hipLaunchKernelGGL(k_ladder_pack, dim3(gx, (unsigned)Q), dim3(NT), 0, st, no, nv, tau, S, A, ladder_ld(P),
                     ladder_ld(Ps));
}

void launch_rows_vpm(long nrows, int nv, const double *in, double *vp, double *vm, hipStream_t st) {
  const long P = (long)nv * (nv + 1) / 2, Ps = (long)nv * (nv - 1) / 2;
  const int gx = (int)std::min<long>((ladder_ld(P) + NT - 1) / NT, 64);
  const unsigned gy = (unsigned)std::min<long>(nrows, 65535);
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_rows_vpm, dim3(gx, gy), dim3(NT), 0, st, nrows, nv, in, vp, vm, ladder_ld(P), ladder_ld(Ps));
}

void launch_pairs_unpack(long nrows, int no, const double *ZP, const double *ZM, double beta, double *Z,
                         hipStream_t st) {
  const long n = nrows * no * no;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pairs_unpack, dim3(nblocks(n)), dim3(NT), 0, st, nrows, no, ZP, ZM, beta, Z);
}

void launch_ladder_unpack(int no, int nv, const double *LP, const double *LM, double scale, double *X,
                          hipStream_t st) {
  const long tb = (nv + 31) / 32, ntiles = (long)no * no * tb * tb;
  hipLaunchKernelGGL(k_ladder_unpack, dim3((unsigned)std::min<long>(std::max<long>(ntiles, 1), 1 << 16)), dim3(NT), 0,
                     st, no, nv, LP, LM, scale, X);
}

void launch_pack_panel(const double *in, long sp, long sr, long sk, int np, long R, int K, unsigned short *out,
                       hipStream_t st) {
  launch_pack_panel2(in, sp, 0, sr, R, sk, np, R, K, out, st);
}

void launch_pack_panel_rows(const double *in, long sp, long sr1, long sr2, long R2v, long R2, long sk, int np,
                            long R, int K, unsigned short *out, hipStream_t st) {
  const long n = panel_size(np, R, K);
  if (n <= 0) return;
  if (R2 <= 0 || R2v <= 0 || R2v > R2) throw std::runtime_error("ugx: bad panel row split");
  hipLaunchKernelGGL(k_pack_panel, dim3(nblocks(n)), dim3(NT), 0, st, in, sp, sr1, sr2, R2, R2v, sk, R, K,
                     (int)((R + 15) / 16), (K + 31) / 32, n, out);
}

void launch_pack_panel2(const double *in, long sp, long sr1, long sr2, long R2, long sk, int np, long R, int K,
                        unsigned short *out, hipStream_t st) {
  launch_pack_panel_rows(in, sp, sr1, sr2, R2, R2, sk, np, R, K, out, st);
}

void launch_permute(const PermDesc &p, hipStream_t st) {
  if (p.rank < 1 || p.rank > 6) throw std::runtime_error("ugx: permute rank out of range");
  PermArgs a;
  a.rank = p.rank;
  a.n = 1;
  for (int d = 0; d < 6; d++) {
    a.od[d] = d < p.rank ? p.od[d] : 1;
    a.istr[d] = d < p.rank ? p.istr[d] : 0;
    a.n *= a.od[d];
  }
  a.in = p.in; a.out = p.out; a.alpha = p.alpha; a.beta = p.beta;
  a.pk = p.pk; a.pblk = p.pblk; a.plo = p.plo;
  a.i32 = a.n < (1L << 31);
  for (int d = 0; d < 6; d++) magic31(a.od[d], a.dm[d], a.ds[d]);
  magic31(p.out_t == 3 ? p.pk : 1, a.pkm, a.pks);
  const int L = p.rank - 1;
  a.vec8t = (p.out_t == 2 || p.out_t == 3) && a.i32 && p.beta == 0.0 && a.od[L] % 8 == 0 && a.n % 8 == 0 &&
            ((uintptr_t)p.out & 15) == 0 && (p.out_t != 3 || (p.pk % 8 == 0 && p.plo % 8 == 0));
  a.vec8 = a.vec8t && a.istr[L] == 1;
  if (p.out_t >= 2 && p.beta != 0.0) throw std::runtime_error("ugx: BF16 pair output cannot accumulate");
  if (p.out_t == 3 && (p.pk <= 0 || a.n % p.pk != 0 || p.pblk != 16 * ((p.pk + 31) / 32 * 32)))
    throw std::runtime_error("ugx: bad MFMA panel output shape");
  if (p.in_t == 0) {
    if (p.out_t == 0) perm_launch<double, OutD>(a, st);
    else if (p.out_t == 1) perm_launch<double, OutF>(a, st);
    else if (p.out_t == 2) perm_launch<double, OutB3>(a, st);
    else perm_launch<double, OutP>(a, st);
  } else {
    if (p.out_t == 0) perm_launch<float, OutD>(a, st);
    else if (p.out_t == 1) perm_launch<float, OutF>(a, st);
    else if (p.out_t == 2) perm_launch<float, OutB3>(a, st);
    else perm_launch<float, OutP>(a, st);
  }
}

void launch_fill(long n, double v, double *y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(NT), 0, st, n, v, y);
}

void launch_interleave(long n, const double *x, const double *y, void *out, int f32, hipStream_t st) {
  if (n <= 0) return;
  if (f32) hipLaunchKernelGGL(k_interleave<float2>, dim3(nblocks(n)), dim3(NT), 0, st, n, x, y, (float2 *)out);
  else hipLaunchKernelGGL(k_interleave<double2>, dim3(nblocks(n)), dim3(NT), 0, st, n, x, y, (double2 *)out);
}

void launch_t1_divide(int no, int nv, const double *eps, const double *R1, double *t1, hipStream_t st) {
  const long n = (long)no * nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_t1_divide, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, eps, R1, t1);
}

void launch_lambda_outer(int no, int nv, const double *l1, const double *hov, double *Z, hipStream_t st) {
  const long n = (long)no * no * nv * nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lambda_outer, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, l1, hov, Z);
}

void launch_nonfinite(long n, const double *x, double *flag, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_nonfinite, dim3(nblocks(n, NT * 8, 4096)), dim3(NT), 0, st, n, x, flag);
}

void launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(NT), 0, st, n, a, x, b, y);
}

void launch_div_d3(int no, int nv, int p, const double *eps, double *x, hipStream_t st) {
  const long n = (long)no * no * nv * nv * nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_div_d3, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, p, eps, x);
}

void launch_diff(long n, const double *x, const double *y, double *z, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_diff, dim3(nblocks(n)), dim3(NT), 0, st, n, x, y, z);
}

void launch_tau(int no, int nv, const double *t1, const double *t2, double *tau, double *ttau, hipStream_t st) {
  long n = (long)no * no * nv * nv;
  hipLaunchKernelGGL(k_tau, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, t1, t2, tau, ttau);
}

void launch_build_U(int no, int nv, const double *t2, double *U, hipStream_t st) {
  const long nplane = (long)no * no, nt = (nv + 31) / 32, ntile = nplane * nt * nt;
  if (ntile <= 0) return;
  hipLaunchKernelGGL(k_build_U, dim3((unsigned)std::min<long>(ntile, 65536)), dim3(NT), 0, st, nv, nplane, t2, U);
}

void launch_build_Y(int no, int nv, const double *t1, const double *t2, double *Y, hipStream_t st) {
  long n = (long)no * no * nv * nv;
  hipLaunchKernelGGL(k_build_Y, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, t1, t2, Y);
}

int update_nparts(int no, int nv) {
  long n = (long)no * nv * (1 + (long)no * nv);
  return nblocks(n, NT, 2048);
}

void launch_update(int no, int nv, const double *eps, const double *R1, const double *X, const double *t1old,
                   const double *t2old, double *t1, double *t2, double *R2out, double *partials, int npart,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_update, dim3(npart), dim3(NT), 0, st, no, nv, eps, R1, X, t1old, t2old, t1, t2, R2out,
                     partials, 0.0);
}

void launch_update_shift(int no, int nv, const double *eps, double shift, const double *R1, const double *X,
                         const double *t1old, const double *t2old, double *t1, double *t2, double *partials, int npart,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_update, dim3(npart), dim3(NT), 0, st, no, nv, eps, R1, X, t1old, t2old, t1, t2, nullptr,
                     partials, shift);
}

int energy_nparts(int no, int nv) {
  long n = (long)no * no * nv * nv;
  return nblocks(n, NT, 2048);
}

void launch_energy(int no, int nv, const double *fov, const double *t1, const double *tau, const double *Loovv,
                   double *partials, int npart, hipStream_t st) {
  hipLaunchKernelGGL(k_energy, dim3(npart), dim3(NT), 0, st, no, nv, fov, t1, tau, Loovv, partials);
}

void launch_guess(int no, int nv, const double *fock, const double *oovv, double *t1, double *t2,
                  hipStream_t st) {
  long n = (long)no * nv * (1 + (long)no * nv);
  hipLaunchKernelGGL(k_guess, dim3(nblocks(n)), dim3(NT), 0, st, no, nv, fock, oovv, t1, t2);
}

void launch_splitk_reduce(int S, int M, int N, const double *W, double alpha, double beta, double *C, long scm,
                          long scn, hipStream_t st) {
  long n = (long)M * N;
  hipLaunchKernelGGL(k_splitk_reduce, dim3(nblocks(n)), dim3(NT), 0, st, S, M, N, W, alpha, beta, C, scm, scn);
}

void launch_dot(long n, const double *x, const double *y, double *partials, int npart, hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3(npart), dim3(NT), 0, st, n, x, y, partials);
}

void launch_reduce(const double *partials, int n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(NT), 0, st, partials, n, out);
}

void launch_diis_store(long n1, long n2, const double *t1, const double *t1old, const double *t2,
                       const double *t2old, double *err, double *amp, hipStream_t st) {
  hipLaunchKernelGGL(k_diis_store, dim3(nblocks(n1 + n2)), dim3(NT), 0, st, n1, n2, t1, t1old, t2, t2old, err,
                     amp);
}

void launch_diis_dots(long len, const double *e, const double *const *vecs, int nvec, double *partials,
                      int npart, hipStream_t st) {
  if (nvec > kDiisMax) throw std::runtime_error("ugx: DIIS subspace above 16");
  VecList vl{};
  for (int p = 0; p < nvec; p++) vl.v[p] = vecs[p];
  hipLaunchKernelGGL(k_diis_dots, dim3(npart), dim3(NT), 0, st, len, e, vl, nvec, partials, npart);
}

void launch_diis_combine(long n1, long n2, const double *const *amps, const double *coef, int nvec, double *t1,
                         double *t2, hipStream_t st) {
  if (nvec > kDiisMax) throw std::runtime_error("ugx: DIIS subspace above 16");
  VecList vl{};
  for (int p = 0; p < nvec; p++) { vl.v[p] = amps[p]; vl.c[p] = coef[p]; }
  hipLaunchKernelGGL(k_diis_combine, dim3(nblocks(n1 + n2)), dim3(NT), 0, st, n1, n2, vl, nvec, t1, t2);
}

}  // namespace ugx
