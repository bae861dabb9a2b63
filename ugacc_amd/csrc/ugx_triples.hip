// (T) energy epilogue, ccwfn.cc:795-852.
//
// For a batch of i>=j>=k triplets whose connected W3 (v^3, [a][b][c]) has
// been built by the MFMA contractions, the bracket is summed over a>=b>=c:
//   V_abc = (W_abc + <ij|ab> t1_kc + <ik|ac> t1_jb + <jk|bc> t1_ia
//            [+ t2_ij^ab f_kc + t2_ik^ac f_jb + t2_jk^bc f_ia])
//           / (1 + d_ab + d_ac + d_bc)
//   X = sum over the six permutations of W*V,  Y = V_abc + V_bca + V_cab,
//   Z = V_acb + V_bac + V_cba
//   E += [(Y - 2Z)(W_abc + W_bca + W_cab) + (Z - 2Y)(W_acb + W_bac + W_cba)
//         + 3X] (2 - d_ij - d_ik - d_jk) / D_ijk^abc
// V is formed on the fly from W, the oovv block and t1, so only W is stored.
// The bracketed f_ov terms are the f_kc t3 part of X2 in tcorr / tcorr_ooc
// (ccwfn.cc:706-709, 725), the values the reference returns: since t2 has
// the pair symmetry of <ij|ab>, that part of the energy is the bracket's
// disconnected term with <pq|xy> -> t2_pq^xy and t1 -> f_ov.  The FOV
// variant of the kernel adds it (any f_ov != 0); it reads <pq|xy> with its
// t2 and t1 with its f_ov as one vector pair (arrays interleaved once per
// (T) call in the kernel's element type, so a prefetch is one wide load
// with no conversion), and stages them in LDS as such, so the numbers of
// loads and LDS accesses stay those of the plain kernel.
//
// Every point needs W at all six index permutations.  A workgroup owns one
// triplet and a pair of E-wide index blocks A >= B, and walks the blocks
// C <= B: for each (A, B, C) it stages the six permuted E^3 sub-cubes
// W[A][B][C], W[A][C][B], W[B][A][C], W[B][C][A], W[C][A][B], W[C][B][A]
// into LDS.  W is stored in the blocked layout the W3 GEMMs write (E^3
// sub-cubes, each contiguous), so every staged sub-cube is one contiguous
// read (16-byte loads) and every W element is read from memory about once.
// The sub-cubes and <pq|xy> blocks of step C+1 are loaded into registers
// while step C computes, and stored to LDS after the step's barrier.
// W is FP64 (FP64 form) or FP32 (low-precision forms); all sums are FP64 and
// the per-(triplet, block pair) partials are reduced per triplet in a fixed
// order, so the result is bitwise repeatable.
#include "ugx_internal.h"

#include <algorithm>
#include <type_traits>

namespace ugx {
namespace {

constexpr int NT = 256;
// Threads per workgroup of the energy kernel: with an FP64 W its LDS
// footprint admits one workgroup per CU, so eight waves (two per SIMD)
// instead of four; with an FP32 W two workgroups of four waves fit, and
// eight-wave ones would not fit their registers in the budget of four
// waves per SIMD.
template <typename WT>
constexpr int energy_nt() { return sizeof(WT) == 4 ? 256 : 512; }

template <int NW>
__device__ inline double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __shared__ double red[NW];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NW; k++) s += red[k];
  return s;
}

// A load through the global address space.  The per-problem arrays are
// reached through pointers read from the problem table, which the compiler
// cannot place in an address space: plain dereferences become FLAT loads,
// which count against lgkmcnt as well as vmcnt, so the compute loop's first
// LDS wait would also wait for the next step's prefetch from HBM.
// (Moved as a native integer of the same size: a struct copy through the
// cast pointer would still go through a generic reference.)
template <typename T>
__device__ inline T gload(const T *p) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "4-, 8- or 16-byte elements");
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) unsigned *)p);
  else if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) unsigned long long *)p);
  else return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) u64x2 *)p);
}

// 1 / d from the hardware reciprocal estimate refined by two Newton-Raphson
// steps r <- r + r (1 - d r) (each squares the relative error): within an
// ulp or two of the quotient, without the division's scaling sequence.
__device__ inline double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// Block pair index p -> (A, B) with B <= A.
__device__ inline void pair_of(int p, int &A, int &B) {
  A = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
  while ((A + 1) * (A + 2) / 2 <= p) A++;
  while (A * (A + 1) / 2 > p) A--;
  B = p - A * (A + 1) / 2;
}

// (<pq|xy>, t2_pq^xy) and (t1, f_ov) pairs of the FOV variant
// Local points (x, y, z) of an E^3 step on a block diagonal, in order:
// pair[p] = x << 4 | y for y <= x, tri[q] = x << 8 | y << 4 | z for
// z <= y <= x (E <= 15)
template <int E>
struct DiagPoints {
  unsigned short pair[E * (E + 1) / 2];
  unsigned short tri[E * (E + 1) * (E + 2) / 6];
};
template <int E>
constexpr DiagPoints<E> make_diag_points() {
  static_assert(E <= 15, "four bits per local index");
  DiagPoints<E> t{};
  int p = 0, q = 0;
  for (int x = 0; x < E; x++)
    for (int y = 0; y <= x; y++) {
      t.pair[p++] = (unsigned short)(x << 4 | y);
      for (int z = 0; z <= y; z++) t.tri[q++] = (unsigned short)(x << 8 | y << 4 | z);
    }
  return t;
}
__constant__ DiagPoints<12> kDiag12 = make_diag_points<12>();

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

template <typename WT, int E, typename OT, bool FOV>
__global__ __launch_bounds__(energy_nt<WT>()) __attribute__((amdgpu_waves_per_eu(2))) void k_triples_energy(int no, int nv, int ntrip, const int *__restrict__ trip,
                                                       const TripProb *__restrict__ probs,
                                                       const WT *__restrict__ Wall, long wstride,
                                                       double *__restrict__ part, int npairs,
                                                       int *__restrict__ counter) {
  constexpr int E3 = E * E * E, E2 = E * E, NTE = energy_nt<WT>();
  // padded cube layout idx(a,b,c) = a*SA + b*SB + c: the strides are odd, so
  // the stores along a or b (and the reads along c) do not collide in banks
  constexpr int SB = E + 1, SA = E * (E + 1) + 1;
  // per-thread shares of what one step stages: the six sub-cubes (PW
  // elements each), the <pq|xy> blocks of the five block pairs that touch C
  // (PC), and, once per item, the four pairs within {A, B} (PF)
  // W moves in 16-byte chunks (EPC elements; sub-cubes are contiguous and
  // 16-byte aligned)
  typedef typename std::conditional<sizeof(WT) == 4, float4, double2>::type W16;
  constexpr int EPC = 16 / (int)sizeof(WT), NCH = E3 / EPC;
  static_assert(E3 % EPC == 0, "sub-cubes are whole 16-byte chunks");
  constexpr int NPC = 3 * 5 * E2, NPF = 3 * 4 * E2;
  constexpr int PW = (NCH + NTE - 1) / NTE, PC = (NPC + NTE - 1) / NTE, PF = (NPF + NTE - 1) / NTE;
  static_assert(4 * E <= NTE, "t1 / eps staging takes one element per thread");
  __shared__ WT S[6][E * SA];
  // disconnected-term operands of the three blocks X, Y in {A, B, C}:
  // O[s][X][Y][x][y] = <pq|xy> for the slices s = (ij), (ik), (jk) (OT),
  // with t2_pq^xy beside it in the FOV variant;
  // T1[s][X][x] = t1 of k, j, i (the amplitude that multiplies slice s),
  // with f_ov of the same row beside it in the FOV variant
  typedef typename std::conditional<FOV, typename Vec2<OT>::type, OT>::type OE;
  typedef typename std::conditional<FOV, typename Vec2<OT>::type, double>::type TE;
  __shared__ OE O[3][3][3][E2];
  __shared__ TE T1[3][3][E];
  __shared__ double EPS[3][E];
  __shared__ int item_sh;
  // the diagonal point tables in LDS (per-lane reads of the constant copy
  // would be vector memory loads)
  __shared__ unsigned short sPair[E * (E + 1) / 2], sTri[E * (E + 1) * (E + 2) / 6];
  const int tid = threadIdx.x;
  const long total = (long)ntrip * npairs;
  static_assert(E == 12, "diagonal point tables for E = 12");
  for (int e = tid; e < E * (E + 1) / 2; e += NTE) sPair[e] = kDiag12.pair[e];
  for (int e = tid; e < E * (E + 1) * (E + 2) / 6; e += NTE) sTri[e] = kDiag12.tri[e];
  // persistent: each workgroup takes (entry, block pair) items from the
  // queue until it is empty; every workgroup reaches the exit test
  for (;;) {
  __syncthreads();  // the previous item is done with LDS and item_sh
  if (tid == 0) item_sh = atomicAdd(counter, 1);
  __syncthreads();
  const long item = item_sh;
  if (item >= total) break;
  const int t = (int)(item / npairs);
  const int pidx = (int)(item % npairs);
  int Ab, Bb;
  pair_of(pidx, Ab, Bb);
  const int i = trip[4 * t], j = trip[4 * t + 1], k = trip[4 * t + 2];
  const TripProb pr = probs[trip[4 * t + 3]];
  const double *__restrict__ eps = pr.eps, *__restrict__ oovv = pr.oovv, *__restrict__ t1 = pr.t1;
  const OE *__restrict__ ot2 = (const OE *)pr.ot2;
  const TE *__restrict__ t1f = (const TE *)pr.t1f;
  const WT *W = Wall + (long)t * wstride;
  const long v = nv, v2 = (long)nv * nv;
  // slices selected by value (a runtime index into a pointer array would
  // place the array in scratch memory)
  const double *O0 = oovv + ((long)i * no + j) * v2, *O1 = oovv + ((long)i * no + k) * v2,
               *O2 = oovv + ((long)j * no + k) * v2;
  const double *T10 = t1 + (long)k * nv, *T11 = t1 + (long)j * nv, *T12 = t1 + (long)i * nv;
  auto osl = [&](int sl) { return sl == 0 ? O0 : (sl == 1 ? O1 : O2); };
  auto t1sl = [&](int sl) { return sl == 0 ? T10 : (sl == 1 ? T11 : T12); };
  // FOV: the (<pq|xy>, t2) pair slices (ij), (ik), (jk) and the (t1, f_ov)
  // pair rows k, j, i, in the kernel's pair types (one load each)
  const OE *Q0 = nullptr, *Q1 = nullptr, *Q2 = nullptr;
  const TE *F0 = nullptr, *F1 = nullptr, *F2 = nullptr;
  if constexpr (FOV) {
    Q0 = ot2 + ((long)i * no + j) * v2; Q1 = ot2 + ((long)i * no + k) * v2; Q2 = ot2 + ((long)j * no + k) * v2;
    F0 = t1f + (long)k * nv; F1 = t1f + (long)j * nv; F2 = t1f + (long)i * nv;
  }
  auto qsl = [&](int sl) { return sl == 0 ? Q0 : (sl == 1 ? Q1 : Q2); };
  auto fsl = [&](int sl) { return sl == 0 ? F0 : (sl == 1 ? F1 : F2); };
  const double eo = gload(eps + i) + gload(eps + j) + gload(eps + k);
  const double fac = 2.0 - ((i == j) + (i == k) + (j == k));
  const int A0 = Ab * E, B0 = Bb * E;

  // V at a permutation (x, y, z) of the point, given as block slots (bx, by,
  // bz) in {0: A, 1: B, 2: C} and local offsets (lx, ly, lz)
  auto Vx = [&](int bx, int lx, int by, int ly, int bz, int lz, double w, double inv) {
    if constexpr (FOV) {
      // the disconnected sum in OT (FP32 pairs: FP32 like their operands)
      const OE o0 = O[0][bx][by][lx * E + ly], o1 = O[1][bx][bz][lx * E + lz], o2 = O[2][by][bz][ly * E + lz];
      const TE u0 = T1[0][bz][lz], u1 = T1[1][by][ly], u2 = T1[2][bx][lx];
      const OT d = o0.x * u0.x + o0.y * u0.y + o1.x * u1.x + o1.y * u1.y + o2.x * u2.x + o2.y * u2.y;
      return (w + (double)d) * inv;
    } else {
      return (w + (double)O[0][bx][by][lx * E + ly] * T1[0][bz][lz] +
              (double)O[1][bx][bz][lx * E + lz] * T1[1][by][ly] + (double)O[2][by][bz][ly * E + lz] * T1[2][bx][lx]) *
             inv;
    }
  };
  // <pq|xy> (and t2_pq^xy) of slice s at block-local (lx, ly) of blocks
  // with origins gx0, gy0
  auto oget = [&](int s, int gx0, int lx, int gy0, int ly) {
    const int gx = gx0 + lx, gy = gy0 + ly;
    const bool in = gx < nv && gy < nv;
    if constexpr (FOV) {
      return in ? gload(qsl(s) + gx * v + gy) : OE{};
    } else {
      return (OE)(in ? gload(osl(s) + gx * v + gy) : 0.0);
    }
  };
  // t1 (and f_ov) of slice s at global virtual index gx (< nv)
  auto tget = [&](int s, int gx) {
    if constexpr (FOV) {
      return gload(fsl(s) + gx);
    } else {
      return gload(t1sl(s) + gx);
    }
  };
  auto tzero = []() { return TE{}; };
  // an orbital energy carried in a TE register, bit for bit (an FP32 pair
  // holds the FP64 value's eight bytes: the denominators stay FP64)
  auto tval = [](double x) {
    if constexpr (FOV && sizeof(OT) == 4) return __builtin_bit_cast(TE, x);
    else if constexpr (FOV) { TE r{}; r.x = x; return r; }
    else return x;
  };
  auto tfirst = [](const TE &x) {
    if constexpr (FOV && sizeof(OT) == 4) return __builtin_bit_cast(double, x);
    else if constexpr (FOV) return (double)x.x;
    else return x;
  };

  // Once per item: the pairs within {A, B} and the t1 / eps entries of A and
  // B.  Every load is issued before any LDS store, so the loads overlap.
  {
    OE rf[PF];
#pragma unroll
    for (int q = 0; q < PF; q++) {
      const int e = tid + q * NTE;
      rf[q] = OE{};
      if (e < NPF) {
        const int sl = e / (4 * E2), r = e % (4 * E2), pp = r / E2, xy = r % E2;
        const int bx = pp >> 1, by = pp & 1;
        rf[q] = oget(sl, bx ? B0 : A0, xy / E, by ? B0 : A0, xy % E);
      }
    }
    TE rt = tzero();
    if (tid < 4 * E) {  // T1 of slices 0 and 1 for blocks A, B
      const int sl = tid / (2 * E), bx = (tid / E) & 1, lx = tid % E, gx = (bx ? B0 : A0) + lx;
      if (gx < nv) rt = tget(sl, gx);
    }
    TE rt2 = tzero();
    double re = 0.0;
    if (tid < 2 * E) {  // T1 of slice 2 and eps for blocks A, B
      const int bx = tid / E, lx = tid % E, gx = (bx ? B0 : A0) + lx;
      if (gx < nv) { rt2 = tget(2, gx); re = gload(eps + no + gx); }
    }
#pragma unroll
    for (int q = 0; q < PF; q++) {
      const int e = tid + q * NTE;
      if (e < NPF) {
        const int sl = e / (4 * E2), r = e % (4 * E2), pp = r / E2, xy = r % E2;
        O[sl][pp >> 1][pp & 1][xy] = rf[q];
      }
    }
    if (tid < 4 * E) {
      const int sl = tid / (2 * E), bx = (tid / E) & 1, lx = tid % E;
      T1[sl][bx][lx] = rt;
    }
    if (tid < 2 * E) {
      const int bx = tid / E, lx = tid % E;
      T1[2][bx][lx] = rt2;
      EPS[bx][lx] = re;
    }
  }

  // W is blocked (E^3 sub-cubes, each contiguous in its own [x][y][z]
  // order): sub-cube (X, Y, Z) starts at ((X nb + Y) nb + Z) E^3
  const long nbk = (nv + E - 1) / E;
  auto cube = [&](int X, int Y, int Z) { return W + ((X * nbk + Y) * nbk + Z) * (long)E3; };
  auto at = [](int a, int b, int c) { return a * SA + b * SB + c; };
  // the five block pairs that touch C: (A,C), (B,C), (C,A), (C,B), (C,C)
  auto cpair = [](int p, int &bx, int &by) {
    bx = p < 2 ? p : 2;
    by = p < 2 ? 2 : (p == 2 ? 0 : (p == 3 ? 1 : 2));
  };
  // Step data in registers: loaded for step C + 1 while step C computes,
  // committed to LDS after the step's closing barrier.
  W16 rw[6][PW];
  OE rc[PC];
  TE rtc = tzero();
  auto fetch = [&](int Cb) {
    const int C0 = Cb * E;
    const WT *c0 = cube(Ab, Bb, Cb), *c1 = cube(Ab, Cb, Bb), *c2 = cube(Bb, Ab, Cb);
    const WT *c3 = cube(Bb, Cb, Ab), *c4 = cube(Cb, Ab, Bb), *c5 = cube(Cb, Bb, Ab);
#pragma unroll
    for (int q = 0; q < PW; q++) {
      const int ch = tid + q * NTE;
      if (ch < NCH) {
        rw[0][q] = reinterpret_cast<const W16 *>(c0)[ch]; rw[1][q] = reinterpret_cast<const W16 *>(c1)[ch];
        rw[2][q] = reinterpret_cast<const W16 *>(c2)[ch]; rw[3][q] = reinterpret_cast<const W16 *>(c3)[ch];
        rw[4][q] = reinterpret_cast<const W16 *>(c4)[ch]; rw[5][q] = reinterpret_cast<const W16 *>(c5)[ch];
      }
    }
#pragma unroll
    for (int q = 0; q < PC; q++) {
      const int e = tid + q * NTE;
      rc[q] = OE{};
      if (e < NPC) {
        const int sl = e / (5 * E2), r = e % (5 * E2), p = r / E2, xy = r % E2;
        int bx, by;
        cpair(p, bx, by);
        const int gx0 = bx == 0 ? A0 : (bx == 1 ? B0 : C0), gy0 = by == 0 ? A0 : (by == 1 ? B0 : C0);
        rc[q] = oget(sl, gx0, xy / E, gy0, xy % E);
      }
    }
    rtc = tzero();
    if (tid < 4 * E) {  // T1 of slices 0..2 at block C, then eps
      const int sl = tid / E, lx = tid % E, gx = C0 + lx;
      if (gx < nv) rtc = sl < 3 ? tget(sl, gx) : tval(gload(eps + no + gx));
    }
  };
  auto commit = [&]() {
    // S[q] = sub-cube q in its own index order (one contiguous E^3 chunk of
    // W, read in order), all six at the same LDS offsets: sub-cube 0 is
    // W[A][B][C], 1 W[A][C][B], 2 W[B][A][C], 3 W[B][C][A], 4 W[C][A][B],
    // 5 W[C][B][A], and the compute reads each at its permutation of the
    // point.  (Permuting on the store side would give every thread 6x the
    // LDS addresses to hold.)  Entries past nv are never used: the points
    // that would need them are skipped.
#pragma unroll
    for (int q = 0; q < PW; q++) {
      const int ch = tid + q * NTE;
      if (ch < NCH) {
        const WT *w0 = reinterpret_cast<const WT *>(&rw[0][q]), *w1 = reinterpret_cast<const WT *>(&rw[1][q]);
        const WT *w2 = reinterpret_cast<const WT *>(&rw[2][q]), *w3 = reinterpret_cast<const WT *>(&rw[3][q]);
        const WT *w4 = reinterpret_cast<const WT *>(&rw[4][q]), *w5 = reinterpret_cast<const WT *>(&rw[5][q]);
#pragma unroll
        for (int u = 0; u < EPC; u++) {
          const int e = ch * EPC + u;
          const int x = at(e / (E * E), (e / E) % E, e % E);  // slowest, middle, fastest
          S[0][x] = w0[u];
          S[1][x] = w1[u];
          S[2][x] = w2[u];
          S[3][x] = w3[u];
          S[4][x] = w4[u];
          S[5][x] = w5[u];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < PC; q++) {
      const int e = tid + q * NTE;
      if (e < NPC) {
        const int sl = e / (5 * E2), r = e % (5 * E2), p = r / E2, xy = r % E2;
        int bx, by;
        cpair(p, bx, by);
        O[sl][bx][by][xy] = rc[q];
      }
    }
    if (tid < 4 * E) {
      const int sl = tid / E, lx = tid % E;
      if (sl < 3) T1[sl][2][lx] = rtc;
      else EPS[2][lx] = tfirst(rtc);
    }
  };

  double s = 0.0;
  fetch(0);
  for (int Cb = 0; Cb <= Bb; Cb++) {
    const int C0 = Cb * E;
    commit();
    __syncthreads();
    if (Cb < Bb) fetch(Cb + 1);  // in flight while this step computes
    // the step's points: the whole E^3 cube when A > B > C; on a block
    // diagonal only the local points in the order it demands (cc <= bb when
    // B = C, bb <= aa when A = B), enumerated compactly so that no lane idles
    // on a point outside a >= b >= c
    static_assert(E == 12, "diagonal point tables for E = 12");
    constexpr int P2 = E * (E + 1) / 2;
    const int type = (Cb == Bb ? 1 : 0) | (Bb == Ab ? 2 : 0);
    // only block A can end past nv (A >= B >= C), and every enumeration
    // runs aa slowest: stopping at the first aa past nv skips those points
    const int nA = min(E, nv - A0);
    const int npts = type == 0   ? nA * E * E
                     : type == 1 ? nA * P2
                     : type == 2 ? nA * (nA + 1) / 2 * E
                                 : nA * (nA + 1) * (nA + 2) / 6;
#pragma unroll 1
    for (int e = tid; e < npts; e += NTE) {
      int aa, bb, cc;
      if (type == 0) {
        cc = e % E; bb = (e / E) % E; aa = e / (E * E);
      } else if (type == 1) {  // B = C: any aa, cc <= bb
        const int pr = sPair[e % P2];
        aa = e / P2; bb = pr >> 4; cc = pr & 15;
      } else if (type == 2) {  // A = B: bb <= aa, any cc
        const int pr = sPair[e / E];
        aa = pr >> 4; bb = pr & 15; cc = e % E;
      } else {  // A = B = C: cc <= bb <= aa
        const int t3 = sTri[e];
        aa = t3 >> 8; bb = (t3 >> 4) & 15; cc = t3 & 15;
      }
      int a = A0 + aa, b = B0 + bb, c = C0 + cc;
      if (a >= nv || b > a || c > b) continue;
      // 1 / (1 + d_ab + d_ac + d_bc): 1, 1/2 or 1/4, exact
      const double inv = (a == b && b == c) ? 0.25 : ((a == b || b == c || a == c) ? 0.5 : 1.0);
      double wabc = (double)S[0][at(aa, bb, cc)], wacb = (double)S[1][at(aa, cc, bb)];
      double wbac = (double)S[2][at(bb, aa, cc)], wbca = (double)S[3][at(bb, cc, aa)];
      double wcab = (double)S[4][at(cc, aa, bb)], wcba = (double)S[5][at(cc, bb, aa)];
      double vabc = Vx(0, aa, 1, bb, 2, cc, wabc, inv), vacb = Vx(0, aa, 2, cc, 1, bb, wacb, inv);
      double vbac = Vx(1, bb, 0, aa, 2, cc, wbac, inv), vbca = Vx(1, bb, 2, cc, 0, aa, wbca, inv);
      double vcab = Vx(2, cc, 0, aa, 1, bb, wcab, inv), vcba = Vx(2, cc, 1, bb, 0, aa, wcba, inv);
      double X = wabc * vabc + wacb * vacb + wbac * vbac + wbca * vbca + wcab * vcab + wcba * vcba;
      double Y = vabc + vbca + vcab;
      double Z = vacb + vbac + vcba;
      double denom = eo - (EPS[0][aa] + EPS[1][bb] + EPS[2][cc]);
      s += ((Y - 2.0 * Z) * (wabc + wbca + wcab) + (Z - 2.0 * Y) * (wacb + wbac + wcba) + 3.0 * X) * fac * recip(denom);
    }
    __syncthreads();
  }
  double tot = block_sum<NTE / 64>(s);
  if (tid == 0) part[(long)t * npairs + pidx] = tot;
  }
}

__global__ void k_zero_counter(int *counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

__global__ void k_triples_reduce(int npairs, int ntrip, const double *part, double *etrip) {
  int t = blockIdx.x * NT + threadIdx.x;
  if (t >= ntrip) return;
  double s = 0.0;
  for (int p = 0; p < npairs; p++) s += part[(long)t * npairs + p];
  etrip[t] = s;
}

template <typename WT, int E, typename OT, bool FOV>
void launch_e(int no, int nv, int ntrip, const int *trip4, const TripProb *probs, const void *W, long wstride,
              double *etrip, double *part, int *counter, hipStream_t st) {
  int nb = (nv + E - 1) / E;
  int npairs = nb * (nb + 1) / 2;
  // as many workgroups as can be resident at once
  int dev = 0, ncu = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  constexpr int NTE = energy_nt<WT>();
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_triples_energy<WT, E, OT, FOV>, NTE, 0);
  long grid = (long)std::max(ncu, 1) * std::max(per, 1);
  grid = std::min<long>(grid, (long)ntrip * npairs);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, counter);
  hipLaunchKernelGGL((k_triples_energy<WT, E, OT, FOV>), dim3((unsigned)grid), dim3(NTE), 0, st, no, nv, ntrip, trip4,
                     probs, (const WT *)W, wstride, part, npairs, counter);
  hipLaunchKernelGGL(k_triples_reduce, dim3((ntrip + NT - 1) / NT), dim3(NT), 0, st, npairs, ntrip, part, etrip);
}

}  // namespace

// Sub-cube edge = the W block edge, 12: six padded 12^3 sub-cubes (45 KB
// FP32, 90 KB FP64) plus 31 KB of FP64 <pq|xy> blocks.
int triples_energy_npairs(int nv, int w_f32) {
  const int E = tw_blk(w_f32);
  int nb = (nv + E - 1) / E;
  return nb * (nb + 1) / 2;
}

void launch_triples_energy(int no, int nv, int ntrip, const int *trip4, const TripProb *probs, const void *W,
                           int w_f32, long wstride, int fov, double *etrip, double *part, int *counter,
                           hipStream_t st) {
  if (ntrip <= 0) return;
  // FOV with an FP32 W: the (<pq|xy>, t2) pairs in FP32, so that the LDS
  // footprint (and two workgroups per CU) is that of the plain kernel
  if (w_f32) {
    if (fov) launch_e<float, 12, float, true>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
    else launch_e<float, 12, double, false>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
  } else {
    if (fov) launch_e<double, 12, double, true>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
    else launch_e<double, 12, double, false>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
  }
}

}  // namespace ugx
