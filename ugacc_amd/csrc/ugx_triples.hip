// (T) energy epilogue, ccwfn.cc:795-852.
//
// For a batch of i>=j>=k triplets whose connected W3 (v^3, [a][b][c]) has
// been built by the MFMA contractions, the bracket is summed over a>=b>=c:
//   V_abc = (W_abc + <ij|ab> t1_kc + <ik|ac> t1_jb + <jk|bc> t1_ia)
//           / (1 + d_ab + d_ac + d_bc)
//   X = sum over the six permutations of W*V,  Y = V_abc + V_bca + V_cab,
//   Z = V_acb + V_bac + V_cba
//   E += [(Y - 2Z)(W_abc + W_bca + W_cab) + (Z - 2Y)(W_acb + W_bac + W_cba)
//         + 3X] (2 - d_ij - d_ik - d_jk) / D_ijk^abc
// V is formed on the fly from W, the oovv block and t1, so only W is stored.
//
// Every point needs W at all six index permutations.  A workgroup owns one
// triplet and a pair of E-wide index blocks A >= B, and walks the blocks
// C <= B: for each (A, B, C) it stages the six permuted E^3 sub-cubes
// W[A][B][C], W[A][C][B], W[B][A][C], W[B][C][A], W[C][A][B], W[C][B][A]
// into LDS.  W is stored in the blocked layout the W3 GEMMs write (E^3
// sub-cubes, each contiguous), so every staged sub-cube is one contiguous
// read and every W element is read from memory about once.
// W is FP64 (FP64 form) or FP32 (low-precision forms); all sums are FP64 and
// the per-(triplet, block pair) partials are reduced per triplet in a fixed
// order, so the result is bitwise repeatable.
#include "ugx_internal.h"

#include <algorithm>

namespace ugx {
namespace {

constexpr int NT = 256;

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
  return s;
}

// Block pair index p -> (A, B) with B <= A.
__device__ inline void pair_of(int p, int &A, int &B) {
  A = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
  while ((A + 1) * (A + 2) / 2 <= p) A++;
  while (A * (A + 1) / 2 > p) A--;
  B = p - A * (A + 1) / 2;
}

template <typename WT, int E, typename OT>
__global__ __launch_bounds__(NT) void k_triples_energy(int no, int nv, int ntrip, const int *__restrict__ trip,
                                                       const TripProb *__restrict__ probs,
                                                       const WT *__restrict__ Wall, long wstride,
                                                       double *__restrict__ part, int npairs,
                                                       int *__restrict__ counter) {
  constexpr int E3 = E * E * E, E2 = E * E;
  // padded cube layout idx(a,b,c) = a*SA + b*SB + c: the strides are odd, so
  // the stores along a or b (and the reads along c) do not collide in banks
  constexpr int SB = E + 1, SA = E * (E + 1) + 1;
  __shared__ WT S[6][E * SA];
  // disconnected-term operands of the three blocks X, Y in {A, B, C}:
  // O[s][X][Y][x][y] = <pq|xy> for the slices s = (ij), (ik), (jk) (OT,
  // FP64 for both W types);
  // T1[s][X][x] = t1 of k, j, i (the amplitude that multiplies slice s)
  __shared__ OT O[3][3][3][E2];
  __shared__ double T1[3][3][E];
  __shared__ double EPS[3][E];
  __shared__ int item_sh;
  const long total = (long)ntrip * npairs;
  // persistent: each workgroup takes (entry, block pair) items from the
  // queue until it is empty; every workgroup reaches the exit test
  for (;;) {
  __syncthreads();  // the previous item is done with LDS and item_sh
  if (threadIdx.x == 0) item_sh = atomicAdd(counter, 1);
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
  const WT *W = Wall + (long)t * wstride;
  const long v = nv, v2 = (long)nv * nv;
  const double *Osl[3] = {oovv + ((long)i * no + j) * v2, oovv + ((long)i * no + k) * v2,
                          oovv + ((long)j * no + k) * v2};
  const double *T1sl[3] = {t1 + (long)k * nv, t1 + (long)j * nv, t1 + (long)i * nv};
  const double eo = eps[i] + eps[j] + eps[k];
  const double fac = 2.0 - ((i == j) + (i == k) + (j == k));
  int X0[3] = {Ab * E, Bb * E, 0};  // block origins of A, B, C (C set per step)

  // V at a permutation (x, y, z) of the point, given as block slots (bx, by,
  // bz) in {0: A, 1: B, 2: C} and local offsets (lx, ly, lz)
  auto Vx = [&](int bx, int lx, int by, int ly, int bz, int lz, double w, double inv) {
    return (w + (double)O[0][bx][by][lx * E + ly] * T1[0][bz][lz] + (double)O[1][bx][bz][lx * E + lz] * T1[1][by][ly] +
            (double)O[2][by][bz][ly * E + lz] * T1[2][bx][lx]) * inv;
  };
  auto stage_pairs = [&](bool all) {
    // O blocks for pairs touching C (or all pairs on the first step) and the
    // t1 / eps entries of the blocks
    for (int e = threadIdx.x; e < 3 * 9 * E2; e += NT) {
      int s = e / (9 * E2), r = e % (9 * E2), pr = r / E2, xy = r % E2;
      int bx = pr / 3, by = pr % 3;
      if (!all && bx != 2 && by != 2) continue;
      int gx = X0[bx] + xy / E, gy = X0[by] + xy % E;
      O[s][bx][by][xy] = (OT)((gx < nv && gy < nv) ? Osl[s][gx * v + gy] : 0.0);
    }
    for (int e = threadIdx.x; e < 3 * 3 * E; e += NT) {
      int s = e / (3 * E), r = e % (3 * E), bx = r / E, lx = r % E;
      if (!all && bx != 2) continue;
      int gx = X0[bx] + lx;
      T1[s][bx][lx] = gx < nv ? T1sl[s][gx] : 0.0;
      if (s == 0) EPS[bx][lx] = gx < nv ? eps[no + gx] : 0.0;
    }
  };

  double s = 0.0;
  const int A0 = X0[0], B0 = X0[1];
  // W is blocked (E^3 sub-cubes, each contiguous in its own [x][y][z]
  // order): sub-cube (X, Y, Z) starts at ((X nb + Y) nb + Z) E^3
  const long nbk = (nv + E - 1) / E;
  auto cube = [&](int X, int Y, int Z) { return W + ((X * nbk + Y) * nbk + Z) * (long)E3; };
  for (int Cb = 0; Cb <= Bb; Cb++) {
    const int C0 = Cb * E;
    X0[2] = C0;
    auto at = [](int a, int b, int c) { return a * SA + b * SB + c; };
    // Stage the six sub-cubes, S[q][at(a,b,c)] = W at permutation q of
    // (A0+a, B0+b, C0+c): each is one contiguous E^3 chunk of W, read in
    // order (entries past nv are never used: the points that would need
    // them are skipped below).
    const WT *c0 = cube(Ab, Bb, Cb), *c1 = cube(Ab, Cb, Bb), *c2 = cube(Bb, Ab, Cb);
    const WT *c3 = cube(Bb, Cb, Ab), *c4 = cube(Cb, Ab, Bb), *c5 = cube(Cb, Bb, Ab);
    for (int e = threadIdx.x; e < E3; e += NT) {
      int f = e % E, m = (e / E) % E, sl = e / (E * E);  // fastest, middle, slowest
      S[0][at(sl, m, f)] = c0[e];   // W[a][b][c]
      S[1][at(sl, f, m)] = c1[e];   // W[a][c][b]
      S[2][at(m, sl, f)] = c2[e];   // W[b][a][c]
      S[3][at(f, sl, m)] = c3[e];   // W[b][c][a]
      S[4][at(m, f, sl)] = c4[e];   // W[c][a][b]
      S[5][at(f, m, sl)] = c5[e];   // W[c][b][a]
    }
    stage_pairs(Cb == 0);
    __syncthreads();
    for (int e = threadIdx.x; e < E3; e += NT) {
      int cc = e % E, bb = (e / E) % E, aa = e / (E * E);
      int a = A0 + aa, b = B0 + bb, c = C0 + cc;
      if (a >= nv || b > a || c > b) continue;
      const int x = at(aa, bb, cc);
      // 1 / (1 + d_ab + d_ac + d_bc): 1, 1/2 or 1/4, exact
      const double inv = (a == b && b == c) ? 0.25 : ((a == b || b == c || a == c) ? 0.5 : 1.0);
      double wabc = (double)S[0][x], wacb = (double)S[1][x], wbac = (double)S[2][x];
      double wbca = (double)S[3][x], wcab = (double)S[4][x], wcba = (double)S[5][x];
      double vabc = Vx(0, aa, 1, bb, 2, cc, wabc, inv), vacb = Vx(0, aa, 2, cc, 1, bb, wacb, inv);
      double vbac = Vx(1, bb, 0, aa, 2, cc, wbac, inv), vbca = Vx(1, bb, 2, cc, 0, aa, wbca, inv);
      double vcab = Vx(2, cc, 0, aa, 1, bb, wcab, inv), vcba = Vx(2, cc, 1, bb, 0, aa, wcba, inv);
      double X = wabc * vabc + wacb * vacb + wbac * vbac + wbca * vbca + wcab * vcab + wcba * vcba;
      double Y = vabc + vbca + vcab;
      double Z = vacb + vbac + vcba;
      double denom = eo - (EPS[0][aa] + EPS[1][bb] + EPS[2][cc]);
      s += ((Y - 2.0 * Z) * (wabc + wbca + wcab) + (Z - 2.0 * Y) * (wacb + wbac + wcba) + 3.0 * X) * fac / denom;
    }
    __syncthreads();
  }
  double tot = block_sum(s);
  if (threadIdx.x == 0) part[(long)t * npairs + pidx] = tot;
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

template <typename WT, int E, typename OT>
void launch_e(int no, int nv, int ntrip, const int *trip4, const TripProb *probs, const void *W, long wstride,
              double *etrip, double *part, int *counter, hipStream_t st) {
  int nb = (nv + E - 1) / E;
  int npairs = nb * (nb + 1) / 2;
  // as many workgroups as can be resident at once
  int dev = 0, ncu = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_triples_energy<WT, E, OT>, NT, 0);
  long grid = (long)std::max(ncu, 1) * std::max(per, 1);
  grid = std::min<long>(grid, (long)ntrip * npairs);
  hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, counter);
  hipLaunchKernelGGL((k_triples_energy<WT, E, OT>), dim3((unsigned)grid), dim3(NT), 0, st, no, nv, ntrip, trip4,
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
                           int w_f32, long wstride, double *etrip, double *part, int *counter, hipStream_t st) {
  if (ntrip <= 0) return;
  if (w_f32) launch_e<float, 12, double>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
  else launch_e<double, 12, double>(no, nv, ntrip, trip4, probs, W, wstride, etrip, part, counter, st);
}

}  // namespace ugx
