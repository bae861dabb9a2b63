// (T) energy epilogue, ccwfn.cc:795-852.
//
// For a batch of i>=j>=k triplets whose connected W3 (v^3, [a][b][c]) has
// been built by the six permuted MFMA contractions, each workgroup takes one
// (triplet, a) row and sums the bracket over b<=a, c<=b:
//   V_abc = (W_abc + <ij|ab> t1_kc + <ik|ac> t1_jb + <jk|bc> t1_ia)
//           / (1 + d_ab + d_ac + d_bc)
//   X = sum over the six permutations of W*V,  Y = V_abc + V_bca + V_cab,
//   Z = V_acb + V_bac + V_cba
//   E += [(Y - 2Z)(W_abc + W_bca + W_cab) + (Z - 2Y)(W_acb + W_bac + W_cba)
//         + 3X] (2 - d_ij - d_ik - d_jk) / D_ijk^abc
// V is formed on the fly from W, the oovv block and t1, so only W is stored.
// Sums are FP64; the per-(triplet, a) partials are reduced per triplet in a
// fixed order, so the result is bitwise repeatable.
#include "ugx_internal.h"

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

__global__ __launch_bounds__(NT) void k_triples_energy(int no, int nv, const int *__restrict__ trip,
                                                       const double *__restrict__ Wall, long wstride,
                                                       const double *__restrict__ eps,
                                                       const double *__restrict__ oovv,
                                                       const double *__restrict__ t1, double *__restrict__ part) {
  const int t = blockIdx.x;
  const int a = blockIdx.y;
  const int i = trip[3 * t], j = trip[3 * t + 1], k = trip[3 * t + 2];
  const double *W = Wall + (long)t * wstride;
  const long v = nv, v2 = (long)nv * nv;
  const double *Oij = oovv + ((long)i * no + j) * v2;  // <ij|xy>
  const double *Oik = oovv + ((long)i * no + k) * v2;
  const double *Ojk = oovv + ((long)j * no + k) * v2;
  const double *t1i = t1 + (long)i * nv, *t1j = t1 + (long)j * nv, *t1k = t1 + (long)k * nv;
  const double eo = eps[i] + eps[j] + eps[k];
  const double fac = 2.0 - ((i == j) + (i == k) + (j == k));
  const double ea = eps[no + a];

  auto Wx = [&](int x, int y, int z) { return W[x * v2 + y * v + z]; };
  // V at (x,y,z) for x,y,z in the roles of a,b,c of the reference
  auto Vx = [&](int x, int y, int z, double w) {
    double d = 1.0 + (x == y) + (x == z) + (y == z);
    return (w + Oij[x * v + y] * t1k[z] + Oik[x * v + z] * t1j[y] + Ojk[y * v + z] * t1i[x]) / d;
  };

  double s = 0.0;
  const int npair = (a + 1) * (a + 2) / 2;  // (b, c) with c <= b <= a
  for (int p = threadIdx.x; p < npair; p += NT) {
    // p -> (b, c): b = floor((sqrt(8p+1)-1)/2)
    int b = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while ((b + 1) * (b + 2) / 2 <= p) b++;
    while (b * (b + 1) / 2 > p) b--;
    int c = p - b * (b + 1) / 2;
    double wabc = Wx(a, b, c), wacb = Wx(a, c, b), wbac = Wx(b, a, c);
    double wbca = Wx(b, c, a), wcab = Wx(c, a, b), wcba = Wx(c, b, a);
    double vabc = Vx(a, b, c, wabc), vacb = Vx(a, c, b, wacb), vbac = Vx(b, a, c, wbac);
    double vbca = Vx(b, c, a, wbca), vcab = Vx(c, a, b, wcab), vcba = Vx(c, b, a, wcba);
    double X = wabc * vabc + wacb * vacb + wbac * vbac + wbca * vbca + wcab * vcab + wcba * vcba;
    double Y = vabc + vbca + vcab;
    double Z = vacb + vbac + vcba;
    double denom = eo - (ea + eps[no + b] + eps[no + c]);
    s += ((Y - 2.0 * Z) * (wabc + wbca + wcab) + (Z - 2.0 * Y) * (wacb + wbac + wcba) + 3.0 * X) * fac / denom;
  }
  double tot = block_sum(s);
  if (threadIdx.x == 0) part[(long)t * nv + a] = tot;
}

__global__ void k_triples_reduce(int nv, int ntrip, const double *part, double *etrip) {
  int t = blockIdx.x * NT + threadIdx.x;
  if (t >= ntrip) return;
  double s = 0.0;
  for (int a = 0; a < nv; a++) s += part[(long)t * nv + a];
  etrip[t] = s;
}

}  // namespace

void launch_triples_energy(int no, int nv, int ntrip, const int *trip_ijk, const double *W, long wstride,
                           const double *eps, const double *oovv, const double *t1, double *etrip, double *part,
                           hipStream_t st) {
  if (ntrip <= 0) return;
  hipLaunchKernelGGL(k_triples_energy, dim3(ntrip, nv), dim3(NT), 0, st, no, nv, trip_ijk, W, wstride, eps, oovv,
                     t1, part);
  hipLaunchKernelGGL(k_triples_reduce, dim3((ntrip + NT - 1) / NT), dim3(NT), 0, st, nv, ntrip, part, etrip);
}

}  // namespace ugx
