// Learned T2 guess: kernels of the per-amplitude network
//   x (8 features) -> 128 -> 128 -> 128 -> 1,  GELU after the hidden layers,
// which replaces the MP2 doubles guess of CCWfn::CCWfn (ccwfn.cc:81-87) by
//   t2_ijab = t2(MP2)_ijab + y_scale * (f(x_ijab) + f(x_jiba)) / 2.
//
// Features of amplitude (i,j,a,b), computed in FP64 and stored as FP32:
//   x0 = asinh(1e3 t_MP2),  t_MP2 = <ij|ab> / D,  D = f_ii + f_jj - f_aa - f_bb
//   x1 = log(-D)
//   x2 = asinh(1e2 <ij|ab>),  x3 = asinh(1e2 <ij|ba>)
//   x4..x7 = f_ii, f_jj, f_aa, f_bb
// The network standardises them with a fixed shift/scale (from the training
// set) before the first layer.
//
// GELU is the tanh form, gelu(z) = z * sigmoid(2u), u = sqrt(2/pi)(z + 0.044715 z^3)
// (Hendrycks & Gimpel), written here as z / (1 + exp(-2u)).
//
// Inference (k_mlp_infer) is one fused kernel on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulate).  Each wave forms the
// features of 64 amplitudes (one per lane), then takes them 16 at a time
// and computes the transposed layers
//   Z^T[f][n] = sum_k W[f][k] H^T[k][n]
// so the weights are the A operand (from LDS) and the activations the B
// operand (in registers).  The accumulator of lane l holds rows
// f = 16F + 4(l/16) + r of column n = l%16; the K index of the next layer is
// permuted so that exactly these values form the lane's B fragment:
// fragment slot j of lane group q in K-step s stands for input feature
//   pi_s(q, j) = 32s + 4q + j          (j < 4)
//              = 32s + 16 + 4q + j - 4  (j >= 4),
// and the host packs W[f][pi_s(q, j)] into the A fragments in the same order.
// Activations therefore never leave the registers between layers.  The
// first layer (K = 8) is one MFMA per 16 rows with the features and W1
// split into BF16 hi + lo (three of the four k-groups: W1hi xhi, W1hi xlo,
// W1lo xhi), accurate to about 2^-16 before its BF16 output rounding.
//
// Training runs in FP32 on the library's MFMA GEMM (forward Z = H W^T,
// backward dH = dZ W and dW = dZ^T H) with the element-wise kernels below
// and Adam; every reduction is in a fixed order, so training is bitwise
// repeatable.
#include "ugx_internal.h"

namespace ugx {
namespace {

constexpr int NT = 256;
constexpr int WID = MLP_WIDTH;
constexpr int NF = MLP_NFEAT;

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline float gelu_f(float z) {
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  return z / (1.0f + __expf(-2.0f * u));
}

// Inference form: exp(-2u) as 2^(z (c1 + c3 z^2)) with c1 = -2 sqrt(2/pi)
// log2(e) and c3 = 0.044715 c1 folded, and the hardware reciprocal (1 ulp)
// instead of an IEEE division; its output is rounded to BF16 (or summed in
// FP32 by the output layer), far coarser than the difference.
__device__ inline float gelu_infer_f(float z) {
  const float e = __builtin_amdgcn_exp2f(z * (-2.302208198144325f + -0.1029432395800235f * (z * z)));
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ inline float gelu_grad_f(float z) {
  const float z2 = z * z;
  const float u = 0.7978845608028654f * (z + 0.044715f * z2 * z);
  const float s = 1.0f / (1.0f + __expf(-2.0f * u));
  const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * z2);
  return s + 2.0f * z * s * (1.0f - s) * du;
}

__device__ inline unsigned int bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// F32: the transcendental functions in FP32 (the inference kernel; the
// features feed a BF16 network), else in FP64 (the training features)
template <bool F32 = false>
__device__ inline void amp_features(int no, int nv, const double *__restrict__ eps,
                                    const double *__restrict__ oovv, long g, float *x) {
  const long v = nv;
  const int b = (int)(g % v);
  long r = g / v;
  const int a = (int)(r % v);
  r /= v;
  const int j = (int)(r % no);
  const int i = (int)(r / no);
  const double ei = eps[i], ej = eps[j], ea = eps[no + a], eb = eps[no + b];
  const double D = ei + ej - ea - eb;
  const double I = oovv[g];
  const double K = oovv[(((long)i * no + j) * v + b) * v + a];
  if (F32) {
    x[0] = asinhf((float)(1e3 * I / D));
    x[1] = logf(fmaxf((float)-D, 1e-6f));
    x[2] = asinhf((float)(1e2 * I));
    x[3] = asinhf((float)(1e2 * K));
  } else {
    x[0] = (float)asinh(1e3 * (I / D));
    x[1] = (float)log(fmax(-D, 1e-6));
    x[2] = (float)asinh(1e2 * I);
    x[3] = (float)asinh(1e2 * K);
  }
  x[4] = (float)ei;
  x[5] = (float)ej;
  x[6] = (float)ea;
  x[7] = (float)eb;
}

__global__ void k_features(int no, int nv, const double *__restrict__ eps, const double *__restrict__ oovv, long n,
                           const long *__restrict__ idx, float *__restrict__ X) {
  const long r = blockIdx.x * (long)NT + threadIdx.x;
  if (r >= n) return;
  float x[NF];
  amp_features(no, nv, eps, oovv, idx ? idx[r] : r, x);
#pragma unroll
  for (int k = 0; k < NF; k++) X[r * NF + k] = x[k];
}

// One wave per 64 amplitudes at a time (four 16-amplitude MFMA tiles),
// eight waves per workgroup, grid-stride over the 64-amplitude groups (the
// trip count is uniform across a wave, so every lane reaches every MFMA).
// 512 threads: the 64 KB of weight fragments in LDS are shared by eight
// waves (two workgroups per CU, four waves per SIMD).
constexpr int NTI = 512;

template <bool HAM>
__global__ __launch_bounds__(NTI) __attribute__((amdgpu_waves_per_eu(4))) void k_mlp_infer(long N, const float *__restrict__ Xraw, int no, int nv,
                                                  const double *__restrict__ eps, const double *__restrict__ oovv,
                                                  MlpDevice P, float *__restrict__ out) {
  __shared__ u32x4 sW1[8 * 64], sW2[8 * 4 * 64], sW3[8 * 4 * 64];
  __shared__ float sb1[WID], sb2[WID], sb3[WID], sW4[WID];
  const u32x4 *gW2 = (const u32x4 *)P.w2p, *gW3 = (const u32x4 *)P.w3p;
  for (int e = threadIdx.x; e < 8 * 4 * 64; e += NTI) {
    sW2[e] = gW2[e];
    sW3[e] = gW3[e];
  }
  // layer 1 (K = 8) as one MFMA per 16 rows in split BF16: the fragment of
  // lane (r, g) for rows 16F + r holds W1 hi in k-groups g = 0, 1, W1 lo in
  // g = 2 and zeros in g = 3; the feature fragments hold x hi, x lo, x hi,
  // zeros, so the products are W1hi xhi + W1hi xlo + W1lo xhi
  static_assert(8 * 64 <= NTI && WID == 128 && NF == 8, "one layer-1 fragment per thread");
  if (threadIdx.x < 8 * 64) {
    const int l = threadIdx.x & 63, gq = l >> 4, fr = 16 * (threadIdx.x >> 6) + (l & 15);
    unsigned int h[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const float w = P.w1f[fr * NF + k];
      const unsigned int hi = bf16_rne(w), lo = bf16_rne(w - __uint_as_float(hi << 16));
      h[k] = gq < 2 ? hi : (gq == 2 ? lo : 0u);
    }
    sW1[threadIdx.x] = u32x4{h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16)};
  }
  for (int e = threadIdx.x; e < WID; e += NTI) {
    sb1[e] = P.b1f[e];
    sb2[e] = P.b2[e];
    sb3[e] = P.b3[e];
    sW4[e] = P.w4[e];
  }
  __syncthreads();
  const float b4 = P.b4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 15, q = lane >> 4;
  // a wave takes 64 amplitudes at a time: each lane forms the features of
  // one, and the four 16-amplitude tiles take theirs by lane shuffles
  const long ngroup = (N + 63) / 64;
  for (long t = (long)blockIdx.x * (NTI / 64) + wave; t < ngroup; t += (long)gridDim.x * (NTI / 64)) {
    // features of amplitude t * 64 + lane, split into BF16 hi and lo pairs
    unsigned int xh[NF / 2], xl[NF / 2];
    {
      float xo[NF];
      const long g = t * 64 + lane;
      if (g < N) {
        if (HAM) amp_features<true>(no, nv, eps, oovv, g, xo);
        else {
#pragma unroll
          for (int k = 0; k < NF; k++) xo[k] = Xraw[g * NF + k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < NF; k++) xo[k] = 0.0f;
      }
#pragma unroll
      for (int m = 0; m < NF / 2; m++) {
        const unsigned int h0 = bf16_rne(xo[2 * m]), h1 = bf16_rne(xo[2 * m + 1]);
        xh[m] = h0 | (h1 << 16);
        xl[m] = bf16_rne(xo[2 * m] - __uint_as_float(h0 << 16)) |
                (bf16_rne(xo[2 * m + 1] - __uint_as_float(h1 << 16)) << 16);
      }
    }
#pragma unroll 1
    for (int u = 0; u < 4; u++) {
      // the weight fragments are re-read from LDS every tile: held across
      // tiles they would take 256 registers and leave one wave per SIMD
      asm volatile("" ::: "memory");
      const long g = t * 64 + 16 * u + n;
      const bool valid = g < N;
      // the tile's feature fragments: x hi for k-groups 0 and 2, x lo for 1
      unsigned int bf[NF / 2];
#pragma unroll
      for (int m = 0; m < NF / 2; m++) {
        const unsigned int hv = __shfl(xh[m], 16 * u + n, 64), lv = __shfl(xl[m], 16 * u + n, 64);
        bf[m] = q == 1 ? lv : (q == 3 ? 0u : hv);
      }
      const u32x4 bx = u32x4{bf[0], bf[1], bf[2], bf[3]};
      // GELU of rows 32 s + (the lane's 4 + 4 rows) of an accumulator pair,
      // packed as the next layer's B fragment s
      auto pack = [&](const flt4 *acc, const float *sb, u32x4 *bo) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
          unsigned int hb[8];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            hb[r] = bf16_rne(gelu_infer_f(acc[2 * s][r] + sb[32 * s + 4 * q + r]));
            hb[4 + r] = bf16_rne(gelu_infer_f(acc[2 * s + 1][r] + sb[32 * s + 16 + 4 * q + r]));
          }
          bo[s] = u32x4{hb[0] | (hb[1] << 16), hb[2] | (hb[3] << 16), hb[4] | (hb[5] << 16), hb[6] | (hb[7] << 16)};
        }
      };
      u32x4 bop[4];
      {
        flt4 acc[8];
#pragma unroll
        for (int F = 0; F < 8; F++)
          acc[F] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, sW1[F * 64 + lane]),
                                                           __builtin_bit_cast(bf16x8, bx), flt4{0.f, 0.f, 0.f, 0.f},
                                                           0, 0, 0);
        pack(acc, sb1, bop);
      }
      // layers 2 and 3 on MFMA
      float part = 0.0f;
#pragma unroll
      for (int layer = 0; layer < 2; layer++) {
        const u32x4 *sW = layer == 0 ? sW2 : sW3;
        const float *sb = layer == 0 ? sb2 : sb3;
        flt4 acc[8];
#pragma unroll
        for (int F = 0; F < 8; F++) {
          acc[F] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int s = 0; s < 4; s++)
            acc[F] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, sW[(F * 4 + s) * 64 + lane]),
                                                             __builtin_bit_cast(bf16x8, bop[s]), acc[F], 0, 0, 0);
        }
        if (layer == 0) {
          pack(acc, sb, bop);
        } else {
          // output layer (128 -> 1): this lane's 32 features, then the four
          // lane groups of column n
#pragma unroll
          for (int F = 0; F < 8; F++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int f = 16 * F + 4 * q + r;
              part += sW4[f] * gelu_infer_f(acc[F][r] + sb[f]);
            }
        }
      }
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      if (q == 0 && valid) out[g] = part + b4;
    }
  }
}

// t2_ijab += y_scale * (G_ijab + G_jiba) / 2 (t2 holds the MP2 guess)
__global__ void k_apply_guess(int no, int nv, const float *__restrict__ G, double ys, double *__restrict__ t2) {
  const long v = nv, n = (long)no * no * v * v;
  const long g = blockIdx.x * (long)NT + threadIdx.x;
  if (g >= n) return;
  const int b = (int)(g % v);
  long r = g / v;
  const int a = (int)(r % v);
  r /= v;
  const int j = (int)(r % no);
  const int i = (int)(r / no);
  const long gt = (((long)j * no + i) * v + b) * v + a;
  t2[g] += ys * 0.5 * ((double)G[g] + (double)G[gt]);
}

// ------------------------------------------------------------ training
__global__ void k_normalize(long n, const float *__restrict__ X, const float *__restrict__ shift,
                            const float *__restrict__ inv, float *__restrict__ Xn) {
  const long e = blockIdx.x * (long)NT + threadIdx.x;
  if (e >= n * NF) return;
  const int k = (int)(e % NF);
  Xn[e] = (X[e] - shift[k]) * inv[k];
}

__global__ void k_gather(int B, const float *__restrict__ X, const float *__restrict__ y, const int *__restrict__ perm,
                         float *__restrict__ Xb, float *__restrict__ yb) {
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r >= B) return;
  const long src = perm[r];
#pragma unroll
  for (int k = 0; k < NF; k++) Xb[(long)r * NF + k] = X[src * NF + k];
  yb[r] = y[src];
}

// Z += b (row broadcast), H = gelu(Z)
__global__ void k_bias_gelu(long n, const float *__restrict__ b, float *__restrict__ Z, float *__restrict__ H) {
  const long e = blockIdx.x * (long)NT + threadIdx.x;
  if (e >= n * WID) return;
  const float z = Z[e] + b[e % WID];
  Z[e] = z;
  H[e] = gelu_f(z);
}

// y = H3 w4 + b4 (one wave per row); with a target: dy = 2 (y - t) / B and
// the squared error per row
__global__ void k_head(int B, const float *__restrict__ H, const float *__restrict__ w4, const float *__restrict__ b4,
                       const float *__restrict__ target, float *__restrict__ y, float *__restrict__ dy,
                       double *__restrict__ sq) {
  const int row = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= B) return;
  const float *h = H + (long)row * WID;
  float s = h[lane] * w4[lane] + h[lane + 64] * w4[lane + 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) {
    const float yv = s + b4[0];
    y[row] = yv;
    if (target) {
      const float d = yv - target[row];
      dy[row] = 2.0f * d / (float)B;
      sq[row] = (double)d * (double)d;
    }
  }
}

// dZ3 = dy (x) w4 * gelu'(Z3)
__global__ void k_head_back(long B, const float *__restrict__ dy, const float *__restrict__ w4,
                            const float *__restrict__ Z, float *__restrict__ dZ) {
  const long e = blockIdx.x * (long)NT + threadIdx.x;
  if (e >= B * WID) return;
  dZ[e] = dy[e / WID] * w4[e % WID] * gelu_grad_f(Z[e]);
}

// dZ = dH * gelu'(Z)
__global__ void k_gelu_back(long n, const float *__restrict__ dH, const float *__restrict__ Z, float *__restrict__ dZ) {
  const long e = blockIdx.x * (long)NT + threadIdx.x;
  if (e >= n) return;
  dZ[e] = dH[e] * gelu_grad_f(Z[e]);
}

// P[c][f] = sum over rows r of chunk c (CH rows) of w[r] * M[r][f]
// (w = null: weight 1); width <= NT columns, one thread per column
constexpr int CH = 256;
__global__ void k_colsum_part(int B, int width, const float *__restrict__ M, const float *__restrict__ w,
                              float *__restrict__ P) {
  const int f = threadIdx.x;
  if (f >= width) return;
  const int r0 = blockIdx.x * CH, r1 = min(B, r0 + CH);
  float s = 0.0f;
  for (int r = r0; r < r1; r++) s += (w ? w[r] : 1.0f) * M[(long)r * width + f];
  P[(long)blockIdx.x * width + f] = s;
}

// out[i] = sum_s P[s][i]
__global__ void k_sum_parts(int S, long n, const float *__restrict__ P, float *__restrict__ out) {
  const long i = blockIdx.x * (long)NT + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int k = 0; k < S; k++) s += P[(long)k * n + i];
  out[i] = s;
}

// out[slot] = sum of the n squared errors / n
__global__ void k_loss(int n, const double *__restrict__ sq, double *__restrict__ out) {
  __shared__ double red[NT];
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += NT) s += sq[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] / (double)n;
}

__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const long i = blockIdx.x * (long)NT + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks(long n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

void launch_features(int no, int nv, const double *eps, const double *oovv, long n, const long *idx, float *X,
                     hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_features, dim3(blocks(n)), dim3(NT), 0, st, no, nv, eps, oovv, n, idx, X);
}

static unsigned infer_grid(long N) {
  const long waves = (N + 63) / 64;
  const long b = (waves + NTI / 64 - 1) / (NTI / 64);
  return (unsigned)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

void launch_mlp_infer_x(long N, const float *X, const MlpDevice &P, float *out, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_mlp_infer<false>, dim3(infer_grid(N)), dim3(NTI), 0, st, N, X, 0, 0, (const double *)nullptr,
                     (const double *)nullptr, P, out);
}

void launch_mlp_infer_ham(int no, int nv, const double *eps, const double *oovv, const MlpDevice &P, float *out,
                          hipStream_t st) {
  const long N = (long)no * no * nv * nv;
  if (N <= 0) return;
  hipLaunchKernelGGL(k_mlp_infer<true>, dim3(infer_grid(N)), dim3(NTI), 0, st, N, (const float *)nullptr, no, nv, eps,
                     oovv, P, out);
}

void launch_apply_guess(int no, int nv, const float *G, double ys, double *t2, hipStream_t st) {
  const long n = (long)no * no * nv * nv;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_apply_guess, dim3(blocks(n)), dim3(NT), 0, st, no, nv, G, ys, t2);
}

void launch_mlp_normalize(long n, const float *X, const float *shift, const float *inv, float *Xn, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_normalize, dim3(blocks(n * NF)), dim3(NT), 0, st, n, X, shift, inv, Xn);
}

void launch_mlp_gather(int B, const float *X, const float *y, const int *perm, float *Xb, float *yb, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(blocks(B)), dim3(NT), 0, st, B, X, y, perm, Xb, yb);
}

void launch_mlp_bias_gelu(long rows, const float *b, float *Z, float *H, hipStream_t st) {
  hipLaunchKernelGGL(k_bias_gelu, dim3(blocks(rows * WID)), dim3(NT), 0, st, rows, b, Z, H);
}

void launch_mlp_head(int B, const float *H, const float *w4, const float *b4, const float *target, float *y, float *dy,
                     double *sq, hipStream_t st) {
  hipLaunchKernelGGL(k_head, dim3((B + NT / 64 - 1) / (NT / 64)), dim3(NT), 0, st, B, H, w4, b4, target, y, dy, sq);
}

void launch_mlp_head_back(int B, const float *dy, const float *w4, const float *Z, float *dZ, hipStream_t st) {
  hipLaunchKernelGGL(k_head_back, dim3(blocks((long)B * WID)), dim3(NT), 0, st, (long)B, dy, w4, Z, dZ);
}

void launch_mlp_gelu_back(long n, const float *dH, const float *Z, float *dZ, hipStream_t st) {
  hipLaunchKernelGGL(k_gelu_back, dim3(blocks(n)), dim3(NT), 0, st, n, dH, Z, dZ);
}

int mlp_colsum_chunks(int B) { return (B + CH - 1) / CH; }

void launch_mlp_colsum(int B, int width, const float *M, const float *w, float *P, float *out, hipStream_t st) {
  const int S = mlp_colsum_chunks(B);
  hipLaunchKernelGGL(k_colsum_part, dim3(S), dim3(NT), 0, st, B, width, M, w, P);
  hipLaunchKernelGGL(k_sum_parts, dim3(blocks(width)), dim3(NT), 0, st, S, (long)width, P, out);
}

void launch_mlp_sum_parts(int S, long n, const float *P, float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_parts, dim3(blocks(n)), dim3(NT), 0, st, S, n, P, out);
}

void launch_mlp_loss(int n, const double *sq, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(NT), 0, st, n, sq, out);
}

void launch_mlp_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     float bc1, float bc2, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(NT), 0, st, n, p, g, m, v, lr, b1, b2, eps, bc1, bc2);
}

}  // namespace ugx
