// C-ABI layer of the learned T2 guess network (the kernels are in
// ugx_mlp.hip).  A network owns its parameters, its Adam state and its work
// buffers on one device and one stream; every call leaves the stream idle,
// so a context on the same device can read the inference weights.
//
// Parameter vector (FP32, this order): W1[128][8], b1[128], W2[128][128],
// b2[128], W3[128][128], b3[128], W4[128], b4.  Linear layers compute
// z = W h + b.  The loss of a batch of n rows is (1/n) sum (f(x) - y)^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/ugx.h"
#include "ugx_host.h"
#include "ugx_internal.h"

using namespace ugx;

namespace {

constexpr int W = MLP_WIDTH, NF = MLP_NFEAT;
constexpr long OFF_W1 = 0, OFF_B1 = OFF_W1 + (long)W * NF, OFF_W2 = OFF_B1 + W, OFF_B2 = OFF_W2 + (long)W * W,
               OFF_W3 = OFF_B2 + W, OFF_B3 = OFF_W3 + (long)W * W, OFF_W4 = OFF_B3 + W, OFF_B4 = OFF_W4 + W,
               NPARAM = OFF_B4 + 1;
constexpr int MAX_BATCH = 1 << 16;

}  // namespace

struct ugx_mlp {
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<float> host;  // parameters (host mirror, kept in sync)
  std::vector<float> shift = std::vector<float>(NF, 0.0f), scale = std::vector<float>(NF, 1.0f);
  double y_scale = 1.0;
  long adam_t = 0;
  DBuf p, g, m, v, dshift, dinv;
  // inference view
  DBuf w1f, b1f, w2p, w3p;
  // work buffers (sized for one batch)
  DBuf Xn, Xb, yb, Z[3], H[3], dZ, dH, yo, dy, sq, part, wpart, loss;
  // training set
  DBuf DX, DY, perm;
};

namespace {

template <typename F>
int mguard(ugx_mlp *n, F &&f) {
  if (!n) return UGX_BAD_INPUT;
  try {
    hipError_t e = hipSetDevice(n->device);
    if (e != hipSuccess) throw UgxError(UGX_HIP_ERROR, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return f();
  } catch (const UgxError &e) {
    n->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    n->err = "host allocation failed";
    return UGX_OOM;
  } catch (const std::exception &e) {
    n->err = e.what();
    return UGX_HIP_ERROR;
  }
}

void h2d(DBuf &b, const void *h, size_t bytes, hipStream_t st) {
  b.alloc(bytes);
  check(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st), "host-to-device copy");
}

unsigned short bf16_host(float x) {
  unsigned int u;
  std::memcpy(&u, &x, 4);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Input feature of the fused kernel's K-step s, lane group q, slot j.
int kperm(int s, int q, int j) { return 32 * s + (j < 4 ? 4 * q + j : 16 + 4 * q + j - 4); }

// Packs W[out][in] (128 x 128) into the MFMA A fragments of the fused kernel:
// fragment (F, s), lane l holds W[16F + l%16][kperm(s, l/16, 0..7)].
std::vector<unsigned short> pack_hidden(const float *Wm) {
  std::vector<unsigned short> out((size_t)8 * 4 * 64 * 8);
  for (int F = 0; F < 8; F++)
    for (int s = 0; s < 4; s++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 8; j++)
          out[(((size_t)F * 4 + s) * 64 + l) * 8 + j] = bf16_host(Wm[(long)(16 * F + l % 16) * W + kperm(s, l / 16, j)]);
  return out;
}

// Uploads parameters and rebuilds the inference view (folded first layer,
// packed hidden layers); resets Adam.
void upload_params(ugx_mlp *n) {
  const float *P = n->host.data();
  h2d(n->p, P, NPARAM * 4, n->st);
  n->g.alloc(NPARAM * 4);
  n->m.alloc(NPARAM * 4);
  n->v.alloc(NPARAM * 4);
  check(hipMemsetAsync(n->m.p, 0, NPARAM * 4, n->st), "memset");
  check(hipMemsetAsync(n->v.p, 0, NPARAM * 4, n->st), "memset");
  n->adam_t = 0;
  std::vector<float> w1f((size_t)W * NF), b1f(W);
  for (int f = 0; f < W; f++) {
    double b = P[OFF_B1 + f];
    for (int k = 0; k < NF; k++) {
      const double w = (double)P[OFF_W1 + (long)f * NF + k] / (double)n->scale[k];
      w1f[(size_t)f * NF + k] = (float)w;
      b -= w * (double)n->shift[k];
    }
    b1f[f] = (float)b;
  }
  h2d(n->w1f, w1f.data(), w1f.size() * 4, n->st);
  h2d(n->b1f, b1f.data(), b1f.size() * 4, n->st);
  std::vector<unsigned short> w2 = pack_hidden(P + OFF_W2), w3 = pack_hidden(P + OFF_W3);
  h2d(n->w2p, w2.data(), w2.size() * 2, n->st);
  h2d(n->w3p, w3.data(), w3.size() * 2, n->st);
  std::vector<float> inv(NF);
  for (int k = 0; k < NF; k++) inv[k] = 1.0f / n->scale[k];
  h2d(n->dshift, n->shift.data(), NF * 4, n->st);
  h2d(n->dinv, inv.data(), NF * 4, n->st);
  check(hipStreamSynchronize(n->st), "stream sync");
}

void download_params(ugx_mlp *n) {
  check(hipMemcpyAsync(n->host.data(), n->p.p, NPARAM * 4, hipMemcpyDeviceToHost, n->st), "params copy");
  check(hipStreamSynchronize(n->st), "stream sync");
}

MlpDevice view(const ugx_mlp *n) {
  MlpDevice d;
  const float *P = (const float *)n->p.p;
  d.w1f = n->w1f.f();
  d.b1f = n->b1f.f();
  d.w2p = n->w2p.p;
  d.b2 = P + OFF_B2;
  d.w3p = n->w3p.p;
  d.b3 = P + OFF_B3;
  d.w4 = P + OFF_W4;
  d.b4 = n->host[OFF_B4];
  return d;
}

// C[m][n] (FP32) = A(m,k) B(k,n) on the FP32 MFMA GEMM
void gemm32(hipStream_t st, int M, int N, int K, const float *A, long sam, long sak, const float *B, long sbk,
            long sbn, float *C, long scm, long scn) {
  GemmDesc d;
  std::memset(&d, 0, sizeof(d));
  d.M = M; d.N = N; d.nseg = 1;
  d.seg[0].A = A; d.seg[0].B = B;
  d.seg[0].sam = sam; d.seg[0].sak = sak; d.seg[0].sbk = sbk; d.seg[0].sbn = sbn;
  d.seg[0].K = K; d.seg[0].sign = 1.0;
  d.C = (double *)C; d.scm = scm; d.scn = scn;
  d.alpha = 1.0; d.beta = 0.0; d.batch = 1; d.c_f32 = 1;
  launch_gemm(d, nullptr, PREC_FP32, -1, st);
}

// dW[out][in] (FP32) = sum_r dZ[r][out] Hp[r][in] over the B rows, split over
// chunks of 256 rows into partials that are then summed in a fixed order
void gemm_wgrad(ugx_mlp *n, int B, int out, int in, const float *dZ, const float *Hp, float *dW) {
  const int CHK = 256, S = (B + CHK - 1) / CHK;
  n->wpart.alloc((size_t)S * out * in * 4);
  GemmDesc d;
  std::memset(&d, 0, sizeof(d));
  d.M = out; d.N = in; d.nseg = 1;
  d.seg[0].A = dZ; d.seg[0].sam = 1; d.seg[0].sak = out; d.seg[0].sAz = (long)CHK * out;
  d.seg[0].B = Hp; d.seg[0].sbk = in; d.seg[0].sbn = 1; d.seg[0].sBz = (long)CHK * in;
  d.seg[0].K = CHK; d.seg[0].sign = 1.0;
  d.ktotal = B;
  d.C = (double *)n->wpart.p; d.scm = in; d.scn = 1; d.sCz = (long)out * in;
  d.alpha = 1.0; d.beta = 0.0; d.batch = S; d.c_f32 = 1;
  launch_gemm(d, nullptr, PREC_FP32, -1, n->st);
  launch_mlp_sum_parts(S, (long)out * in, n->wpart.f(), dW, n->st);
}

void ensure_work(ugx_mlp *n, int B) {
  const size_t act = (size_t)B * W * 4;
  for (int l = 0; l < 3; l++) { n->Z[l].alloc(act); n->H[l].alloc(act); }
  n->dZ.alloc(act); n->dH.alloc(act);
  n->yo.alloc((size_t)B * 4); n->dy.alloc((size_t)B * 4); n->sq.alloc((size_t)B * 8);
  n->part.alloc((size_t)mlp_colsum_chunks(B) * W * 4);
  n->loss.alloc(8);
}

// Forward of B normalised rows X (device); with a target, also the loss
// (into loss_out, device) and the full gradient into n->g.
void fwd_bwd(ugx_mlp *n, int B, const float *X, const float *target, double *loss_out) {
  if (target && !loss_out) throw UgxError(UGX_BAD_INPUT, "internal: loss output missing");
  ensure_work(n, B);
  hipStream_t st = n->st;
  float *P = n->p.f(), *G = n->g.f();
  const float *Hin = X;
  int kin = NF;
  const long offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
  for (int l = 0; l < 3; l++) {
    // Z = Hin W^T: A(m,k) = Hin[m][k], B(k,n) = W[n][k]
    gemm32(st, B, W, kin, Hin, kin, 1, P + offW[l], 1, kin, n->Z[l].f(), W, 1);
    launch_mlp_bias_gelu(B, P + offB[l], n->Z[l].f(), n->H[l].f(), st);
    Hin = n->H[l].f();
    kin = W;
  }
  launch_mlp_head(B, n->H[2].f(), P + OFF_W4, P + OFF_B4, target, n->yo.f(), n->dy.f(), (double *)n->sq.p, st);
  if (!target) return;
  launch_mlp_loss(B, (const double *)n->sq.p, loss_out, st);
  // output layer: dW4 = sum_r dy_r H3[r], db4 = sum_r dy_r
  launch_mlp_colsum(B, W, n->H[2].f(), n->dy.f(), n->part.f(), G + OFF_W4, st);
  launch_mlp_colsum(B, 1, n->dy.f(), nullptr, n->part.f(), G + OFF_B4, st);
  launch_mlp_head_back(B, n->dy.f(), P + OFF_W4, n->Z[2].f(), n->dZ.f(), st);
  for (int l = 2; l >= 0; l--) {
    const float *Hp = l > 0 ? n->H[l - 1].f() : X;
    const int in = l > 0 ? W : NF;
    gemm_wgrad(n, B, W, in, n->dZ.f(), Hp, G + offW[l]);
    launch_mlp_colsum(B, W, n->dZ.f(), nullptr, n->part.f(), G + offB[l], st);
    if (l > 0) {
      // dH = dZ W: A(m,k) = dZ[m][k], B(k,n) = W[k][n]
      gemm32(st, B, W, W, n->dZ.f(), W, 1, P + offW[l], W, 1, n->dH.f(), W, 1);
      launch_mlp_gelu_back((long)B * W, n->dH.f(), n->Z[l - 1].f(), n->dZ.f(), st);
    }
  }
}

void adam_step(ugx_mlp *n, double lr) {
  const double b1 = 0.9, b2 = 0.999;
  n->adam_t++;
  const double bc1 = 1.0 - std::pow(b1, (double)n->adam_t), bc2 = 1.0 - std::pow(b2, (double)n->adam_t);
  launch_mlp_adam(NPARAM, n->p.f(), n->g.f(), n->m.f(), n->v.f(), (float)lr, (float)b1, (float)b2, 1e-8f,
                  (float)bc1, (float)bc2, n->st);
}

void check_rows(const float *X, long n) {
  if (!X || n <= 0) throw UgxError(UGX_BAD_INPUT, "empty feature matrix");
}

// Copies n raw feature rows to the device and standardises them into n->Xn.
void upload_normalized(ugx_mlp *n, const float *X, long rows) {
  h2d(n->Xb, X, (size_t)rows * NF * 4, n->st);
  n->Xn.alloc((size_t)rows * NF * 4);
  launch_mlp_normalize(rows, n->Xb.f(), n->dshift.f(), n->dinv.f(), n->Xn.f(), n->st);
}

}  // namespace

// Inference view for the CCSD guess (same device as the network).
bool ugx::mlp_inference_view(const ugx_mlp *n, int device, MlpDevice *out, double *y_scale) {
  if (!n || n->device != device || !n->p.p) return false;
  *out = view(n);
  *y_scale = n->y_scale;
  return true;
}

extern "C" {

long ugx_mlp_num_params(void) { return NPARAM; }

int ugx_mlp_create(ugx_mlp **out, int device, unsigned long long seed) {
  if (!out) return UGX_BAD_INPUT;
  *out = nullptr;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) {
    (void)hipGetLastError();
    return UGX_NO_DEVICE;
  }
  if (device < 0 || device >= cnt) return UGX_NO_DEVICE;
  ugx_mlp *n = new (std::nothrow) ugx_mlp();
  if (!n) return UGX_OOM;
  n->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) {
    delete n;
    return UGX_HIP_ERROR;
  }
  // default initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights
  // and biases of each layer, from a seeded 64-bit Mersenne twister
  n->host.assign(NPARAM, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](long off, long cnt_, int fan_in) {
    std::uniform_real_distribution<double> u(-1.0 / std::sqrt((double)fan_in), 1.0 / std::sqrt((double)fan_in));
    for (long i = 0; i < cnt_; i++) n->host[off + i] = (float)u(rng);
  };
  fill(OFF_W1, (long)W * NF, NF); fill(OFF_B1, W, NF);
  fill(OFF_W2, (long)W * W, W); fill(OFF_B2, W, W);
  fill(OFF_W3, (long)W * W, W); fill(OFF_B3, W, W);
  fill(OFF_W4, W, W); fill(OFF_B4, 1, W);
  int rc = mguard(n, [&]() { upload_params(n); return UGX_OK; });
  if (rc != UGX_OK) {
    ugx_mlp_destroy(n);
    return rc;
  }
  *out = n;
  return UGX_OK;
}

void ugx_mlp_destroy(ugx_mlp *n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  if (n->st) (void)hipStreamSynchronize(n->st);
  hipStream_t st = n->st;
  delete n;
  if (st) (void)hipStreamDestroy(st);
}

const char *ugx_mlp_last_error(const ugx_mlp *n) { return n ? n->err.c_str() : "null network"; }

int ugx_mlp_set_params(ugx_mlp *n, const float *params) {
  return mguard(n, [&]() {
    if (!params) throw UgxError(UGX_BAD_INPUT, "null parameter vector");
    for (long i = 0; i < NPARAM; i++)
      if (!std::isfinite(params[i])) throw UgxError(UGX_NONFINITE, "non-finite parameter");
    std::memcpy(n->host.data(), params, NPARAM * 4);
    upload_params(n);
    return UGX_OK;
  });
}

int ugx_mlp_get_params(ugx_mlp *n, float *params) {
  return mguard(n, [&]() {
    if (!params) throw UgxError(UGX_BAD_INPUT, "null parameter vector");
    std::memcpy(params, n->host.data(), NPARAM * 4);
    return UGX_OK;
  });
}

int ugx_mlp_set_scaling(ugx_mlp *n, const float *x_shift, const float *x_scale, double y_scale) {
  return mguard(n, [&]() {
    if (!x_shift || !x_scale || !(y_scale > 0) || !std::isfinite(y_scale))
      throw UgxError(UGX_BAD_INPUT, "ugx_mlp_set_scaling: bad arguments");
    for (int k = 0; k < NF; k++)
      if (!(x_scale[k] > 0) || !std::isfinite(x_scale[k]) || !std::isfinite(x_shift[k]))
        throw UgxError(UGX_BAD_INPUT, "ugx_mlp_set_scaling: scales must be positive and finite");
    n->shift.assign(x_shift, x_shift + NF);
    n->scale.assign(x_scale, x_scale + NF);
    n->y_scale = y_scale;
    upload_params(n);  // refolds the first layer (and resets Adam)
    return UGX_OK;
  });
}

int ugx_mlp_forward(ugx_mlp *n, const float *X, long rows, int precision, float *y) {
  return mguard(n, [&]() {
    check_rows(X, rows);
    if (!y) throw UgxError(UGX_BAD_INPUT, "null output");
    if (precision != UGX_PREC_FP32 && precision != UGX_PREC_BF16)
      throw UgxError(UGX_BAD_INPUT, "ugx_mlp_forward: precision must be UGX_PREC_FP32 or UGX_PREC_BF16");
    if (precision == UGX_PREC_BF16) {
      h2d(n->Xb, X, (size_t)rows * NF * 4, n->st);
      n->yb.alloc((size_t)rows * 4);
      launch_mlp_infer_x(rows, n->Xb.f(), view(n), n->yb.f(), n->st);
      check(hipMemcpyAsync(y, n->yb.p, (size_t)rows * 4, hipMemcpyDeviceToHost, n->st), "output copy");
    } else {
      if (rows > MAX_BATCH) throw UgxError(UGX_BAD_INPUT, "ugx_mlp_forward: FP32 form takes at most 65536 rows");
      upload_normalized(n, X, rows);
      fwd_bwd(n, (int)rows, n->Xn.f(), nullptr, nullptr);
      check(hipMemcpyAsync(y, n->yo.p, (size_t)rows * 4, hipMemcpyDeviceToHost, n->st), "output copy");
    }
    check(hipStreamSynchronize(n->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_mlp_loss_grad(ugx_mlp *n, const float *X, const float *y, long rows, double *loss, float *grad) {
  return mguard(n, [&]() {
    check_rows(X, rows);
    if (!y || !loss || !grad) throw UgxError(UGX_BAD_INPUT, "null argument");
    if (rows > MAX_BATCH) throw UgxError(UGX_BAD_INPUT, "ugx_mlp_loss_grad: at most 65536 rows");
    upload_normalized(n, X, rows);
    h2d(n->yb, y, (size_t)rows * 4, n->st);
    n->loss.alloc(8);  // before its address is taken
    fwd_bwd(n, (int)rows, n->Xn.f(), n->yb.f(), n->loss.d());
    check(hipMemcpyAsync(loss, n->loss.p, 8, hipMemcpyDeviceToHost, n->st), "loss copy");
    check(hipMemcpyAsync(grad, n->g.p, NPARAM * 4, hipMemcpyDeviceToHost, n->st), "gradient copy");
    check(hipStreamSynchronize(n->st), "stream sync");
    return UGX_OK;
  });
}

int ugx_mlp_train(ugx_mlp *n, const float *X, const float *y, long rows, int epochs, int batch, double lr,
                  unsigned long long seed, double *epoch_loss) {
  return mguard(n, [&]() {
    check_rows(X, rows);
    if (!y || epochs < 1 || batch < 1 || batch > MAX_BATCH || !(lr > 0) || rows > (1L << 31) - 1)
      throw UgxError(UGX_BAD_INPUT, "ugx_mlp_train: bad arguments");
    const int B = (int)std::min<long>(batch, rows);
    const long steps = rows / B;  // whole batches per epoch; the remainder rows are reshuffled next epoch
    // the training set stays on the device, standardised once
    h2d(n->Xb, X, (size_t)rows * NF * 4, n->st);
    n->DX.alloc((size_t)rows * NF * 4);
    launch_mlp_normalize(rows, n->Xb.f(), n->dshift.f(), n->dinv.f(), n->DX.f(), n->st);
    h2d(n->DY, y, (size_t)rows * 4, n->st);
    n->perm.alloc((size_t)rows * 4);
    n->Xn.alloc((size_t)B * NF * 4);
    n->yb.alloc((size_t)B * 4);
    DBuf losses;
    losses.alloc((size_t)steps * 8);
    std::vector<int> order(rows);
    std::mt19937_64 rng(seed);
    std::vector<double> hl(steps);
    for (int ep = 0; ep < epochs; ep++) {
      for (long i = 0; i < rows; i++) order[i] = (int)i;
      std::shuffle(order.begin(), order.end(), rng);
      check(hipMemcpyAsync(n->perm.p, order.data(), (size_t)rows * 4, hipMemcpyHostToDevice, n->st), "perm copy");
      for (long s = 0; s < steps; s++) {
        launch_mlp_gather(B, n->DX.f(), n->DY.f(), (const int *)n->perm.p + s * B, n->Xn.f(), n->yb.f(), n->st);
        fwd_bwd(n, B, n->Xn.f(), n->yb.f(), losses.d() + s);
        adam_step(n, lr);
      }
      check(hipMemcpyAsync(hl.data(), losses.p, (size_t)steps * 8, hipMemcpyDeviceToHost, n->st), "loss copy");
      check(hipStreamSynchronize(n->st), "stream sync");
      double sum = 0.0;
      for (long s = 0; s < steps; s++) sum += hl[s];
      if (!std::isfinite(sum)) throw UgxError(UGX_NONFINITE, "training loss is not finite");
      if (epoch_loss) epoch_loss[ep] = sum / (double)steps;
    }
    download_params(n);
    // refresh the inference view without resetting the optimiser state
    const long t = n->adam_t;
    std::vector<float> m(NPARAM), v(NPARAM);
    check(hipMemcpyAsync(m.data(), n->m.p, NPARAM * 4, hipMemcpyDeviceToHost, n->st), "copy");
    check(hipMemcpyAsync(v.data(), n->v.p, NPARAM * 4, hipMemcpyDeviceToHost, n->st), "copy");
    check(hipStreamSynchronize(n->st), "stream sync");
    upload_params(n);
    check(hipMemcpyAsync(n->m.p, m.data(), NPARAM * 4, hipMemcpyHostToDevice, n->st), "copy");
    check(hipMemcpyAsync(n->v.p, v.data(), NPARAM * 4, hipMemcpyHostToDevice, n->st), "copy");
    check(hipStreamSynchronize(n->st), "stream sync");
    n->adam_t = t;
    return UGX_OK;
  });
}

}  // extern "C"
