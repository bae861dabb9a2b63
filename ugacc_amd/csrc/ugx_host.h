// Host-side helpers shared by the C-ABI translation units: the error type
// every ABI call turns into a status code, and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../include/ugx.h"
#include "ugx_internal.h"

namespace ugx {

struct UgxError : std::runtime_error {
  int code;
  UgxError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw UgxError(UGX_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device allocation; alloc() only grows.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (bytes >= b && p) return;
    release();
    if (b == 0) b = 8;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw UgxError(UGX_OOM, "device allocation of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
  }
  double *d() const { return (double *)p; }
  float *f() const { return (float *)p; }
};

// Owning pinned host allocation (device-to-host copies land in it directly,
// with no staging copy); alloc() only grows.
struct HBuf {
  double *p = nullptr;
  size_t n = 0;  // doubles
  HBuf() = default;
  HBuf(const HBuf &) = delete;
  HBuf &operator=(const HBuf &) = delete;
  ~HBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (n >= count && p) return;
    release();
    void *q = nullptr;
    if (hipHostMalloc(&q, std::max<size_t>(count, 1) * 8, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      throw UgxError(UGX_OOM, "pinned host allocation of " + std::to_string(count * 8) + " bytes failed");
    }
    p = (double *)q;
    n = count;
  }
};

// Inference view of a learned-guess network for a context on `device`
// (false when the network lives on another device or has no parameters).
bool mlp_inference_view(const ugx_mlp *net, int device, MlpDevice *out, double *y_scale);

}  // namespace ugx
