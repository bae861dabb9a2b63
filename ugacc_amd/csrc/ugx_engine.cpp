// Contraction engine (see ugx_engine.h).
#include "ugx_engine.h"

#include <rocblas/rocblas.h>

#include <cstdlib>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace ugx {

namespace {
// Profiling label of a contraction: the low-precision forms get a suffix so
// that their device time is reported apart from the FP64 polish.
std::string prec_label(const char *label, int prec) {
  std::string s(label);
  if (prec == PREC_FP32) s += ":fp32";
  else if (prec == PREC_BF16X3) s += ":bf16x3";
  return s;
}
}  // namespace


void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

namespace {

struct Spec {
  std::string a, b, c;
};

Spec parse(const std::string &s) {
  Spec r;
  size_t comma = s.find(','), arrow = s.find("->");
  if (arrow == std::string::npos) throw std::runtime_error("ugx: contraction spec needs '->': " + s);
  if (comma == std::string::npos || comma > arrow) {
    r.a = s.substr(0, arrow);
    r.c = s.substr(arrow + 2);
  } else {
    r.a = s.substr(0, comma);
    r.b = s.substr(comma + 1, arrow - comma - 1);
    r.c = s.substr(arrow + 2);
  }
  return r;
}

long extent_of(const std::string &letters, const std::map<char, int> &ext) {
  long n = 1;
  for (char ch : letters) n *= ext.at(ch);
  return n;
}

// Input strides (elements) of a row-major tensor with the given letters.
std::map<char, long> strides_of(const std::string &letters, const std::map<char, int> &ext) {
  std::map<char, long> s;
  long st = 1;
  for (int d = (int)letters.size() - 1; d >= 0; d--) {
    s[letters[d]] = st;
    st *= ext.at(letters[d]);
  }
  return s;
}

bool is_perm_of(const std::string &x, const std::string &y) {
  std::string a = x, b = y;
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  return a == b;
}

// bytes per element: FP64, FP32, BF16 hi/lo pair
size_t esize(int dtype) { return dtype == 0 ? 8 : 4; }


}  // namespace

Engine::~Engine() {
  if (blas_) (void)rocblas_destroy_handle(blas_);
  if (blas_ws_) (void)hipFree(blas_ws_);
  for (auto &s : slots_)
    if (s.first) (void)hipFree(s.first);
  for (void *p : retired_) (void)hipFree(p);
  for (auto &kv : cache_) (void)hipFree(kv.second.first);
  for (auto &p : pending_) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : free_events_) (void)hipEventDestroy(e);
}

void Engine::drop_cache() {
  hip_check(hipStreamSynchronize(st_), "drop_cache sync");
  for (auto &kv : cache_) (void)hipFree(kv.second.first);
  cache_.clear();
  cache_pd_.clear();
  for (void *p : retired_) (void)hipFree(p);
  retired_.clear();
  for (auto &s : slots_) {
    if (s.first) (void)hipFree(s.first);
    s = {nullptr, 0};
  }
  gen_++;
}

void Engine::refresh(const void *p, int dmin, int dmax) {
  if (sizing_) return;
  for (auto &kv : cache_pd_) {
    const int dt = std::get<2>(kv.first);
    if (std::get<0>(kv.first) == p && dt >= dmin && dt <= dmax) launch_permute(kv.second, st_);
  }
}

size_t Engine::cached_bytes() const {
  size_t n = 0;
  for (auto &kv : cache_) n += kv.second.second;
  return n;
}

void *Engine::slot(int i, size_t bytes) {
  auto &s = slots_.at(i);
  if (s.second < bytes) {
    // the old workspace may be addressed by queued or captured work: retire
    // it (freed by drop_cache) instead of freeing it here, which would also
    // synchronise the device under other threads' streams
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes), "workspace alloc");
    if (s.first) retired_.push_back(s.first);
    s.first = p;
    s.second = bytes;
    gen_++;
    if (!sizing_) allocs_++;
  }
  return s.first;
}

hipEvent_t Engine::event() {
  if (!free_events_.empty()) {
    hipEvent_t e = free_events_.back();
    free_events_.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "event create");
  return e;
}

void Engine::collect_stats() {
  if (pending_.empty()) return;
  hip_check(hipStreamSynchronize(st_), "stats sync");
  for (auto &p : pending_) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, p.a, p.b), "event elapsed");
    auto &k = stats_[p.label];
    k.launches++;
    k.ms += ms;
    k.flops += p.flops;
    free_events_.push_back(p.a);
    free_events_.push_back(p.b);
  }
  pending_.clear();
}

// Returns a device pointer holding tensor T (letters) in the order `want`
// with element type dtype.  No copy when T already is that; constants are
// permuted/converted once and cached; variables go to workspace slot_id.
const void *Engine::prepared(const TRef &T, const std::string &letters, const std::string &want, int dtype,
                             int slot_id, long pk) {
  if (letters == want && T.dtype == dtype) return T.p;
  std::map<char, int> ext;
  for (size_t d = 0; d < letters.size(); d++) ext[letters[d]] = T.dims[d];
  long n = extent_of(letters, ext);
  size_t bytes = n * esize(dtype);
  long pblk = 0, plo = 0;
  if (dtype == 3) {
    if (pk <= 0 || n % pk) throw std::runtime_error("ugx: bad panel operand shape");
    const long R = n / pk, Rp = (R + 15) / 16 * 16, Kp = (pk + 31) / 32 * 32;
    pblk = 16 * Kp;
    plo = Rp * Kp;
    bytes = (size_t)plo * 2 * 2;  // hi and lo halves
  }
  void *dst;
  if (T.constant) {
    // the key names the permutation (source letters -> wanted order), not
    // just the wanted order: one stored block is read under several letter
    // assignments (e.g. <mn|ef> as "mnef" and as "mnfe")
    auto key = std::make_tuple(T.p, letters + "->" + want, dtype);
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second.first;
    hip_check(hipMalloc(&dst, bytes), "cache alloc");
    cache_[key] = {dst, bytes};
    if (!sizing_) allocs_++;
  } else {
    dst = slot(slot_id, bytes);
    if (sizing_) return dst;  // sized; its contents come with the real evaluation
  }
  auto is = strides_of(letters, ext);
  PermDesc pd{};
  pd.rank = (int)want.size();
  if (pd.rank == 0) throw std::runtime_error("ugx: empty operand");
  for (int d = 0; d < pd.rank; d++) { pd.od[d] = ext.at(want[d]); pd.istr[d] = is.at(want[d]); }
  pd.in = T.p; pd.out = dst; pd.in_t = T.dtype; pd.out_t = dtype; pd.alpha = 1.0; pd.beta = 0.0;
  pd.pk = pk; pd.pblk = pblk; pd.plo = plo;
  launch_permute(pd, st_);
  if (T.constant) cache_pd_[std::make_tuple(T.p, letters + "->" + want, dtype)] = pd;
  return dst;
}

void Engine::permute(const std::string &s, double alpha, const TRef &A, double beta, const TRef &C) {
  if (sizing_) return;
  Spec sp = parse(s);
  if (!sp.b.empty()) throw std::runtime_error("ugx: permute takes one operand");
  if (!is_perm_of(sp.a, sp.c)) throw std::runtime_error("ugx: permute letters differ: " + s);
  if (A.dims.size() != sp.a.size() || C.dims.size() != sp.c.size())
    throw std::runtime_error("ugx: permute rank mismatch: " + s);
  std::map<char, int> ext;
  for (size_t d = 0; d < sp.a.size(); d++) ext[sp.a[d]] = A.dims[d];
  for (size_t d = 0; d < sp.c.size(); d++)
    if (ext.at(sp.c[d]) != C.dims[d]) throw std::runtime_error("ugx: permute extent mismatch: " + s);
  auto is = strides_of(sp.a, ext);
  PermDesc pd{};
  pd.rank = (int)sp.c.size();
  for (int d = 0; d < pd.rank; d++) { pd.od[d] = ext.at(sp.c[d]); pd.istr[d] = is.at(sp.c[d]); }
  pd.in = A.p; pd.out = const_cast<void *>(C.p); pd.in_t = A.dtype; pd.out_t = C.dtype;
  pd.alpha = alpha; pd.beta = beta;
  launch_permute(pd, st_);
}

void Engine::timed(const char *label, double flops, const std::function<void()> &fn) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (profiling_) { e0 = event(); e1 = event(); hip_check(hipEventRecord(e0, st_), "event record"); }
  fn();
  flops_ += flops;
  if (profiling_) {
    hip_check(hipEventRecord(e1, st_), "event record");
    pending_.push_back(Pending{label, e0, e1, flops});
  }
}

struct Engine::Plan {
  double cost;
  bool swap;          // GEMM computes with (B, A) in the (A, B) roles
  std::string m, n, k;
};

void Engine::contract(const std::string &s, double alpha, const TRef &A, const TRef &B, double beta, const TRef &C,
                      int prec, const char *label) {
  Spec sp = parse(s);
  if (sp.b.empty()) throw std::runtime_error("ugx: contraction needs two operands: " + s);
  if (A.dims.size() != sp.a.size() || B.dims.size() != sp.b.size() || C.dims.size() != sp.c.size())
    throw std::runtime_error("ugx: rank mismatch in " + s);
  if (C.dtype != 0) throw std::runtime_error("ugx: contraction output must be FP64");
  std::map<char, int> ext;
  auto add = [&](const std::string &l, const TRef &T) {
    for (size_t d = 0; d < l.size(); d++) {
      auto it = ext.find(l[d]);
      if (it != ext.end() && it->second != T.dims[d]) throw std::runtime_error("ugx: extent clash in " + s);
      ext[l[d]] = T.dims[d];
    }
  };
  add(sp.a, A); add(sp.b, B); add(sp.c, C);

  auto in = [](const std::string &x, char ch) { return x.find(ch) != std::string::npos; };
  std::string K, Ma, Nb;  // K in A's order, M in A's order, N in B's order
  for (char ch : sp.a) {
    if (in(sp.b, ch) && !in(sp.c, ch)) K += ch;
    else if (in(sp.c, ch) && !in(sp.b, ch)) Ma += ch;
    else throw std::runtime_error(std::string("ugx: unsupported index '") + ch + "' in " + s);
  }
  for (char ch : sp.b) {
    if (in(sp.a, ch)) continue;
    if (in(sp.c, ch)) Nb += ch;
    else throw std::runtime_error(std::string("ugx: unsupported index '") + ch + "' in " + s);
  }
  for (char ch : sp.c)
    if (!in(sp.a, ch) && !in(sp.b, ch)) throw std::runtime_error("ugx: output index not in operands: " + s);
  auto order_in = [&](const std::string &group, const std::string &ref) {
    std::string r;
    for (char ch : ref)
      if (in(group, ch)) r += ch;
    return r;
  };
  std::string Mc = order_in(Ma, sp.c), Nc = order_in(Nb, sp.c), Kb = order_in(K, sp.b);

  const double sA = (double)extent_of(sp.a, ext), sB = (double)extent_of(sp.b, ext), sC = (double)extent_of(sp.c, ext);
  const bool fp64 = (prec == PREC_FP64);
  auto opcost = [&](const TRef &T, const std::string &letters, const std::string &g1, const std::string &g2,
                    double size) {
    bool ok = (letters == g1 + g2) || (letters == g2 + g1);
    bool needs = !ok || (!fp64 && T.dtype == 0);
    if (!needs) return 0.0;
    return T.constant ? 0.01 * size : 2.0 * size;
  };
  // C written straight from the GEMM epilogue: each of the GEMM's row and
  // column index groups maps to C with at most two stride levels (outer
  // letters, inner letters; GemmDesc cdiv / ndiv), the columns' innermost
  // stride being 1 so that a wave's stores stay contiguous.  Otherwise the
  // GEMM writes a workspace that is permuted into C.
  struct OutMap {
    bool ok = false;
    long div = 0, s1 = 0, s0 = 0;
  };
  const auto cstr = strides_of(sp.c, ext);
  auto flat = [&](const std::string &g) {
    for (size_t q = 0; q + 1 < g.size(); q++)
      if (cstr.at(g[q]) != cstr.at(g[q + 1]) * ext.at(g[q + 1])) return false;
    return true;
  };
  auto outmap = [&](const std::string &g) {
    OutMap o;
    if (g.empty()) { o.ok = true; return o; }
    if (flat(g)) { o.ok = true; o.s0 = cstr.at(g.back()); return o; }
    for (size_t p = 1; p < g.size(); p++) {
      const std::string outer = g.substr(0, p), inner = g.substr(p);
      if (flat(outer) && flat(inner)) {
        o.ok = true;
        o.div = extent_of(inner, ext);
        o.s0 = cstr.at(inner.back());
        o.s1 = cstr.at(outer.back());
        o.ok = o.div <= 0x7fffffffL;
        return o;
      }
    }
    return o;
  };
  auto two_level = [&](const std::string &rows, const std::string &cols) {
    const OutMap rm = outmap(rows), cm = outmap(cols);
    return rm.ok && cm.ok && (cols.empty() || cm.s0 == 1);
  };
  Plan best{1e300, false, "", "", ""};
  for (int sw = 0; sw < 2; sw++)
    for (int mi = 0; mi < 2; mi++)
      for (int ni = 0; ni < 2; ni++)
        for (int ki = 0; ki < 2; ki++) {
          std::string m = mi ? Ma : Mc, n = ni ? Nb : Nc, k = ki ? Kb : K;
          double c = opcost(A, sp.a, m, k, sA) + opcost(B, sp.b, k, n, sB);
          // C = A B with rows m, cols n (sw = 0) or rows n, cols m (sw = 1)
          std::string rows = sw ? n : m, cols = sw ? m : n;
          if (sp.c == rows + cols || two_level(rows, cols)) c += 0.0;
          else c += 3.0 * sC;
          if (c < best.cost) best = Plan{c, sw == 1, m, n, k};
        }

  // operand element type: FP64, FP32, or (BF16x3) pre-split BF16 hi/lo pairs
  // with k contiguous
  const bool pair = (prec == PREC_BF16X3);
  const int dt = fp64 ? 0 : (pair ? 2 : 1);
  const int kprec = pair ? (int)PREC_BF16X3_PRESPLIT : prec;
  const long Me = extent_of(best.m, ext), Ne = extent_of(best.n, ext), Ke = extent_of(best.k, ext);
  // Operand A as matrix (m, k); B as (k, n).
  const long Re = best.swap ? Ne : Me, Ce = best.swap ? Me : Ne;
  // BF16x3 operands as MFMA panels (LDS-copy kernel, GemmDesc::dma) when K
  // is whole panel slices (no k padding to clear) and the GEMM does not
  // take the split-K path below
  bool panel = false;
  if (pair && panels_ && Ke % BK_PANEL == 0) {
    GemmDesc t{};
    t.M = (int)std::min<long>(Re, 0x7fffffffL); t.N = (int)std::min<long>(Ce, 0x7fffffffL); t.batch = 1; t.nseg = 1;
    int bm, bn;
    gemm_tile_dims(gemm_pick_tile(t), bm, bn);
    const long tiles = ((Re + bm - 1) / bm) * ((Ce + bn - 1) / bn);
    panel = !(tiles < 256 && Ke >= 2048);
  }
  auto mat = [&](const TRef &T, const std::string &letters, const std::string &g1, const std::string &g2,
                 long e1, long e2, int slot_id, long &s1, long &s2) -> const void * {
    // returns pointer; strides s1 (along g1 index), s2 (along g2 index)
    if (panel) {
      s1 = BK_PANEL; s2 = 1;
      return prepared(T, letters, g1 + g2, 3, slot_id, e2);
    }
    if (pair) {
      s1 = e2; s2 = 1;
      if (letters == g1 + g2 && T.dtype == 2) return T.p;
      return prepared(T, letters, g1 + g2, 2, slot_id);
    }
    if (letters == g1 + g2 && T.dtype == dt) { s1 = e2; s2 = 1; return T.p; }
    if (letters == g2 + g1 && T.dtype == dt) { s1 = 1; s2 = e1; return T.p; }
    // permute/convert into g1+g2 unless the g2+g1 layout only needs a conversion
    if (letters == g2 + g1) {
      s1 = 1; s2 = e1;
      return prepared(T, letters, g2 + g1, dt, slot_id);
    }
    s1 = e2; s2 = 1;
    return prepared(T, letters, g1 + g2, dt, slot_id);
  };
  GemmDesc d{};
  long sam, sak, sbk, sbn;
  const void *Ap = mat(A, sp.a, best.m, best.k, Me, Ke, 0, sam, sak);
  long skb, snb;
  const void *Bp = mat(B, sp.b, best.n, best.k, Ne, Ke, 1, snb, skb);
  sbk = skb; sbn = snb;
  const long Kp = (Ke + BK_PANEL - 1) / BK_PANEL * BK_PANEL;
  const long loA = panel ? (Me + 15) / 16 * 16 * Kp : (pair ? Me * Ke : 0);
  const long loB = panel ? (Ne + 15) / 16 * 16 * Kp : (pair ? Ne * Ke : 0);
  const long pblk = panel ? 16 * Kp : 0;

  std::string rows = best.swap ? best.n : best.m, cols = best.swap ? best.m : best.n;
  GemmSeg g{};
  if (!best.swap) {
    g.A = Ap; g.sam = sam; g.sak = sak; g.B = Bp; g.sbk = sbk; g.sbn = sbn; g.loA = loA; g.loB = loB;
  } else {
    // rows = n: A' (n,k) = B(k,n);  B' (k,m) = A(m,k)
    g.A = Bp; g.sam = sbn; g.sak = sbk; g.B = Ap; g.sbk = sak; g.sbn = sam; g.loA = loB; g.loB = loA;
  }
  g.pblkA = pblk; g.pblkB = pblk;
  if (panel) d.dma = 2;
  g.K = (int)Ke; g.sign = 1.0;
  d.M = (int)Re; d.N = (int)Ce; d.nseg = 1; d.seg[0] = g; d.batch = 1;
  if (Re > 0x7fffffffL || Ce > 0x7fffffffL || Ke > 0x7fffffffL) throw std::runtime_error("ugx: GEMM too large");

  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (profiling_ && label && !sizing_) {
    e0 = event(); e1 = event(); hip_check(hipEventRecord(e0, st_), "event record");
  }

  const bool direct = (sp.c == rows + cols);
  // Split-K: a small output with a long contraction (e.g. Foo = ttau.L over
  // K = o v^2) would occupy only a few workgroups; cut K into S chunks that
  // run as a batch into a workspace, then reduce.
  {
    int tile = gemm_pick_tile(d), bm, bn;
    gemm_tile_dims(tile, bm, bn);
    long tiles = ((Re + bm - 1) / bm) * ((Ce + bn - 1) / bn);
    if (tiles < 256 && Ke >= 2048) {
      long S = std::min<long>((512 + tiles - 1) / tiles, Ke / 512);
      S = std::min<long>(S, 4096);
      if (S > 1) {
        if (panel) throw std::runtime_error("ugx: split-K on panel operands");
        long chunk = (Ke + S - 1) / S;
        chunk = (chunk + 31) / 32 * 32;  // whole slices (and 16-byte pair loads) per chunk
        S = (Ke + chunk - 1) / chunk;
        double *W = (double *)slot(3, (size_t)S * Re * Ce * 8);
        GemmDesc d2 = d;
        d2.batch = (int)S;
        d2.seg[0].K = (int)chunk;
        d2.seg[0].sAz = chunk * d.seg[0].sak;
        d2.seg[0].sBz = chunk * d.seg[0].sbk;
        d2.ktotal = Ke;
        d2.C = W; d2.scm = Ce; d2.scn = 1; d2.scm1 = 0; d2.sCz = Re * Ce; d2.alpha = 1.0; d2.beta = 0.0;
        const int t2 = tile_for(d2, nullptr, kprec, tile);
        if (sizing_) {
          if (!direct) (void)slot(2, (size_t)Re * Ce * 8);
          return;
        }
        run_gemm(d2, kprec, t2);
        if (direct) {
          launch_splitk_reduce((int)S, (int)Re, (int)Ce, W, alpha, beta, (double *)C.p, Ce, 1, st_);
        } else {
          double *tmp = (double *)slot(2, (size_t)Re * Ce * 8);
          launch_splitk_reduce((int)S, (int)Re, (int)Ce, W, 1.0, 0.0, tmp, Ce, 1, st_);
          TRef T = vref(tmp, {});
          for (char ch : rows + cols) T.dims.push_back(ext.at(ch));
          permute(rows + cols + "->" + sp.c, alpha, T, beta, C);
        }
        const double fl = 2.0 * (double)Me * (double)Ne * (double)Ke;
        flops_ += fl;
        if (profiling_ && label) {
          hip_check(hipEventRecord(e1, st_), "event record");
          pending_.push_back(Pending{prec_label(label, prec), e0, e1, fl});
        }
        return;
      }
    }
  }
  if (direct) {
    d.C = (double *)C.p; d.scm = Ce; d.scn = 1; d.scm1 = 0; d.alpha = alpha; d.beta = beta;
    const int tile = tile_for(d, nullptr, kprec, -1);
    if (tile == TILE_LIBRARY) blas_ready();
    if (sizing_) return;
    run_gemm(d, kprec, tile);
  } else if (two_level(rows, cols) && !plain_to_library(d, kprec, Re, Ce)) {
    const OutMap rm = outmap(rows), cm = outmap(cols);
    d.C = (double *)C.p; d.alpha = alpha; d.beta = beta;
    d.scm = rm.s0; d.cdiv = (int)rm.div; d.scm1 = rm.s1;
    d.scn = cm.s0; d.ndiv = (int)cm.div; d.scn1 = cm.s1;
    const int tile = tile_for(d, nullptr, kprec, -1);
    if (sizing_) return;
    run_gemm(d, kprec, tile);
  } else {
    double *tmp = (double *)slot(2, (size_t)Re * Ce * 8);
    d.C = tmp; d.scm = Ce; d.scn = 1; d.scm1 = 0; d.alpha = 1.0; d.beta = 0.0;
    const int tile = tile_for(d, nullptr, kprec, -1);
    if (tile == TILE_LIBRARY) blas_ready();
    if (sizing_) return;
    run_gemm(d, kprec, tile);
    TRef T = vref(tmp, {});
    for (char ch : rows + cols) T.dims.push_back(ext.at(ch));
    permute(rows + cols + "->" + sp.c, alpha, T, beta, C);
  }
  const double fl = 2.0 * (double)Me * (double)Ne * (double)Ke;
  flops_ += fl;
  if (profiling_ && label) {
    hip_check(hipEventRecord(e1, st_), "event record");
    pending_.push_back(Pending{prec_label(label, prec), e0, e1, fl});
  }
}

// ------------------------------------------------------------ autotuner
namespace {

// Tuned tiles, shared by every engine of the process (keyed by device and
// shape; the best tile is a property of the device and the shape).
struct Tuned {
  int tile;
  double ms;
};
std::mutex g_tiles_mu;
std::map<std::string, Tuned> g_tiles;

// The tuned tile depends on the GPU model, not on its index: keys start
// with the architecture name and the CU count of the current device.
std::string device_tag() {
  static std::mutex mu;
  static std::map<int, std::string> tags;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "get device");
  std::lock_guard<std::mutex> lk(mu);
  auto it = tags.find(dev);
  if (it != tags.end()) return it->second;
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "device properties");
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  std::string tag = arch + "/" + std::to_string(prop.multiProcessorCount) + "cu";
  tags[dev] = tag;
  return tag;
}

// Largest element offset of one batch entry's output, plus one.
long out_extent(const GemmDesc &d, long table_extent) {
  if (d.offm) return table_extent;
  auto part = [](int n, int div, long s, long s1) -> long {
    if (div > 0 && div < n) return (long)((n - 1) / div) * s1 + (long)(div - 1) * s;
    return (long)(n - 1) * s;
  };
  return part(d.M, d.cdiv, d.scm, d.scm1) + part(d.N, d.ndiv, d.scn, d.scn1) + 1;
}

}  // namespace

const char *tile_name(int tile) { return tile == TILE_LIBRARY ? "rocblas" : gemm_tile_name(tile); }

// A single FP64 GEMM with plain row-major C (unit column stride), unit
// stride along k or along the row / column index of each operand, and at
// least 2^30 multiply-adds (smaller ones stay on the library's kernel,
// where launch cost dominates).
bool Engine::blas_fits(const GemmDesc &d, int kprec) const {
  if (kprec != PREC_FP64 || d.nseg != 1 || d.batch != 1 || d.offm || d.offn || d.offz || d.cdiv || d.ndiv ||
      d.c_f32 || d.tag || d.ktotal || d.zsub || d.scn != 1 || d.scm < d.N)
    return false;
  const GemmSeg &g = d.seg[0];
  if (g.sign != 1.0 || g.pblkA || g.pblkB || !g.A || !g.B) return false;
  const bool a_ok = (g.sak == 1 && g.sam >= g.K) || (g.sam == 1 && g.sak >= d.M);
  const bool b_ok = (g.sbn == 1 && g.sbk >= d.N) || (g.sbk == 1 && g.sbn >= g.K);
  return a_ok && b_ok && (double)d.M * (double)d.N * (double)g.K >= 1073741824.0;
}

void Engine::blas_ready() {
  if (blas_) return;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw std::runtime_error("ugx: rocblas_create_handle failed");
  blas_ = h;
  const size_t ws = 32u << 20;
  hip_check(hipMalloc(&blas_ws_, ws), "rocBLAS workspace");
  if (rocblas_set_stream(h, st_) != rocblas_status_success || rocblas_set_workspace(h, blas_ws_, ws) != rocblas_status_success ||
      rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success ||
      rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success)
    throw std::runtime_error("ugx: rocBLAS handle set-up failed");
}

// Row-major C (M x N, row stride scm) = alpha A (M x K) B (K x N) + beta C
// as the column-major product C^T = B^T A^T.
void Engine::blas_gemm(const GemmDesc &d) {
  const GemmSeg &g = d.seg[0];
  const bool bn = (g.sbn == 1 && g.sbk >= d.N);
  const bool an = (g.sak == 1 && g.sam >= g.K);
  const rocblas_operation opB = bn ? rocblas_operation_none : rocblas_operation_transpose;
  const rocblas_operation opA = an ? rocblas_operation_none : rocblas_operation_transpose;
  const rocblas_int ldb = (rocblas_int)(bn ? g.sbk : g.sbn), lda = (rocblas_int)(an ? g.sam : g.sak);
  const double alpha = d.alpha, beta = d.beta;
  const rocblas_status r = rocblas_dgemm(blas_, opB, opA, d.N, d.M, g.K, &alpha, (const double *)g.B, ldb,
                                         (const double *)g.A, lda, &beta, (double *)d.C, (rocblas_int)d.scm);
  if (r != rocblas_status_success)
    throw std::runtime_error(std::string("ugx: rocblas_dgemm failed: ") + rocblas_status_to_string(r));
}

// A GEMM whose output needs the two-level map takes the plain temporary and
// a permute instead when the tuner picks rocBLAS for that plain form (large
// FP64 GEMMs only; the permute costs a few percent of such a GEMM)
bool Engine::plain_to_library(const GemmDesc &d, int kprec, long Re, long Ce) {
  if (kprec != PREC_FP64 || !autotune_ || (double)Re * (double)Ce * (double)d.seg[0].K < 1073741824.0) return false;
  GemmDesc p = d;
  p.C = (double *)slot(2, (size_t)Re * Ce * 8);
  p.scm = Ce; p.scn = 1; p.scm1 = 0; p.cdiv = 0; p.ndiv = 0; p.scn1 = 0; p.alpha = 1.0; p.beta = 0.0;
  return blas_fits(p, kprec) && tile_for(p, nullptr, kprec, -1) == TILE_LIBRARY;
}

// A tuned tile runs on the library's kernel or, for TILE_LIBRARY, on
// rocBLAS (a table entry that no longer fits falls back to the heuristic tile)
void Engine::run_gemm(const GemmDesc &d, int kprec, int tile) {
  if (tile == TILE_LIBRARY) {
    if (blas_fits(d, kprec)) {
      blas_ready();
      blas_gemm(d);
      return;
    }
    tile = -1;
  }
  launch_gemm(d, nullptr, kprec, tile, st_);
}

int Engine::tile_for_impl(const GemmDesc &d, const GemmPtrs *ptrs_host, int kprec, int fallback) {
  if (!autotune_) return fallback;
  std::string key = "M" + std::to_string(d.M) + " N" + std::to_string(d.N) + " p" + std::to_string(kprec) +
                    (d.c_f32 ? " f" : " d") + (d.beta != 0.0 ? " acc" : "") + (d.cdiv > 0 ? " cd" : "") +
                    (d.ndiv > 0 ? " nd" : "") + (d.offm ? " tab" : "") + (d.tag ? " t" : "") + (ptrs_host ? " grouped" : " b" + std::to_string(d.batch)) +
                    (d.ktotal > 0 ? " kt" + std::to_string(d.ktotal) : "") + (d.dma > 1 ? " dma" + std::to_string(d.dma) : "");
  for (int s = 0; s < d.nseg; s++) {
    const GemmSeg &g = d.seg[s];
    key += " [K" + std::to_string(g.K) + " " + std::to_string(g.sam) + "," + std::to_string(g.sak) + "," +
           std::to_string(g.sbk) + "," + std::to_string(g.sbn) + "]";
  }
  key = device_tag() + " " + key;
  {
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    auto it = g_tiles.find(key);
    if (it != g_tiles.end()) return it->second.tile;
  }
  // a shape first seen while the stream is being captured into a graph
  // keeps the fallback tile: timing needs synchronisation
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return fallback;

  const size_t es = d.c_f32 ? 4 : 8;
  const long ext = out_extent(d, out_extent_hint_);
  GemmDesc t = d;
  const GemmPtrs *tab_dev = nullptr;
  std::vector<GemmPtrs> tab;
  size_t bytes;
  if (ptrs_host) {
    // up to 64 entries of the table (capped at 1 GiB of output), outputs
    // laid end to end in scratch: enough workgroups in flight that a tile
    // which streams an operand twice pays for it in HBM traffic, as it does
    // on a full batch, rather than hitting cache
    const long fit = std::max<long>(1, (1L << 30) / std::max<long>(1, ext * (long)es));
    const int nb = (int)std::min<long>(std::min<long>(d.batch, 64), fit);
    t.batch = nb;
    bytes = ((size_t)nb * ext * es + 7) / 8 * 8;  // whole doubles: the fill below clears it as such
    char *out = (char *)slot(4, bytes);
    tab.assign(ptrs_host, ptrs_host + nb);
    for (int z = 0; z < nb; z++) tab[z].C = (double *)(out + (size_t)z * ext * es);
    void *dt = slot(5, tab.size() * sizeof(GemmPtrs));
    hip_check(hipMemcpyAsync(dt, tab.data(), tab.size() * sizeof(GemmPtrs), hipMemcpyHostToDevice, st_),
              "tuning table copy");
    tab_dev = (const GemmPtrs *)dt;
  } else {
    // the whole batch (strided outputs); cap the scratch at 1 GiB by
    // tuning on fewer batch entries
    long nb = d.batch;
    if (nb > 1 && d.sCz > 0) nb = std::max<long>(1, std::min<long>(nb, ((1L << 30) / (long)es - ext) / d.sCz + 1));
    t.batch = (int)nb;
    if (d.ktotal > 0) t.ktotal = std::min<long>(d.ktotal, (long)nb * d.seg[0].K);
    bytes = (((size_t)(nb - 1) * (size_t)d.sCz + (size_t)ext) * es + 7) / 8 * 8;
    t.C = (double *)slot(4, bytes);
  }
  launch_fill((long)(bytes / 8), 0.0, (double *)slot(4, bytes), st_);
  hipEvent_t e0 = event(), e1 = event();
  int best = fallback >= 0 ? fallback : gemm_pick_tile(d);
  double best_ms = 1e300;
  const int reps = 3;
  const bool lib = !ptrs_host && blas_fits(d, kprec);
  if (lib) blas_ready();
  for (int tile = 0; tile < gemm_num_tiles() + (lib ? 1 : 0); tile++) {
    const int id = tile < gemm_num_tiles() ? tile : TILE_LIBRARY;
    if (id == TILE_LIBRARY) blas_gemm(t);  // warm
    else launch_gemm(t, tab_dev, kprec, id, st_);
    hip_check(hipEventRecord(e0, st_), "event record");
    for (int r = 0; r < reps; r++) {
      if (id == TILE_LIBRARY) blas_gemm(t);
      else launch_gemm(t, tab_dev, kprec, id, st_);
    }
    hip_check(hipEventRecord(e1, st_), "event record");
    hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    if (ms / reps < best_ms) { best_ms = ms / reps; best = id; }
  }
  free_events_.push_back(e0);
  free_events_.push_back(e1);
  {
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    g_tiles[key] = Tuned{best, best_ms};
  }
  return best;
}

std::string Engine::tile_table() const {
  std::string s;
  std::lock_guard<std::mutex> lk(g_tiles_mu);
  for (auto &kv : g_tiles)
    s += kv.first + " -> " + tile_name(kv.second.tile) + " " + std::to_string(kv.second.ms) + " ms\n";
  return s;
}

// File form of the tuned table: one "key<TAB>tile name<TAB>ms" line per
// shape (tiles by name, so a table survives changes of the tile ids).
bool tiles_save(const std::string &path) {
  std::string s;
  {
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    for (auto &kv : g_tiles)
      s += kv.first + "\t" + tile_name(kv.second.tile) + "\t" + std::to_string(kv.second.ms) + "\n";
  }
  const std::string tmp = path + ".tmp";
  FILE *f = std::fopen(tmp.c_str(), "w");
  if (!f) return false;
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  if (std::fclose(f) != 0 || !ok) return false;
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}

long tiles_load(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return -1;
  std::map<std::string, int> by_name;
  for (int t = 0; t < gemm_num_tiles(); t++) by_name[gemm_tile_name(t)] = t;
  by_name[tile_name(TILE_LIBRARY)] = TILE_LIBRARY;
  long n = 0;
  std::string line;
  char buf[4096];
  std::lock_guard<std::mutex> lk(g_tiles_mu);
  while (std::fgets(buf, sizeof(buf), f)) {
    line = buf;
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    const size_t t1 = line.find('\t'), t2 = t1 == std::string::npos ? t1 : line.find('\t', t1 + 1);
    if (t2 == std::string::npos) continue;
    auto it = by_name.find(line.substr(t1 + 1, t2 - t1 - 1));
    if (it == by_name.end()) continue;
    g_tiles[line.substr(0, t1)] = Tuned{it->second, std::atof(line.c_str() + t2 + 1)};
    n++;
  }
  std::fclose(f);
  return n;
}

}  // namespace ugx
