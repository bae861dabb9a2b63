// Contraction engine: einsum-style tensor contractions executed as
// (permute) -> MFMA GEMM -> (permute-accumulate) on one HIP stream.
//
// A contraction "mnfe,nfjb->mejb" names the index letters of A, B and C (all
// row-major).  Letters in A and B but not C are summed (K); letters of A in C
// form the GEMM rows (M), letters of B in C the columns (N).  The planner
// tries the index orders that can make each operand a plain strided matrix
// (M and N in C's or the operand's order, K in A's or B's order, and the
// A/B role swap) and picks the plan with the least permutation traffic.
// Permuted copies of constant operands (the integral blocks) are cached, so
// after the first iteration they cost nothing; FP32 copies for the mixed
// forms are cached the same way.  C is always FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <cstdlib>
#include <vector>

#include "ugx_internal.h"

struct _rocblas_handle;  // rocBLAS handle (rocblas_handle)

namespace ugx {

struct TRef {
  const void *p = nullptr;
  std::vector<int> dims;
  bool constant = false;  // contents never change while cached copies live
  int dtype = 0;          // 0 double, 1 float
};

inline TRef cref(const double *p, std::vector<int> d) { TRef t; t.p = p; t.dims = std::move(d); t.constant = true; return t; }
inline TRef vref(const double *p, std::vector<int> d) { TRef t; t.p = p; t.dims = std::move(d); return t; }

struct KernelStat {
  long launches = 0;
  double ms = 0.0;
  double flops = 0.0;
};

// Tile id of a GEMM the autotuner hands to rocBLAS (plain FP64 GEMMs only,
// when its timing beats every tile of the library's kernel)
constexpr int TILE_LIBRARY = 1000;

class Engine {
 public:
  // UGX_CCSD_PANELS=0 (read here, per engine): BF16x3 operands as hi/lo
  // pair arrays only, never as MFMA panels
  explicit Engine(hipStream_t st) : st_(st) {
    const char *e = std::getenv("UGX_CCSD_PANELS");
    panels_ = !(e && e[0] == '0' && e[1] == 0);
  }
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  // C = alpha * contract(A, B) + beta * C at the given precision.
  void contract(const std::string &spec, double alpha, const TRef &A, const TRef &B, double beta, const TRef &C,
                int prec, const char *label = nullptr);
  // C = alpha * permute(A) + beta * C, spec "mbje->mejb".
  void permute(const std::string &spec, double alpha, const TRef &A, double beta, const TRef &C);

  // Runs fn (which enqueues device work on the stream) between a pair of
  // events recorded under label when profiling is on; flops are credited to
  // the label either way.
  void timed(const char *label, double flops, const std::function<void()> &fn);
  // Per-label device time (HIP events around each contraction) when on.
  void set_profiling(bool on) { profiling_ = on; }
  void collect_stats();  // waits for the stream, folds event pairs into stats_
  std::map<std::string, KernelStat> &stats() { return stats_; }
  // Adds another engine's collected per-label statistics to this one's.
  void merge_stats(const std::map<std::string, KernelStat> &other) {
    for (auto &kv : other) {
      KernelStat &k = stats_[kv.first];
      k.launches += kv.second.launches;
      k.ms += kv.second.ms;
      k.flops += kv.second.flops;
    }
  }
  double flops_total() const { return flops_; }
  void add_flops(double f) { flops_ += f; }
  bool profiling() const { return profiling_; }
  void reset_counters() { flops_ = 0.0; stats_.clear(); }
  // Frees the cached conversions and every workspace (retired ones too).
  void drop_cache();

  // Sizing mode: contract() plans each contraction, allocates its
  // workspaces, converts and caches its constant operands and tunes its
  // tiles (when autotuning is on), but launches no GEMM, no variable
  // operand conversion and no output permute; permute() and refresh() do
  // nothing.  A solve's contractions evaluated once in this mode leave the
  // engine holding everything the same contractions need, so the solve's
  // iterations allocate no device memory.
  void set_sizing(bool on) { sizing_ = on; }
  bool sizing() const { return sizing_; }
  // Device allocations made outside sizing mode since the last call
  // (workspace growth and new cached conversions).
  long take_allocations() { long n = allocs_; allocs_ = 0; return n; }
  // Incremented whenever a workspace is replaced by a larger one.  A
  // replaced workspace is retired, not freed, until drop_cache(), so work
  // captured into a graph before the change still addresses valid memory;
  // callers that cache graphs compare generations to re-capture against the
  // current workspaces.
  unsigned generation() const { return gen_; }
  // Re-converts, in place, every cached conversion of the constant at p
  // with element type in [dmin, dmax] (whose contents the caller has just
  // replaced): buffers keep their addresses, so captured graphs that read
  // them stay valid.  A caller that replaces p's contents must refresh every
  // type it will read before reading it.
  void refresh(const void *p, int dmin = 0, int dmax = 3);
  size_t cached_bytes() const;
  hipStream_t stream() const { return st_; }

  // Tile of a GEMM launch.  Autotuning off: `fallback` (-1 = the size
  // heuristic).  On: the first launch of each shape times every tile into a
  // scratch output (the real operands, the same batch addressing) and the
  // fastest is kept for the context's lifetime.  ptrs_host (grouped
  // launches) is the host copy of the launch's pointer table; tuning then
  // uses its first entries with their outputs redirected to scratch.
  // (outputs addressed through offset tables: `extent` is one entry's
  // largest offset plus one)
  int tile_for(const GemmDesc &d, const GemmPtrs *ptrs_host, int kprec, int fallback, long extent = 0) {
    out_extent_hint_ = extent;
    return tile_for_impl(d, ptrs_host, kprec, fallback);
  }
  void set_autotune(bool on) { autotune_ = on; }
  bool autotune() const { return autotune_; }
  // "key tile ms" lines of the tuned shapes (process-wide: contexts on one
  // device share the tuned tiles)
  std::string tile_table() const;

 private:
  struct Plan;
  void *slot(int i, size_t bytes);
  // dtype 3: BF16 hi/lo MFMA panels of the [rows][pk] view (pk = the k extent)
  const void *prepared(const TRef &T, const std::string &letters, const std::string &want, int dtype,
                       int slot_id, long pk = 0);
  hipStream_t st_;
  bool panels_ = true;
  bool sizing_ = false;
  long allocs_ = 0;
  unsigned gen_ = 0;
  std::vector<std::pair<void *, size_t>> slots_{6, {nullptr, 0}};
  std::vector<void *> retired_;
  bool autotune_ = false;
  long out_extent_hint_ = 0;
  int tile_for_impl(const GemmDesc &d, const GemmPtrs *ptrs_host, int kprec, int fallback);
  // rocBLAS for large single FP64 GEMMs with unit-stride operands: one
  // handle on this engine's stream (created at the first tuning or sizing
  // use, with its workspace, so no allocation happens in a captured or
  // timed region), atomics off (bitwise repeatable)
  ::_rocblas_handle *blas_ = nullptr;
  void *blas_ws_ = nullptr;
  bool blas_fits(const GemmDesc &d, int kprec) const;
  void blas_ready();
  void blas_gemm(const GemmDesc &d);
  void run_gemm(const GemmDesc &d, int kprec, int tile);
  bool plain_to_library(const GemmDesc &d, int kprec, long Re, long Ce);
  std::map<std::tuple<const void *, std::string, int>, std::pair<void *, size_t>> cache_;
  std::map<std::tuple<const void *, std::string, int>, PermDesc> cache_pd_;  // how each entry was filled
  bool profiling_ = false;
  struct Pending {
    std::string label;
    hipEvent_t a, b;
    double flops;
  };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> free_events_;
  std::map<std::string, KernelStat> stats_;
  double flops_ = 0.0;
  hipEvent_t event();
};

void hip_check(hipError_t e, const char *what);
// The process-wide tuned tile table to / from a file (returns false / -1
// when the file cannot be written / read; load returns the entries merged).
bool tiles_save(const std::string &path);
long tiles_load(const std::string &path);

}  // namespace ugx
