// Host-side helpers shared by the translation units of libgmms.so: parameter
// resolution, launch geometry, error text, opt-in stage timing and the dtype
// dispatch. Internal; the public interface is include/gpu_mms/gmms.h.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/gpu_mms/gmms.h"
#include "gmms_basecase.hxx"
#include "gmms_elem.hxx"
#include "gmms_merge.hxx"
#include "gmms_tmerge.hxx"
#include "gmms_wmerge.hxx"

namespace gmms_host {
using namespace gmms;

struct Resolved {
  int K, PL, D, mergers;
  int64_t slack;
  int kernel;  // GMMS_MERGE_* : wavefront tile, pipelined heap, workgroup tile
  int run;     // sorted run length after the base case (1024 or 4096)
  uint32_t flip;  // 0, or all ones for a descending sort
  int W;          // 32-bit words per element
};

enum { GMMS_MERGE_WAVE = 0, GMMS_MERGE_HEAP = 1, GMMS_MERGE_TILE = 2 };

// wavefront-merger window budget (elements): the windows and the transpose
// take 8 KB of LDS per wavefront for 32- and 64-bit keys, 12 KB for pairs
constexpr int wm_win(int W) { return W == 1 ? 1024 : 512; }

constexpr int TM_F = 64;  // tile-merge window per list

constexpr size_t LDS_LIMIT = 160 * 1024;  // gfx950 LDS per workgroup

// resident wavefront mergers per CU for W-word elements: what the LDS holds
// (with the plain launch's progress table), at most WM_PLAIN_MAX = 16 (4
// wavefronts per SIMD: the kernel's register budget), a multiple of 4; the
// plain launch runs them as one workgroup
inline int wm_resident_per_cu(int W) {
  int nw = WM_PLAIN_MAX;
  while (nw > 4 &&
         ((size_t)nw * wmerge_lds_words_per_wave(W, wm_win(W)) + wmerge_peer_words(nw)) * sizeof(uint32_t) > LDS_LIMIT)
    nw -= 4;
  return nw;
}

inline int ilog2(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

inline int cu_count() {
  static int cached = 0;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  if (cached) return cached;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  cached = cus;
  return cached;
}

inline int elem_words(int dtype) {
  switch (dtype) {
    case GMMS_INT32: case GMMS_UINT32: case GMMS_FLOAT32: return 1;
    case GMMS_INT64: case GMMS_UINT64: case GMMS_FLOAT64: return 2;
    case GMMS_PAIR_U64_U32: return 3;
    default: return 0;
  }
}
inline bool valid_dtype(int dtype) { return elem_words(dtype) > 0; }
inline bool has_vals(int dtype) { return dtype == GMMS_PAIR_U64_U32; }

// K = 0 means: the largest power of two <= 64 whose heap + rings fit in LDS
inline int resolve(int dtype, const gmms_params* p, Resolved* r, bool need_device) {
  gmms_params d;
  gmms_get_default_params(&d);
  const gmms_params& q = p ? *p : d;
  int D = q.prefetch ? q.prefetch : d.prefetch;
  if (D < 1 || D > 8) return GMMS_ERROR_INVALID_VALUE;
  const int W = elem_words(dtype);
  if (q.pipeline < 0 || q.pipeline > 2) return GMMS_ERROR_INVALID_VALUE;
  const int kern = q.pipeline;
  auto lds_bytes = [&](int k) -> size_t {
    if (kern == GMMS_MERGE_HEAP) return merge_lds_words(k, W, D) * sizeof(uint32_t);
    if (kern == GMMS_MERGE_TILE) return tmerge_lds_words(k, W, TM_F) * sizeof(uint32_t);
    // wavefront merger: window per list WIN/K must stay >= 16 elements
    return wm_win(W) / k >= 16 ? WM_WAVES * wmerge_lds_words_per_wave(W, wm_win(W)) * sizeof(uint32_t)
                               : LDS_LIMIT + 1;
  };
  int K = q.kway;
  if (K == 0) {
    // wavefront merger and heap: 16 lists (the heap: more resident mergers
    // than with 64); the workgroup tile merger: the largest power of two
    // that fits LDS
    K = kern == GMMS_MERGE_TILE ? MAX_K : 16;
    while (K > 2 && lds_bytes(K) > LDS_LIMIT) K >>= 1;
  }
  if (K < 2 || K > MAX_K || (K & (K - 1))) return GMMS_ERROR_INVALID_VALUE;
  if (lds_bytes(K) > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
  r->kernel = kern;
  if (q.mergers < 0 || q.slack < 0) return GMMS_ERROR_INVALID_VALUE;
  r->K = K;
  r->W = W;
  r->PL = ilog2(K);
  r->D = D;
  // resident mergers per CU: as many as the LDS holds (heap: one workgroup
  // of log2(K) wavefronts per merger, at most 32 wavefronts per CU; workgroup
  // tile 3; wavefront tile: workgroups of WM_WAVES mergers)
  int per_cu;
  if (kern == GMMS_MERGE_HEAP) {
    const int by_lds = (int)std::max<size_t>(1, LDS_LIMIT / lds_bytes(K));
    const int by_waves = std::max(1, 32 / std::max(1, ilog2(K)));
    per_cu = std::min(by_lds, by_waves);
  } else {
    per_cu = kern == GMMS_MERGE_TILE ? 3 : wm_resident_per_cu(W);
  }
  r->mergers = q.mergers ? q.mergers : per_cu * (need_device ? cu_count() : 256);
  r->slack = q.slack;
  if (q.base_run != 0 && q.base_run != TILE && q.base_run != TILE * BASE_WAVES) return GMMS_ERROR_INVALID_VALUE;
  r->run = q.base_run ? q.base_run : TILE * BASE_WAVES;
  if (q.descending != 0 && q.descending != 1) return GMMS_ERROR_INVALID_VALUE;
  r->flip = q.descending ? 0xFFFFFFFFu : 0u;
  return GMMS_SUCCESS;
}

// mergers per task for a round
inline int mergers_per_task(int64_t n, int64_t L, int K, int target) {
  const int64_t TK = (int64_t)K * L;
  const int64_t tasks = (n + TK - 1) / TK;
  if (tasks >= target) return 1;
  // target = the mergers resident at once. A merger past a multiple of it
  // waits for a free slot and then runs a whole merger's time at the end of
  // the round, so fill whole generations: the smallest m with
  // tasks*m = k*target (k = 1..4) to within 10 %, else ceil(target/tasks).
  int64_t m = (target + tasks - 1) / tasks;
  for (int k = 1; k <= 4; ++k) {
    const int64_t mk = (int64_t)k * target / tasks;
    if (mk < 1) continue;
    const int64_t used = tasks * mk, slots = (used + target - 1) / target * target;
    if (used * 10 >= slots * 9) {
      m = mk;
      break;
    }
  }
  // keep every merger at least a few heap-builds worth of output
  const int64_t minsz = (int64_t)WAVE * K * 4;
  const int64_t tsz = std::min<int64_t>(n, TK);
  m = std::min<int64_t>(m, std::max<int64_t>(1, tsz / minsz));
  return (int)std::max<int64_t>(1, m);
}

inline int64_t slack_for(const Resolved& r, int64_t n, int64_t L, int m) {
  if (r.slack) return r.slack;
  const int64_t TK = (int64_t)r.K * L;
  const int64_t tsz = std::min<int64_t>(n, TK);
  return std::max<int64_t>(WAVE, tsz / m / 32);
}

// Workspace of a round with more than one merger per task: the sample array
// of its input (every SAMP-th element: keys, then values for pairs; 256-byte
// aligned), then the cut rows and skips.
inline size_t samp_bytes(int64_t n, int W) {
  const size_t cnt = (size_t)((n + SAMP - 1) >> SAMP_SHIFT);
  const size_t b = cnt * (W == 1 ? 4 : 8) + (W == 3 ? cnt * 4 : 0);
  return (b + 255) & ~(size_t)255;
}
inline size_t cut_bytes(int64_t n, int64_t L, const Resolved& r) {
  const int64_t TK = (int64_t)r.K * L;
  const int64_t tasks = (n + TK - 1) / TK;
  const int m = mergers_per_task(n, L, r.K, r.mergers);
  return m > 1 ? (size_t)(tasks * m) * (size_t)(r.K + 1) * sizeof(int64_t) : 0;
}
inline size_t ws_bytes_for(int64_t n, const Resolved& r) {
  size_t need = 0;
  for (int64_t L = r.run; L < n; L *= r.K) need = std::max(need, cut_bytes(n, L, r));
  return need ? samp_bytes(n, r.W) + need : 0;
}

extern thread_local char g_last_error[256];

inline int hip_fail(hipError_t e, const char* what) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  return GMMS_ERROR_HIP;
}

#define HIP_OK(x)                                    \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #x);   \
  } while (0)

// Opt-in per-launch timing: HIP events recorded on the launch stream around
// every kernel, read back by gmms_stage_times().
struct StageRec {
  hipEvent_t a, b;
  int kind;
  int64_t elems;
};
struct StageTimer {
  std::mutex mu;
  bool on = false;
  std::vector<StageRec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
};
extern StageTimer g_timer;

struct StageScope {
  hipEvent_t b = nullptr;
  hipStream_t s;
  StageScope(int kind, int64_t elems, hipStream_t st) : s(st) {
    if (!g_timer.on) return;
    std::lock_guard<std::mutex> g(g_timer.mu);
    hipEvent_t a = g_timer.get();
    b = g_timer.get();
    if (!a || !b) {
      b = nullptr;
      return;
    }
    (void)hipEventRecord(a, s);
    g_timer.recs.push_back({a, b, kind, elems});
  }
  ~StageScope() {
    if (b) (void)hipEventRecord(b, s);
  }
};

template <class F>
inline int dispatch(int dtype, F&& f) {
  switch (dtype) {
    case GMMS_INT32: return f(std::integral_constant<int, DT_I32>{});
    case GMMS_UINT32: return f(std::integral_constant<int, DT_U32>{});
    case GMMS_FLOAT32: return f(std::integral_constant<int, DT_F32>{});
    case GMMS_INT64: return f(std::integral_constant<int, DT_I64>{});
    case GMMS_UINT64: return f(std::integral_constant<int, DT_U64>{});
    case GMMS_FLOAT64: return f(std::integral_constant<int, DT_F64>{});
    case GMMS_PAIR_U64_U32: return f(std::integral_constant<int, DT_PAIR>{});
    default: return GMMS_ERROR_INVALID_VALUE;
  }
}

// Batched-sort internals used by the server (gmms_serve.hip): the plan's
// host table block, where it goes in the workspace, and the launches alone
// (the caller has copied the table block to ws + *ws_offset on the stream).
const void* seg_plan_table(const gmms_seg_plan* plan, size_t* bytes, size_t* ws_offset);
size_t seg_plan_workspace(const gmms_seg_plan* plan);
// The workspace bytes of the plan of these segments (default parameters),
// from its counts alone: no tables are built.
int seg_plan_workspace_for(int dtype, const int64_t* offsets, int64_t n_seg, size_t* bytes);
int seg_plan_launch(const gmms_seg_plan* plan, void* keys, void* keys_alt, void* vals, void* vals_alt, void* ws,
                    hipStream_t s, int* alt);

}  // namespace gmms_host
