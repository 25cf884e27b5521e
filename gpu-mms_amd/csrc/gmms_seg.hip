// Batched segmented sort (declarations: include/gpu_mms/gmms.h).
//
// Every segment is sorted by the schedule of the plain sort
// (src/multimergesort.hxx:56-141: 1024-element tiles, then rounds with lists
// of 1024*K^r, tasks of K lists, mergers splitting a task at consistent cuts),
// but all segments share the launches:
//   k_seg_setup     expands the segment table into tile descriptors and
//                   clears the per-round work-queue counters,
//   k_basecase_seg  sorts every tile of every segment,
//   per round g     k_partition over the round's partition items, then
//                   k_wmerge_persistent over its merger items.
// A segment with R_s rounds takes part in the last R_s of the R global
// rounds, so every segment finishes in round R-1 and all results land in the
// same buffer (R & 1); its tiles are written to buffer (R - R_s) & 1, where
// its first round reads them. Merger items of a round are ordered largest
// first, so the persistent launch schedules longest-processing-time first.
#include <memory>

#include "gmms_basecase.hxx"
#include "gmms_host.hxx"

using namespace gmms;
using namespace gmms_host;

namespace {

// Device-side segment descriptor (32 bytes).
struct SegDesc {
  int64_t off;    // first element
  int64_t n;      // elements (> 0)
  int64_t tile0;  // index of its first run descriptor
  int32_t flags;  // TileDesc flags of all its runs
  int32_t run;    // run length: TILE or 4 * TILE
};
static_assert(sizeof(MergerInfo) == 48, "MergerInfo layout");
static_assert(sizeof(TileDesc) == 16, "TileDesc layout");

constexpr int SEG_SETUP_WAVES = 4;

// One wavefront per segment writes the segment's run descriptors (runs of
// sd.run elements); wavefront 0 also clears the round counters.
__global__ __launch_bounds__(SEG_SETUP_WAVES* WAVE) void k_seg_setup(const SegDesc* segs, int64_t n_segs,
                                                                      TileDesc* tiles, unsigned int* queue,
                                                                      int n_queue) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t w = (int64_t)blockIdx.x * SEG_SETUP_WAVES + threadIdx.x / WAVE;
  if (w == 0) {
    for (int i = lane; i < n_queue; i += WAVE) queue[i] = 0u;
  }
  if (w >= n_segs) return;
  const SegDesc sd = segs[w];
  const int64_t nt = (sd.n + sd.run - 1) / sd.run;
  for (int64_t t = lane; t < nt; t += WAVE) {
    TileDesc td;
    td.off = sd.off + t * sd.run;
    const int64_t rest = sd.n - t * sd.run;
    td.cnt = (int32_t)(rest < sd.run ? rest : sd.run);
    td.flags = sd.flags;
    tiles[sd.tile0 + t] = td;
  }
}

// ------------------------------------------------------------ tuning table
struct Tuning {
  int32_t K[GMMS_SEG_CLASSES];
  int64_t merger[GMMS_SEG_CLASSES];
};
std::mutex g_tune_mu;
Tuning g_tune[7];  // per dtype; zero = defaults

int size_class(int64_t n) {
  int c = 0;
  while (c < GMMS_SEG_CLASSES - 1 && ((int64_t)1 << c) < n) ++c;
  return c;
}

// largest K the wavefront merger runs for W-word elements (window >= 16 per list)
int max_seg_k(int W) { return std::min(MAX_K, wm_win(W) / 16); }

constexpr int SEG_DEFAULT_K = 16;

struct SegRound {
  int64_t items_off, n_items;  // merger items (in MergerInfo units from the table start)
  int64_t part_off, n_part;    // partition items
  int64_t elems;               // elements merged
  int64_t rows;                // cut rows used
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct gmms_seg_plan {
  int dtype = 0;
  int W = 1;
  int R = 0;
  int slots = 0;  // resident mergers (for the persistent grid)
  uint32_t flip = 0;  // descending: XOR of the internal bits
  int64_t n_seg = 0, elements = 0, extent = 0, tiles = 0;
  int64_t runs4 = 0, runs1 = 0;  // run descriptors: 4096-element (first), 1024-element
  std::vector<int32_t> seg_K, seg_R;  // per segment (all, empties included)
  std::vector<SegRound> rounds;
  int64_t n_segs_dev = 0;  // non-empty segments
  int64_t n_items_all = 0;
  int64_t max_rows = 0;
  // tables: [SegDesc x n_segs_dev][MergerInfo x n_items_all], copied as one
  // block; `host` is their pinned copy, made by the first sort
  std::vector<char> table;
  void* host = nullptr;
  size_t host_bytes = 0;
  size_t items_byte_off = 0;  // offset of the first MergerInfo in the block
  // workspace layout
  size_t off_queue = 0, off_table = 0, off_tiles = 0, off_cuts = 0, ws_bytes = 0;
  ~gmms_seg_plan() {
    if (host) (void)hipHostFree(host);
  }
};

struct gmms_seg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int alt = 0;
  ~gmms_seg_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

// sizing_only: counts, offsets and the workspace size, without the tables.
int plan_build(gmms_seg_plan* P, int dtype, const int64_t* offsets, int64_t n_seg, const gmms_params* p,
               bool sizing_only = false) {
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  if (r.kernel != GMMS_MERGE_WAVE) return GMMS_ERROR_UNSUPPORTED;
  const int W = elem_words(dtype);
  P->dtype = dtype;
  P->W = W;
  P->n_seg = n_seg;
  P->slots = (p && p->mergers > 0) ? p->mergers : wm_resident_per_cu(W) * cu_count();
  P->flip = r.flip;
  const int kmax = max_seg_k(W);
  const bool forced = p && p->kway > 0;
  Tuning tune;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune = g_tune[dtype];
  }
  // per-segment K and rounds
  P->seg_K.assign(n_seg, 0);
  P->seg_R.assign(n_seg, 0);
  std::vector<int64_t> seg_merger(n_seg, 0);
  int R = 0;
  int64_t tiles = 0, elements = 0, nonempty = 0;
  for (int64_t s = 0; s < n_seg; ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    const int c = size_class(n);
    int K = forced ? r.K : (tune.K[c] ? tune.K[c] : SEG_DEFAULT_K);
    if (K > kmax) K = kmax;
    // segments longer than one tile start from 4096-element runs
    const int64_t run = n > TILE ? (int64_t)TILE * BASE_WAVES : TILE;
    int rs = 0;
    for (int64_t L = run; L < n; L *= K) ++rs;
    P->seg_K[s] = K;
    P->seg_R[s] = rs;
    seg_merger[s] = tune.merger[c];
    R = std::max(R, rs);
    tiles += (n + TILE - 1) / TILE;
    elements += n;
    nonempty += n > 0;
  }
  P->R = R;
  P->tiles = tiles;
  P->elements = elements;
  P->extent = n_seg ? offsets[n_seg] : 0;
  P->n_segs_dev = nonempty;

  // segment table: the 4096-element run descriptors of all long segments
  // come first, then the tile descriptors of the short ones
  std::vector<SegDesc> segs;
  if (!sizing_only) segs.reserve(nonempty);
  int64_t r4 = 0, r1 = 0;
  for (int64_t s = 0; s < n_seg; ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    if (n > TILE) r4 += (n + TILE * BASE_WAVES - 1) / (TILE * BASE_WAVES);
    else if (n > 0) r1 += 1;
  }
  P->runs4 = r4;
  P->runs1 = r1;
  int64_t a4 = 0, a1 = r4;
  for (int64_t s = 0; s < (sizing_only ? 0 : n_seg); ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    if (n == 0) continue;
    SegDesc d{};
    d.off = offsets[s];
    d.n = n;
    const int rs = P->seg_R[s];
    d.flags = ((R - rs) & 1) | (rs == 0 ? 2 : 0);
    if (n > TILE) {
      d.run = TILE * BASE_WAVES;
      d.tile0 = a4;
      a4 += (n + d.run - 1) / d.run;
    } else {
      d.run = TILE;
      d.tile0 = a1++;
    }
    segs.push_back(d);
  }

  // rounds: round volume first (for the automatic merger size)
  std::vector<int64_t> vol(R, 0);
  for (int64_t s = 0; s < n_seg; ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    for (int g = R - P->seg_R[s]; g < R; ++g) vol[g] += n;
  }
  // Every task of every segment's rounds, in (segment, round, task) order:
  // emit(g, s, tbase, T, L, K, m) with the task's merger count m.
  auto walk = [&](auto&& emit) {
    for (int64_t s = 0; s < n_seg; ++s) {
      const int64_t n = offsets[s + 1] - offsets[s];
      const int rs = P->seg_R[s];
      const int K = P->seg_K[s];
      int64_t L = n > TILE ? (int64_t)TILE * BASE_WAVES : TILE;
      for (int rr = 0; rr < rs; ++rr, L *= K) {
        const int g = R - rs + rr;
        const int64_t TK = (int64_t)K * L;
        const int64_t tasks = (n + TK - 1) / TK;
        const int64_t minsz = (int64_t)WAVE * K * 4;
        const int64_t target =
            seg_merger[s] > 0 ? seg_merger[s] : std::max<int64_t>(minsz, (vol[g] + P->slots - 1) / P->slots);
        for (int64_t t = 0; t < tasks; ++t) {
          const int64_t tb = t * TK;
          const int64_t T = std::min(TK, n - tb);
          int64_t m = (T + target - 1) / target;
          m = std::min<int64_t>(m, std::max<int64_t>(1, T / minsz));
          emit(g, s, tb, T, L, K, std::max<int64_t>(1, m));
        }
      }
    }
  };
  // pass 1: item and cut-row counts per round (all that sizing needs)
  std::vector<int64_t> n_items(R, 0), n_part(R, 0), rows(R, 0);
  walk([&](int g, int64_t, int64_t, int64_t, int64_t, int, int64_t m) {
    n_items[g] += m;
    n_part[g] += m - 1;
    if (m > 1) rows[g] += m;
  });
  P->rounds.assign(R, SegRound{});
  int64_t off = 0;
  for (int g = 0; g < R; ++g) {
    SegRound& sr = P->rounds[g];
    sr.items_off = off;
    sr.n_items = n_items[g];
    off += sr.n_items;
    sr.part_off = off;
    sr.n_part = n_part[g];
    off += sr.n_part;
    sr.elems = vol[g];
    sr.rows = rows[g];
    P->max_rows = std::max(P->max_rows, rows[g]);
  }
  P->n_items_all = off;
  P->items_byte_off = align256((size_t)nonempty * sizeof(SegDesc));
  P->host_bytes = P->items_byte_off + (size_t)off * sizeof(MergerInfo);

  // pass 2: the staging block (host; copied to pinned memory on the first
  // sort): segment table, then per round its merger items and partition items
  if (!sizing_only) {
    P->table.assign(P->host_bytes, 0);
    char* h = P->table.data();
    if (!segs.empty()) std::memcpy(h, segs.data(), segs.size() * sizeof(SegDesc));
    MergerInfo* it = (MergerInfo*)(h + P->items_byte_off);
    std::vector<int64_t> mi_at(R), pi_at(R), row_at(R, 0);
    for (int g = 0; g < R; ++g) {
      mi_at[g] = P->rounds[g].items_off;
      pi_at[g] = P->rounds[g].part_off;
    }
    walk([&](int g, int64_t s, int64_t tb, int64_t T, int64_t L, int K, int64_t m) {
      MergerInfo mi{};
      mi.tbase = offsets[s] + tb;
      mi.T = T;
      mi.L = L;
      mi.cut_row = m > 1 ? row_at[g] : 0;
      mi.m = (int32_t)m;
      mi.K = K;
      if (m > 1) row_at[g] += m;
      for (int64_t j = 0; j < m; ++j) {
        mi.j = (int32_t)j;
        it[mi_at[g]++] = mi;
        if (j > 0) it[pi_at[g]++] = mi;
      }
    });
    // LPT order inside each round: larger mergers first (stable for equal
    // sizes; a round of one segment's equal tasks is in that order already)
    const auto larger = [](const MergerInfo& x, const MergerInfo& y) { return x.T / x.m > y.T / y.m; };
    for (int g = 0; g < R; ++g) {
      MergerInfo* b = it + P->rounds[g].items_off;
      MergerInfo* e = b + P->rounds[g].n_items;
      if (!std::is_sorted(b, e, larger)) std::stable_sort(b, e, larger);
    }
  }
  // workspace: queue counters | table block | tiles | cuts
  P->off_queue = 0;
  P->off_table = align256(sizeof(unsigned int) * (size_t)std::max(R, 1));
  P->off_tiles = P->off_table + align256(P->host_bytes);
  P->off_cuts = P->off_tiles + align256((size_t)(P->runs4 + P->runs1) * sizeof(TileDesc));
  P->ws_bytes = P->off_cuts + align256((size_t)P->max_rows * MAX_K * sizeof(int64_t));
  return GMMS_SUCCESS;
}

int plan_launches(const gmms_seg_plan* P) {
  if (P->n_segs_dev == 0) return 0;
  int n = 1 + (P->runs4 > 0) + (P->runs1 > 0);
  for (const SegRound& sr : P->rounds) n += 1 + (sr.n_part > 0);
  return n;
}

// Opt the persistent merge kernel into the full LDS (once per process;
// called before any stream capture, since it is not a stream operation).
template <int DT, bool CONV>
void persistent_attr() {
  using E = typename DType<DT>::E;
  constexpr int WIN = wm_win(E::W);
  static std::once_flag once;
  std::call_once(once, [] {
    (void)hipFuncSetAttribute((const void*)k_wmerge_persistent<DT, CONV, WIN>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_LIMIT);
  });
}

template <int DT, bool CONV>
int launch_persistent(RoundArgs& a, int64_t grid, hipStream_t s) {
  using E = typename DType<DT>::E;
  constexpr int WIN = wm_win(E::W);
  const size_t lds = WM_WAVES * wmerge_lds_words_per_wave(E::W, WIN) * sizeof(uint32_t);
  if (lds > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
  persistent_attr<DT, CONV>();
  hipLaunchKernelGGL((k_wmerge_persistent<DT, CONV, WIN>), dim3((unsigned)grid), dim3(WM_WAVES * WAVE), lds, s, a);
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

// The launches of one batched sort (tables already in the workspace).
template <int DT>
int seg_enqueue(const gmms_seg_plan* P, void* k0, void* k1, void* v0, void* v1, char* ws, hipStream_t s,
                bool timed) {
  if (P->n_segs_dev == 0) return GMMS_SUCCESS;
  unsigned int* queue = (unsigned int*)(ws + P->off_queue);
  const SegDesc* segs = (const SegDesc*)(ws + P->off_table);
  const MergerInfo* items = (const MergerInfo*)(ws + P->off_table + P->items_byte_off);
  TileDesc* tiles = (TileDesc*)(ws + P->off_tiles);
  int64_t* cuts = (int64_t*)(ws + P->off_cuts);
  {
    const int64_t grid = (P->n_segs_dev + SEG_SETUP_WAVES - 1) / SEG_SETUP_WAVES;
    hipLaunchKernelGGL(k_seg_setup, dim3((unsigned)grid), dim3(SEG_SETUP_WAVES * WAVE), 0, s, segs, P->n_segs_dev,
                       tiles, queue, std::max(P->R, 1));
    HIP_OK(hipGetLastError());
  }
  {
    std::unique_ptr<StageScope> sc(timed ? new StageScope(GMMS_STAGE_BASECASE, P->elements, s) : nullptr);
    if (P->runs4 > 0) {
      if constexpr (DType<DT>::E::W == 1) {
        // 32-bit keys: one wavefront per run descriptor
        hipLaunchKernelGGL((k_basecase_run_seg<DT>), dim3((unsigned)((P->runs4 + BASE_WAVES - 1) / BASE_WAVES)),
                           dim3(BASE_WAVES * WAVE), 0, s, k0, k1, (const TileDesc*)tiles, P->runs4, P->flip);
      } else {
        hipLaunchKernelGGL((k_basecase_seg<DT, true>), dim3((unsigned)P->runs4), dim3(BASE_WAVES * WAVE), 0, s, k0,
                           (uint32_t*)v0, k1, (uint32_t*)v1, (const TileDesc*)tiles, P->runs4, P->flip);
      }
      HIP_OK(hipGetLastError());
    }
    if (P->runs1 > 0) {
      const int64_t grid = (P->runs1 + BASE_WAVES - 1) / BASE_WAVES;
      hipLaunchKernelGGL((k_basecase_seg<DT, false>), dim3((unsigned)grid), dim3(BASE_WAVES * WAVE), 0, s, k0,
                         (uint32_t*)v0, k1, (uint32_t*)v1, (const TileDesc*)tiles + P->runs4, P->runs1, P->flip);
      HIP_OK(hipGetLastError());
    }
  }
  void* kb[2] = {k0, k1};
  void* vb[2] = {v0, v1};
  const int per_cu = wm_resident_per_cu(P->W);
  for (int g = 0; g < P->R; ++g) {
    const SegRound& sr = P->rounds[g];
    RoundArgs a{};
    a.in_keys = kb[g & 1];
    a.in_vals = (const uint32_t*)vb[g & 1];
    a.out_keys = kb[(g + 1) & 1];
    a.out_vals = (uint32_t*)vb[(g + 1) & 1];
    a.n = P->extent;
    a.cuts = cuts;
    a.skips = nullptr;
    a.cut_stride = MAX_K;
    a.flip = P->flip;
    a.slack = 0;  // per task
    a.queue = queue + g;
    if (sr.n_part > 0) {
      std::unique_ptr<StageScope> sc(timed ? new StageScope(GMMS_STAGE_PARTITION, sr.elems, s) : nullptr);
      a.items = items + sr.part_off;
      a.n_items = sr.n_part;
      const int64_t grid = (sr.n_part + PART_WAVES - 1) / PART_WAVES;
      hipLaunchKernelGGL((k_partition<DT>), dim3((unsigned)grid), dim3(PART_WAVES * WAVE), 0, s, a);
      HIP_OK(hipGetLastError());
    }
    std::unique_ptr<StageScope> sc(timed ? new StageScope(GMMS_STAGE_MERGE, sr.elems, s) : nullptr);
    a.items = items + sr.items_off;
    a.n_items = sr.n_items;
    const int64_t want = (sr.n_items + WM_WAVES - 1) / WM_WAVES;
    const int64_t cap = (int64_t)std::max(1, per_cu / WM_WAVES) * cu_count();
    const int64_t grid = std::max<int64_t>(1, std::min(want, cap));
    const int st = g == P->R - 1 ? launch_persistent<DT, true>(a, grid, s) : launch_persistent<DT, false>(a, grid, s);
    if (st) return st;
  }
  return GMMS_SUCCESS;
}

int check_buffers(const gmms_seg_plan* P, void* keys, void* keys_alt, void* vals, void* vals_alt) {
  if (P->elements == 0) return GMMS_SUCCESS;
  if (!keys || !keys_alt) return GMMS_ERROR_INVALID_VALUE;
  if (has_vals(P->dtype) != (vals != nullptr) || has_vals(P->dtype) != (vals_alt != nullptr))
    return GMMS_ERROR_INVALID_VALUE;
  return GMMS_SUCCESS;
}

int pin_tables(gmms_seg_plan* P) {
  if (P->host || !P->host_bytes) return GMMS_SUCCESS;
  const hipError_t e = hipHostMalloc(&P->host, P->host_bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    P->host = nullptr;
    hip_fail(e, "hipHostMalloc(segment tables)");
    return GMMS_ERROR_OUT_OF_MEMORY;
  }
  std::memcpy(P->host, P->table.data(), P->host_bytes);
  return GMMS_SUCCESS;
}

int upload(gmms_seg_plan* P, char* ws, hipStream_t s) {
  if (!P->host_bytes) return GMMS_SUCCESS;
  const int st = pin_tables(P);
  if (st) return st;
  HIP_OK(hipMemcpyAsync(ws + P->off_table, P->host, P->host_bytes, hipMemcpyHostToDevice, s));
  return GMMS_SUCCESS;
}

// splitmix64 finaliser: random keys for the autotuner's batches
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_fill_random(uint32_t* words, int64_t n_words, uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += stride)
    words[i] = (uint32_t)(mix64(seed + (uint64_t)i) >> 16);
}

}  // namespace

namespace gmms_host {

const void* seg_plan_table(const gmms_seg_plan* P, size_t* bytes, size_t* ws_offset) {
  *bytes = P->host_bytes;
  *ws_offset = P->off_table;
  return P->table.data();
}

size_t seg_plan_workspace(const gmms_seg_plan* P) { return P->ws_bytes; }

int seg_plan_workspace_for(int dtype, const int64_t* offsets, int64_t n_seg, size_t* bytes) {
  gmms_seg_plan P;
  const int st = plan_build(&P, dtype, offsets, n_seg, nullptr, true);
  *bytes = st ? 0 : P.ws_bytes;
  return st;
}

int seg_plan_launch(const gmms_seg_plan* P, void* keys, void* keys_alt, void* vals, void* vals_alt, void* ws,
                    hipStream_t s, int* alt) {
  if (alt) *alt = P->R & 1;
  if (P->n_segs_dev == 0) return GMMS_SUCCESS;
  return dispatch(P->dtype, [&](auto dt) {
    return seg_enqueue<decltype(dt)::value>(P, keys, keys_alt, vals, vals_alt, (char*)ws, s, false);
  });
}

}  // namespace gmms_host

extern "C" {

int gmms_seg_plan_create(int dtype, const int64_t* offsets, int64_t n_seg, const gmms_params* p,
                         gmms_seg_plan** plan) {
  if (!plan) return GMMS_ERROR_INVALID_VALUE;
  *plan = nullptr;
  if (!valid_dtype(dtype) || n_seg < 0 || (!offsets)) return GMMS_ERROR_INVALID_VALUE;
  if (offsets[0] < 0) return GMMS_ERROR_INVALID_VALUE;
  for (int64_t s = 0; s < n_seg; ++s)
    if (offsets[s + 1] < offsets[s]) return GMMS_ERROR_INVALID_VALUE;
  gmms_seg_plan* P = new (std::nothrow) gmms_seg_plan();
  if (!P) return GMMS_ERROR_OUT_OF_MEMORY;
  const int st = plan_build(P, dtype, offsets, n_seg, p);
  if (st) {
    delete P;
    return st;
  }
  *plan = P;
  return GMMS_SUCCESS;
}

int gmms_seg_plan_info(const gmms_seg_plan* P, gmms_seg_info* info) {
  if (!P || !info) return GMMS_ERROR_INVALID_VALUE;
  std::memset(info, 0, sizeof(*info));
  info->n_seg = P->n_seg;
  info->elements = P->elements;
  info->extent = P->extent;
  info->tiles = P->tiles;
  for (const SegRound& sr : P->rounds) {
    info->merge_items += sr.n_items;
    info->partition_items += sr.n_part;
  }
  info->workspace_bytes = P->ws_bytes;
  info->rounds = P->R;
  info->launches = plan_launches(P);
  return GMMS_SUCCESS;
}

int gmms_seg_plan_segment(const gmms_seg_plan* P, int64_t seg, int* kway, int* rounds) {
  if (!P || seg < 0 || seg >= P->n_seg) return GMMS_ERROR_INVALID_VALUE;
  if (kway) *kway = P->seg_K[seg];
  if (rounds) *rounds = P->seg_R[seg];
  return GMMS_SUCCESS;
}

void gmms_seg_plan_destroy(gmms_seg_plan* P) { delete P; }

int gmms_seg_sort(gmms_seg_plan* P, void* keys, void* keys_alt, void* vals, void* vals_alt, void* ws,
                  size_t wsb, void* stream, int* alt) {
  if (!P) return GMMS_ERROR_INVALID_VALUE;
  if (alt) *alt = P->R & 1;
  int st = check_buffers(P, keys, keys_alt, vals, vals_alt);
  if (st) return st;
  if (P->n_segs_dev == 0) return GMMS_SUCCESS;
  if (!ws || wsb < P->ws_bytes) return GMMS_ERROR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  st = upload(P, (char*)ws, s);
  if (st) return st;
  return dispatch(P->dtype, [&](auto dt) {
    return seg_enqueue<decltype(dt)::value>(P, keys, keys_alt, vals, vals_alt, (char*)ws, s, g_timer.on);
  });
}

int gmms_seg_graph_create(gmms_seg_plan* P, void* keys, void* keys_alt, void* vals, void* vals_alt, void* ws,
                          size_t wsb, gmms_seg_graph** out) {
  if (!P || !out) return GMMS_ERROR_INVALID_VALUE;
  *out = nullptr;
  int st = check_buffers(P, keys, keys_alt, vals, vals_alt);
  if (st) return st;
  if (P->n_segs_dev > 0 && (!ws || wsb < P->ws_bytes)) return GMMS_ERROR_WORKSPACE;
  std::unique_ptr<gmms_seg_graph> G(new (std::nothrow) gmms_seg_graph());
  if (!G) return GMMS_ERROR_OUT_OF_MEMORY;
  G->alt = P->R & 1;
  if (P->host_bytes && P->n_segs_dev > 0)
    HIP_OK(hipMemcpy((char*)ws + P->off_table, P->table.data(), P->host_bytes, hipMemcpyHostToDevice));
  // kernel attributes are set outside the capture
  if (P->n_segs_dev > 0) {
    hipStream_t cs;
    HIP_OK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    st = dispatch(P->dtype, [&](auto dt) {
      constexpr int DT = decltype(dt)::value;
      persistent_attr<DT, true>();
      persistent_attr<DT, false>();
      hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
      if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
      int r = seg_enqueue<DT>(P, keys, keys_alt, vals, vals_alt, (char*)ws, cs, false);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(cs, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
      G->graph = g;
      e = hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
      return (int)GMMS_SUCCESS;
    });
    (void)hipStreamDestroy(cs);
    if (st) return st;
  }
  *out = G.release();
  return GMMS_SUCCESS;
}

int gmms_seg_graph_launch(gmms_seg_graph* G, void* stream, int* alt) {
  if (!G) return GMMS_ERROR_INVALID_VALUE;
  if (alt) *alt = G->alt;
  if (!G->exec) return GMMS_SUCCESS;
  HIP_OK(hipGraphLaunch(G->exec, (hipStream_t)stream));
  return GMMS_SUCCESS;
}

void gmms_seg_graph_destroy(gmms_seg_graph* G) { delete G; }

int gmms_segmented_sort(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt, const int64_t* offsets,
                        int64_t n_seg, const gmms_params* p, void* stream, int* alt) {
  gmms_seg_plan* P = nullptr;
  int st = gmms_seg_plan_create(dtype, offsets, n_seg, p, &P);
  if (st) return st;
  std::unique_ptr<gmms_seg_plan> hold(P);
  void* ws = nullptr;
  if (P->ws_bytes) {
    const hipError_t e = hipMalloc(&ws, P->ws_bytes);
    if (e != hipSuccess) {
      hip_fail(e, "hipMalloc(segment workspace)");
      return GMMS_ERROR_OUT_OF_MEMORY;
    }
  }
  st = gmms_seg_sort(P, keys, keys_alt, vals, vals_alt, ws, P->ws_bytes, stream, alt);
  const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess && !st) st = hip_fail(e, "hipStreamSynchronize");
  if (ws) (void)hipFree(ws);
  return st;
}

int gmms_seg_tuning_get(int dtype, int32_t* kway, int64_t* merger_elems, int classes) {
  if (!valid_dtype(dtype) || classes < 0 || classes > GMMS_SEG_CLASSES) return GMMS_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (int c = 0; c < classes; ++c) {
    if (kway) kway[c] = g_tune[dtype].K[c];
    if (merger_elems) merger_elems[c] = g_tune[dtype].merger[c];
  }
  return GMMS_SUCCESS;
}

int gmms_seg_tuning_set(int dtype, const int32_t* kway, const int64_t* merger_elems, int classes) {
  if (!valid_dtype(dtype) || classes < 0 || classes > GMMS_SEG_CLASSES) return GMMS_ERROR_INVALID_VALUE;
  const int kmax = max_seg_k(elem_words(dtype));
  for (int c = 0; c < classes; ++c) {
    if (kway && kway[c] != 0 && (kway[c] < 2 || kway[c] > kmax || (kway[c] & (kway[c] - 1))))
      return GMMS_ERROR_INVALID_VALUE;
    if (merger_elems && merger_elems[c] < 0) return GMMS_ERROR_INVALID_VALUE;
  }
  std::lock_guard<std::mutex> g(g_tune_mu);
  for (int c = 0; c < classes; ++c) {
    if (kway) g_tune[dtype].K[c] = kway[c];
    if (merger_elems) g_tune[dtype].merger[c] = merger_elems[c];
  }
  return GMMS_SUCCESS;
}

int gmms_seg_autotune(int dtype, int min_class, int max_class, int64_t batch_elems, int reps, void* stream,
                      int32_t* kway_out, int64_t* merger_out, float* ms_out) {
  if (!valid_dtype(dtype) || min_class < 0 || max_class >= GMMS_SEG_CLASSES || min_class > max_class ||
      batch_elems < 1 || reps < 1)
    return GMMS_ERROR_INVALID_VALUE;
  const int W = elem_words(dtype);
  const int kmax = max_seg_k(W);
  size_t kb = 0, vb = 0;
  gmms_element_size(dtype, &kb, &vb);
  const int64_t cap = std::max<int64_t>(batch_elems, (int64_t)1 << max_class);
  hipStream_t s = (hipStream_t)stream;
  void* buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // keys, alt, vals, vals_alt, source
  auto release = [&]() {
    for (void*& b : buf)
      if (b) (void)hipFree(b), b = nullptr;
  };
  const size_t kbytes = (size_t)cap * kb, vbytes = (size_t)cap * vb;
  const size_t sizes[5] = {kbytes, kbytes, vbytes, vbytes, kbytes + vbytes};
  for (int i = 0; i < 5; ++i) {
    if (!sizes[i]) continue;
    const hipError_t e = hipMalloc(&buf[i], sizes[i]);
    if (e != hipSuccess) {
      release();
      hip_fail(e, "hipMalloc(autotune buffers)");
      return GMMS_ERROR_OUT_OF_MEMORY;
    }
  }
  {
    const int64_t words = (int64_t)((kbytes + vbytes) / 4);
    hipLaunchKernelGGL(k_fill_random, dim3(1024), dim3(256), 0, s, (uint32_t*)buf[4], words,
                       (uint64_t)0x6D6D7304ull);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = GMMS_SUCCESS;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = GMMS_ERROR_HIP;
  gmms_params base;
  gmms_get_default_params(&base);
  for (int c = min_class; c <= max_class && !st; ++c) {
    const int64_t n = (int64_t)1 << c;
    const int64_t nseg = std::max<int64_t>(1, batch_elems / n);
    std::vector<int64_t> offs(nseg + 1);
    for (int64_t i = 0; i <= nseg; ++i) offs[i] = i * n;
    float best = 1e30f, t_def = 1e30f;
    int32_t bestK = 0;
    int64_t bestM = 0;
    const int32_t defK = std::min(SEG_DEFAULT_K, kmax);
    // K stays at the default; the candidates are the merger sizes. A class
    // timed alone in a small batch gains from every round a K saves (one
    // launch less), but in a batch of mixed sizes the launches are shared
    // and another K's windows cost more per element: on the 528-segment
    // batch, K tuned per class (2..64) took 3.52 ms, K tuned within 2..16
    // 3.23 ms, the default table 3.2 ms and K = 16 with tuned merger sizes
    // 3.04 ms.
    for (int K = defK; K <= defK && !st; K <<= 1) {
      int rounds = 0;
      for (int64_t L = TILE; L < n; L *= K) ++rounds;
      if (rounds == 0 && K != defK) continue;  // one tile: K is irrelevant
      const int64_t minsz = (int64_t)WAVE * K * 4;
      const int64_t ms_cand[3] = {0, minsz * 4, minsz * 16};
      for (int mi = 0; mi < 3 && !st; ++mi) {
        if (mi > 0 && (rounds == 0 || ms_cand[mi] >= n)) continue;
        gmms_params q = base;
        q.kway = K;
        gmms_seg_plan* P = nullptr;
        // a per-call merger size goes through the table entry of this class
        int32_t kt[GMMS_SEG_CLASSES];
        int64_t mt[GMMS_SEG_CLASSES];
        gmms_seg_tuning_get(dtype, kt, mt, GMMS_SEG_CLASSES);
        const int64_t saved = mt[c];
        mt[c] = ms_cand[mi];
        gmms_seg_tuning_set(dtype, nullptr, mt, GMMS_SEG_CLASSES);
        st = gmms_seg_plan_create(dtype, offs.data(), nseg, &q, &P);
        mt[c] = saved;
        gmms_seg_tuning_set(dtype, nullptr, mt, GMMS_SEG_CLASSES);
        if (st) break;
        std::unique_ptr<gmms_seg_plan> hold(P);
        void* ws = nullptr;
        hipError_t he = hipMalloc(&ws, std::max<size_t>(P->ws_bytes, 256));
        if (he != hipSuccess) {
          st = hip_fail(he, "hipMalloc(autotune workspace)");
          break;
        }
        float t_best = 1e30f;
        auto ok = [&](hipError_t e, const char* what) {
          if (e != hipSuccess && !st) st = hip_fail(e, what);
          return st == GMMS_SUCCESS;
        };
        for (int it = 0; it <= reps && !st; ++it) {
          // restore the unsorted input, then time one batched sort
          if (!ok(hipMemcpyAsync(buf[0], buf[4], (size_t)(nseg * n) * kb, hipMemcpyDeviceToDevice, s), "copy keys"))
            break;
          if (vb && !ok(hipMemcpyAsync(buf[2], (char*)buf[4] + kbytes, (size_t)(nseg * n) * vb,
                                       hipMemcpyDeviceToDevice, s),
                        "copy values"))
            break;
          if (!ok(hipEventRecord(e0, s), "hipEventRecord")) break;
          int alt = 0;
          st = gmms_seg_sort(P, buf[0], buf[1], buf[2], buf[3], ws, P->ws_bytes, s, &alt);
          if (st) break;
          if (!ok(hipEventRecord(e1, s), "hipEventRecord") || !ok(hipEventSynchronize(e1), "hipEventSynchronize"))
            break;
          float t = 0.f;
          if (!ok(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime")) break;
          if (it > 0) t_best = std::min(t_best, t);  // the first run is a warm-up
        }
        (void)hipFree(ws);
        if (K == defK && mi == 0) t_def = t_best;
        if (!st && t_best < best) {
          best = t_best;
          bestK = K;
          bestM = ms_cand[mi];
        }
      }
    }
    if (st) break;
    // the default entry stays unless a candidate beats it by 3 %
    if (t_def < 1e30f && best > 0.97f * t_def) {
      best = t_def;
      bestK = defK;
      bestM = 0;
    }
    if (kway_out) kway_out[c] = bestK;
    if (merger_out) merger_out[c] = bestM;
    if (ms_out) ms_out[c] = best;
    int32_t kt[GMMS_SEG_CLASSES];
    int64_t mt[GMMS_SEG_CLASSES];
    gmms_seg_tuning_get(dtype, kt, mt, GMMS_SEG_CLASSES);
    kt[c] = bestK;
    mt[c] = bestM;
    gmms_seg_tuning_set(dtype, kt, mt, GMMS_SEG_CLASSES);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  release();
  return st;
}

}  // extern "C"
