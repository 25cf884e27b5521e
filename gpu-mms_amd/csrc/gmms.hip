// Host driver and C ABI of libgmms.so (declarations: include/gpu_mms/gmms.h).
//
// The round schedule follows the reference driver src/multimergesort.hxx:
// 56-141: base case over 1024-element tiles, then
// `for (L = 1024; L < n; L *= K)` rounds, tasks = ceil(n / (K*L)) (:89),
// buffers ping-pong (:128) and the caller learns which one holds the result
// (:140). Differences by design: the workspace is caller-owned (no
// cudaMalloc in the timed path), sizes are 64-bit, all launches go to the
// caller's stream, and errors come back as status codes.
#include "gmms_basecase.hxx"
#include "gmms_host.hxx"

using namespace gmms;

namespace gmms_host {
thread_local char g_last_error[256] = "";
StageTimer g_timer;
}  // namespace gmms_host

using namespace gmms_host;

namespace {

template <int DT>
int launch_basecase(void* keys, uint32_t* vals, int64_t n, bool cin, bool cout, bool run4, uint32_t flip,
                    hipStream_t s) {
  if (n <= 0) return GMMS_SUCCESS;
  StageScope scope(GMMS_STAGE_BASECASE, n, s);
  const int64_t tiles = (n + TILE - 1) / TILE;
  const dim3 grid((unsigned)((tiles + BASE_WAVES - 1) / BASE_WAVES));
  const dim3 block(BASE_WAVES * WAVE);
  if constexpr (DType<DT>::E::W == 1) {
    if (run4) {
      // 32-bit keys: one wavefront per run of 4096 (64 keys per lane)
      const dim3 rgrid((unsigned)(((n + RUN - 1) / RUN + BASE_WAVES - 1) / BASE_WAVES));
      if (cout) hipLaunchKernelGGL((k_basecase_run<DT, true, true>), rgrid, block, 0, s, keys, n, flip);
      else hipLaunchKernelGGL((k_basecase_run<DT, true, false>), rgrid, block, 0, s, keys, n, flip);
      HIP_OK(hipGetLastError());
      return GMMS_SUCCESS;
    }
  }
  if (run4) {
    // the sort's own base case: maps the caller's bits in, and out only if no round follows
    if (cout) hipLaunchKernelGGL((k_basecase<DT, true, true, true>), grid, block, 0, s, keys, vals, n, flip);
    else hipLaunchKernelGGL((k_basecase<DT, true, false, true>), grid, block, 0, s, keys, vals, n, flip);
  } else if (cin && cout) {
    hipLaunchKernelGGL((k_basecase<DT, true, true>), grid, block, 0, s, keys, vals, n, flip);
  } else if (cin) {
    hipLaunchKernelGGL((k_basecase<DT, true, false>), grid, block, 0, s, keys, vals, n, flip);
  } else if (cout) {
    hipLaunchKernelGGL((k_basecase<DT, false, true>), grid, block, 0, s, keys, vals, n, flip);
  } else {
    hipLaunchKernelGGL((k_basecase<DT, false, false>), grid, block, 0, s, keys, vals, n, flip);
  }
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

template <int DT>
int launch_partition(RoundArgs& a, hipStream_t s, void* samp_k = nullptr, uint32_t* samp_v = nullptr) {
  const int64_t waves = a.tasks * (a.m - 1);
  if (waves <= 0) return GMMS_SUCCESS;
  const dim3 grid((unsigned)((waves + PART_WAVES - 1) / PART_WAVES));
  StageScope scope(GMMS_STAGE_PARTITION, a.n, s);
  if (samp_k) {
    // gather the sample array of the round's input; the searches start there
    const int64_t cnt = (a.n + SAMP - 1) >> SAMP_SHIFT;
    hipLaunchKernelGGL((k_sample<DT>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, a.in_keys, a.in_vals,
                       a.n, samp_k, samp_v);
    HIP_OK(hipGetLastError());
    a.samp_keys = samp_k;
    a.samp_vals = samp_v;
  }
  hipLaunchKernelGGL((k_partition<DT>), grid, dim3(PART_WAVES * WAVE), 0, s, a);
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

template <int DT>
int launch_merge(RoundArgs& a, bool cout, hipStream_t s) {
  using E = typename DType<DT>::E;
  const size_t lds = merge_lds_words(a.K, E::W, a.D) * sizeof(uint32_t);
  if (lds > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
  const dim3 grid((unsigned)(a.tasks * a.m));
  const dim3 block(a.PL * WAVE);
  StageScope scope(GMMS_STAGE_MERGE, a.n, s);
  if (cout) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_merge<DT, true>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr = true;
    }
    hipLaunchKernelGGL((k_merge<DT, true>), grid, block, lds, s, a);
  } else {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_merge<DT, false>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      attr = true;
    }
    hipLaunchKernelGGL((k_merge<DT, false>), grid, block, lds, s, a);
  }
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

template <int DT>
int launch_tmerge(RoundArgs& a, bool cout, hipStream_t s) {
  using E = typename DType<DT>::E;
  const size_t lds = tmerge_lds_words(a.K, E::W, TM_F) * sizeof(uint32_t);
  if (lds > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
  const dim3 grid((unsigned)(a.tasks * a.m));
  StageScope scope(GMMS_STAGE_MERGE, a.n, s);
  if (cout) {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_tmerge<DT, true, TM_F>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)LDS_LIMIT);
      attr = true;
    }
    hipLaunchKernelGGL((k_tmerge<DT, true, TM_F>), grid, dim3(TM_THREADS), lds, s, a);
  } else {
    static bool attr = false;
    if (!attr) {
      (void)hipFuncSetAttribute((const void*)k_tmerge<DT, false, TM_F>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)LDS_LIMIT);
      attr = true;
    }
    hipLaunchKernelGGL((k_tmerge<DT, false, TM_F>), grid, dim3(TM_THREADS), lds, s, a);
  }
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

template <int DT, bool CONV, int WIN, bool PEERS>
int launch_wmerge_kernel(RoundArgs& a, int nw, size_t lds, hipStream_t s) {
  const int64_t mergers = a.tasks * a.m;
  const dim3 grid((unsigned)((mergers + nw - 1) / nw));
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_wmerge<DT, CONV, WIN, PEERS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)LDS_LIMIT);
    attr = true;
  }
  hipLaunchKernelGGL((k_wmerge<DT, CONV, WIN, PEERS>), grid, dim3(nw * WAVE), lds, s, a);
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

template <int DT, bool CONV, int WIN>
int launch_wmerge_win(RoundArgs& a, hipStream_t s) {
  using E = typename DType<DT>::E;
  const int per_cu = wm_resident_per_cu(E::W);
  const int64_t mergers = a.tasks * a.m;
  if (mergers <= (int64_t)per_cu * cu_count()) {
    // every merger resident at once: one workgroup per CU's worth of mergers
    const size_t lds = ((size_t)per_cu * wmerge_lds_words_per_wave(E::W, WIN) + wmerge_peer_words(per_cu)) *
                       sizeof(uint32_t);
    if (lds > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
    return launch_wmerge_kernel<DT, CONV, WIN, true>(a, per_cu, lds, s);
  }
  // later mergers start as earlier ones finish: small workgroups
  const size_t lds = (size_t)WM_WAVES * wmerge_lds_words_per_wave(E::W, WIN) * sizeof(uint32_t);
  if (lds > LDS_LIMIT) return GMMS_ERROR_UNSUPPORTED;
  return launch_wmerge_kernel<DT, CONV, WIN, false>(a, WM_WAVES, lds, s);
}

template <int DT>
int launch_wmerge(RoundArgs& a, bool cout, hipStream_t s) {
  using E = typename DType<DT>::E;
  constexpr int WIN = wm_win(E::W);
  StageScope scope(GMMS_STAGE_MERGE, a.n, s);
  return cout ? launch_wmerge_win<DT, true, WIN>(a, s) : launch_wmerge_win<DT, false, WIN>(a, s);
}

RoundArgs round_args(const void* ik, const void* iv, void* ok, void* ov, int64_t n, int64_t L,
                     const Resolved& r, int m, void* ws) {
  RoundArgs a{};
  a.in_keys = ik;
  a.in_vals = (const uint32_t*)iv;
  a.out_keys = ok;
  a.out_vals = (uint32_t*)ov;
  a.n = n;
  a.L = L;
  a.K = r.K;
  a.PL = r.PL;
  a.D = r.D;
  a.m = m;
  const int64_t TK = (int64_t)r.K * L;
  a.tasks = (n + TK - 1) / TK;
  a.slack = slack_for(r, n, L, m);
  const int64_t b = a.tasks * m;
  a.cuts = (int64_t*)ws;
  a.skips = ws ? (int64_t*)ws + b * r.K : nullptr;
  a.cut_stride = r.K;
  a.flip = r.flip;
  return a;
}

// One merge round. The partition searches the sample array of the input
// (every SAMP-th element, at the start of ws) when the lists start at
// multiples of SAMP: gathered here unless `samp_in` says the previous round
// wrote it. `samp_out`: the wavefront merger writes the samples of its
// output for the next round (*wrote says whether it did).
template <int DT>
int run_round(const void* ik, const void* iv, void* ok, void* ov, int64_t n, int64_t L, const Resolved& r,
              bool cout, void* ws, size_t wsb, hipStream_t s, bool samp_in = false, bool samp_out = false,
              bool* wrote = nullptr) {
  const int m = mergers_per_task(n, L, r.K, r.mergers);
  const size_t sb = samp_bytes(n, r.W);
  RoundArgs a = round_args(ik, iv, ok, ov, n, L, r, m, ws ? (char*)ws + sb : nullptr);
  const size_t cnt = (size_t)((n + SAMP - 1) >> SAMP_SHIFT);
  uint32_t* sv = (ws && r.W == 3) ? (uint32_t*)((char*)ws + cnt * 8) : nullptr;
  if (m > 1) {
    if (!ws || wsb < sb + cut_bytes(n, L, r)) return GMMS_ERROR_WORKSPACE;
    int st;
    if (L % SAMP != 0) {
      st = launch_partition<DT>(a, s);
    } else if (samp_in) {
      a.samp_keys = ws;
      a.samp_vals = sv;
      st = launch_partition<DT>(a, s);
    } else {
      st = launch_partition<DT>(a, s, ws, sv);
    }
    if (st) return st;
  }
  if (wrote) *wrote = false;
  if (samp_out && !cout && r.kernel == GMMS_MERGE_WAVE && ws && wsb >= sb) {
    a.samp_out_keys = ws;
    a.samp_out_vals = sv;
    if (wrote) *wrote = true;
  }
  if (r.kernel == GMMS_MERGE_HEAP) return launch_merge<DT>(a, cout, s);
  if (r.kernel == GMMS_MERGE_TILE) return launch_tmerge<DT>(a, cout, s);
  return launch_wmerge<DT>(a, cout, s);
}

template <int DT>
int run_sort(void* k0, void* k1, void* v0, void* v1, int64_t n, const Resolved& r, void* ws, size_t wsb,
             hipStream_t s, int* alt) {
  int rounds = 0;
  for (int64_t L = r.run; L < n; L *= r.K) ++rounds;
  int st = launch_basecase<DT>(k0, (uint32_t*)v0, n, true, rounds == 0, r.run > TILE, r.flip, s);
  if (st) return st;
  void* kb[2] = {k0, k1};
  void* vb[2] = {v0, v1};
  int cur = 0, done = 0;
  bool samp = false;  // the sample array of the current input is in ws
  for (int64_t L = r.run; L < n; L *= r.K, ++done) {
    // the next round partitions (and its lists are SAMP-aligned): this
    // round's merge writes the samples of its output
    const int64_t L2 = L * r.K;
    const bool next = L2 < n && L2 % SAMP == 0 && mergers_per_task(n, L2, r.K, r.mergers) > 1;
    st = run_round<DT>(kb[cur], vb[cur], kb[cur ^ 1], vb[cur ^ 1], n, L, r, done == rounds - 1, ws, wsb, s, samp,
                       next, &samp);
    if (st) return st;
    cur ^= 1;
  }
  if (alt) *alt = cur;
  return GMMS_SUCCESS;
}


// Stream copy with 128-bit loads/stores: the measured HBM peak the roofline
// is quoted against, and the calibration kernel for FETCH_SIZE/WRITE_SIZE.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Copy-probe variants: U 16-byte loads per lane in flight before the U
// stores (lane-consecutive within each of the U slices of a block's
// stretch), optionally non-temporal stores; grid-stride over stretches.
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_u(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16) {
  const int64_t span = (int64_t)U * blockDim.x;
  const int64_t stride = (int64_t)gridDim.x * span;
  for (int64_t b = (int64_t)blockIdx.x * span; b < n16; b += stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + (int64_t)u * blockDim.x + threadIdx.x;
      if (i < n16) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + (int64_t)u * blockDim.x + threadIdx.x;
      if (i < n16) {
        if constexpr (NT) __builtin_nontemporal_store(v[u], dst + i);
        else dst[i] = v[u];
      }
    }
  }
}

template <int DT>
__global__ void k_pad(void* keys, uint32_t* vals, int64_t count) {
  using E = typename DType<DT>::E;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) gstore<DT, true>(keys, vals, i, Limits<E>::hi());
}

// Workspace of gmms_sort_noalloc: one array per device for the life of the
// process, so that call allocates nothing. Host code must never pass the
// array's name to a kernel (in host code it is a stand-in, not a device
// address): scratch_ptr() returns this device's copy, looked up once per
// device with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[GMMS_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_ptr(void** out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return GMMS_ERROR_UNSUPPORTED;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return GMMS_ERROR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return GMMS_SUCCESS;
}

}  // namespace

extern "C" {

const char* gmms_version(void) { return "gpu-mms-amd 0.1.0 gfx950"; }

const char* gmms_last_error(void) { return g_last_error; }

int gmms_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return GMMS_ERROR_INVALID_VALUE;
  if (!bytes) return GMMS_SUCCESS;
  const int64_t n16 = (int64_t)(bytes / 16);
  const int64_t blocks = std::min<int64_t>((n16 + 255) / 256, (int64_t)cu_count() * 32);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (const uint4*)src,
                     (uint4*)dst, n16);
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

int gmms_copy_probe(void* dst, const void* src, size_t bytes, int unroll, int nontemporal, int blocks_per_cu,
                    void* stream) {
  if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return GMMS_ERROR_INVALID_VALUE;
  if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8) return GMMS_ERROR_INVALID_VALUE;
  if (blocks_per_cu < 0 || blocks_per_cu > 1024 || (nontemporal != 0 && nontemporal != 1))
    return GMMS_ERROR_INVALID_VALUE;
  if (!bytes) return GMMS_SUCCESS;
  const int64_t n16 = (int64_t)(bytes / 16);
  const int64_t need = (n16 + 256LL * unroll - 1) / (256LL * unroll);
  const int64_t blocks = blocks_per_cu ? std::min<int64_t>(need, (int64_t)cu_count() * blocks_per_cu) : need;
  const dim3 g((unsigned)blocks), b(256);
  hipStream_t s = (hipStream_t)stream;
  const u32x4* in = (const u32x4*)src;
  u32x4* out = (u32x4*)dst;
  switch (unroll * 2 + nontemporal) {
    case 2: hipLaunchKernelGGL((k_copy_u<1, false>), g, b, 0, s, in, out, n16); break;
    case 3: hipLaunchKernelGGL((k_copy_u<1, true>), g, b, 0, s, in, out, n16); break;
    case 4: hipLaunchKernelGGL((k_copy_u<2, false>), g, b, 0, s, in, out, n16); break;
    case 5: hipLaunchKernelGGL((k_copy_u<2, true>), g, b, 0, s, in, out, n16); break;
    case 8: hipLaunchKernelGGL((k_copy_u<4, false>), g, b, 0, s, in, out, n16); break;
    case 9: hipLaunchKernelGGL((k_copy_u<4, true>), g, b, 0, s, in, out, n16); break;
    case 16: hipLaunchKernelGGL((k_copy_u<8, false>), g, b, 0, s, in, out, n16); break;
    default: hipLaunchKernelGGL((k_copy_u<8, true>), g, b, 0, s, in, out, n16); break;
  }
  HIP_OK(hipGetLastError());
  return GMMS_SUCCESS;
}

int gmms_stage_timing(int enable) {
  std::lock_guard<std::mutex> g(g_timer.mu);
  g_timer.on = enable != 0;
  g_timer.recs.clear();
  g_timer.used = 0;
  return GMMS_SUCCESS;
}

int gmms_stage_times(int* kinds, float* ms, int64_t* elems, int cap, int* count) {
  if (!count || cap < 0) return GMMS_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> g(g_timer.mu);
  const int n = (int)g_timer.recs.size();
  *count = n;
  for (int i = 0; i < n && i < cap; ++i) {
    const StageRec& r = g_timer.recs[i];
    HIP_OK(hipEventSynchronize(r.b));
    float t = 0.f;
    HIP_OK(hipEventElapsedTime(&t, r.a, r.b));
    if (kinds) kinds[i] = r.kind;
    if (ms) ms[i] = t;
    if (elems) elems[i] = r.elems;
  }
  return GMMS_SUCCESS;
}

const char* gmms_status_string(int s) {
  switch (s) {
    case GMMS_SUCCESS: return "success";
    case GMMS_ERROR_INVALID_VALUE: return "invalid value";
    case GMMS_ERROR_UNSUPPORTED: return "unsupported";
    case GMMS_ERROR_WORKSPACE: return "workspace missing or too small";
    case GMMS_ERROR_HIP: return "HIP runtime error";
    case GMMS_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case GMMS_ERROR_BUDGET: return "request exceeds the HBM budget";
    default: return "unknown status";
  }
}

int gmms_get_default_params(gmms_params* p) {
  if (!p) return GMMS_ERROR_INVALID_VALUE;
  std::memset(p, 0, sizeof(*p));
  p->kway = 0;      // auto (per merge kernel and dtype)
  p->pipeline = 0;  // wavefront tile merge
  p->mergers = 0;
  p->prefetch = 2;
  p->slack = 0;
  return GMMS_SUCCESS;
}

int gmms_resolve_params(int dtype, const gmms_params* in, gmms_params* out) {
  if (!valid_dtype(dtype) || !out) return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, in, &r, true);
  if (st) return st;
  gmms_get_default_params(out);
  if (in) *out = *in;
  out->kway = r.K;
  out->pipeline = r.kernel;
  out->mergers = r.mergers;
  out->prefetch = r.D;
  out->slack = (int32_t)r.slack;
  out->base_run = r.run;
  out->descending = r.flip ? 1 : 0;
  return GMMS_SUCCESS;
}

int gmms_element_size(int dtype, size_t* kb, size_t* vb) {
  if (!valid_dtype(dtype) || !kb || !vb) return GMMS_ERROR_INVALID_VALUE;
  const int w = elem_words(dtype);
  *kb = w == 1 ? 4 : 8;
  *vb = has_vals(dtype) ? 4 : 0;
  return GMMS_SUCCESS;
}

int gmms_num_rounds(int64_t n, int K, int* rounds) {
  if (!rounds || n < 0 || K < 2) return GMMS_ERROR_INVALID_VALUE;
  int r = 0;
  for (int64_t L = TILE; L < n; L *= K) ++r;
  *rounds = r;
  return GMMS_SUCCESS;
}

int gmms_sort_rounds(int dtype, int64_t n, const gmms_params* p, int* rounds) {
  if (!valid_dtype(dtype) || !rounds || n < 0) return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  int k = 0;
  for (int64_t L = r.run; L < n; L *= r.K) ++k;
  *rounds = k;
  return GMMS_SUCCESS;
}

int gmms_workspace_size(int dtype, int64_t n, const gmms_params* p, size_t* bytes) {
  if (!valid_dtype(dtype) || !bytes || n < 0) return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  *bytes = ws_bytes_for(n, r);
  return GMMS_SUCCESS;
}

int gmms_round_workspace_size(int dtype, int64_t n, int64_t L, const gmms_params* p, size_t* bytes) {
  if (!valid_dtype(dtype) || !bytes || n < 0 || L < 1) return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  *bytes = 0;
  if (n == 0) return GMMS_SUCCESS;
  if (mergers_per_task(n, L, r.K, r.mergers) > 1) *bytes = samp_bytes(n, r.W) + cut_bytes(n, L, r);
  return GMMS_SUCCESS;
}

int gmms_sort(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt, int64_t n,
              const gmms_params* p, void* ws, size_t wsb, void* stream, int* alt) {
  if (!valid_dtype(dtype) || n < 0) return GMMS_ERROR_INVALID_VALUE;
  if (alt) *alt = 0;
  if (n == 0) return GMMS_SUCCESS;
  if (!keys || !keys_alt) return GMMS_ERROR_INVALID_VALUE;
  if (has_vals(dtype) != (vals != nullptr) || has_vals(dtype) != (vals_alt != nullptr))
    return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  if (wsb < ws_bytes_for(n, r) || (ws_bytes_for(n, r) > 0 && !ws)) return GMMS_ERROR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  return dispatch(dtype, [&](auto dt) {
    return run_sort<decltype(dt)::value>(keys, keys_alt, vals, vals_alt, n, r, ws, wsb, s, alt);
  });
}

int gmms_noalloc_scratch_bytes(size_t* bytes) {
  if (!bytes) return GMMS_ERROR_INVALID_VALUE;
  *bytes = GMMS_NOALLOC_SCRATCH_BYTES;
  return GMMS_SUCCESS;
}

int gmms_sort_noalloc(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt, int64_t n,
                      const gmms_params* p, void* stream, int* alt) {
  if (!valid_dtype(dtype) || n < 0) return GMMS_ERROR_INVALID_VALUE;
  if (alt) *alt = 0;
  if (n == 0) return GMMS_SUCCESS;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  const size_t need = ws_bytes_for(n, r);
  if (need > GMMS_NOALLOC_SCRATCH_BYTES) {
    std::snprintf(g_last_error, sizeof(g_last_error),
                  "gmms_sort_noalloc: n=%lld needs %zu workspace bytes, the static scratch holds %zu",
                  (long long)n, need, (size_t)GMMS_NOALLOC_SCRATCH_BYTES);
    return GMMS_ERROR_WORKSPACE;
  }
  void* ws = nullptr;
  if (need) {
    st = scratch_ptr(&ws);
    if (st) return st;
  }
  // The sort writes every workspace word it reads (samples, then cuts) in
  // the same call, so the scratch needs no clearing between calls.
  return gmms_sort(dtype, keys, keys_alt, vals, vals_alt, n, p, ws, need, stream, alt);
}

int gmms_multimergesort_params(int dtype, void* input, void* output, void* host_array, const gmms_params* pp,
                               int64_t n, void** result) {
  (void)host_array;
  if (!result) return GMMS_ERROR_INVALID_VALUE;
  *result = input;
  if (!valid_dtype(dtype) || n < 0) return GMMS_ERROR_INVALID_VALUE;
  if (n == 0) return GMMS_SUCCESS;
  if (!input || !output) return GMMS_ERROR_INVALID_VALUE;
  if (has_vals(dtype)) return GMMS_ERROR_UNSUPPORTED;  // pairs need the value buffers: use gmms_sort
  gmms_params p;
  if (pp) p = *pp; else gmms_get_default_params(&p);
  size_t wsb = 0;
  int st = gmms_workspace_size(dtype, n, &p, &wsb);
  if (st) return st;
  void* ws = nullptr;
  if (wsb) {
    const hipError_t e = hipMalloc(&ws, wsb);
    if (e != hipSuccess) {
      hip_fail(e, "hipMalloc(workspace)");
      return GMMS_ERROR_OUT_OF_MEMORY;
    }
  }
  int alt = 0;
  st = gmms_sort(dtype, input, output, nullptr, nullptr, n, &p, ws, wsb, nullptr, &alt);
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess && !st) st = hip_fail(e, "hipDeviceSynchronize");
  if (ws) (void)hipFree(ws);
  if (!st) *result = alt ? output : input;
  return st;
}

int gmms_multimergesort(int dtype, void* input, void* output, void* host_array, int thread_blocks, int64_t n,
                        void** result) {
  gmms_params p;
  gmms_get_default_params(&p);
  if (thread_blocks > 0) p.mergers = thread_blocks;
  return gmms_multimergesort_params(dtype, input, output, host_array, &p, n, result);
}

int gmms_squaresort(int dtype, void* keys, void* vals, int64_t n, int to_internal, int to_external, void* stream) {
  if (!valid_dtype(dtype) || n < 0) return GMMS_ERROR_INVALID_VALUE;
  if (n == 0) return GMMS_SUCCESS;
  if (!keys || has_vals(dtype) != (vals != nullptr)) return GMMS_ERROR_INVALID_VALUE;
  hipStream_t s = (hipStream_t)stream;
  return dispatch(dtype, [&](auto dt) {
    return launch_basecase<decltype(dt)::value>(keys, (uint32_t*)vals, n, to_internal != 0, to_external != 0, false,
                                                0u, s);
  });
}

int gmms_merge_round(int dtype, const void* ik, const void* iv, void* ok, void* ov, int64_t n, int64_t L,
                     const gmms_params* p, int to_external, void* ws, size_t wsb, void* stream) {
  if (!valid_dtype(dtype) || n < 0 || L < 1) return GMMS_ERROR_INVALID_VALUE;
  if (n == 0) return GMMS_SUCCESS;
  if (!ik || !ok || has_vals(dtype) != (iv != nullptr) || has_vals(dtype) != (ov != nullptr))
    return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  return dispatch(dtype, [&](auto dt) {
    return run_round<decltype(dt)::value>(ik, iv, ok, ov, n, L, r, to_external != 0, ws, wsb, s);
  });
}

int gmms_find_partitions(int dtype, const void* keys, const void* vals, int64_t n, int64_t L, const gmms_params* p,
                         int64_t* cuts, int64_t* skips, int64_t cap, int* mpt, void* stream) {
  if (!valid_dtype(dtype) || n < 0 || L < 1 || !mpt) return GMMS_ERROR_INVALID_VALUE;
  if (!keys || !cuts || !skips || has_vals(dtype) != (vals != nullptr)) return GMMS_ERROR_INVALID_VALUE;
  Resolved r;
  int st = resolve(dtype, p, &r, true);
  if (st) return st;
  const int m = mergers_per_task(n, L, r.K, r.mergers);
  *mpt = m;
  const int64_t TK = (int64_t)r.K * L;
  const int64_t tasks = (n + TK - 1) / TK;
  if (tasks * m > cap) return GMMS_ERROR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  HIP_OK(hipMemsetAsync(cuts, 0, (size_t)(tasks * m * r.K) * sizeof(int64_t), s));
  HIP_OK(hipMemsetAsync(skips, 0, (size_t)(tasks * m) * sizeof(int64_t), s));
  if (m <= 1) return GMMS_SUCCESS;
  RoundArgs a = round_args(keys, vals, nullptr, nullptr, n, L, r, m, nullptr);
  a.cuts = cuts;
  a.skips = skips;
  return dispatch(dtype, [&](auto dt) { return launch_partition<decltype(dt)::value>(a, s); });
}

int gmms_pad(int dtype, void* keys, void* vals, int64_t count, void* stream) {
  if (!valid_dtype(dtype) || count < 0) return GMMS_ERROR_INVALID_VALUE;
  if (count == 0) return GMMS_SUCCESS;
  if (!keys || has_vals(dtype) != (vals != nullptr)) return GMMS_ERROR_INVALID_VALUE;
  hipStream_t s = (hipStream_t)stream;
  return dispatch(dtype, [&](auto dt) {
    constexpr int DT = decltype(dt)::value;
    hipLaunchKernelGGL((k_pad<DT>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, keys, (uint32_t*)vals,
                       count);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? (int)GMMS_SUCCESS : hip_fail(e, "k_pad launch");
  });
}

}  // extern "C"
