// Sort server (declarations: include/gpu_mms/gmms.h): admission against an
// HBM budget, one resident device arena, continuous batching of pending
// requests into batched segmented sorts, several batches in flight on their
// own streams. No device code here; all HIP calls are made by the scheduler
// thread without holding the server lock (stream callbacks take that lock).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

#include "gmms_host.hxx"

using namespace gmms_host;

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t ALIGN = 256;
size_t al(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// Per-slot pinned staging for the batch tables (copied to the device with
// the batch's other inputs); larger tables are copied from pageable memory.
constexpr size_t STAGING_BYTES = 16u << 20;

size_t key_bytes(int dtype) { return elem_words(dtype) == 1 ? 4 : 8; }
size_t val_bytes(int dtype) { return has_vals(dtype) ? 4 : 0; }

// Data bytes of one request: both key buffers and, for pairs, both value
// buffers, each rounded to the arena alignment.
size_t data_bytes(int dtype, int64_t n) {
  return 2 * al((size_t)n * key_bytes(dtype)) + 2 * al((size_t)n * val_bytes(dtype));
}

// First-fit allocator over [0, size) with coalescing; offsets and lengths
// are multiples of ALIGN.
struct Arena {
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(size_t sz) {
    size = sz;
    free_.clear();
    if (sz) free_[0] = sz;
  }
  size_t largest() const {
    size_t m = 0;
    for (const auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  bool alloc(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t noff = it->first + bytes;
      free_.erase(it);
      if (rest) free_[noff] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  uint64_t id = 0;
  int dtype = 0;
  const void* keys = nullptr;
  const void* vals = nullptr;
  int64_t n = 0;
  void* out_keys = nullptr;
  void* out_vals = nullptr;
  Clock::time_point t_submit, t_done;
  int status = GMMS_SUCCESS;
  bool done = false;
  int64_t batch = -1;
  // recorded on the caller's stream at submission (gmms_server_submit_after):
  // the batch's copies wait for it, so they follow the work that produces
  // the inputs and that last used the result buffers
  hipEvent_t after = nullptr;
};

struct Batch;

}  // namespace

struct gmms_server {
  gmms_server_config cfg{};
  int device = 0;
  void* arena_dev = nullptr;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<void*> staging;  // pinned, per slot
  std::vector<bool> slot_busy;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  std::vector<hipEvent_t> free_events;  // ordering events of finished requests, for reuse
  std::mutex ev_mu;                     // guards free_events (taken without mu)
  std::vector<std::unique_ptr<Batch>> inflight;
  uint64_t next_id = 1;
  int64_t next_batch = 0;
  bool stop = false;
  gmms_server_stats st{};
  std::thread sched;
};

namespace {

struct Batch {
  gmms_server* S = nullptr;
  int64_t id = 0;
  int slot = 0;
  int dtype = 0;
  std::vector<std::shared_ptr<Request>> reqs;
  gmms_seg_plan* plan = nullptr;
  size_t off = 0, bytes = 0;
  bool done = false;
  int status = GMMS_SUCCESS;
  Clock::time_point t_done;
};

void batch_done_cb(void* p) {
  Batch* b = (Batch*)p;
  std::lock_guard<std::mutex> g(b->S->mu);
  b->done = true;
  b->t_done = Clock::now();
  b->S->cv.notify_all();
}

int policy(const int32_t* dtype, const int64_t* n, int64_t n_pending, size_t free_bytes, int64_t max_elems,
           int32_t max_reqs, std::vector<int64_t>& chosen) {
  chosen.clear();
  if (n_pending <= 0) return GMMS_SUCCESS;
  const int d0 = dtype[0];
  size_t bytes = 0;
  int64_t elems = 0;
  for (int64_t i = 0; i < n_pending; ++i) {
    if (dtype[i] != d0) continue;  // other dtypes wait for their own batch
    const size_t b = data_bytes(dtype[i], n[i]);
    if (bytes + b > free_bytes || elems + n[i] > max_elems || (int64_t)chosen.size() >= max_reqs) {
      if (i == 0) return GMMS_SUCCESS;  // the oldest request does not fit yet
      break;                            // keep arrival order: stop at the first misfit
    }
    bytes += b;
    elems += n[i];
    chosen.push_back(i);
  }
  return GMMS_SUCCESS;
}

// Device layout of a batch inside its arena block.
struct Layout {
  size_t k0, k1, v0, v1, ws, total;
};
Layout layout_for(int dtype, int64_t elems, size_t ws_bytes) {
  Layout L{};
  const size_t kb = al((size_t)elems * key_bytes(dtype)), vb = al((size_t)elems * val_bytes(dtype));
  L.k0 = 0;
  L.k1 = kb;
  L.v0 = 2 * kb;
  L.v1 = 2 * kb + vb;
  L.ws = 2 * kb + 2 * vb;
  L.total = L.ws + al(ws_bytes);
  return L;
}

// Issue one batch on its slot's stream (no lock held).
int launch_batch(gmms_server* S, Batch* b) {
  hipStream_t s = S->streams[b->slot];
  const int dt = b->dtype;
  int64_t elems = 0;
  for (auto& r : b->reqs) elems += r->n;
  const Layout L = layout_for(dt, elems, seg_plan_workspace(b->plan));
  char* base = (char*)S->arena_dev + b->off;
  const size_t kb = key_bytes(dt), vb = val_bytes(dt);
  int64_t o = 0;
  for (auto& r : b->reqs)
    if (r->after) HIP_OK(hipStreamWaitEvent(s, r->after, 0));
  for (auto& r : b->reqs) {
    // request buffers in host or device memory (unified addressing picks the copy)
    HIP_OK(hipMemcpyAsync(base + L.k0 + o * kb, r->keys, (size_t)r->n * kb, hipMemcpyDefault, s));
    if (vb) HIP_OK(hipMemcpyAsync(base + L.v0 + o * vb, r->vals, (size_t)r->n * vb, hipMemcpyDefault, s));
    o += r->n;
  }
  size_t tb = 0, toff = 0;
  const void* table = seg_plan_table(b->plan, &tb, &toff);
  if (tb) {
    const void* src = table;
    if (tb <= STAGING_BYTES && S->staging[b->slot]) {
      std::memcpy(S->staging[b->slot], table, tb);  // the slot's previous batch has finished
      src = S->staging[b->slot];
    }
    HIP_OK(hipMemcpyAsync(base + L.ws + toff, src, tb, hipMemcpyHostToDevice, s));
  }
  int alt = 0;
  int st = seg_plan_launch(b->plan, base + L.k0, base + L.k1, vb ? base + L.v0 : nullptr,
                           vb ? base + L.v1 : nullptr, base + L.ws, s, &alt);
  if (st) return st;
  const size_t rk = alt ? L.k1 : L.k0, rv = alt ? L.v1 : L.v0;
  o = 0;
  for (auto& r : b->reqs) {
    HIP_OK(hipMemcpyAsync(r->out_keys, base + rk + o * kb, (size_t)r->n * kb, hipMemcpyDefault, s));
    if (vb) HIP_OK(hipMemcpyAsync(r->out_vals, base + rv + o * vb, (size_t)r->n * vb, hipMemcpyDefault, s));
    o += r->n;
  }
  HIP_OK(hipLaunchHostFunc(s, batch_done_cb, b));
  return GMMS_SUCCESS;
}

// Plan the chosen requests; drop the newest until the batch with its tables
// fits the largest free block. Returns the plan (nullptr: nothing fits).
gmms_seg_plan* fit_batch(gmms_server* S, std::vector<std::shared_ptr<Request>>& reqs, size_t free_bytes,
                         size_t* bytes) {
  while (!reqs.empty()) {
    std::vector<int64_t> offs(reqs.size() + 1, 0);
    for (size_t i = 0; i < reqs.size(); ++i) offs[i + 1] = offs[i] + reqs[i]->n;
    size_t ws = 0;  // sized from the plan's counts; the tables are built once it fits
    if (seg_plan_workspace_for(reqs[0]->dtype, offs.data(), (int64_t)reqs.size(), &ws) != GMMS_SUCCESS)
      return nullptr;
    const Layout L = layout_for(reqs[0]->dtype, offs.back(), ws);
    if (L.total <= free_bytes) {
      gmms_seg_plan* P = nullptr;
      if (gmms_seg_plan_create(reqs[0]->dtype, offs.data(), (int64_t)reqs.size(), nullptr, &P) != GMMS_SUCCESS)
        return nullptr;
      *bytes = layout_for(reqs[0]->dtype, offs.back(), seg_plan_workspace(P)).total;
      return P;
    }
    reqs.pop_back();
  }
  return nullptr;
}

void finish(gmms_server* S, Batch* b) {
  // under S->mu
  for (auto& r : b->reqs) {
    r->status = b->status;
    r->t_done = b->done ? b->t_done : Clock::now();
    r->done = true;
    r->batch = b->id;
    if (b->status == GMMS_SUCCESS) {
      S->st.completed++;
      S->st.keys_sorted += r->n;
    } else {
      S->st.failed++;
    }
  }
  S->arena.release(b->off, b->bytes);
  S->st.reserved_bytes -= b->bytes;
  S->slot_busy[b->slot] = false;
}

void scheduler(gmms_server* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // 1. retire finished batches
    std::vector<gmms_seg_plan*> dead;
    std::vector<hipEvent_t> dead_ev;
    for (size_t i = 0; i < S->inflight.size();) {
      Batch* b = S->inflight[i].get();
      if (b->done) {
        finish(S, b);
        dead.push_back(b->plan);
        for (auto& r : b->reqs)
          if (r->after) {
            dead_ev.push_back(r->after);
            r->after = nullptr;
          }
        S->inflight.erase(S->inflight.begin() + i);
      } else {
        ++i;
      }
    }
    if (!dead.empty()) S->cv.notify_all();  // waiters
    // 2. admit and launch while a slot and arena space are free
    for (;;) {
      int slot = -1;
      for (size_t i = 0; i < S->slot_busy.size(); ++i)
        if (!S->slot_busy[i]) {
          slot = (int)i;
          break;
        }
      if (slot < 0 || S->pending.empty()) break;
      std::vector<int32_t> dts;
      std::vector<int64_t> ns;
      for (auto& r : S->pending) {
        dts.push_back(r->dtype);
        ns.push_back(r->n);
      }
      std::vector<int64_t> chosen;
      const size_t free_bytes = S->arena.largest();
      policy(dts.data(), ns.data(), (int64_t)ns.size(), free_bytes, S->cfg.max_batch_elems,
             S->cfg.max_batch_requests, chosen);
      if (chosen.empty()) break;
      std::vector<std::shared_ptr<Request>> reqs;
      for (int64_t i : chosen) reqs.push_back(S->pending[i]);
      size_t bytes = 0;
      gmms_seg_plan* P = fit_batch(S, reqs, free_bytes, &bytes);
      if (!P) break;
      size_t off = 0;
      if (!S->arena.alloc(bytes, &off)) {  // cannot happen: bytes <= largest free block
        gmms_seg_plan_destroy(P);
        break;
      }
      // remove the admitted requests from the pending queue (in order)
      {
        std::deque<std::shared_ptr<Request>> rest;
        size_t k = 0;
        for (size_t i = 0; i < S->pending.size(); ++i) {
          if (k < reqs.size() && S->pending[i] == reqs[k]) {
            ++k;
            continue;
          }
          rest.push_back(S->pending[i]);
        }
        S->pending.swap(rest);
      }
      auto b = std::make_unique<Batch>();
      b->S = S;
      b->id = S->next_batch++;
      b->slot = slot;
      b->dtype = reqs[0]->dtype;
      b->reqs = std::move(reqs);
      b->plan = P;
      b->off = off;
      b->bytes = bytes;
      S->slot_busy[slot] = true;
      S->st.batches++;
      S->st.max_batch_size = std::max<int64_t>(S->st.max_batch_size, (int64_t)b->reqs.size());
      S->st.reserved_bytes += bytes;
      S->st.peak_reserved_bytes = std::max(S->st.peak_reserved_bytes, S->st.reserved_bytes);
      S->inflight.push_back(std::move(b));
      S->st.max_inflight_seen = std::max<int64_t>(S->st.max_inflight_seen, (int64_t)S->inflight.size());
      Batch* bp = S->inflight.back().get();
      lk.unlock();
      const int st = launch_batch(S, bp);
      lk.lock();
      if (st) {
        // the stream may still run part of the batch: wait for it without the lock
        lk.unlock();
        (void)hipStreamSynchronize(S->streams[bp->slot]);
        lk.lock();
        bp->status = st;
        bp->t_done = Clock::now();
        bp->done = true;
      }
    }
    if (!dead.empty()) {
      lk.unlock();
      for (gmms_seg_plan* P : dead) gmms_seg_plan_destroy(P);
      if (!dead_ev.empty()) {  // their batches have finished: reusable
        std::lock_guard<std::mutex> g(S->ev_mu);
        S->free_events.insert(S->free_events.end(), dead_ev.begin(), dead_ev.end());
      }
      lk.lock();
      continue;  // re-check: the state may have changed meanwhile
    }
    if (S->stop && S->pending.empty() && S->inflight.empty()) break;
    bool any_done = false;
    for (auto& b : S->inflight) any_done |= b->done;
    if (!any_done) S->cv.wait(lk);
  }
}

}  // namespace

extern "C" {

int gmms_server_default_config(gmms_server_config* c) {
  if (!c) return GMMS_ERROR_INVALID_VALUE;
  std::memset(c, 0, sizeof(*c));
  c->hbm_budget = (size_t)64 * 1000 * 1000 * 1000;  // 64 GB
  c->max_batch_elems = (int64_t)1 << 28;
  c->max_batch_requests = 4096;
  c->max_inflight = 3;
  return GMMS_SUCCESS;
}

int gmms_server_request_bytes(int dtype, int64_t n, size_t* bytes) {
  if (!valid_dtype(dtype) || n < 0 || !bytes) return GMMS_ERROR_INVALID_VALUE;
  *bytes = data_bytes(dtype, n);
  return GMMS_SUCCESS;
}

int gmms_server_batch_policy(const int32_t* dtype, const int64_t* n, int64_t n_pending, size_t free_bytes,
                             int64_t max_batch_elems, int32_t max_batch_requests, int64_t* chosen,
                             int64_t* n_chosen) {
  if (n_pending < 0 || !n_chosen || (n_pending > 0 && (!dtype || !n || !chosen)) || max_batch_elems < 1 ||
      max_batch_requests < 1)
    return GMMS_ERROR_INVALID_VALUE;
  for (int64_t i = 0; i < n_pending; ++i)
    if (!valid_dtype(dtype[i]) || n[i] < 0) return GMMS_ERROR_INVALID_VALUE;
  std::vector<int64_t> c;
  policy(dtype, n, n_pending, free_bytes, max_batch_elems, max_batch_requests, c);
  for (size_t i = 0; i < c.size(); ++i) chosen[i] = c[i];
  *n_chosen = (int64_t)c.size();
  return GMMS_SUCCESS;
}

int gmms_server_create(const gmms_server_config* cfg, gmms_server** out) {
  if (!out) return GMMS_ERROR_INVALID_VALUE;
  *out = nullptr;
  gmms_server_config c;
  gmms_server_default_config(&c);
  if (cfg) {
    if (cfg->hbm_budget) c.hbm_budget = cfg->hbm_budget;
    if (cfg->max_batch_elems) c.max_batch_elems = cfg->max_batch_elems;
    if (cfg->max_batch_requests) c.max_batch_requests = cfg->max_batch_requests;
    if (cfg->max_inflight) c.max_inflight = cfg->max_inflight;
  }
  if (c.max_batch_elems < 1 || c.max_batch_requests < 1 || c.max_inflight < 1 || c.max_inflight > 16)
    return GMMS_ERROR_INVALID_VALUE;
  c.hbm_budget &= ~(ALIGN - 1);
  std::unique_ptr<gmms_server> S(new (std::nothrow) gmms_server());
  if (!S) return GMMS_ERROR_OUT_OF_MEMORY;
  S->cfg = c;
  (void)cu_count();  // cached now: the planner later runs under the server lock
  HIP_OK(hipGetDevice(&S->device));
  hipError_t e = hipMalloc(&S->arena_dev, c.hbm_budget);
  if (e != hipSuccess) {
    S->arena_dev = nullptr;
    hip_fail(e, "hipMalloc(server arena)");
    return GMMS_ERROR_OUT_OF_MEMORY;
  }
  S->arena.init(c.hbm_budget);
  S->st.budget = c.hbm_budget;
  S->st.arena_bytes = c.hbm_budget;
  auto cleanup = [&]() {
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    for (void* p : S->staging)
      if (p) (void)hipHostFree(p);
    (void)hipFree(S->arena_dev);
  };
  for (int i = 0; i < c.max_inflight; ++i) {
    hipStream_t s;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);  // default priority
    if (e != hipSuccess) {
      cleanup();
      return hip_fail(e, "hipStreamCreateWithFlags");
    }
    S->streams.push_back(s);
    void* p = nullptr;
    if (hipHostMalloc(&p, STAGING_BYTES, hipHostMallocDefault) != hipSuccess) p = nullptr;
    S->staging.push_back(p);
  }
  S->slot_busy.assign(c.max_inflight, false);
  gmms_server* raw = S.get();
  S->sched = std::thread(scheduler, raw);
  *out = S.release();
  return GMMS_SUCCESS;
}

int gmms_server_submit(gmms_server* S, int dtype, const void* keys, const void* vals, int64_t n, void* out_keys,
                       void* out_vals, uint64_t* ticket) {
  return gmms_server_submit_after(S, dtype, keys, vals, n, out_keys, out_vals, nullptr, 0, ticket);
}

int gmms_server_submit_after(gmms_server* S, int dtype, const void* keys, const void* vals, int64_t n,
                             void* out_keys, void* out_vals, void* stream, int ordered, uint64_t* ticket) {
  if (!S || !ticket || !valid_dtype(dtype) || n < 0) return GMMS_ERROR_INVALID_VALUE;
  if (n > 0 && (!keys || !out_keys || has_vals(dtype) != (vals != nullptr) ||
                has_vals(dtype) != (out_vals != nullptr)))
    return GMMS_ERROR_INVALID_VALUE;
  auto r = std::make_shared<Request>();
  r->dtype = dtype;
  r->keys = keys;
  r->vals = vals;
  r->n = n;
  r->out_keys = out_keys;
  r->out_vals = out_vals;
  r->t_submit = Clock::now();
  if (n > 0) {
    // a request alone, with its tables, must fit the budget
    int64_t offs[2] = {0, n};
    size_t ws = 0;
    int st = seg_plan_workspace_for(dtype, offs, 1, &ws);
    if (st) return st;
    const size_t need = layout_for(dtype, n, ws).total;
    if (need > S->cfg.hbm_budget) {
      std::lock_guard<std::mutex> g(S->mu);
      S->st.rejected++;
      return GMMS_ERROR_BUDGET;
    }
  }
  if (n > 0 && ordered) {
    {
      std::lock_guard<std::mutex> g(S->ev_mu);
      if (!S->free_events.empty()) {
        r->after = S->free_events.back();
        S->free_events.pop_back();
      }
    }
    if (!r->after) HIP_OK(hipEventCreateWithFlags(&r->after, hipEventDisableTiming));
    const hipError_t e = hipEventRecord(r->after, (hipStream_t)stream);
    if (e != hipSuccess) {
      (void)hipEventDestroy(r->after);
      return hip_fail(e, "hipEventRecord(request order)");
    }
  }
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stop) {
    if (r->after) (void)hipEventDestroy(r->after);
    return GMMS_ERROR_INVALID_VALUE;
  }
  r->id = S->next_id++;
  *ticket = r->id;
  S->st.submitted++;
  S->live[r->id] = r;
  if (n == 0) {
    r->done = true;
    r->t_done = r->t_submit;
    S->st.completed++;
  } else {
    S->pending.push_back(r);
  }
  S->cv.notify_all();
  return GMMS_SUCCESS;
}

int gmms_server_wait(gmms_server* S, uint64_t ticket, double timeout_s, int* done, int* status, double* latency_ms,
                     int64_t* batch_id) {
  if (!S || !done) return GMMS_ERROR_INVALID_VALUE;
  *done = 0;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->live.find(ticket);
  if (it == S->live.end()) return GMMS_ERROR_INVALID_VALUE;
  std::shared_ptr<Request> r = it->second;
  auto ready = [&] { return r->done; };
  if (timeout_s < 0) {
    S->cv.wait(lk, ready);
  } else {
    S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
  }
  if (!r->done) return GMMS_SUCCESS;
  *done = 1;
  if (status) *status = r->status;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (batch_id) *batch_id = r->batch;
  S->live.erase(ticket);
  return GMMS_SUCCESS;
}

int gmms_server_drain(gmms_server* S) {
  if (!S) return GMMS_ERROR_INVALID_VALUE;
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv.wait(lk, [&] { return S->pending.empty() && S->inflight.empty(); });
  return GMMS_SUCCESS;
}

int gmms_server_stats_get(gmms_server* S, gmms_server_stats* out) {
  if (!S || !out) return GMMS_ERROR_INVALID_VALUE;
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->st;
  return GMMS_SUCCESS;
}

void gmms_server_destroy(gmms_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
    S->cv.notify_all();
  }
  if (S->sched.joinable()) S->sched.join();
  for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
  for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
  for (void* p : S->staging)
    if (p) (void)hipHostFree(p);
  if (S->arena_dev) (void)hipFree(S->arena_dev);
  delete S;
}

}  // extern "C"
