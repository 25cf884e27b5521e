// GPU-MMS oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for what pins it).
// Plain C++17 for g++; no GPU code. Loaded by tests/, smoke() and bench.py's
// cpu_baseline leg only.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <vector>
#include <parallel/algorithm>
#include <omp.h>

namespace {

inline double u53(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

struct Pair {
  uint64_t k;
  uint32_t v;
};
inline bool pair_less(const Pair& a, const Pair& b) {
  return a.k < b.k || (a.k == b.k && a.v < b.v);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- generators
uint64_t oracle_mt64_nth(uint64_t seed, uint64_t nth) {
  std::mt19937_64 g(seed);
  if (nth > 1) g.discard(nth - 1);
  return g();
}

void oracle_gen_u64(uint64_t seed, int64_t n, uint64_t* out) {
  std::mt19937_64 g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g();
}

void oracle_gen_i32(uint64_t seed, int64_t n, int shift, int32_t* out) {
  std::mt19937_64 g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)(uint32_t)(g() >> shift);
}

void oracle_gen_f32_normal(uint64_t seed, int64_t n, float* out) {
  std::mt19937_64 g(seed);
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < n; ++i) {
    double u1 = 1.0 - u53(g());  // (0, 1]
    double u2 = u53(g());
    out[i] = (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2));
  }
}

void oracle_gen_pairs(uint64_t seed, int64_t n, int dist, double zipf_s,
                      int64_t distinct, uint64_t* keys, uint32_t* vals) {
  std::mt19937_64 g(seed);
  for (int64_t i = 0; i < n; ++i) vals[i] = (uint32_t)i;
  if (dist == 0) {
    for (int64_t i = 0; i < n; ++i) keys[i] = g();
  } else if (dist == 1 || dist == 2) {
    for (int64_t i = 0; i < n; ++i) keys[i] = g();
    __gnu_parallel::sort(keys, keys + n);  // the same result as std::sort
    if (dist == 2) std::reverse(keys, keys + n);
  } else {
    std::vector<double> cdf((size_t)distinct);
    double acc = 0.0;
    for (int64_t r = 0; r < distinct; ++r) {
      acc += std::pow((double)(r + 1), -zipf_s);
      cdf[(size_t)r] = acc;
    }
    for (auto& c : cdf) c /= acc;
    // the draws in stream order, then the inverse-CDF lookups in parallel
    for (int64_t i = 0; i < n; ++i) keys[i] = g();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double u = u53(keys[i]);
      int64_t r = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
      if (r >= distinct) r = distinct - 1;
      const uint64_t rank = (uint64_t)r + 1;
      keys[i] = (rank * 0x9E3779B97F4A7C15ull) ^ 0xD1B54A32D192ED03ull;
    }
  }
}

void oracle_gen_seglens(uint64_t seed, int64_t count, double lo_log2,
                        double hi_log2, int64_t* lens) {
  std::mt19937_64 g(seed);
  for (int64_t i = 0; i < count; ++i)
    lens[i] = (int64_t)std::floor(std::exp2(lo_log2 + (hi_log2 - lo_log2) * u53(g())));
}

// glibc rand(): TYPE_3 additive feedback generator r[i] = r[i-3] + r[i-31],
// state seeded by the Lehmer step 16807 mod (2^31-1), 310 outputs discarded,
// result r >> 1; srand(0) is treated as srand(1).
void oracle_glibc_rand_list(uint32_t seed, int64_t n, int32_t range, int32_t* out) {
  std::vector<uint32_t> r(34);
  int64_t s = seed == 0 ? 1 : (int64_t)(int32_t)seed;
  r[0] = (uint32_t)s;
  for (int i = 1; i < 31; ++i) {
    int64_t v = (16807LL * (int64_t)(int32_t)r[i - 1]) % 2147483647LL;
    if (v < 0) v += 2147483647LL;
    r[i] = (uint32_t)v;
  }
  for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
  // ring buffer of the last 34 values
  std::vector<uint32_t> ring(r);
  int64_t idx = 34;
  auto next = [&]() -> uint32_t {
    uint32_t v = ring[(idx - 31) % 34] + ring[(idx - 3) % 34];
    ring[idx % 34] = v;
    ++idx;
    return v;
  };
  for (int i = 34; i < 344; ++i) next();
  // buildData.h:59-63: data[i] = rand()%RANGE; temp = rand()%RANGE; the
  // second draw is shifted out of an int key, but still consumed.
  for (int64_t i = 0; i < n; ++i) {
    int32_t a = (int32_t)(next() >> 1);
    (void)next();
    out[i] = a % range;
  }
}

// -------------------------------------------------------------- comparators
uint32_t oracle_f32_order_key(uint32_t u) {
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}
uint64_t oracle_f64_order_key(uint64_t u) {
  return u ^ ((uint64_t)((int64_t)u >> 63) | 0x8000000000000000ull);
}

// ---------------------------------------------------------------- the oracle
void oracle_sort_i32(int32_t* a, int64_t n) { std::sort(a, a + n); }
void oracle_sort_u32(uint32_t* a, int64_t n) { std::sort(a, a + n); }
void oracle_sort_i64(int64_t* a, int64_t n) { std::sort(a, a + n); }
void oracle_sort_u64(uint64_t* a, int64_t n) { std::sort(a, a + n); }

void oracle_sort_f32(float* a, int64_t n) {
  uint32_t* u = reinterpret_cast<uint32_t*>(a);
  std::sort(u, u + n, [](uint32_t x, uint32_t y) {
    return oracle_f32_order_key(x) < oracle_f32_order_key(y);
  });
}
void oracle_sort_f64(double* a, int64_t n) {
  uint64_t* u = reinterpret_cast<uint64_t*>(a);
  std::sort(u, u + n, [](uint64_t x, uint64_t y) {
    return oracle_f64_order_key(x) < oracle_f64_order_key(y);
  });
}

void oracle_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n) {
  std::vector<Pair> p((size_t)n);
  for (int64_t i = 0; i < n; ++i) p[(size_t)i] = {keys[i], vals[i]};
  std::sort(p.begin(), p.end(), pair_less);
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = p[(size_t)i].k;
    vals[i] = p[(size_t)i].v;
  }
}

void oracle_psort_i32(int32_t* a, int64_t n, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  __gnu_parallel::sort(a, a + n);
}

void oracle_psort_u64(uint64_t* a, int64_t n, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  __gnu_parallel::sort(a, a + n);
}

void oracle_psort_f32(float* a, int64_t n, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  uint32_t* u = reinterpret_cast<uint32_t*>(a);
  __gnu_parallel::sort(u, u + n, [](uint32_t x, uint32_t y) {
    return oracle_f32_order_key(x) < oracle_f32_order_key(y);
  });
}

void oracle_psort_u64pairs(uint64_t* keys, uint32_t* vals, int64_t n, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  std::vector<Pair> p((size_t)n);
  for (int64_t i = 0; i < n; ++i) p[(size_t)i] = {keys[i], vals[i]};
  __gnu_parallel::sort(p.begin(), p.end(), pair_less);
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = p[(size_t)i].k;
    vals[i] = p[(size_t)i].v;
  }
}

// ----------------------------------------------------------------- base case
void oracle_basecase_i32(int32_t* a, int64_t n, int64_t tile) {
  for (int64_t t = 0; t < n; t += tile) std::sort(a + t, a + std::min(n, t + tile));
}

// ------------------------------------------------------------------ schedule
int oracle_rounds(int64_t n, int64_t K, int64_t tile) {
  int r = 0;
  for (int64_t L = tile; L < n; L *= K) ++r;
  return r;
}
int64_t oracle_tasks(int64_t n, int64_t K, int64_t L) {
  return n / (K * L) + ((n % (K * L)) > 0);
}
int64_t oracle_edge_task_size(int64_t n, int64_t K, int64_t L) {
  return (n % (K * L)) > 0 ? n % (K * L) : K * L;
}
int64_t oracle_merger_target(int64_t j, int64_t task_size, int64_t mergers) {
  return j * task_size / mergers;
}

// --------------------------------------------------------------- merge round
void oracle_merge_round_i32(const int32_t* in, int32_t* out, int64_t n,
                            int64_t K, int64_t L) {
  std::memcpy(out, in, (size_t)n * sizeof(int32_t));
  for (int64_t t = 0; t < n; t += K * L)
    std::stable_sort(out + t, out + std::min(n, t + K * L));
}

// ------------------------------------------------------------------- co-rank
void oracle_corank_i32(const int32_t* base, int64_t stride, const int64_t* lens,
                       int K, int64_t r, int64_t* s) {
  for (int k = 0; k < K; ++k) s[k] = 0;
  for (int64_t step = 0; step < r; ++step) {
    int best = -1;
    for (int k = 0; k < K; ++k) {
      if (s[k] >= lens[k]) continue;
      // strict (key, list) order: a lower list index wins ties
      if (best < 0 || base[k * stride + s[k]] < base[best * stride + s[best]]) best = k;
    }
    if (best < 0) break;
    ++s[best];
  }
}

// ------------------------------------------------------------- CPU baseline
int oracle_time_sort_i32_runs(const int32_t* a, int64_t n, int threads, int warmups, int reps, double* secs) {
  std::vector<int32_t> buf((size_t)n);
  for (int it = 0; it < warmups + reps; ++it) {
    std::memcpy(buf.data(), a, (size_t)n * sizeof(int32_t));  // a fresh copy per run, not timed
    auto t0 = std::chrono::steady_clock::now();
    if (threads == 1) std::sort(buf.begin(), buf.end());
    else oracle_psort_i32(buf.data(), n, threads);
    auto t1 = std::chrono::steady_clock::now();
    if (it >= warmups) secs[it - warmups] = std::chrono::duration<double>(t1 - t0).count();
  }
  return std::is_sorted(buf.begin(), buf.end()) ? 0 : 1;
}

double oracle_time_sort_i32(const int32_t* a, int64_t n, int threads, int reps) {
  std::vector<int32_t> buf((size_t)n);
  std::vector<double> t;
  for (int it = 0; it <= reps; ++it) {
    std::memcpy(buf.data(), a, (size_t)n * sizeof(int32_t));
    auto t0 = std::chrono::steady_clock::now();
    if (threads == 1) std::sort(buf.begin(), buf.end());
    else oracle_psort_i32(buf.data(), n, threads);
    auto t1 = std::chrono::steady_clock::now();
    if (it > 0) t.push_back(std::chrono::duration<double>(t1 - t0).count());
  }
  std::sort(t.begin(), t.end());
  return t.empty() ? 0.0 : t[t.size() / 2];
}

int oracle_hw_threads(void) { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
