// custom_orders -- sorts under the example comparators of the reference's
// src/cmp.hxx (an L1-norm order on the low 32 bits, :41-42; an exact fraction
// order on (numerator << 32 | denominator), :44-45; a key/value struct
// compared by key, :59-73) plus a plain descending order written by hand,
// through the drop-in multimergesort<T, f> and gpu_mms::multimergesort_custom
// (gpu_mms/multimergesort_custom.hxx, built with hipcc), and checks each
// result against std::stable_sort with the same comparator, bit for bit.
//
// usage: custom_orders N [example]   example: l1, frac, kv, desc (default all)
// Prints one line per example: "<name> N <ms> MATCHES std::stable_sort" (or
// DIFFERS ... at index i); exits nonzero on any difference.
#define DATATYPE int64_t
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "gpu_mms/multimergesort_custom.hxx"

// src/cmp.hxx:41-42: "L1-norm": compare the low 32 bits as signed ints
__host__ __device__ int l1_cmp(int64_t a, int64_t b) { return ((int)a) < ((int)b); }
// src/cmp.hxx:44-45: fractions a = num << 32 | den, compared without division
__host__ __device__ int frac_cmp(int64_t a, int64_t b) {
  return ((int)(a >> 32) * (int)b) < ((int)(b >> 32) * (int)a);
}
// a descending order written as a caller comparator (not greater<>)
__host__ __device__ int desc_cmp(int64_t a, int64_t b) { return a > b; }

// src/cmp.hxx:59-73: a key/value pair sorted by key
struct KV {
  int key;
  int val;
};
struct KeyLess {
  __host__ __device__ bool operator()(const KV& a, const KV& b) const { return a.key < b.key; }
};

template <class T, class HostLess, class Sort>
static bool run(const char* name, std::vector<T>& h, HostLess hl, Sort sorter) {
  const long n = (long)h.size();
  T *d_in = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_in, n * sizeof(T)) != hipSuccess || hipMalloc(&d_out, n * sizeof(T)) != hipSuccess) {
    std::printf("%s: hipMalloc failed\n", name);
    return false;
  }
  (void)hipMemcpy(d_in, h.data(), n * sizeof(T), hipMemcpyHostToDevice);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  (void)hipEventRecord(a, 0);
  T* res = sorter(d_in, d_out, n);
  (void)hipEventRecord(b, 0);
  (void)hipEventSynchronize(b);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  if (!res) {
    std::printf("%s %ld failed: status %d\n", name, n, multimergesort_status());
    return false;
  }
  std::vector<T> got(n);
  (void)hipMemcpy(got.data(), res, n * sizeof(T), hipMemcpyDeviceToHost);
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  std::vector<T> want(h);
  std::stable_sort(want.begin(), want.end(), hl);
  for (long i = 0; i < n; ++i) {
    if (std::memcmp(&got[i], &want[i], sizeof(T)) != 0) {
      std::printf("%s %ld %.3f DIFFERS from std::stable_sort at index %ld\n", name, n, (double)ms, i);
      return false;
    }
  }
  std::printf("%s %ld %.3f MATCHES std::stable_sort\n", name, n, (double)ms);
  return true;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: custom_orders N [l1|frac|kv|desc]\n");
    return 1;
  }
  const long n = std::atol(argv[1]);
  const std::string which = argc > 2 ? argv[2] : "all";
  if (n < 1) return 1;
  std::mt19937_64 g(0x6D6D7330 + (uint64_t)n);
  bool ok = true;
  if (which == "all" || which == "l1") {
    std::vector<int64_t> h(n);
    for (auto& x : h) x = (int64_t)((g() & ~0xFFFFFFFFull) | (uint32_t)(int32_t)((int64_t)(g() % 8192) - 4096));
    ok &= run("l1", h, [](int64_t a, int64_t b) { return l1_cmp(a, b) != 0; },
              [](int64_t* i, int64_t* o, long m) { return multimergesort<DATATYPE, l1_cmp>(i, o, nullptr, m); });
  }
  if (which == "all" || which == "frac") {
    std::vector<int64_t> h(n);
    for (auto& x : h) x = (int64_t)(((g() % 1024) << 32) | (1 + g() % 1023));
    ok &= run("frac", h, [](int64_t a, int64_t b) { return frac_cmp(a, b) != 0; },
              [](int64_t* i, int64_t* o, long m) { return multimergesort<DATATYPE, frac_cmp>(i, o, nullptr, m); });
  }
  if (which == "all" || which == "kv") {
    std::vector<KV> h(n);
    for (long i = 0; i < n; ++i) h[i] = {(int)(g() % 1024), (int)i};
    ok &= run("kv", h, KeyLess{},
              [](KV* i, KV* o, long m) { return gpu_mms::multimergesort_custom(i, o, (int64_t)m, KeyLess{}); });
  }
  if (which == "all" || which == "desc") {
    std::vector<int64_t> h(n);
    for (auto& x : h) x = (int64_t)g();
    ok &= run("desc", h, [](int64_t a, int64_t b) { return desc_cmp(a, b) != 0; },
              [](int64_t* i, int64_t* o, long m) { return multimergesort<DATATYPE, desc_cmp>(i, o, nullptr, m); });
  }
  return ok ? 0 : 4;
}
