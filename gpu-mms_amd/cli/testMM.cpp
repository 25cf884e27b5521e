// testMM -- command-line harness with the reference's interface and output
// (src/testMM.cu:49-63,133-218): `./testMM N [iters]` builds N keys
// rand() % RANGE after srand(0) (src/buildData.h:54-65), copies them to the
// GPU, times multimergesort<int,cmp> with HIP events (iters runs, like ITERS),
// prints "avg min max" in ms (testMM.cu:192), copies the result back and
// prints "SORTED!" or "NOT SORTED! Item at index i is less than its
// predecessor." (testMM.cu:198-213). Unlike the reference it also checks the
// output is the input's permutation (std::sort), and says so on a third line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpu_mms/multimergesort.hxx"

static void create_random_list(int* data, long n) {
  for (long i = 0; i < n; ++i) {
    data[i] = rand() % RANGE;
    long temp = rand() % RANGE;  // the reference's second draw (value half of a pair)
    (void)temp;
  }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("Please provide size of input N. Usage: ./testMM <N> [iters]\n");
    return 1;
  }
  const long N = std::atol(argv[1]);
  const int iters = argc > 2 ? std::atoi(argv[2]) : 1;
  if (N < 1 || iters < 1) {
    std::printf("N and iters must be positive\n");
    return 1;
  }
  std::vector<int> h(N), ref(N), out(N);
  srand(0);
  int *d_data = nullptr, *d_output = nullptr;
  if (hipMalloc(&d_data, N * sizeof(int)) != hipSuccess || hipMalloc(&d_output, N * sizeof(int)) != hipSuccess) {
    std::printf("hipMalloc failed\n");
    return 2;
  }
  hipEvent_t start, stop;
  (void)hipEventCreate(&start);
  (void)hipEventCreate(&stop);
  float total = 0.f, mn = 1e30f, mx = 0.f;
  int* result = nullptr;
  for (int it = 0; it < iters; ++it) {
    create_random_list(h.data(), N);
    if (it == iters - 1) ref = h;
    (void)hipMemcpy(d_data, h.data(), N * sizeof(int), hipMemcpyHostToDevice);
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(start, 0);
    result = multimergesort<int, cmp<int>>(d_data, d_output, h.data(), BLOCKS, N);
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(stop, 0);
    (void)hipEventSynchronize(stop);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, start, stop);
    if (!result) {
      std::printf("multimergesort failed: status %d (%s) %s\n", multimergesort_status(),
                  gmms_status_string(multimergesort_status()), gmms_last_error());
      return 3;
    }
    total += ms;
    mn = std::min(mn, ms);
    mx = std::max(mx, ms);
  }
  std::printf("%lf %lf %lf\n", (double)(total / iters), (double)mn, (double)mx);
  (void)hipMemcpy(out.data(), result, N * sizeof(int), hipMemcpyDeviceToHost);
  long bad = -1;
  for (long i = 1; i < N; ++i) {
    if (host_cmp(out[i], out[i - 1])) {
      bad = i;
      break;
    }
  }
  if (bad >= 0) std::printf("NOT SORTED! Item at index %ld is less than its predecessor.\n", bad);
  else std::printf("SORTED!\n");
  std::sort(ref.begin(), ref.end());
  std::printf(ref == out ? "MATCHES std::sort\n" : "DIFFERS from std::sort\n");
  (void)hipFree(d_data);
  (void)hipFree(d_output);
  return bad >= 0 || ref != out ? 4 : 0;
}
