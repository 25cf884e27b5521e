/*
 * multimergesort_custom.hxx -- caller-defined orders for the drop-in entry
 * point: multimergesort<T, f>(input, output, host_array, threadBlocks, N)
 * with any comparator f, as the reference's src/cmp.hxx:35-46 invites ("Edit
 * this to be whatever comparison function is needed": an L1-norm order, an
 * exact fraction order, a key/value struct compared by key, :41-45, :59-73).
 *
 * Header-only HIP: include this header instead of multimergesort.hxx and
 * compile the translation unit with hipcc for gfx950. The built-in orders
 * (cmp<DATATYPE>, greater<DATATYPE>) still run libgmms.so; any other
 * comparator instantiates the kernels below with it. A comparator is a
 * __host__ __device__ function int f(T a, T b) returning nonzero when a
 * strictly precedes b (a strict weak order; ties need not be bitwise
 * equal), or, through gpu_mms::multimergesort_custom, any callable object
 * with that meaning and any trivially copyable T.
 *
 * The sort is stable: elements that the order does not distinguish keep
 * their input order, so the result equals std::stable_sort with the same
 * comparator, bit for bit. The round schedule is the reference's
 * (src/multimergesort.hxx:56-141): 1024-element base tiles, then rounds that
 * merge K lists of length L into runs of K*L, the buffers ping-ponging and
 * the returned pointer naming the one that holds the result.
 *
 *   base case  one 256-lane workgroup per 1024-element tile: a bitonic sort
 *              in LDS of (element, input index) under "a precedes b, or
 *              neither precedes the other and a's index is smaller" -- a
 *              strict total order, so the network's result is the stable one;
 *              the ragged last tile treats missing slots as after everything.
 *   round      every element computes its output rank in its task: its
 *              position in its list plus, in every other list of the task,
 *              the number of elements placed before it (binary searches: in
 *              a lower list the elements not after it, in a higher list the
 *              elements strictly before it), and is written there. This is
 *              the K-way merge with ties broken by (list, position), i.e.
 *              stable, without a partition step or workspace.
 *
 * Errors return nullptr (multimergesort_status() tells why), as in
 * multimergesort.hxx. Synchronous, like the reference.
 */
#ifndef GPU_MMS_MULTIMERGESORT_CUSTOM_HXX
#define GPU_MMS_MULTIMERGESORT_CUSTOM_HXX

#ifdef GPU_MMS_MULTIMERGESORT_HXX
#error "include gpu_mms/multimergesort_custom.hxx instead of (not after) gpu_mms/multimergesort.hxx"
#endif

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#define GPU_MMS_CUSTOM_ORDERS 1
#include "multimergesort.hxx"

namespace gpu_mms {
namespace custom {

constexpr int TILE = 1024;  // the reference's M
constexpr int TILE_THREADS = 256;

// a precedes b under `less`, ties by input index; idx >= n marks a missing
// slot of the ragged last tile, after every real element
template <class T, class Less>
__device__ __forceinline__ bool before(const T& a, int64_t ia, const T& b, int64_t ib, int64_t n, const Less& less) {
  if (ia >= n || ib >= n) return ia < ib;  // missing slots last (their indexes are the largest)
  if (less(a, b)) return true;
  if (less(b, a)) return false;
  return ia < ib;
}

// Base case: sort each 1024-element tile of data[0, n) in place, stably.
template <class T, class Less>
__global__ __launch_bounds__(TILE_THREADS) void k_tile_sort(T* data, int64_t n, Less less) {
  __shared__ T v[TILE];
  __shared__ int64_t ix[TILE];
  const int64_t base = (int64_t)blockIdx.x * TILE;
  for (int t = threadIdx.x; t < TILE; t += TILE_THREADS) {
    const int64_t g = base + t;
    ix[t] = g;
    if (g < n) v[t] = data[g];
  }
  __syncthreads();
  // bitonic sort: block size `size`, stage distance `d`; pair (i, i ^ d) with
  // i the lower index, ascending where (i & size) == 0
  for (int size = 2; size <= TILE; size <<= 1) {
    for (int d = size >> 1; d > 0; d >>= 1) {
      for (int p = threadIdx.x; p < TILE / 2; p += TILE_THREADS) {
        const int i = 2 * p - (p & (d - 1));  // lower element of the pair
        const int j = i + d;
        const bool up = (i & size) == 0;
        const bool jfirst = before(v[j], ix[j], v[i], ix[i], n, less);
        if (jfirst == up) {
          const T tv = v[i];
          v[i] = v[j];
          v[j] = tv;
          const int64_t ti = ix[i];
          ix[i] = ix[j];
          ix[j] = ti;
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < TILE; t += TILE_THREADS) {
    const int64_t g = base + t;
    if (g < n) data[g] = v[t];
  }
}

// count of list[0, len) elements that come before x: elements not after x
// (inclusive: a lower list wins ties) or strictly before x
template <class T, class Less>
__device__ __forceinline__ int64_t count_before(const T* list, int64_t len, const T& x, bool inclusive,
                                                const Less& less) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    const T y = list[mid];
    const bool b = inclusive ? !less(x, y) : less(y, x);
    lo = b ? mid + 1 : lo;
    hi = b ? hi : mid;
  }
  return lo;
}

// One round: runs of L elements -> runs of kway*L (kway lists per task),
// each element written at its rank in its task. (Parameter names avoid the
// params.h knob macros K, M, W, B, PL.)
template <class T, class Less>
__global__ __launch_bounds__(256) void k_rank_merge(const T* in, T* out, int64_t n, int64_t L, int kway, Less less) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t TK = (int64_t)kway * L;
  const int64_t tb = i / TK * TK;
  const int k = (int)((i - tb) / L);
  const int64_t pos = i - tb - (int64_t)k * L;
  const T x = in[i];
  int64_t rank = pos;
  for (int q = 0; q < kway; ++q) {
    if (q == k) continue;
    const int64_t s = tb + (int64_t)q * L;
    if (s >= n) break;
    const int64_t len = n - s < L ? n - s : L;
    rank += count_before(in + s, len, x, q < k, less);
  }
  out[tb + rank] = x;
}

template <class T, class Less>
T* sort(T* input, T* output, int64_t N, const Less& less, int kway, hipStream_t stream) {
  static_assert(std::is_trivially_copyable<T>::value, "elements are moved as bytes");
  gpu_mms::last_status_storage(GMMS_SUCCESS);
  if (N <= 0) return input;
  if (!input || !output || kway < 2 || kway > 64 || (kway & (kway - 1))) {
    gpu_mms::last_status_storage(GMMS_ERROR_INVALID_VALUE);
    return nullptr;
  }
  const int64_t tiles = (N + TILE - 1) / TILE;
  hipLaunchKernelGGL((k_tile_sort<T, Less>), dim3((unsigned)tiles), dim3(TILE_THREADS), 0, stream, input, N, less);
  hipError_t e = hipGetLastError();
  T* cur = input;
  T* nxt = output;
  for (int64_t L = TILE; L < N && e == hipSuccess; L *= kway) {
    hipLaunchKernelGGL((k_rank_merge<T, Less>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream, cur, nxt,
                       N, L, kway, less);
    e = hipGetLastError();
    T* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    gpu_mms::last_status_storage(GMMS_ERROR_HIP);
    return nullptr;
  }
  return cur;
}

// a comparator function pointer (the reference's fptr_t) as a callable
template <class T, int (*F)(T, T)>
struct fptr_less {
  __host__ __device__ bool operator()(const T& a, const T& b) const { return F(a, b) != 0; }
};

}  // namespace custom

// Sort input[0, N) by `less` (a callable: less(a, b) true when a strictly
// precedes b), stably, K lists per merge round (params.h's K by default).
// Same buffer contract as multimergesort: both buffers are clobbered and the
// result is in the returned one; nullptr on error.
template <class T, class Less>
T* multimergesort_custom(T* input, T* output, int64_t N, Less less, int kway = K, hipStream_t stream = nullptr) {
  return custom::sort<T, Less>(input, output, N, less, kway, stream);
}

template <class T, int (*F)(T, T)>
T* custom_order_sort(T* input, T* output, int64_t N) {
  return custom::sort<T, custom::fptr_less<T, F>>(input, output, N, custom::fptr_less<T, F>{}, K, nullptr);
}

}  // namespace gpu_mms

#endif
