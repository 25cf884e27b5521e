/*
 * multimergesort.hxx -- the drop-in entry point of the reference,
 * multimergesort<T,cmp>(input, output, host_array, threadBlocks, N)
 * (README.md:16-25) and its 4-argument code form
 * (src/multimergesort.hxx:56-57), as host C++ over the C ABI of libgmms.so.
 *
 * Contract kept from the reference: `input` and `output` are device arrays of
 * at least N elements, both are clobbered, and the returned pointer is the
 * one holding the sorted data (the base case sorts `input` in place and the
 * merge rounds ping-pong, src/multimergesort.hxx:128,140). The call is
 * synchronous and allocates its partition workspace itself, like the
 * reference (:61-63,137-138). host_array is unused, as on the reference's
 * live path. On error it returns nullptr (the reference cannot fail
 * visibly); gmms_last_error() names the HIP failure.
 *
 * Callers that need a stream, a caller-owned workspace, pairs or 64-bit N use
 * gmms_sort() from gmms.h directly.
 *
 * Link with -lgmms (gpu-mms_amd/gmms/libgmms.so); no HIP compiler is needed
 * to use this header.
 */
#ifndef GPU_MMS_MULTIMERGESORT_HXX
#define GPU_MMS_MULTIMERGESORT_HXX

#include <cstdint>
#include <type_traits>

#include "cmp.hxx"
#include "gmms.h"
#include "params.h"

namespace gpu_mms {

template <typename T>
struct dtype_of;
template <> struct dtype_of<int32_t> { static constexpr int value = GMMS_INT32; };
template <> struct dtype_of<uint32_t> { static constexpr int value = GMMS_UINT32; };
template <> struct dtype_of<float> { static constexpr int value = GMMS_FLOAT32; };
template <> struct dtype_of<int64_t> { static constexpr int value = GMMS_INT64; };
template <> struct dtype_of<uint64_t> { static constexpr int value = GMMS_UINT64; };
template <> struct dtype_of<double> { static constexpr int value = GMMS_FLOAT64; };

// Which built-in order a comparator names: 0 cmp<DATATYPE> (ascending),
// 1 greater<DATATYPE> (descending), -1 anything else.
template <fptr_t F> struct order_of { static constexpr int value = -1; };
template <> struct order_of<&cmp<DATATYPE>> { static constexpr int value = 0; };
template <> struct order_of<&greater<DATATYPE>> { static constexpr int value = 1; };

inline int last_status_storage(int set = -1) {
  static thread_local int s = GMMS_SUCCESS;
  if (set >= 0) s = set;
  return s;
}

// The knobs of params.h, checked as the reference's build assumes them
// (PL = log2 K, src/params.h:19-20): K a power of two in 2..64.
static_assert(K >= 2 && K <= 64 && (K & (K - 1)) == 0, "params.h: K must be a power of two in 2..64");
static_assert((1 << (PL)) == K, "params.h: PL must be log2(K)");
static_assert(MERGE_KERNEL >= 0 && MERGE_KERNEL <= 2, "params.h: MERGE_KERNEL is 0, 1 or 2");

// The gmms_params the drop-in template runs with: params.h's K, merge
// kernel (PIPELINE / MERGE_KERNEL) and BLOCKS, or threadBlocks when > 0.
inline gmms_params drop_in_params(int threadBlocks, int descending = 0) {
  gmms_params p;
  gmms_get_default_params(&p);
  p.kway = K;
  p.pipeline = MERGE_KERNEL;
  p.mergers = threadBlocks > 0 ? threadBlocks : BLOCKS;
  p.descending = descending;
  return p;
}

}  // namespace gpu_mms

/* Status of the calling thread's last multimergesort call. */
inline int multimergesort_status() { return gpu_mms::last_status_storage(); }

#ifdef GPU_MMS_CUSTOM_ORDERS
namespace gpu_mms {
// defined in multimergesort_custom.hxx (hipcc): the sort under a caller order
template <class T, int (*F)(T, T)>
T* custom_order_sort(T* input, T* output, int64_t N);
}  // namespace gpu_mms
#endif

template <typename T, fptr_t f>
T* multimergesort(T* input, T* output, T* h_data, int threadBlocks, int64_t N) {
  static_assert(std::is_same<T, DATATYPE>::value, "T must be DATATYPE (cmp.hxx)");
#ifdef GPU_MMS_CUSTOM_ORDERS
  if constexpr (gpu_mms::order_of<f>::value < 0) {
    // a caller-defined order: the header-only kernels of multimergesort_custom.hxx
    (void)h_data;
    (void)threadBlocks;
    return gpu_mms::custom_order_sort<T, f>(input, output, N);
  } else {
#else
  static_assert(gpu_mms::order_of<f>::value >= 0,
                "libgmms.so runs cmp<DATATYPE> (a < b) or greater<DATATYPE> (a > b); for another order "
                "include gpu_mms/multimergesort_custom.hxx instead and compile with hipcc");
  {
#endif
  void* result = nullptr;
  const gmms_params p = gpu_mms::drop_in_params(threadBlocks, gpu_mms::order_of<f>::value);
  const int st = gmms_multimergesort_params(gpu_mms::dtype_of<T>::value, input, output, h_data, &p, N, &result);
  gpu_mms::last_status_storage(st);
  return st == GMMS_SUCCESS ? static_cast<T*>(result) : nullptr;
  }
}

template <typename T, fptr_t f>
T* multimergesort(T* input, T* output, T* h_data, int N) {
  return multimergesort<T, f>(input, output, h_data, BLOCKS, (int64_t)N);
}

#endif
