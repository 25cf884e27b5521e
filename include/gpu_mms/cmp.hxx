/*
 * cmp.hxx -- the comparator contract of the drop-in template, after the
 * reference's src/cmp.hxx:21-54:
 *
 *   DATATYPE  element type being sorted (default int)
 *   CASTTYPE  type an element travels as between lanes (here: the element's
 *             32-bit word planes; kept for source compatibility)
 *   MAXVAL / MINVAL  the largest / smallest key of DATATYPE
 *   fptr_t    comparator type, int (*)(DATATYPE, DATATYPE)
 *   cmp<T>    "a strictly precedes b": a < b
 *   host_cmp<T>  the same order for host-side checks
 *   greater<T>, host_greater<T>  the reversed order, a > b (descending)
 *
 * cmp<DATATYPE> and greater<DATATYPE> name libgmms.so's built-in orders (integers
 * by '<', floats by IEEE total order, pairs by (key, value), or the exact
 * reverse); the GPU never calls them. Any other order -- the reference's
 * "edit this to be whatever comparison function is needed" -- is a
 * comparator of the caller's own, a __host__ __device__ function
 * int f(DATATYPE a, DATATYPE b) that is nonzero when a strictly precedes b,
 * passed as multimergesort<DATATYPE, f>: include multimergesort_custom.hxx
 * and compile with hipcc, and the header instantiates its kernels with it
 * (stable: the result equals std::stable_sort with f). Through
 * multimergesort.hxx alone such a comparator fails to compile, so a caller
 * never silently gets an order it did not ask for. Editing cmp below does not
 * change the library's order. Unlike the reference, MAXVAL is not a
 * sentinel: keys equal to MAXVAL sort like any other key.
 */
#ifndef GPU_MMS_CMP_HXX
#define GPU_MMS_CMP_HXX

#include <limits>

#ifndef DATATYPE
#define DATATYPE int
#endif
#ifndef CASTTYPE
#define CASTTYPE DATATYPE
#endif
#define MAXVAL (std::numeric_limits<DATATYPE>::max())
#define MINVAL (std::numeric_limits<DATATYPE>::lowest())

typedef int (*fptr_t)(DATATYPE, DATATYPE);

template <typename T>
inline int cmp(T a, T b) {
  return a < b;
}

template <typename T>
inline int host_cmp(T a, T b) {
  return a < b;
}

template <typename T>
inline int greater(T a, T b) {
  return b < a;
}

template <typename T>
inline int host_greater(T a, T b) {
  return b < a;
}

#endif
