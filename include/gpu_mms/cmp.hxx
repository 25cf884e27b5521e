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
 * The GPU never calls these functions: libgmms.so runs the built-in order of
 * the element type (integers by '<', floats by IEEE total order, pairs by
 * (key, value)), or its exact reverse. multimergesort.hxx accepts cmp<DATATYPE>
 * and greater<DATATYPE> as its comparator argument and rejects any other at
 * compile time, so a caller can never get an order it did not ask for.
 * Unlike the reference, MAXVAL is not a sentinel: keys equal to MAXVAL sort
 * like any other key.
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
