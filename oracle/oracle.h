/*
 * GPU-MMS oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference multiway mergesort's observable
 * behaviour (algoparc/GPU-MMS). Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the reported CPU baseline -- never as the product path.
 *
 * Every function below cites the reference file:line whose contract it
 * restates. The sort itself is pinned by a closed form (the output is the
 * input permutation ordered by the comparator, unique once ties are broken
 * by the full element bits), so std::sort with the same comparator is the
 * bit-exact oracle; the reference's own GPU output is not (scripts/testAll.txt
 * records 45/1257 "NOT SORTED!" runs).
 *
 * Pins (see tests/test_oracle.py):
 *   gen_*          : mt19937_64 10000th-output KAT from the C++ standard;
 *                    glibc rand() stream compared against libc itself.
 *   sort_*         : numpy.sort / numpy.lexsort on the same bits; brute force
 *                    over all permutations for n <= 7; float total-order
 *                    special values (-nan < -inf < -0 < +0 < +inf < +nan).
 *   schedule       : closed forms checked against the integer round loop and
 *                    the sizes quoted for the reference (R=3 at 2^28, K=64).
 *   corank         : brute-force stable merge in Python (heapq.merge).
 */
#ifndef GMMS_ORACLE_H
#define GMMS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators (tests and bench inputs; bits are portable) ---------------
 * mt19937_64 with an explicit seed; values derived with bit operations only
 * (std::uniform_*_distribution is implementation-defined). */
uint64_t oracle_mt64_nth(uint64_t seed, uint64_t nth);          /* nth output, 1-based */
void oracle_gen_u64(uint64_t seed, int64_t n, uint64_t* out);   /* raw 64-bit outputs */
/* int32 keys = (int32)(u64 >> shift) + offset-free: shift 32 -> full int32
 * range, 34 -> [0,2^30), 54 -> [0,1024). */
void oracle_gen_i32(uint64_t seed, int64_t n, int shift, int32_t* out);
/* float32 normal(0,1) via Box-Muller on 53-bit doubles, rounded to float. */
void oracle_gen_f32_normal(uint64_t seed, int64_t n, float* out);
/* Key distributions for (u64 key, u32 value) pairs: 0 uniform, 1 sorted,
 * 2 reverse sorted, 3 Zipf(s) over `distinct` values (own inverse-CDF table,
 * rank mapped through a fixed odd-multiplier bijection). value = index. */
void oracle_gen_pairs(uint64_t seed, int64_t n, int dist, double zipf_s,
                      int64_t distinct, uint64_t* keys, uint32_t* vals);
/* Segment lengths floor(2^(lo + (hi-lo)*U)), U uniform in [0,1). */
void oracle_gen_seglens(uint64_t seed, int64_t count, double lo_log2,
                        double hi_log2, int64_t* lens);
/* The reference harness generator, src/buildData.h:54-65 with
 * src/testMM.cu:156 (srand(0)) -- glibc rand() % range. Restated from
 * glibc's published TYPE_3 additive feedback generator (random_r.c). */
void oracle_glibc_rand_list(uint32_t seed, int64_t n, int32_t range, int32_t* out);

/* ---- comparators (src/cmp.hxx:37-54) --------------------------------------
 * float32 total order: key(u) = u ^ ((int32)u >> 31 | 0x80000000), i.e.
 * negatives reversed, -0 before +0, NaNs ordered by sign and payload. */
uint32_t oracle_f32_order_key(uint32_t bits);
uint64_t oracle_f64_order_key(uint64_t bits);

/* ---- the sort oracle: std::sort with the comparator, in place ------------- */
void oracle_sort_i32(int32_t* a, int64_t n);
void oracle_sort_u32(uint32_t* a, int64_t n);
void oracle_sort_f32(float* a, int64_t n);       /* total order on bits */
void oracle_sort_i64(int64_t* a, int64_t n);
void oracle_sort_u64(uint64_t* a, int64_t n);
void oracle_sort_f64(double* a, int64_t n);      /* total order on bits */
/* (key, value) lexicographic; keys/vals permuted together. */
void oracle_sort_pairs(uint64_t* keys, uint32_t* vals, int64_t n);
/* parallel std::sort baseline (libstdc++ parallel mode), threads <= 0: all */
void oracle_psort_i32(int32_t* a, int64_t n, int threads);
void oracle_psort_u64pairs(uint64_t* keys, uint32_t* vals, int64_t n, int threads);
void oracle_psort_u64(uint64_t* a, int64_t n, int threads);
void oracle_psort_f32(float* a, int64_t n, int threads);        /* total order on bits */

/* ---- base case (src/basecase/squareSort.hxx:62-131) -----------------------
 * Sort every contiguous tile of `tile` elements in place; the ragged tail
 * N mod tile is sorted as its own tile (the reference's serial tail sort,
 * squareSort.hxx:69-88). */
void oracle_basecase_i32(int32_t* a, int64_t n, int64_t tile);

/* ---- schedule (src/multimergesort.hxx:87-90, warp-partition.hxx:526-531) ---
 * rounds = number of iterations of `for (L = tile; L < n; L *= K)`.
 * For round r (0-based): list length L = tile*K^r, tasks = ceil(n / (K*L)),
 * edge task size = n mod (K*L) or K*L. */
int     oracle_rounds(int64_t n, int64_t K, int64_t tile);
int64_t oracle_tasks(int64_t n, int64_t K, int64_t L);
int64_t oracle_edge_task_size(int64_t n, int64_t K, int64_t L);
/* Merger target ranks inside a task, src/warp-partition.hxx:234:
 * target(j) = j * taskSize / mergersPerTask (integer division). */
int64_t oracle_merger_target(int64_t j, int64_t task_size, int64_t mergers);

/* ---- one merge round (src/multimergesort.hxx:87-135) ----------------------
 * in: data whose consecutive runs of length L are sorted; out: data whose
 * runs of K*L are sorted (each task = stable K-way merge of its lists). */
void oracle_merge_round_i32(const int32_t* in, int32_t* out, int64_t n,
                            int64_t K, int64_t L);

/* ---- co-rank (src/warp-partition.hxx:179-366 contract) --------------------
 * K sorted int32 lists given by (base, lens[k]) laid out back to back at
 * stride `stride`. For rank r returns s[k] = number of elements of list k
 * among the r smallest under the stable order (key, list index, position).
 * Computed by a sequential K-way merge. */
void oracle_corank_i32(const int32_t* base, int64_t stride, const int64_t* lens,
                       int K, int64_t r, int64_t* s);

/* ---- CPU baseline timing (reported, not the target) ----------------------
 * Returns median seconds of `reps` runs (after one warm-up) of std::sort
 * (threads == 1) or parallel sort on a fresh copy of `a`. */
double oracle_time_sort_i32(const int32_t* a, int64_t n, int threads, int reps);
/* The same with every run's seconds: `warmups` untimed runs, then `reps`
 * timed ones into secs[0..reps); returns 0 when the last run is sorted. */
int    oracle_time_sort_i32_runs(const int32_t* a, int64_t n, int threads, int warmups, int reps,
                                 double* secs);
int    oracle_hw_threads(void);

#ifdef __cplusplus
}
#endif
#endif
