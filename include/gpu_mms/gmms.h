/*
 * gmms.h -- C ABI of the MI355X-native multiway mergesort (libgmms.so).
 *
 * Plain pointers and sizes only: device pointers are HIP device memory on the
 * current device, `stream` is a hipStream_t passed as void* (NULL = the null
 * stream). Every entry point returns a gmms_status; none prints or aborts.
 *
 * Each declaration names the reference interface it replaces
 * (algoparc/GPU-MMS, paths relative to the reference repository root). The
 * reference has no FFI of its own: its boundary is the header-only template
 * multimergesort<T,f> (src/multimergesort.hxx:56-57, documented with a
 * threadBlocks argument at README.md:16-25). These functions are what a
 * foreign caller (ctypes, cgo, JNI) binds instead; the C++ template in
 * multimergesort.hxx is built on top of them.
 */
#ifndef GPU_MMS_GMMS_H
#define GPU_MMS_GMMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference returns nothing and only prints under DEBUG
 * (src/multimergesort.hxx:102-108,129-134). */
typedef enum gmms_status {
  GMMS_SUCCESS = 0,
  GMMS_ERROR_INVALID_VALUE = 1, /* bad pointer, size, dtype or parameter */
  GMMS_ERROR_UNSUPPORTED = 2,   /* valid request the library does not run */
  GMMS_ERROR_WORKSPACE = 3,     /* workspace missing or too small         */
  GMMS_ERROR_HIP = 4,           /* a HIP runtime call or launch failed    */
  GMMS_ERROR_OUT_OF_MEMORY = 5, /* device allocation failed               */
  GMMS_ERROR_BUDGET = 6         /* serving: request exceeds the HBM budget */
} gmms_status;

/* Element types with their built-in comparators. Replaces the compile-time
 * DATATYPE / cmp<T> pair of src/cmp.hxx:21-46: integers order by '<',
 * floats by IEEE total order (-NaN < -inf < ... < -0 < +0 < ... < +inf <
 * +NaN), pairs by (key, value). */
typedef enum gmms_dtype {
  GMMS_INT32 = 0,
  GMMS_UINT32 = 1,
  GMMS_FLOAT32 = 2,
  GMMS_INT64 = 3,
  GMMS_UINT64 = 4,
  GMMS_FLOAT64 = 5,
  GMMS_PAIR_U64_U32 = 6 /* keys: uint64_t[n], values: uint32_t[n] */
} gmms_dtype;

/* Run-time knobs. Replaces the compile-time macros of src/params.h:18-38:
 * K (branching factor, a power of two, 2..64), PIPELINE (which K-way merge
 * kernel runs a round) and BLOCKS (= threadBlocks, the merger count when a
 * round has few tasks). Zero fields take the defaults of
 * gmms_get_default_params. */
enum {
  GMMS_MERGE_WAVEFRONT = 0, /* tile merge, one merger per wavefront (default) */
  GMMS_MERGE_MINBLOCKHEAP = 1, /* the reference's PIPELINE minBlockHeap:
                                  one wavefront per heap level            */
  GMMS_MERGE_WORKGROUP = 2  /* tile merge, one merger per workgroup        */
};
typedef struct gmms_params {
  int32_t kway;      /* K: lists merged per round; 0: auto (16 for the
                        wavefront and heap mergers; for the workgroup tile
                        merger the largest power of two <= 64 that fits LDS) */
  int32_t pipeline;  /* merge kernel, GMMS_MERGE_*                           */
  int32_t mergers;   /* mergers that run at once when tasks are few: a
                        round splits its tasks into whole multiples of it
                        (up to 4x, to within 10 %); 0: auto, the resident
                        count of the merge kernel                          */
  int32_t prefetch;  /* blocks prefetched per list into LDS (1..8)           */
  int32_t slack;     /* partition tolerance in elements; 0: auto             */
  int32_t base_run;  /* sorted run length the base case leaves: 1024 (one
                        wavefront tile, the reference's M) or 4096 (four tiles
                        merged in a workgroup); 0: auto (4096)               */
  int32_t descending; /* 1: reverse order (a custom cmp(a,b) = a > b, as the
                         reference's cmp.hxx allows); 0: ascending          */
  int32_t reserved[1];
} gmms_params;

/* Library version string, e.g. "gpu-mms-amd 0.1.0 gfx950". */
const char* gmms_version(void);
/* Human-readable name of a status code. */
const char* gmms_status_string(int status);
/* The HIP call and error behind the calling thread's last GMMS_ERROR_HIP
 * ("" if none). */
const char* gmms_last_error(void);
/* Fill *p with the defaults (K auto, wavefront merger, auto mergers). */
int gmms_get_default_params(gmms_params* p);
/* The knobs a sort of `dtype` would actually run with: zero/auto fields of
 * *in (NULL: defaults) resolved (K, merge kernel, merger target). */
int gmms_resolve_params(int dtype, const gmms_params* in, gmms_params* out);
/* Bytes per key and per value (0 for key-only dtypes) of a dtype. */
int gmms_element_size(int dtype, size_t* key_bytes, size_t* value_bytes);
/* Number of merge rounds for n elements: iterations of the reference's
 * `for (listSize = M; listSize < N; listSize *= K)` (multimergesort.hxx:87),
 * M = 1024. */
int gmms_num_rounds(int64_t n, int kway, int* rounds);
/* Merge rounds gmms_sort runs for n elements with these knobs: the same loop
 * from listSize = the resolved base_run (so one round fewer than
 * gmms_num_rounds whenever the 4096-element base runs save one). */
int gmms_sort_rounds(int dtype, int64_t n, const gmms_params* p, int* rounds);

/* Device workspace that gmms_sort needs for n elements: the partition's
 * sample array (every 256th element of a round's input) and its cut rows;
 * 0 when no round splits a task. The reference cudaMalloc's its pivot array
 * inside every call (multimergesort.hxx:61-63,138); here the caller owns it
 * so the timed path never allocates. */
int gmms_workspace_size(int dtype, int64_t n, const gmms_params* p, size_t* bytes);
/* Workspace for one gmms_merge_round over n elements in lists of list_len. */
int gmms_round_workspace_size(int dtype, int64_t n, int64_t list_len, const gmms_params* p,
                              size_t* bytes);

/* Sort n elements. Replaces multimergesort<T,f>(input, output, h_data,
 * threadBlocks, N) (src/multimergesort.hxx:56-141, README.md:16-25) with the
 * same buffer contract: both key buffers (and both value buffers for pairs)
 * are clobbered, the base case sorts `keys` in place and the rounds
 * ping-pong; *result_in_alt is set to 0 if the sorted data ends in
 * keys/vals and 1 if it ends in keys_alt/vals_alt (the reference returns
 * that pointer, :140). vals/vals_alt must be NULL for key-only dtypes.
 * Asynchronous on `stream`; no allocation, no host synchronisation. */
int gmms_sort(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt,
              int64_t n, const gmms_params* p, void* workspace,
              size_t workspace_bytes, void* stream, int* result_in_alt);

/* gmms_sort without a workspace argument and without any allocation. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The workspace is a static device array
 * of GMMS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a
 * sort whose workspace (gmms_workspace_size) exceeds it returns
 * GMMS_ERROR_WORKSPACE (16 MiB with the default knobs on 256 compute units:
 * 32-bit keys up to about 1.03e9 elements, 64-bit keys about 5.1e8, pairs
 * about 3.4e8). Every call on one GPU shares
 * that array, so calls that overlap on one GPU (two streams, two threads)
 * race; serialise them. Otherwise as gmms_sort. */
#define GMMS_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
int gmms_sort_noalloc(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt,
                      int64_t n, const gmms_params* p, void* stream, int* result_in_alt);
/* The static scratch size in bytes (GMMS_NOALLOC_SCRATCH_BYTES). */
int gmms_noalloc_scratch_bytes(size_t* bytes);

/* The reference's entry point with its own argument list
 * (src/multimergesort.hxx:56-57 plus README.md's threadBlocks): sorts the
 * device array `input` of n elements using `output` as the second buffer,
 * allocates and frees its workspace itself and synchronises before it
 * returns, like the reference; *result receives `input` or `output`,
 * whichever holds the sorted data. host_array is accepted and unused, as in
 * the reference (it is a debugging aid there). thread_blocks <= 0: auto. */
int gmms_multimergesort(int dtype, void* input, void* output, void* host_array,
                        int thread_blocks, int64_t n, void** result);
/* The same with every knob of src/params.h:18-38 (K, PIPELINE, BLOCKS as
 * p->mergers) instead of threadBlocks alone; used by multimergesort.hxx. */
int gmms_multimergesort_params(int dtype, void* input, void* output, void* host_array,
                               const gmms_params* p, int64_t n, void** result);

/* Device-to-device copy with 128-bit accesses (bytes and both pointers
 * 16-byte aligned). No counterpart in the reference: it is the stream-copy
 * probe that the HBM roofline of the sort is measured against. */
int gmms_copy(void* dst, const void* src, size_t bytes, void* stream);
/* Variants of that probe, to validate its figure: `unroll` (1, 2, 4, 8)
 * 16-byte loads per lane in flight before their stores, non-temporal stores
 * when `nontemporal` is 1, and at most blocks_per_cu workgroups of 256 lanes
 * per compute unit (grid-stride beyond; 0: one workgroup per stretch of
 * the array, no stride). No counterpart in the reference. */
int gmms_copy_probe(void* dst, const void* src, size_t bytes, int unroll, int nontemporal, int blocks_per_cu,
                    void* stream);

/* ---- per-launch timing (the reference times the whole call with CUDA
 * events, src/testMM.cu:171-186) ---------------------------------------- */
enum { GMMS_STAGE_BASECASE = 0, GMMS_STAGE_PARTITION = 1, GMMS_STAGE_MERGE = 2 };
/* Start (enable=1) or stop (0) recording HIP events around every kernel the
 * library launches, on the launch's own stream; either call clears the list. */
int gmms_stage_timing(int enable);
/* Read the recorded launches (waits for their events): kinds[i] is a
 * GMMS_STAGE_*, ms[i] its duration, elems[i] the elements it covered.
 * *count receives the number recorded (entries beyond cap are skipped). */
int gmms_stage_times(int* kinds, float* ms, int64_t* elems, int cap, int* count);

/* ---- single stages, for tests and for callers that drive rounds ------- */

/* squareSort (src/basecase/squareSort.hxx:62-131): sort every 1024-element
 * tile of keys[0..n) in place, the ragged tail as its own tile. to_internal /
 * to_external select the order-preserving bit maps on load / store (pass 1/1
 * for a plain tile sort). */
int gmms_squaresort(int dtype, void* keys, void* vals, int64_t n, int to_internal,
                    int to_external, void* stream);

/* One merge round (findPartitions + multimergeLevel,
 * src/multimergesort.hxx:89-128): in holds sorted runs of list_len elements;
 * out receives sorted runs of K*list_len. Elements must be in the internal
 * order (as left by gmms_squaresort(..., to_internal=1, to_external=0));
 * to_external=1 maps the output back to the user's bits. */
int gmms_merge_round(int dtype, const void* in_keys, const void* in_vals,
                     void* out_keys, void* out_vals, int64_t n, int64_t list_len,
                     const gmms_params* p, int to_external, void* workspace,
                     size_t workspace_bytes, void* stream);

/* Partition only (findPartitions, src/warp-partition.hxx:487-532): for the
 * round with lists of list_len, writes for every merger b = task*m + j the
 * per-list cut offsets cuts[b*K + k] (consistent: all elements before the
 * cut precede all after it under (key, list, position)) and skips[b] =
 * target rank - rank of the cut, 0 <= skips[b] <= slack. j = 0 entries are
 * written as zero. *mergers_per_task receives m. Internal order. */
int gmms_find_partitions(int dtype, const void* keys, const void* vals, int64_t n,
                         int64_t list_len, const gmms_params* p, int64_t* cuts,
                         int64_t* skips, int64_t cap_mergers,
                         int* mergers_per_task, void* stream);

/* pad (src/pad.hxx:1-4): write `count` copies of the dtype's maximum
 * element (all-ones internal bits) at data[0..count). The sort itself never
 * needs padding (ragged tiles are padded in registers). */
int gmms_pad(int dtype, void* keys, void* vals, int64_t count, void* stream);

/* ------------------------------------------------------------------------
 * Batched segmented sort. No counterpart in the reference, whose driver
 * (src/multimergesort.hxx:56-141) sorts one array per call; this runs that
 * same schedule for many arrays at once. Segment s is
 * keys[offsets[s] .. offsets[s+1]) (offsets non-decreasing, offsets[0] >= 0,
 * n_seg + 1 entries); each segment comes out exactly as gmms_sort would
 * leave it on its own. Every segment uses the K of its size class (the
 * tuning table below); the segments are aligned so that all of them finish
 * in the same last round, and one launch sorts every 1024-element tile of
 * every segment, then each round is one partition launch and one persistent
 * merge launch whose wavefronts take (segment, task, merger) items from a
 * device work queue, largest first. Launch count: 2 + at most 2 per round of
 * the largest segment, independent of n_seg.
 * ------------------------------------------------------------------------ */
typedef struct gmms_seg_plan gmms_seg_plan;
typedef struct gmms_seg_graph gmms_seg_graph;

typedef struct gmms_seg_info {
  int64_t n_seg;            /* segments (empty ones included)                */
  int64_t elements;         /* sum of the segment lengths                    */
  int64_t extent;           /* offsets[n_seg]: buffers hold at least this    */
  int64_t tiles;            /* 1024-element tiles over all segments          */
  int64_t merge_items;      /* mergers over all rounds                       */
  int64_t partition_items;  /* partition searches over all rounds            */
  size_t workspace_bytes;   /* device workspace gmms_seg_sort needs          */
  int32_t rounds;           /* merge rounds (of the largest segment)         */
  int32_t launches;         /* kernel launches per gmms_seg_sort             */
} gmms_seg_info;

/* Plans a batch from host-side offsets (host work only: no HIP call; the
 * first gmms_seg_sort copies the plan's tables into pinned host memory that
 * the plan keeps for its later sorts).
 * p: kway = 0 takes K per size class from the tuning table, kway > 0 forces
 * one K for all segments; p->mergers > 0 sets the resident merger count the
 * rounds are split for. Only the wavefront merge kernel (pipeline 0) runs
 * batches. */
int gmms_seg_plan_create(int dtype, const int64_t* offsets, int64_t n_seg, const gmms_params* p,
                         gmms_seg_plan** plan);
int gmms_seg_plan_info(const gmms_seg_plan* plan, gmms_seg_info* info);
/* K and number of merge rounds chosen for segment `seg`. */
int gmms_seg_plan_segment(const gmms_seg_plan* plan, int64_t seg, int* kway, int* rounds);
void gmms_seg_plan_destroy(gmms_seg_plan* plan);

/* Sorts the planned batch on `stream`: copies the plan's tables into the
 * workspace (workspace_bytes >= info.workspace_bytes) and launches. Buffers
 * as for gmms_sort, extent elements each; *alt = 1 when the result is in
 * keys_alt/vals_alt (the same for every segment). Asynchronous; the plan and
 * the workspace must stay alive until the stream has finished the call.
 * Elements before offsets[0] are neither read nor written. */
int gmms_seg_sort(gmms_seg_plan* plan, void* keys, void* keys_alt, void* vals, void* vals_alt, void* workspace,
                  size_t workspace_bytes, void* stream, int* alt);

/* The same launches captured once into a HIP graph bound to these buffers
 * and this workspace (the tables are copied at creation); every
 * gmms_seg_graph_launch re-sorts whatever the buffers hold. */
int gmms_seg_graph_create(gmms_seg_plan* plan, void* keys, void* keys_alt, void* vals, void* vals_alt,
                          void* workspace, size_t workspace_bytes, gmms_seg_graph** graph);
int gmms_seg_graph_launch(gmms_seg_graph* graph, void* stream, int* alt);
void gmms_seg_graph_destroy(gmms_seg_graph* graph);

/* One-shot: plan, allocate the workspace, sort, synchronize, free. */
int gmms_segmented_sort(int dtype, void* keys, void* keys_alt, void* vals, void* vals_alt, const int64_t* offsets,
                        int64_t n_seg, const gmms_params* p, void* stream, int* alt);

/* Tuning table of the batched sort, per dtype and size class c = 0..40
 * (segments with 2^(c-1) < n <= 2^c; class 0 holds n <= 1): the K of the
 * segments of that class (0: the default, 16) and the merger size in
 * elements the planner splits their tasks into (0: by the round's volume
 * over the resident mergers). */
#define GMMS_SEG_CLASSES 41
int gmms_seg_tuning_get(int dtype, int32_t* kway, int64_t* merger_elems, int classes);
int gmms_seg_tuning_set(int dtype, const int32_t* kway, const int64_t* merger_elems, int classes);
/* Autotuner: for each class in [min_class, max_class], sorts batches of
 * random segments of that size (about batch_elems elements per batch) with
 * K = 16 and every candidate merger size, times them on `stream` with HIP
 * events (best of `reps`) and installs the fastest into the tuning table;
 * the default entry (merger size 0) stays unless a candidate is at least
 * 3 % faster. (Another K saves a round for a class alone but costs more
 * per element in a batch of mixed sizes; set it with gmms_seg_tuning_set.) kway_out/merger_out/ms_out (may be NULL) receive
 * the chosen entries and their time per batch, indexed by class.
 * Allocates its own buffers; synchronous. */
int gmms_seg_autotune(int dtype, int min_class, int max_class, int64_t batch_elems, int reps, void* stream,
                      int32_t* kway_out, int64_t* merger_out, float* ms_out);

/* ------------------------------------------------------------------------
 * Sort server. No counterpart in the reference (its harness, src/testMM.cu,
 * sorts one array per process). Requests (host buffers of any supported
 * dtype) are admitted against an HBM budget: the server allocates one device
 * arena of exactly hbm_budget bytes at creation and carves every batch out
 * of it, so device memory in use never exceeds the budget and no request
 * allocates. A scheduler thread batches pending requests continuously: as
 * soon as an in-flight slot and arena space are free it takes the oldest
 * pending request plus the following pending requests of the same dtype
 * (up to max_batch_requests / max_batch_elems, as long as the batch fits in
 * the arena's largest free block), copies them in, sorts them as segments
 * of one batched segmented sort, copies the results out and frees the
 * space. Up to max_inflight batches run at once, each on its own stream.
 * A request that can never fit (needs more than the budget) is rejected
 * with GMMS_ERROR_BUDGET at submission. Host buffers should be pinned
 * (hipHostMalloc) for copies to overlap; pageable ones work.
 * ------------------------------------------------------------------------ */
typedef struct gmms_server gmms_server;

typedef struct gmms_server_config {
  size_t hbm_budget;           /* device bytes the server may use (default 64 GB)  */
  int64_t max_batch_elems;     /* elements per batch (default 2^28)                */
  int32_t max_batch_requests;  /* requests per batch (default 4096)                */
  int32_t max_inflight;        /* batches on the device at once (default 3)        */
  int32_t reserved[4];
} gmms_server_config;

typedef struct gmms_server_stats {
  int64_t submitted, completed, rejected, failed;
  int64_t batches;              /* batches launched                                 */
  int64_t max_batch_size;       /* most requests in one batch                       */
  int64_t max_inflight_seen;    /* most batches on the device at once               */
  int64_t keys_sorted;          /* elements of completed requests                   */
  size_t budget;                /* hbm_budget                                       */
  size_t arena_bytes;           /* device bytes allocated by the server (= budget)  */
  size_t reserved_bytes;        /* arena bytes held by in-flight batches now        */
  size_t peak_reserved_bytes;   /* most arena bytes ever held at once               */
} gmms_server_stats;

int gmms_server_default_config(gmms_server_config* cfg);
/* Device bytes one request of n elements holds while in flight (both key
 * buffers, both value buffers for pairs; the batch's tables come on top). */
int gmms_server_request_bytes(int dtype, int64_t n, size_t* bytes);
int gmms_server_create(const gmms_server_config* cfg, gmms_server** server);
/* Queue a request: sort n elements of keys (and vals for pairs) into
 * out_keys (out_vals); the buffers may be device memory or (pinned) host
 * memory, and must stay valid until the request completes. *ticket
 * identifies it. Returns at once. */
int gmms_server_submit(gmms_server* server, int dtype, const void* keys, const void* vals, int64_t n,
                       void* out_keys, void* out_vals, uint64_t* ticket);
/* The same, ordered after the caller's stream when `ordered` is 1: the
 * server records an event on `stream` (a hipStream_t; NULL is the null
 * stream) at submission, and the request's copies into and out of the arena
 * wait for it. Use it for device buffers that work still queued on that
 * stream writes (the inputs) or still reads (the result buffers).
 * gmms_server_submit is this call with ordered = 0: the caller has finished
 * every such work before it submits. */
int gmms_server_submit_after(gmms_server* server, int dtype, const void* keys, const void* vals, int64_t n,
                             void* out_keys, void* out_vals, void* stream, int ordered, uint64_t* ticket);
/* Wait up to timeout_s seconds (< 0: forever) for a request; *done = 1 once
 * it has finished with status *status; latency_ms (may be NULL) receives
 * submission-to-completion time and batch_id (may be NULL) the batch that
 * sorted it. A finished request is forgotten after the first wait that
 * reports it done. */
int gmms_server_wait(gmms_server* server, uint64_t ticket, double timeout_s, int* done, int* status,
                     double* latency_ms, int64_t* batch_id);
/* Wait until every submitted request has finished. Finished requests stay
 * known to the server until a gmms_server_wait reports them done, so wait
 * on each ticket once (timeout 0 after a drain) to release it. */
int gmms_server_drain(gmms_server* server);
int gmms_server_stats_get(gmms_server* server, gmms_server_stats* stats);
void gmms_server_destroy(gmms_server* server);

/* The batching policy alone (host logic, no device): given the pending
 * requests in arrival order (dtype[i], n[i]), the bytes of the largest free
 * arena block and the limits, writes the indices of the next batch to
 * chosen[] (*n_chosen of them; 0 if the oldest request does not fit now):
 * the oldest request, then later ones of its dtype, in order, while the
 * batch's data bytes (gmms_server_request_bytes) stay within free_bytes and
 * the limits. Requests of other dtypes are skipped, not blocked. */
int gmms_server_batch_policy(const int32_t* dtype, const int64_t* n, int64_t n_pending, size_t free_bytes,
                             int64_t max_batch_elems, int32_t max_batch_requests, int64_t* chosen,
                             int64_t* n_chosen);

#ifdef __cplusplus
}
#endif
#endif /* GPU_MMS_GMMS_H */
