/*
 * params.h -- compile-time defaults of the drop-in template, same knob names
 * as the reference's src/params.h:18-38. They seed the gmms_params that
 * multimergesort<T,cmp> passes to the library; every value can also be
 * chosen at run time through the C ABI (gmms.h).
 *
 *   K         lists merged per round (reference: 64; here 16, the fastest
 *             on MI355X for every measured size and dtype)
 *   PL        log2(K): heap levels, i.e. wavefronts per merger of the
 *             pipelined minBlockHeap kernel; derived from K unless given,
 *             and checked against it (multimergesort.hxx)
 *   PIPELINE  as in the reference: defined -> the pipelined minBlockHeap
 *             merge, one wavefront per heap level (PL wavefronts per
 *             merger, src/io-merge-gen.hxx:305-355); undefined -> one
 *             wavefront per merger (src/io-merge-gen.hxx:357-381), here the
 *             register-tile merger. Unlike the reference it is NOT defined
 *             by default: on CDNA4 the one-wavefront merger is the fast one
 *             (2^28 int32: 3.9 ms against 14.7 ms for the heap). -DPIPELINE
 *             selects the heap; -DNO_PIPELINE is accepted and keeps it off.
 *   MERGE_KERNEL  the merge kernel by number, overriding PIPELINE when
 *             given: 0 one-wavefront tile merger, 1 pipelined minBlockHeap,
 *             2 tile merge with one merger per workgroup
 *   RANGE     range of the harness's random keys, rand() % RANGE
 *   BLOCKS    mergers that run at once when a round has few tasks
 *             (reference: 482); a round splits its tasks into a whole
 *             multiple of it (up to 4x, filled to within 10 %). 0 = auto:
 *             the resident merger count of the merge kernel, for the default
 *             one 16 per compute unit (4096 on MI355X), 12 for pairs
 *   M         base-case tile (1024 elements, one wavefront)
 *   W         wavefront width on CDNA4
 *   B         elements per heap node (one per lane)
 */
#ifndef GPU_MMS_PARAMS_H
#define GPU_MMS_PARAMS_H

#if defined(PIPELINE) && defined(NO_PIPELINE)
#error "params.h: define at most one of PIPELINE and NO_PIPELINE"
#endif

#ifndef MERGE_KERNEL
#ifdef PIPELINE
#define MERGE_KERNEL 1
#else
#define MERGE_KERNEL 0
#endif
#endif
#ifndef K
#define K 16
#endif
#ifndef PL
#define PL ((K) >= 64 ? 6 : (K) >= 32 ? 5 : (K) >= 16 ? 4 : (K) >= 8 ? 3 : (K) >= 4 ? 2 : 1)
#endif
#ifndef RANGE
#define RANGE 1024
#endif
#ifndef BLOCKS
#define BLOCKS 0
#endif
#define M 1024
#define W 64
#define B 64

#endif
