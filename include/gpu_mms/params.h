/*
 * params.h -- compile-time defaults of the drop-in template, same knob names
 * as the reference's src/params.h:18-38. They only seed gmms_params; every
 * value can also be chosen at run time through the C ABI (gmms.h).
 *
 *   K         lists merged per round (reference: 64)
 *   PL        log2(K): heap levels = wavefronts per merger of the heap kernel
 *   PIPELINE  defined as in the reference (the minBlockHeap kernel is always
 *             the pipelined one: one wavefront per heap level)
 *   MERGE_KERNEL  which K-way merge runs the rounds: 0 tile merge with one
 *             merger per wavefront (default, fastest), 1 the pipelined
 *             minBlockHeap, 2 tile merge with one merger per workgroup
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

#ifndef MERGE_KERNEL
#define MERGE_KERNEL 0
#endif
#ifndef K
#define K 16
#endif
#ifndef PL
#define PL 4
#endif
#ifndef NO_PIPELINE
#define PIPELINE
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
