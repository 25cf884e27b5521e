/*
 * params.h -- compile-time defaults of the drop-in template, same knob names
 * as the reference's src/params.h:18-38. They only seed gmms_params; every
 * value can also be chosen at run time through the C ABI (gmms.h).
 *
 *   K         lists merged per round (reference: 64)
 *   PL        log2(K): heap levels = wavefronts per merger when PIPELINE is on
 *   PIPELINE  one wavefront per heap level (the reference's pipelined heap);
 *             the MI355X build has only the pipelined kernel
 *   RANGE     range of the harness's random keys, rand() % RANGE
 *   BLOCKS    mergers per round when a round has few tasks (reference: 482;
 *             here 0 = two per compute unit, 512 on MI355X)
 *   M         base-case tile (1024 elements, one wavefront)
 *   W         wavefront width on CDNA4
 *   B         elements per heap node (one per lane)
 */
#ifndef GPU_MMS_PARAMS_H
#define GPU_MMS_PARAMS_H

#ifndef K
#define K 64
#endif
#ifndef PL
#define PL 6
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
