/* halloc_batch.h -- batched device calls of the MI355X halloc allocator.
 *
 * No counterpart in the reference (src/halloc.h has single-block calls only).
 * hamalloc_batch gives each calling lane up to HA_MAX_BATCH blocks of one size
 * per call: the lanes of one size class take one ticket and add all their
 * chunks to the slab counter with one atomic, instead of one of each per block.
 */
#ifndef HALLOC_BATCH_H_
#define HALLOC_BATCH_H_

#include <stddef.h>

#define HA_MAX_BATCH 16

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
/** out[0..cnt) receive blocks of nbytes (NULL where memory ran out); cnt is
    capped at HA_MAX_BATCH; returns the number of blocks obtained */
__device__ int hamalloc_batch(size_t nbytes, int cnt, void **out);
/** frees ptrs[0..cnt); NULL entries are ignored */
__device__ void hafree_batch(void *const *ptrs, int cnt);
#endif

#endif /* HALLOC_BATCH_H_ */
