/* halloc.h -- drop-in API of the MI355X halloc allocator.
 *
 * Same declarations as the reference's public header (src/halloc.h:15-55):
 * the options struct with the same fields and defaults, device hamalloc /
 * hafree, host ha_init / ha_shutdown / ha_extfrag.  Device code that includes
 * this header is built with `hipcc -fgpu-rdc --offload-arch=gfx950` and linked
 * with `-fgpu-rdc --hip-link -lhalloc`.  Allocations persist across kernel
 * launches until hafree.
 */
#ifndef HALLOC_H_
#define HALLOC_H_

#include <stddef.h>

/** options of the allocator; replaces src/halloc.h:15-34 field for field */
struct halloc_opts_t {
  /** total memory for allocation, bytes; capped at 3/4 of device memory */
  size_t memory;
  /** fraction of memory given to the slab heap; the rest is the large
      region that serves requests above 3072 bytes (the reference gives it to
      the CUDA device heap) */
  double halloc_fraction;
  /** occupancy fraction at which a slab is busy (head slab is replaced) */
  double busy_fraction;
  /** occupancy fraction at or below which a slab is roomy */
  double roomy_fraction;
  /** occupancy fraction at or below which a slab is sparse and may be reused by
      another size class with the same chunk size */
  double sparse_fraction;
  /** log2 of the slab size in bytes, 21..26 */
  int sb_sz_sh;
  /** defaults of src/halloc.h:30-32 */
  halloc_opts_t(size_t memory = 512 * 1024 * 1024)
      : memory(memory), halloc_fraction(0.75), busy_fraction(0.835), roomy_fraction(0.6),
        sparse_fraction(0.012), sb_sz_sh(22) {}
};

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
/** allocates nbytes on the device; NULL when no memory is left
    (replaces src/halloc.h:38) */
__device__ void *hamalloc(size_t nbytes);
/** frees memory from hamalloc; NULL is ignored (replaces src/halloc.h:41) */
__device__ void hafree(void *p);

#if defined(HALLOC_CPP)
/* Typed device allocation, the counterpart of the reference's HALLOC_CPP
   operator new / delete (src/halloc.h:67-80, src/halloc.cu:201-208,258-265).
   HIP already defines the plain device operator new / delete (over the
   device heap's malloc / free), so the halloc forms take a tag, as the
   reference's commented-out halloc_tag_t sketches:
     T *p = new (halloc) T(args);    a hamalloc block; when the heap is
                                     exhausted, NULL and no constructor runs
     U *a = new (halloc) U[n];       U trivially destructible
     hadelete(p);  hadelete_array(a);   destructor, then hafree; NULL ignored
   p must point at the object allocated (its most-derived type); blocks are
   8-byte aligned at least (16 for classes that are not a multiple of 24). */
#include <new>
#include <type_traits>
struct halloc_tag_t {};
inline constexpr halloc_tag_t halloc{};
__device__ inline void *operator new(size_t nbytes, halloc_tag_t) noexcept { return hamalloc(nbytes); }
__device__ inline void *operator new[](size_t nbytes, halloc_tag_t) noexcept { return hamalloc(nbytes); }
/* called only when a constructor throws, which device code does not */
__device__ inline void operator delete(void *p, halloc_tag_t) noexcept { hafree(p); }
__device__ inline void operator delete[](void *p, halloc_tag_t) noexcept { hafree(p); }
template <class T>
__device__ inline void hadelete(T *p) {
  if (p) {
    p->~T();
    hafree((void *)p);
  }
}
/* an array of a trivially destructible type carries no element count ahead
   of its first element, so the block starts at a */
template <class T>
__device__ inline void hadelete_array(T *a) {
  static_assert(std::is_trivially_destructible<T>::value, "hadelete_array: trivially destructible types only");
  hafree((void *)a);
}
#endif /* HALLOC_CPP */
#endif

/** initializes the allocator on the current device (replaces src/halloc.h:46).
    A failure prints the reason and exits, as the reference's cucheck does;
    ha_init_c in halloc_amd.h returns an error code instead. */
void ha_init(halloc_opts_t opts = halloc_opts_t());

/** releases the allocator's device memory (src/halloc.h:49; the reference's
    is a no-op).  ha_init may be called again afterwards. */
void ha_shutdown(void);

/** external fragmentation 1 - max_alloc / free_mem (src/halloc.h:55,
    src/statistics.cuh:50-69) */
double ha_extfrag(bool ideal);

#endif /* HALLOC_H_ */
