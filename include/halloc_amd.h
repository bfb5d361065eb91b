/* halloc_amd.h -- C ABI of the MI355X halloc allocator (libhalloc_amd.so).
 *
 * Host code (C, C++, or a ctypes / cgo / JNI binding) reaches HIP only through
 * these functions: plain pointers, sizes and return codes.  Device code calls
 * hamalloc / hafree from halloc.h.  Every function returns 0 (or a count) on
 * success and a negative HA_E* code on failure; ha_last_error() describes it.
 * The workload drivers (ha_run_*) are the counterparts of the reference test
 * harness kernels (tst/common/common.h:361-412 and the tst/perf tests); they take host
 * buffers, keep their inputs in device memory for the timed launches, and
 * time those launches with HIP events on the stream they launch on.
 */
#ifndef HALLOC_AMD_H_
#define HALLOC_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HA_OK = 0,
  HA_E_INVALID = -1,   /* bad argument or options */
  HA_E_STATE = -2,     /* not initialized / already initialized */
  HA_E_HIP = -3,       /* a HIP runtime call failed */
  HA_E_NOMEM = -4      /* device allocation of the pool failed */
};

/* Options: the fields of halloc_opts_t (src/halloc.h:15-34) plus tuning. */
typedef struct ha_opts_c {
  uint64_t memory;           /* bytes; capped at 3/4 of device memory */
  double halloc_fraction;    /* default 0.75 */
  double busy_fraction;      /* default 0.835 */
  double roomy_fraction;     /* default 0.6 */
  double sparse_fraction;    /* default 0.012 */
  int32_t sb_sz_sh;          /* default 22 (4 MiB slabs), 21..26 */
  int32_t nheads;            /* head-slab sets (1..64), default 32, cut at init to the largest
                                power of two s with 3 * s <= slabs where it is above (what one
                                class can hold heads in: 32 at 512 MiB, 8 at 128 MiB;
                                ha_geometry reports the count): each (set, size class) pair
                                has its own head slab, ticket and lock.  A wave's full-spread set
                                is its XCD's (xcc_id % nheads); above 8 (a multiple of 8), one of
                                its XCD's nheads / 8 sets, picked by the wave's CU.  0: the
                                largest power of two up to 16 with nheads * 16 classes * 3 <=
                                slabs (2 at 512 MiB, 4 at 1 GiB, 16 from 4 GiB) */
  double hash_step_frac;     /* 0 (default): the probe stride's prime is the largest below
                                nblocks/256 + nblocks/64 (src/halloc.cu:378-381); else below
                                hash_step_frac * nblocks, 0 < hash_step_frac < 1 */
  int32_t cached_arena;      /* 0 (default): the arena is uncached device memory, coherent
                                across the XCDs' L2s, and kept for reuse after ha_shutdown;
                                1: L2-cached (hipMalloc) arena, and every hafree performs an
                                agent-scope release before the blocks can move to another XCD */
  int32_t fixed_sets;        /* 0 (default): adaptive sets.  Waves use the first `spread` sets
                                only; the spread starts at nheads and shrinks, as size classes take
                                their first head, to the largest power of two s with
                                3 * s * classes seen <= slabs, and the pairs above it give their
                                head slabs back to the slab sets.  So one hot class keeps every
                                set, and many live classes on a small heap do not pin the slabs
                                the others need (each pair holds a partly filled head slab).
                                1: every wave keeps its full-spread set. */
} ha_opts_c;

/* Fills *o with the defaults of src/halloc.h:30-32, nheads = 32 and adaptive
   sets. */
void ha_default_opts(ha_opts_c *o);

/* Replaces ha_init (src/halloc.cu:267-424).  Initializes the current device. */
int ha_init_c(const ha_opts_c *o);
/* Replaces ha_shutdown (src/halloc.cu:426-428); frees the pool. */
int ha_shutdown_c(void);
/* Replaces ha_extfrag (src/statistics.cuh:50-69).  Writes the value to *out. */
int ha_extfrag_c(int ideal, double *out);
/* Free bytes of the slab heap, the counterpart of find_free_mem_k
   (src/statistics.cuh:34-48), summed on the device. */
int ha_free_bytes_c(uint64_t *out);
/* No counterpart in the reference.  ha_free_bytes_c without allocation, for
   callers that cannot allocate device memory (a serving loop polling the
   heap between launches, harnesses that forbid hipMalloc): the same value,
   reduced in a 1 KiB device array the library owns (per device, for the life
   of the process) instead of a hipMalloc buffer; it calls no hipMalloc or
   hipFree.  The scratch need is fixed (66 words) for any heap size, and the
   call fails with HA_E_STATE when the current device is not the one
   ha_init_c initialized.  Calls on one GPU share the array; the library lock
   serializes them within a process. */
int ha_free_bytes_noalloc_c(uint64_t *out);
/* Description of the last failure on this thread. */
const char *ha_last_error(void);

/* Geometry of the initialized heap: nsbs, sb_sz, nheads, nsw, bit_words,
   nib_words, the 16 x 12 size table (nchunks chunk_sz chunk_id ncib sparse
   hash_step roomy busy block_sz nblocks tmod chunk_bytes), the device address
   of the slab arena (the large region of large_bytes follows its arena_bytes),
   and the large region's 4 KiB units and classes. */
typedef struct ha_geometry {
  uint32_t nsbs, sb_sz, sb_sz_sh, nheads, nsw, bit_words, nib_words, pad;
  uint64_t arena_bytes, large_bytes;
  uint32_t size_table[16 * 12];
  uint64_t arena_base;
  uint32_t large_units, large_classes;
} ha_geometry;
int ha_geometry_c(ha_geometry *g);

/* ---------------------------------------------------------------- harness */

/* Replay nsteps x lanes operations on ONE wavefront (lanes <= 64).
   ops[2*i] = kind (0 none, 1 malloc, 2 free, 3 free of the pointer 4096 bytes
   into the block: an interior pointer of a large block), ops[2*i+1] = bytes
   for malloc or the index of the earlier malloc op for free (a second free of
   the same op is a double free).  Per step the malloc lanes call hamalloc
   together, then the free lanes call hafree together.
   recs[4*i..]: malloc -> (size class, slab, chunk), a large block
   (16 + its class, 0xfffffffe, 4 KiB unit), NULL (0xffffffff x 3); free ->
   (chunks freed, slab, chunk) or (units freed, 0xfffffffe, unit), read before
   the free; (0, 0xffffffff, 0xffffffff) when the free changes nothing.  iheads (may be NULL) forces the
   head set of each step's mallocs; otherwise the wave's XCD id picks it.
   xcc[step] receives the step's full-spread head set (given, or the wave's
   own), which the allocator folds onto the sets in use. */
int ha_replay(const uint32_t *ops, int nsteps, int lanes, const uint32_t *iheads, uint32_t *recs,
              uint32_t *xcc);

/* Batched replay on ONE wavefront (include/halloc_batch.h): ops hold 4 u32
   per (step, lane): kind (0 none, 1 hamalloc_batch, 2 hafree_batch, 3
   hafree_batch that keeps the op's pointers, so that a later op frees the same
   blocks again), arg (bytes, or the index of an earlier kind-1 op whose blocks
   are freed), count (kind 1), 0.  recs hold 4 u32 per (op, j < 16), as for ha_replay.  xcc[step]
   receives the wave's full-spread head set. */
int ha_replay_batch(const uint32_t *ops, int nsteps, int lanes, uint32_t *recs, uint32_t *xcc);

/* Canonical state dump (layout documented in oracle/model_capi.cpp). */
int64_t ha_dump_words(void);
int ha_dump(uint64_t *out, int64_t nwords);

/* Plain 16 B throughput (tst/perf/priv-throughput/priv-throughput.cu:14-27):
   nthreads lanes in blocks of bs (at most 256); each lane does iters x { p = hamalloc(size);
   write a 16 B pattern {tid, ~tid, tid ^ 0x9E3779B9, iter}; read it back;
   hafree(p) }; `launches` timed launches after `warmup` untimed ones, queued
   back to back on one stream (no host synchronization and no event between
   them).  total_ms is the time between HIP events recorded before the first
   and after the last, wall_ms the host clock around them, the stream
   synchronized before and after.  If ms is not NULL, a second pass of
   `launches` launches follows, each between its own pair of events, and
   ms[i] receives launch i's duration.  errors counts pattern mismatches and
   nulls NULL returns, over both passes. */
typedef struct ha_plain_result {
  uint64_t errors;
  uint64_t nulls;
  double total_ms;
  double wall_ms;
} ha_plain_result;
int ha_run_plain(int nthreads, int bs, int size, int iters, int warmup, int launches, float *ms,
                 ha_plain_result *res);

/* Allocate-only kernel (tst/common/common.h:375-388): lane i stores
   hamalloc(sizes[i]) into ptrs[i] (host array, device addresses returned) and
   writes the 32-bit size into the block's first word. */
int ha_malloc_sizes(const uint32_t *sizes, int n, int bs, uint64_t *ptrs);
/* Free kernel (tst/common/common.h:391-399): lane i calls hafree(ptrs[i]). */
int ha_free_ptrs(const uint64_t *ptrs, int n, int bs);
/* Reads back the first 32-bit word of each block (0 for NULL). */
int ha_read_words(const uint64_t *ptrs, int n, uint32_t *out);

/* A caller's own kernel over heap blocks, to measure what the arena's memory
   type costs it (no counterpart in the reference).  n lanes each hamalloc a
   block of 8 * words bytes (words in 1..384); lane i then makes `iters`
   passes over its block's 64-bit words in three modes, each launched once
   untimed and then `reps` times: 2 stores it * words + w + i into word w,
   0 loads and sums the words, 1 replaces each word x by
   x * 6364136223846793005 + it (mod 2^64).  ms[m] is mode m's mean launch
   time (HIP events), sums[m] the sum (mod 2^64) over lanes of the block's
   words after mode m (mode 0: of the per-lane sums of its last launch);
   *nulls counts the lanes whose hamalloc returned NULL (they skip the rest).
   The blocks are freed at the end. */
int ha_run_touch(int n, int words, int iters, int reps, double *ms, uint64_t *sums, uint64_t *nulls);

/* tst/corr/test/test.cu test 0 (:84-132): nthreads threads (blocks of bs)
   each allocate nmallocs 16 B blocks in the grid-stride order of its malloc_k
   (:47-55); every block gets two u64 holding its own address, read back after
   the whole grid wrote (bad_contents counts mismatches); sorted, consecutive
   blocks must lie at least 16 B apart (bad_gaps); then free_k frees them all
   (free_ms, HIP events). */
typedef struct ha_corr0_result {
  uint64_t blocks, nulls, bad_contents, bad_gaps;
  double free_ms;
} ha_corr0_result;
int ha_run_corr_test0(int nthreads, int nmallocs, int bs, ha_corr0_result *r);

/* Lists (tst/common/common.h:375-399, check_alloc tst/common/common.cu:468-529):
   lane i allocates sizes[j*n + i] for j < per (0 = none) and writes the size
   into the block's first word; then thread t frees ptrs[order[j*n + t]]
   (order = NULL: the same slots) after checking that word.  Both kernels are
   timed with HIP events; ptrs_out (may be NULL) receives the device addresses
   between the two kernels.  errors counts wrong first words at free time. */
typedef struct ha_list_result {
  uint64_t requests, nulls, errors;
  double malloc_ms, free_ms, wall_ms;
} ha_list_result;
int ha_run_lists(const uint32_t *sizes, int n, int per, const uint32_t *order, int bs, uint64_t *ptrs_out,
                 ha_list_result *res);

/* Phase workload (tst/perf/phase-throughput/phase-throughput.cu:18-50): each
   of nthreads lanes owns `slots` slots and takes one action per step: with
   probability p it allocates a block of 8*U{min/8..max/8} bytes into an empty
   slot (stamping its first word), otherwise it checks and frees a live slot.
   Even phases allocate (p = p_hi) until the live bytes reach fill_hi of the
   slab heap, or more than 1% of a step's requests return NULL; odd phases free
   (p = p_lo) until fill_lo.  Then everything left is freed.  bs is a multiple
   of 64.  trace (3 doubles per step: phase, fill, NULLs) may be NULL. */
typedef struct ha_phase_args {
  int32_t nthreads, bs, slots, nphases, max_steps, min_size, max_size;
  uint32_t seed;
  float p_hi, p_lo;
  double fill_hi, fill_lo;
} ha_phase_args;
typedef struct ha_phase_result {
  uint64_t mallocs, frees, nulls, errors;
  double kernel_ms, wall_ms;
  int32_t steps, trace_len;
  int32_t phase_steps[8];
} ha_phase_result;
int ha_run_phase(const ha_phase_args *a, ha_phase_result *res, double *trace, int trace_cap);

/* Random-graph builds (batching workload): job j builds a graph of nnodes[j]
   nodes, each with a 16 B header and an int32 adjacency array of Poisson(8)
   random targets, checks it and frees it.  Plain per-call form: one thread per
   node calling hamalloc / hafree, jobs launched in chunks of at most
   plain_chunk_nodes nodes.  Batched form: one persistent launch of `waves`
   wavefronts (blocks of 256), captured in a HIP graph, taking chunks of
   64 * batch nodes (batch 4, 8 or 16): every build chunk, and each job's check
   and free chunks after the job is built; a chunk's adjacency arrays of one
   size class come from one hamalloc_batch call and go back through
   hafree_batch.  Each form runs `repeats` times and reports its best time;
   errors counts wrong graph nodes. */
typedef struct ha_graph_args {
  int32_t njobs, bs, batch, waves, plain_chunk_nodes, repeats, run_plain, run_batch;
  uint32_t seed;
  uint32_t cross; /* 1: instead of timing, build with each form and check and free with the other */
} ha_graph_args;
typedef struct ha_graph_result {
  uint64_t nodes, nulls, errors;
  double plain_ms, batch_ms;
  int32_t plain_launches, pad;
} ha_graph_result;
int ha_run_graphs(const ha_graph_args *a, const uint32_t *nnodes, ha_graph_result *res);

/* Serving loop (continuous batching): one persistent launch of `slots` lanes
   (a multiple of 256), one sequence slot each.  Request r arrives
   arrival_ns[r] after the start (sorted, at most 60 s) and decodes length[r]
   tokens (1..32768) of 16 B, one per iteration, in 256 B pages from hamalloc.
   Idle slots admit arrived requests in arrival order while the heap's free
   bytes, refreshed from the slab counters, stay >= watermark.  Token t of a
   request holds {tag, t, h, ~h}; at completion the request's pages are read
   back, the hashes h of the intact tokens summed into checksum[r], and the
   pages freed.  A page malloc that returns NULL ends the request early:
   produced[r] is the number of tokens it holds.  first_ns / done_ns (may be
   NULL) get the first-token and completion times; every page malloc whose
   (request + page index) is a multiple of sample_every is timed into lat_ns
   (at most lat_cap samples, ns).  bad / bad_first (may be NULL) get the
   number of tokens that did not read back intact and words 0-1 of the first
   one.  No counterpart in the reference: it is the
   paged-growth workload built on hamalloc/hafree. */
typedef struct ha_serve_io {
  const uint64_t *arrival_ns;
  const uint32_t *length;
  const uint32_t *tag;
  uint32_t *checksum;
  uint32_t *produced;
  uint64_t *first_ns;
  uint64_t *done_ns;
  uint32_t *lat_ns;
  uint32_t *bad;
  uint64_t *bad_first;
  int nreq, slots, sample_every, lat_cap;
  uint64_t watermark;
} ha_serve_io;
typedef struct ha_serve_result {
  double kernel_ms;
  uint64_t pages, nulls, stalls, samples;
  uint32_t completed, admitted, clock_khz, pad;
} ha_serve_result;
int ha_run_serve(const ha_serve_io *io, ha_serve_result *res);

/* The reference test programs' kernels (tst/corr/checkptr, tst/corr/prob-checkptr,
   tst/perf/throughput, tst/perf/priv-throughput) for the option set of
   tst/common/common.h CommonOpts, ntries times:
     test 0 checkptr: allocate nallocs blocks per active thread (sizes stamped
       into the blocks), check_alloc, free;
     test 1 prob-checkptr: niters steps of the alloc/free chain with
       probabs[phase][allocated], check_alloc of the blocks held; the rest is
       freed at the end;
     test 2 throughput: allocate, check that no pointer is NULL, free (timed);
     test 3 priv-throughput: niters rounds of allocate-then-free (timed);
     test 4 phase-throughput (tst/perf/phase-throughput): the chain of test 1
       without stamps, timed per try (even tries as malloc time, odd as free
       time) with a NULL check of the blocks held; freeing the rest counts as
       free time;
     test 5 latency (tst/perf/latency): every malloc, then every free, timed
       in shader-clock cycles, with a NULL check between;
     test 6 phase-latency (tst/perf/phase-latency): the chain of test 1 with
       every malloc and free timed;
     test 7 phase-alloc-write (tst/perf/phase-alloc-write): test 4 with 123
       written into every whole 8 B word of each block allocated;
     test 8 phase-extfrag (tst/perf/phase-extfrag): test 4 with
       extfrag[2 t] = ha_extfrag(false), extfrag[2 t + 1] = ha_extfrag(true)
       after try t (extfrag: ntries * 2 doubles, or NULL).
   lat_n / lat_avg / lat_min / lat_max [0] mallocs, [1] frees (cycles).
   malloc_ms / free_ms: kernel time (HIP events); malloc_wall_ms /
   free_wall_ms: host wall-clock around launch and synchronization, as the
   reference's tests time (omp_get_wtime).
   Active threads: tid < nthreads and not (tid >> group_sh) & period_mask.
   Sizes: alloc_sz when max_alloc_sz == alloc_sz, else distr 1 uniform,
   2 exponential-uniform, 3 exponential-equal over [alloc_sz, max_alloc_sz]
   with max_alloc_sh = the largest sh with alloc_sz << sh <= max_alloc_sz.
   Draws hash (seed, tid, draw index).  check_alloc: every held pointer is
   non-NULL and, sorted, leaves room for its size; checks_failed counts the
   violations (NULLs and overlaps), failed_try is the first try with one. */
typedef struct ha_test_opts {
  int32_t test, nthreads, bs, ntries, nallocs, niters;
  uint32_t alloc_sz, max_alloc_sz, max_alloc_sh, distr, period_mask, group_sh, seed, pad;
  float probabs[2][2];
  double *extfrag;
} ha_test_opts;
typedef struct ha_test_result {
  uint64_t mallocs, frees, nulls, checked, checks_failed;
  double malloc_ms, free_ms;
  int32_t tries_run, failed_try;
  uint64_t lat_n[2];
  double lat_avg[2], lat_min[2], lat_max[2];
  double malloc_wall_ms, free_wall_ms;
} ha_test_result;
int ha_run_test(const ha_test_opts *o, ha_test_result *r);

/* The reference samples (the samples directory) as application checks.
     sample 0 grid-points: n points uniform in [0,1)^3 are sorted into an
       ncells^3 grid, a 16 B list node from hamalloc per point pushed onto its
       cell's list; the lists are checked (every point once, in its own cell)
       and freed;
     sample 1 add-strings: two arrays of n strings (lengths uniform in
       [min_len, max_len]) allocated and filled, concatenated elementwise into
       a third, checked, and all three freed;
     sample 2 random-graph: n vertices with U{1..max_len} edges each, every
       edge a 16 B list node from hamalloc; the lists are checked (length,
       targets) and freed;
     sample 3 objects (the typed forms of halloc.h under HALLOC_CPP, the
       counterpart of src/halloc.h:67-80): per thread one object by
       new (halloc) with a counting constructor and destructor and one array
       of U{1..max_len} words by new (halloc) u32[k], checked, and released
       by hadelete / hadelete_array; an error if constructions and
       destructions differ.
   alloc_ms / work_ms / free_ms: inserting or allocating / the grid check or
   the concatenation / freeing; errors: points or strings that fail the
   check. */
typedef struct ha_sample_args {
  int32_t sample, n, bs, ncells, min_len, max_len;
  uint32_t seed, pad;
} ha_sample_args;
typedef struct ha_sample_result {
  uint64_t nulls, errors;
  double alloc_ms, work_ms, free_ms;
} ha_sample_result;
int ha_run_sample(const ha_sample_args *a, ha_sample_result *r);

/* Copy probe: `reps` launches of a streaming copy of `bytes` (a multiple of 16,
   1 MiB..64 GiB) between two device buffers; *gbps = 2 * bytes / best launch
   time.  The known-size kernel for calibrating HBM byte counters. */
int ha_probe_copy(uint64_t bytes, int reps, double *gbps);

/* Atomic-throughput probe: 64-bit atomic OR over a 1 GiB buffer from 2^20
   lanes with three access shapes; rates in atomics/s:
   out[0] lane-random (one cache line per lane), out[1] wave-run (a wave
   covers 64 consecutive words), out[2] single hot word. */
int ha_probe_atomics(double *out);

/* Dependent-chain probe: nthreads lanes (a multiple of 256) in waves of 64;
   each wave makes rounds[i] dependent 64-bit atomic round trips (the next
   address depends on the previous result), lane 0 issuing, over a 1 GiB
   buffer.  ms[i] = the best of 5 launch times (HIP events) for rounds[i].
   The slope over rounds is the atomic round-trip latency with every wave of
   the grid waiting on memory: the bound of a kernel that is a chain of
   dependent atomics per wave, as each hamalloc / hafree call is. */
int ha_probe_chain(const int *rounds, int n, int nthreads, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* HALLOC_AMD_H_ */
