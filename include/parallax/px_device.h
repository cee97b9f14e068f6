/*
 * parallax/px_device.h -- device-resident, stream-ordered entry points of the MI355X
 * engine, and the batched (segmented) forms.
 *
 * NO COUNTERPART IN THE REFERENCE: the reference's primitives exist only as the
 * synchronous, host-visible calls of parallax/runtime.h.  These entry points serve
 * callers whose buffers already live in HBM (benchmarks, the request server, other
 * frameworks): they take device pointers, enqueue on the context's HIP stream and
 * return without waiting.  The synchronous C ABI of runtime.h runs on the same
 * kernels.
 *
 * Every function returns 0 on success or a negative PX_E* code; pxr_last_error()
 * describes the last failure of the calling thread.
 */
#ifndef PARALLAX_PX_DEVICE_H
#define PARALLAX_PX_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#include "parallax/descriptor.h"
#include "parallax/runtime.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PX_EOK 0
#define PX_EARG (-1)
#define PX_EDEVICE (-2)
#define PX_ENODEV (-3)
#define PX_ENOMEM (-4)
#define PX_EUNSUPPORTED (-5)

typedef struct pxd_ctx* pxd_ctx_t;

/* ---- runtime information */
int pxr_device_count(void);
/* Fills name (NUL-terminated, <= name_len) and numbers; any pointer may be NULL. */
int pxr_device_info(char* name, size_t name_len, int* compute_units, size_t* hbm_bytes,
                    size_t* lds_per_block);
int pxr_last_error(char* buf, size_t len);
/* Counters of device work launched by this process: launches, reduce, scan, sort,
 * compact, transform, segmented (7 values). */
int pxr_stats(uint64_t* out7);
const char* pxr_version(void);
/* Arena of the C ABI: host mirror base, device base, capacity, bytes in use. */
int pxr_arena_info(void** host_base, void** dev_base, size_t* capacity, size_t* used);
/* The reference's relocation probe (tests/unit/test_physptr.cpp): the kernel receives
 * the arena HOST pointer, relocates it (dev_base + (host_ptr - host_base)) and does
 * r[i] = r[i]*2+1 for count floats. */
int pxr_relocate_probe(void* arena_ptr, unsigned count);
/* The descriptor (12 words) that a funnel key resolves to without registration, or
 * PX_EUNSUPPORTED when the key names a functor funnel. */
int pxr_builtin_descriptor(const char* key, unsigned* words12);
/* Event timing of the C ABI's device work (per call, on its stream), keyed
 * "<op>:<dtype>" with op in reduce, scan, exclusive_scan, sort, copy_if, partition and
 * dtype in i32, u32, i64, u64, f32, f64.  enable(1) clears the table. */
int pxr_prof_enable(int on);
int pxr_prof_get(const char* key, double* total_ms, uint64_t* calls);

/* ---- device memory helpers (for bindings without a GPU framework); each has completed
   when it returns, so work queued afterwards on any stream sees its result */
void* pxd_malloc(size_t bytes);
void pxd_free(void* p);
int pxd_memcpy_h2d(void* dst, const void* src, size_t bytes);
int pxd_memcpy_d2h(void* dst, const void* src, size_t bytes);
int pxd_memcpy_d2d(void* dst, const void* src, size_t bytes);
int pxd_memset(void* dst, int value, size_t bytes);
int pxd_device_sync(void);

/* ---- contexts: one HIP stream (hip_stream == NULL: the context creates its own) plus
 * the scratch its kernels use.  Calls on one context run in order; different contexts
 * may run concurrently. */
pxd_ctx_t pxd_ctx_create(void* hip_stream);
void pxd_ctx_destroy(pxd_ctx_t ctx);
void* pxd_ctx_stream(pxd_ctx_t ctx);
int pxd_sync(pxd_ctx_t ctx);
/* Event timing on the context's stream: begin records a start event, end records a stop
 * event, waits for it and returns the elapsed milliseconds in *ms. */
int pxd_timer_begin(pxd_ctx_t ctx);
int pxd_timer_end(pxd_ctx_t ctx, double* ms);

/* ---- plain primitives (device pointers) */
/* Device-to-device copy of `bytes` on the context's stream: the streaming-copy kernel
 * (16-byte non-temporal loads and stores), i.e. the read-and-write bandwidth that the
 * look-back kernels are compared with.  pxd_memcpy_d2d runs the same kernel. */
int pxd_copy(pxd_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes);
/* *d_out = sum of d_in[0..n) in the element type (one element written). */
int pxd_reduce(pxd_ctx_t ctx, int dtype, const void* d_in, size_t n, void* d_out);
/* Σ map(x) in one pass: `map` is a PX_OP_TRANSFORM descriptor (input dtype, output dtype
 * in aux, map in fn: every map on the same type or widening i32->i64, u32->u64,
 * f32->f64; converting identity between any two types; PX_MAP_PRED counts into an
 * integer type) or a PX_OP_MAP (output type = input type).  One element of the output
 * type to *d_out.  PX_EARG for maps the fused kernel does not cover (gather). */
int pxd_transform_reduce(pxd_ctx_t ctx, const px_desc* map, const void* d_in, size_t n,
                         void* d_out);
int pxd_inclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out, size_t n);
/* pxd_inclusive_scan without any allocation: NO COUNTERPART IN THE REFERENCE.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): no context, no hipMalloc / hipFree; the scan's
 * scratch (about 8 MiB) is a device array of the library that each GPU holds until the
 * process ends.  Same kernel, tiling and results as pxd_inclusive_scan.  Runs on
 * hip_stream (NULL: the null stream) of the current device and returns without waiting.
 * At most 2^31 elements (PX_EARG beyond).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams or threads) race. */
int pxd_inclusive_scan_noalloc(int dtype, const void* d_in, void* d_out, size_t n, void* hip_stream);
/* h_init: host bytes of one element (NULL means 0). */
int pxd_exclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out, size_t n,
                       const void* h_init);
int pxd_sort(pxd_ctx_t ctx, int dtype, void* d_data, size_t n);
/* Stable compaction of elements whose predicate (flags->fn, PX_PRED_UNIQUE for unique)
 * holds; *d_count (device uint64) receives the kept count. */
int pxd_copy_if(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out, size_t n,
                uint64_t* d_count);
/* Partition in std::partition's order (libstdc++'s two-pointer exchange: the partition
 * funnel's CPU fallback), kept first; *d_num_true receives the split.  d_in == d_out
 * partitions in place.  The stable order is scatter mode PX_SCATTER_STABLE_PARTITION of
 * parallax_copy_if. */
int pxd_partition(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out, size_t n,
                  uint64_t* d_num_true);
int pxd_transform(pxd_ctx_t ctx, const px_desc* map, const void* d_in, void* d_out, size_t n,
                  const void* captures, size_t capture_size);
int pxd_find(pxd_ctx_t ctx, const px_desc* pred, const void* d_in, size_t n, uint64_t* d_index);
int pxd_argminmax(pxd_ctx_t ctx, int dtype, const void* d_in, size_t n, int want_max,
                  int want_last, uint64_t* d_index);
int pxd_mismatch(pxd_ctx_t ctx, int dtype, const void* d_a, const void* d_b, size_t n,
                 uint64_t* d_index);
/* Rewrite stored host addresses inside [host_base, host_base+bytes) to dev_base + offset. */
int pxd_relocate_pointers(pxd_ctx_t ctx, uint64_t* d_ptrs, size_t n, uint64_t host_base,
                          uint64_t dev_base, uint64_t bytes);

/* ---- batched (segmented) forms: segment s is [h_offsets[s], h_offsets[s+1]) of the
 * concatenated array (host offsets, nseg+1 entries).  One launch (sort: one fixed
 * launch sequence) covers every segment; a segment's result is the same as the plain
 * call on that segment alone. */
int pxb_seg_reduce(pxd_ctx_t ctx, int dtype, const void* d_in, const uint64_t* h_offsets,
                   size_t nseg, void* d_out /* nseg elements */);
int pxb_seg_inclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out,
                           const uint64_t* h_offsets, size_t nseg);
int pxb_seg_sort(pxd_ctx_t ctx, int dtype, void* d_data, const uint64_t* h_offsets, size_t nseg);
/* Kept elements of segment s go to d_out + h_offsets[s]; d_counts[s] (device uint64)
 * receives the kept count of segment s.  flags->fn == PX_PRED_UNIQUE gives unique. */
int pxb_seg_copy_if(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out,
                    const uint64_t* h_offsets, size_t nseg, uint64_t* d_counts);

/* ---- funnel extension (no counterpart in the reference, which runs transform2 and
 * reduce as separate calls through an arena array, stdpar.hpp:246-293): Σ map(x) in one
 * device pass for the transform_reduce / count_if / all_of / any_of / none_of funnels.
 * map_kernel: a TRANSFORM (or MAP) descriptor handle; reduce_kernel: a REDUCE of the
 * map's output type.  Same conventions as parallax_reduce (void; failures reported on
 * stderr; the result is host-visible on return). */
void parallax_transform_reduce(parallax_kernel_t map_kernel, parallax_kernel_t reduce_kernel,
                               void* input, size_t count, size_t in_elem_size,
                               size_t out_elem_size, void* result);

/* ---- graph plans: capture one batched call into a hipGraph once, replay it many
 * times (same buffers, same offsets).  op: PX_OP_REDUCE, PX_OP_SCAN, PX_OP_SORT,
 * PX_OP_FLAGS (segmented copy_if/unique with `flags`).  For REDUCE, SCAN and FLAGS the
 * tile shape is autotuned at creation: each candidate shape is timed on this batch and
 * the fastest is captured (PARALLAX_AUTOTUNE=0 keeps the size-based default;
 * PARALLAX_TUNE_CACHE=<file> keeps choices across processes).  Creation runs the call
 * several times: d_out and d_counts are overwritten; an in-place call (d_in == d_out) is
 * timed into a scratch output and a sort warms up on a scratch copy of its keys, so the
 * caller's data is intact when the plan is returned. */
typedef struct pxb_plan* pxb_plan_t;
pxb_plan_t pxb_plan_create(pxd_ctx_t ctx, int op, int dtype, const px_desc* flags,
                           const void* d_in, void* d_out, const uint64_t* h_offsets, size_t nseg,
                           uint64_t* d_counts);
/* Force the tile shape (threads per block, 16-byte vectors per thread) of one op and
 * dtype for the calls that follow: PX_OP_SCAN and PX_OP_FLAGS (plain and batched; shapes
 * 256x4, 256x8, 512x8, 1024x4, 1024x8) and PX_OP_REDUCE (batched; 256x4, 256x8, 256x16).
 * block = items = 0 restores the automatic choice.  Not synchronised with calls in
 * flight on other threads.  Plans created while a shape is forced skip the autotuner. */
int pxr_tile_override(int op, int dtype, int block, int items);
/* Tile shape of the captured call (threads per block, 16-byte vectors per thread; 0 for
 * sort, whose shape is fixed) and the autotuner's best timed duration in ms (0 when the
 * default was kept). */
int pxb_plan_tile(pxb_plan_t plan, int* block, int* items, double* tuned_ms);
int pxb_plan_launch(pxb_plan_t plan);
void pxb_plan_destroy(pxb_plan_t plan);

#ifdef __cplusplus
}
#endif

#endif /* PARALLAX_PX_DEVICE_H */
