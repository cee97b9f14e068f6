/*
 * parallax/runtime.h -- the C ABI of the Parallax primitive runtime, served by the
 * MI355X engine (libparallax_amd.so).
 *
 * Every declaration below has the same name, argument order, argument meaning and
 * linkage as the reference runtime's header (include/parallax/runtime.h:12-145), so a
 * program or funnel header compiled against the reference links against this library
 * unchanged.  What changes behind the ABI:
 *
 *   - The `words` payload given to parallax_kernel_load / parallax_kernel_register is a
 *     primitive descriptor (parallax/descriptor.h: op, dtype, predicate or map op,
 *     constants) instead of SPIR-V.  The engine owns hand-written gfx950 kernels for
 *     every descriptor; a payload that is not a valid descriptor loads as NULL, so the
 *     caller takes its CPU loop.
 *   - Pointers may be arena pointers (host-mapped mirror of the device arena), device
 *     pointers (hipMalloc'd / HBM-resident), or ordinary host memory (staged through
 *     device scratch and copied back: results always reach the caller's pointer).
 *   - Every primitive is synchronous, as in the reference: outputs are host-visible
 *     (or device-visible for device pointers) when the call returns.
 *   - Void entry points cannot return an error code; on a failure the engine prints one
 *     "[parallax]" line to stderr and records it for pxr_last_error() (px_device.h).  It
 *     never recomputes on the host: once a call reaches the runtime it runs on the GPU
 *     or fails loudly (a lookup miss, before the call, is what sends a funnel to its
 *     CPU loop).
 */
#ifndef PARALLAX_RUNTIME_H
#define PARALLAX_RUNTIME_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque handles (reference runtime.h:12-13). */
typedef struct parallax_runtime* parallax_runtime_t;
typedef struct parallax_kernel* parallax_kernel_t;

/* General unified allocation (reference runtime.h:16-17).  Backed by the arena; falls
 * back to malloc when no device is present. */
void* parallax_umalloc(size_t size, unsigned flags);
void parallax_ufree(void* ptr);

/* Arena allocation (reference runtime.h:22-24).  Returns a host-dereferenceable pointer
 * into the pinned mirror of the device arena (256-byte aligned), or NULL when the arena
 * is exhausted or no device exists.  Kernels run on the device copy at the same offset;
 * operand ranges migrate on a side stream around each call. */
void* parallax_arena_alloc(size_t size, size_t align);
void parallax_arena_free(void* ptr);
int parallax_arena_contains(const void* ptr);

/* Whole-heap query (reference runtime.h:30).  The engine has no heap pool: always 0. */
int parallax_heap_contains(const void* ptr);

/* Kernel handles (reference runtime.h:33-35).  The variadic launches read
 * (void* buf, size_t count, size_t elem_size) and
 * (void* in, void* out, size_t count, size_t elem_size) respectively. */
parallax_kernel_t parallax_kernel_load(const unsigned int* spirv, size_t words);
void parallax_kernel_launch(parallax_kernel_t kernel, ...);
void parallax_kernel_launch_transform(parallax_kernel_t kernel, ...);

/* out[i] = f(in[i]) with distinct element sizes (reference runtime.h:37-39). */
void parallax_kernel_launch_transform2(parallax_kernel_t kernel, void* in_buffer,
                                       void* out_buffer, size_t count,
                                       size_t in_elem_size, size_t out_elem_size);

/* As transform2, with the functor's capture bytes (reference runtime.h:43-46). */
void parallax_kernel_launch_transform2_captures(parallax_kernel_t kernel, void* in_buffer,
                                                void* out_buffer, size_t count,
                                                size_t in_elem_size, size_t out_elem_size,
                                                void* captures, size_t capture_size);

/* Sum of `count` elements written as elem_size bytes to `result` (reference
 * runtime.h:52-53).  Identity 0; the caller combines its init. */
void parallax_reduce(parallax_kernel_t kernel, void* data, size_t count,
                     size_t elem_size, void* result);

/* Search entry points (reference runtime.h:59-68).  Exported for link compatibility;
 * they run on the device for descriptors of the matching op and return `count` for
 * anything else (the reference's "not found" value). */
size_t parallax_argminmax(parallax_kernel_t kernel, void* data, size_t count,
                          size_t elem_size, int is_float, int want_max, int want_last);
size_t parallax_find(parallax_kernel_t kernel, void* data, size_t count,
                     size_t elem_size, int negate, const void* value);
size_t parallax_mismatch(parallax_kernel_t kernel, void* a, void* b, size_t count, size_t elem_size);

/* Inclusive scan in place (reference runtime.h:72-73).  One single-pass decoupled
 * look-back launch; add_kernel is accepted for ABI compatibility and must be non-NULL. */
void parallax_scan(parallax_kernel_t scan_kernel, parallax_kernel_t add_kernel,
                   void* data, size_t count, size_t elem_size);

/* output[i] = init + sum(input[0..i)) (reference runtime.h:78-80).  `input` is not
 * modified by this engine. */
void parallax_exclusive_scan(parallax_kernel_t scan_kernel, parallax_kernel_t add_kernel,
                             parallax_kernel_t shift_kernel, void* input, void* output,
                             size_t count, size_t elem_size, const void* init);

/* Ascending sort in place (reference runtime.h:85).  Any count is accepted (positions
 * past the next power of two are virtual +infinity pads, never stored). */
void parallax_sort(parallax_kernel_t kernel, void* data, size_t count, size_t elem_size);

/* Stable stream compaction (reference runtime.h:91-94).  The flags descriptor carries
 * the predicate (or adjacent-unique), the scatter descriptor the mode (compact or
 * partition).  Returns the kept count (num_true for partition); 0 on failure. */
size_t parallax_copy_if(parallax_kernel_t flags_kernel, parallax_kernel_t scan_kernel,
                        parallax_kernel_t add_kernel, parallax_kernel_t scatter_kernel,
                        void* input, void* output, size_t count, size_t elem_size,
                        int elem_is_float);

/* Funnel registry (reference runtime.h:102-103).  register records the pointer only
 * (it may run before main); lookup loads lazily and returns NULL on a miss, so the
 * funnel runs its CPU loop. */
void parallax_kernel_register(const char* key, const unsigned int* spirv, size_t words);
parallax_kernel_t parallax_kernel_lookup(const char* key);

/* In-place map with capture bytes (reference runtime.h:106-113). */
void parallax_kernel_launch_with_captures(
    parallax_kernel_t kernel,
    void* buffer,
    size_t count,
    void* captures,
    size_t capture_size,
    size_t elem_size
);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
/* Compile-time contiguity gate used by funnel headers (reference runtime.h:129-139). */
#include <iterator>
#include <type_traits>
namespace parallax {
#if __cplusplus >= 202002L
template <class It>
inline constexpr bool is_contiguous_iterator_v =
    std::contiguous_iterator<std::remove_cvref_t<It>>;
#else
template <class It>
inline constexpr bool is_contiguous_iterator_v =
    std::is_pointer_v<std::remove_reference_t<It>>;
#endif
}  // namespace parallax
#endif /* __cplusplus */

#endif /* PARALLAX_RUNTIME_H */

/* Register an external host buffer (reference runtime.h:145: C++ linkage, outside the
 * include guard).  The engine records the range; later calls on pointers inside it copy
 * in and out on every call (so re-registered or rewritten buffers never go stale).
 * C++ linkage, so it is only visible to C++ translation units. */
#ifdef __cplusplus
bool parallax_register_buffer(void* ptr, size_t size);
#endif
