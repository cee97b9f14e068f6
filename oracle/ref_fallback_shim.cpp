// ref_fallback_shim.cpp -- C entry points around the REFERENCE's own CPU fallback.
//
// Compiled by oracle/build_ref.sh against the reference's header-only funnel header
// (include/parallax/stdpar.hpp under the reference checkout) and linked to
// libparallax_amd.so for the runtime.h symbols.  Run with PARALLAX_FORCE_CPU=1 (or on a
// machine without a GPU) every parallax_kernel_lookup misses, so each call below runs
// the reference's ISO host loop (stdpar.hpp:146-148, 178, 209-210, 237-238, 319-321,
// 343-345, 369, 392).  TEST INFRASTRUCTURE ONLY: a checker and the timed CPU baseline.
#include <cstddef>
#include <cstdint>

#include <parallax/stdpar.hpp>

#define REF_TYPES(X) X(i32, int32_t) X(u32, uint32_t) X(i64, int64_t) X(u64, uint64_t) X(f32, float) X(f64, double)

extern "C" {

#define REF_FUNCS(N, T)                                                                     \
  T ref_reduce_##N(const T* p, size_t n) { return parallax::reduce(parallax::par, p, p + n, T{}); } \
  void ref_inclusive_scan_##N(const T* in, T* out, size_t n) {                              \
    parallax::inclusive_scan(parallax::par, in, in + n, out);                               \
  }                                                                                         \
  void ref_exclusive_scan_##N(const T* in, T* out, size_t n, T init) {                      \
    parallax::exclusive_scan(parallax::par, in, in + n, out, init);                         \
  }                                                                                         \
  void ref_sort_##N(T* p, size_t n) { parallax::sort(parallax::par, p, p + n); }            \
  size_t ref_unique_##N(T* p, size_t n) {                                                   \
    return static_cast<size_t>(parallax::unique(parallax::par, p, p + n) - p);              \
  }                                                                                         \
  size_t ref_copy_if_gt_##N(const T* in, T* out, size_t n, T c) {                           \
    return static_cast<size_t>(                                                             \
        parallax::copy_if(parallax::par, in, in + n, out, [c](T x) { return x > c; }) - out); \
  }                                                                                         \
  size_t ref_remove_if_gt_##N(T* p, size_t n, T c) {                                        \
    return static_cast<size_t>(                                                             \
        parallax::remove_if(parallax::par, p, p + n, [c](T x) { return x > c; }) - p);      \
  }                                                                                         \
  size_t ref_partition_gt_##N(T* p, size_t n, T c) {                                        \
    return static_cast<size_t>(                                                             \
        parallax::partition(parallax::par, p, p + n, [c](T x) { return x > c; }) - p);      \
  }

REF_TYPES(REF_FUNCS)

// the P6 predicate of the plain suite: x % 2 == 0 on int32
size_t ref_copy_if_even_i32(const int32_t* in, int32_t* out, size_t n) {
  return static_cast<size_t>(
      parallax::copy_if(parallax::par, in, in + n, out, [](int32_t x) { return x % 2 == 0; }) - out);
}

}  // extern "C"
