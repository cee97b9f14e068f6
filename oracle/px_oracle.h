/*
 * px_oracle.h -- CPU oracle for the primitive hot path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product library (libparallax_amd.so) never links or calls it.
 *
 * Each function restates the reference runtime's CPU fallback for one primitive
 * (include/parallax/stdpar.hpp), cited per function in px_oracle.cpp.  Element types
 * and predicate / map descriptors use the engine's encoding (parallax/descriptor.h) so
 * the oracle can be driven with exactly the descriptor the GPU path receives.
 *
 * Pinning (what each function is checked against besides retyped formulas):
 *   reduce / scans / sort / copy_if / unique:
 *       the reference's known-answer tests (tests/unit/test_*.cpp) and closed forms
 *       (arithmetic series, sorted permutations, modular sums, numpy equivalents).
 *   partition: libstdc++'s two-pointer exchange restated as a loop and in closed form
 *       (tests/test_oracle.py), both against this function's std::partition.
 *   The reference's fallback cannot be built on its own here: its funnel header needs
 *   the reference runtime library (Vulkan), so none of this is compared with a build
 *   of the reference.
 *   sum_exact / scan_exact: double-double sums checked against Python's math.fsum
 *       (exactly rounded) and integer-valued closed forms.
 *   predicates / maps: brute-force tables in tests/test_oracle.py against Python
 *       integer and IEEE arithmetic (fma via math.fma-free exact rationals).
 */
#ifndef PX_ORACLE_H
#define PX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "parallax/descriptor.h"

#ifdef __cplusplus
extern "C" {
#endif

int pxo_abi_version(void);

/* reduce: the fallback's left fold in the element type, init E{} (stdpar.hpp:146-148). */
void pxo_reduce(int dtype, const void* in, size_t n, void* out);
/* Float reference sum in double-double (hi + lo), plus sum |x|.  Integer dtypes return
 * the value converted to double. */
void pxo_sum_exact(int dtype, const void* in, size_t n, double* hi, double* lo, double* abs_sum);

/* scans (stdpar.hpp:209-210, :237-238) */
void pxo_inclusive_scan(int dtype, const void* in, void* out, size_t n);
void pxo_exclusive_scan(int dtype, const void* in, void* out, size_t n, const void* init);
/* Per-prefix reference values (double-double rounded to double) and per-prefix sum |x|
 * for float error bounds.  exclusive != 0 gives out[i] = init + sum(x[0..i)). */
void pxo_scan_exact(int dtype, const void* in, size_t n, int exclusive, const void* init,
                    double* out_val, double* out_abs);

/* sort (stdpar.hpp:178): std::sort ascending by operator<. */
void pxo_sort(int dtype, void* data, size_t n);

/* compaction family (stdpar.hpp:319-321, :343-345, :369, :392) */
int pxo_pred_eval(const px_desc* d, const void* x);             /* one element */
size_t pxo_copy_if(const px_desc* flags, const void* in, void* out, size_t n);
size_t pxo_remove_if(const px_desc* flags, void* data, size_t n);
size_t pxo_partition(const px_desc* flags, void* data, size_t n);        /* std::partition */
size_t pxo_stable_partition(const px_desc* flags, void* data, size_t n); /* std::stable_partition */
size_t pxo_unique(int dtype, void* data, size_t n);                      /* std::unique */
long pxo_count_if(const px_desc* flags, const void* in, size_t n);       /* stdpar.hpp:290-292 */
/* Search (the reference launcher's host combine, kernel_launcher.cpp:1078-1098, 1180-1189,
   1272-1277): first index passing / differing (n if none); min or max index by <, ties to
   the smaller index (larger with want_last), 0 for n == 0.  NaN inputs: unpinned. */
size_t pxo_find(const px_desc* pred, const void* in, size_t n);
size_t pxo_mismatch(int dtype, const void* a, const void* b, size_t n);
size_t pxo_argminmax(int dtype, const void* in, size_t n, int want_max, int want_last);

/* element-wise skeletons (stdpar.hpp:81, :119) */
void pxo_map(const px_desc* d, void* data, size_t n, const void* captures, size_t capture_size);
void pxo_transform(const px_desc* d, const void* in, void* out, size_t n,
                   const void* captures, size_t capture_size);

/* segmented forms: the plain fallback applied to each [offsets[s], offsets[s+1]) */
void pxo_seg_reduce(int dtype, const void* in, const uint64_t* offsets, size_t nseg, void* out);
void pxo_seg_sum_exact(int dtype, const void* in, const uint64_t* offsets, size_t nseg,
                       double* hi, double* abs_sum);
void pxo_seg_inclusive_scan(int dtype, const void* in, void* out, const uint64_t* offsets,
                            size_t nseg);
/* per-segment pxo_scan_exact (inclusive) into out_val / out_abs at the segment's offsets */
void pxo_seg_scan_exact(int dtype, const void* in, const uint64_t* offsets, size_t nseg,
                        double* out_val, double* out_abs);
void pxo_seg_sort(int dtype, void* data, const uint64_t* offsets, size_t nseg);
/* per-segment std::unique of in -> out packed segment after segment; out_offsets[nseg+1] */
size_t pxo_seg_unique(int dtype, const void* in, void* out, const uint64_t* offsets,
                      size_t nseg, uint64_t* out_offsets);

#ifdef __cplusplus
}
#endif

#endif /* PX_ORACLE_H */
