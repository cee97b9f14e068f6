// kernels.hpp -- host-callable launchers of the gfx950 primitive kernels.
//
// Every launcher is stream-ordered and allocation-free: the caller (a stream context)
// owns the scratch below and guarantees its sizes.  Pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>
#include <vector>

#include "parallax/descriptor.h"

namespace pxk {

// Control block shared by the self-cleaning kernels of one stream context.  Zeroed once
// at allocation with epoch = 1; every launch leaves ticket == done == 0 and advances
// the epoch, so no reset launch or memset is needed between calls (and graph replays
// stay valid).
struct Ctl {
  uint32_t epoch;   // current launch epoch (1 .. 2^30-1)
  uint32_t ticket;  // dynamic tile ticket
  uint32_t done;    // blocks finished in this launch
  uint32_t pad0;
  uint64_t result;  // kept count written by the last compaction tile
  uint64_t radix_vary;  // segmented radix: key bits set in some key and clear in another
  uint64_t radix_nmin;  // signed radix keys: max over blocks of ~enc(x) (zero between calls)
  uint64_t radix_base;  // signed radix keys: (U)min(x), the key of x being (U)x - base
  uint64_t pad1[2];
  uint8_t value[64];  // reduce result slot (one element, written by the last block)
  uint32_t radix_skip;  // radix sorts: bit d set when every key shares digit d
  uint64_t radix_or[16];   // segmented radix: OR of the encoded keys, 16 slots (block % 16)
  uint64_t radix_nor[16];  // ... OR of their complements; all zero between calls
  uint32_t radix_hist[8 * 256];  // radix sort digit counts (zero between calls)
  uint32_t radix_bins[8 * 256];  // radix sort exclusive bin starts per digit position
};

constexpr uint32_t kEpochMax = (1u << 30) - 1;

struct SegGeom;

// Launch shapes.  A tile of the look-back kernels (scan, compaction) is `block` threads x
// `k` 16-byte vectors per thread; the shape is chosen per element type (registers) and
// recorded in the geometry, so the tile count computed on the host and the tile size
// the kernel uses always agree.  Defaults can be overridden for sweeps with
// PARALLAX_TUNE, e.g. "scan_block=1024,scan_k=8,compact_block=1024,compact_k=8,bpc=1,
// red_unroll=8,red_bpc=1,red_nt=1"; only the compiled shapes below are accepted.
struct TileCfg {
  int block;
  int k;
};
struct Tune {
  TileCfg scan[PX_NUM_DTYPES];     // per element type
  TileCfg compact[PX_NUM_DTYPES];
  int bpc = 0;                // resident blocks per CU of the look-back kernels (0: 1024/block)
  TileCfg seg_reduce{256, 4};
  int red_unroll = 8;         // 16-byte loads in flight per thread in the plain reduce
  int red_bpc = 1;            // blocks per CU of the plain reduce
  int red_nt = 1;             // non-temporal loads in the plain reduce
  int sort_c4 = 6;            // stages per HBM register pass of the sort, 4-byte keys (<= 7)
  int sort_c8 = 6;            // ... 8-byte keys (<= 6)
  // plain sorts of at least this many keys use radix (0: never; ~0: 2^17 keys
  // of 4 bytes, 2^20 of 8 bytes, the measured crossovers with the bitonic path)
  uint64_t sort_radix_min = ~0ull;
  int radix_block = 0;        // threads per radix digit-pass block (256, 512, 1024; 0: by size)
  int radix_hist_bpc = 0;     // radix histogram blocks per CU (0: default, 1)
  int radix_hist_block = 0;   // radix histogram block: 256, 512 or 1024 threads (0: default, 512)
  int seg_hist_bpc = 0;       // segmented radix histogram blocks per CU (0: default, 4)
  uint64_t seg_radix_min = 0;  // segments longer than this use the segmented radix sort (0: default, 1,024 / 512 keys)
  int seg_radix_block = 0;     // segmented radix digit-pass block, 512 or 1024 (0: chosen per call)
  int copy_bpc = 0;            // streaming copy: 256-thread blocks per CU (0: default, 1)
  int copy_unroll = 0;         // ... 16-byte loads in flight per thread, 1/2/4/8 (0: default, 4)
};
Tune& tune();
// Compiled tile shapes of the look-back kernels: op PX_OP_SCAN (scans) or PX_OP_FLAGS
// (compaction); scans also have 1024 x 16 (two register sub-tiles per ticket).
bool tile_cfg_supported(TileCfg c, int op);
// Forced tile shape of (op, dtype) for plain and batched calls (op: PX_OP_REDUCE
// (batched only), PX_OP_SCAN, PX_OP_FLAGS); {0, 0} when none is set.
TileCfg tile_override(int op, int dtype);
bool set_tile_override(int op, int dtype, TileCfg c);  // {0, 0} clears
// Persistent grid of a look-back launch: every tile, at most the resident blocks.
uint32_t lookback_grid(uint32_t ntiles, uint32_t max_blocks, int block);

// Scratch of one stream context.  All of it is device memory; ctl is zeroed once with
// epoch = 1 and is left clean by every kernel.
struct Scratch {
  Ctl* ctl = nullptr;
  uint64_t* status = nullptr;   // look-back status words
  size_t status_words = 0;
  void* partials = nullptr;     // per-block / per-tile partial sums
  size_t partial_bytes = 0;
  uint32_t max_blocks = 2048;   // resident-block cap (CUs x 8)
};

#define PXK_DISPATCH(dtype, ...)                                 \
  switch (dtype) {                                               \
    case PX_I32: { using T = int32_t; __VA_ARGS__; } break;      \
    case PX_U32: { using T = uint32_t; __VA_ARGS__; } break;     \
    case PX_I64: { using T = int64_t; __VA_ARGS__; } break;      \
    case PX_U64: { using T = uint64_t; __VA_ARGS__; } break;     \
    case PX_F32: { using T = float; __VA_ARGS__; } break;        \
    case PX_F64: { using T = double; __VA_ARGS__; } break;       \
    default: break;                                              \
  }

template <class T>
constexpr int dtype_of() {
  if constexpr (std::is_same_v<T, int32_t>) return PX_I32;
  else if constexpr (std::is_same_v<T, uint32_t>) return PX_U32;
  else if constexpr (std::is_same_v<T, int64_t>) return PX_I64;
  else if constexpr (std::is_same_v<T, uint64_t>) return PX_U64;
  else if constexpr (std::is_same_v<T, float>) return PX_F32;
  else return PX_F64;
}

inline size_t dtype_size(int dtype) {
  switch (dtype) {
    case PX_I32: case PX_U32: case PX_F32: return 4;
    case PX_I64: case PX_U64: case PX_F64: return 8;
    default: return 0;
  }
}
// Bytes of the carry type (status words per tile = carry_bytes / 4).
inline size_t carry_bytes(int dtype) {
  switch (dtype) {
    case PX_I32: case PX_U32: return 4;
    case PX_I64: case PX_U64: case PX_F32: return 8;
    case PX_F64: return 16;
    default: return 0;
  }
}
// Elements per tile of the tiled kernels.
inline uint32_t tile_elems(int dtype, TileCfg c) {
  return static_cast<uint32_t>(c.block * c.k * (16 / dtype_size(dtype)));
}

// ---- reduce (k_reduce.hip)
hipError_t reduce(int dtype, const void* in, uint64_t n, void* out_dev, const Scratch& s,
                  hipStream_t stream);
hipError_t seg_reduce(int dtype, const void* in, const SegGeom& g, void* out, const Scratch& s,
                      hipStream_t stream);
// Fused Σ map(x): `map` is a PX_OP_TRANSFORM descriptor (input dtype, output dtype in
// aux, map op in fn, predicate in fn2 for PX_MAP_PRED, constants c0 / c1).  Supported:
// same-type maps, converting identity, predicate counts into an integer type.
bool map_reduce_supported(const px_desc& map);
hipError_t map_reduce(const px_desc& map, const void* in, uint64_t n, void* out_dev,
                      const Scratch& s, hipStream_t stream);

// ---- scan (k_scan.hip): inclusive, or exclusive with `init` (host bytes of one element)
hipError_t scan(int dtype, const void* in, void* out, const SegGeom& g, bool exclusive,
                const void* init, const Scratch& s, hipStream_t stream);
// Inclusive scan of segments [lo[i], hi[i]) written from out_lo[i] (lo[i] when out_lo is
// null), one wave per segment: for segments of at most a few hundred elements.
hipError_t seg_scan_short(int dtype, const void* in, void* out, const uint64_t* lo, const uint64_t* hi,
                          const uint64_t* out_lo, uint32_t nshort, const Scratch& s, hipStream_t stream);

// ---- compaction (k_compact.hip)
//   kModeKept: kept elements of segment s to out + seg_out[s] (or seg_lo[s]); kept count
//     to counts[s] (counts may be null for one segment: then Ctl::result).
//   kModeStable (one segment): kept first, the rest after *num_true_dev, both stable.
//   kModeMisplaced (one segment, tile misplaced_tile()): the indices (uint32, or uint64
//     from 2^32 elements) of the elements on the wrong side of *num_true_dev --
//     predicate false below it, true at or above it -- in increasing order to out, their
//     count to counts[0]; with swap_pairs() this gives std::partition's order in place.
enum CompactMode { kModeKept = 0, kModeStable = 1, kModeMisplaced = 2 };
hipError_t compact(const px_desc& pred, int mode, const void* in, void* out, const SegGeom& g,
                   uint64_t* counts, const uint64_t* num_true_dev, const Scratch& s,
                   hipStream_t stream);
TileCfg misplaced_tile(int dtype, uint64_t n);
// Exchanges data[idx[i]] and data[idx[2m-1-i]] for i < m, 2m = *count2_dev.
hipError_t swap_pairs(int dtype, void* data, const void* idx, bool wide_idx,
                      const uint64_t* count2_dev, uint64_t n, const Scratch& s,
                      hipStream_t stream);
// Count of predicate-true elements of one segment into *count_dev (partition's first pass).
hipError_t count_if(const px_desc& pred, const void* in, uint64_t n, uint64_t* count_dev,
                    const Scratch& s, hipStream_t stream);

// Optional per-launch hook of multi-launch primitives: fn(ctx, kernel, 0) right before
// and fn(ctx, kernel, 1) right after each kernel launch (used for per-kernel timing).
struct LaunchHook {
  void (*fn)(void* ctx, const char* kernel, int phase) = nullptr;
  void* ctx = nullptr;
  void operator()(const char* kernel, int phase) const {
    if (fn) fn(ctx, kernel, phase);
  }
};

// ---- sort (k_sort.hip): ascending; positions past n act as +infinity.  Kernels:
// "sort_tile" (LDS sort of each tile), "sort_global" (register pass over HBM) and
// "sort_merge" (LDS half-cleaners of one level); each reads and writes the array once.
hipError_t sort(int dtype, void* data, uint64_t n, hipStream_t stream, const LaunchHook& hook = {});
// ---- radix sort (k_radix.hip): same ascending order for large arrays; kernels
// "sort_hist" (digit histograms, one read) and "sort_radix" (one per 8-bit digit, one
// read + write each).  Scratch: radix_status_words() status words and n * size bytes of
// partials (the second buffer).
// Long segments of a batched sort, sorted by the radix kernels: tiles of
// seg_radix_tile() keys that never span two segments, in segment order.
struct RadixSegs {
  uint32_t tile = 0;                    // keys per tile: seg_radix_tile()
  uint32_t ntiles = 0;                  // tiles of the long segments
  uint32_t nseg = 0;                    // long segments
  const uint32_t* tile_seg = nullptr;   // tile -> long-segment index
  const uint32_t* seg_first = nullptr;  // long segment -> its first tile
  const uint64_t* seg_lo = nullptr;     // long segment -> first element (in the array)
  const uint64_t* seg_hi = nullptr;     // long segment -> end element (hi - lo < 2^32)
  const uint64_t* seg_pack = nullptr;   // long segment -> first element in tmp_b / tmp2
  uint32_t* seg_cnt = nullptr;          // long segment x digit position x 256 counts
  void* tmp_b = nullptr;                // second buffer (segments packed, seg_pack)
  void* tmp2 = nullptr;                 // third buffer (same packing)
};
bool radix_sort_ok(int dtype, uint64_t n);
uint64_t radix_status_words(int dtype, uint64_t n);
int radix_passes(int dtype);
hipError_t radix_sort(int dtype, void* data, uint64_t n, const Scratch& s, hipStream_t stream,
                      const LaunchHook& hook = {});
// Segmented radix sort of the long segments: one counting pass (per-segment digit counts,
// and the digits every key shares, whose passes are skipped on the device), then one pass
// per digit that runs, the last writing the array.  Scratch: 256 status words per tile;
// the caller provides seg.tmp_b and seg.tmp2 (each holding the long segments packed at
// seg.seg_pack: Σ lengths plus alignment) and seg.seg_cnt (seg_radix_cnt_bytes()).
uint64_t seg_radix_min_len(int dtype);  // segments up to this length stay on the bitonic path
// Keys per tile for these long segments ([lo[i], hi[i]), i < n): 1024-thread digit-pass
// blocks (16K keys per tile) unless their tiles hold over 5 % more slots than 512-thread
// ones (8K), i.e. unless many segments fill a small part of a large tile.
uint32_t seg_radix_tile(int dtype, const uint64_t* lo, const uint64_t* hi, uint32_t n);
uint64_t seg_radix_cnt_bytes(int dtype, uint32_t nseg);
// The counting kernels run on `stream`; the digit passes (persistent look-back
// kernels) on `pass_stream` when given, after `prep_done` is recorded on `stream` and
// waited for there, else on `stream`.
hipError_t seg_radix_sort(int dtype, void* data, const RadixSegs& seg, const Scratch& s,
                          hipStream_t stream, hipStream_t pass_stream = nullptr,
                          hipEvent_t prep_done = nullptr);
// Segmented sort of [seg_off[s], seg_off[s+1]) for every segment: a host plan (block
// tables for each launch kind) uploaded by the caller, then a fixed launch sequence.
struct SegSortPlan {
  int dtype = 0;
  uint32_t nseg = 0;
  std::vector<uint32_t> tables;  // local_seg | local_first | per c: glob_seg[c] | glob_first[c]
  size_t o_local_first = 0;
  size_t o_gseg[8] = {};
  size_t o_gfirst[8] = {};
  uint32_t nlocal = 0;
  uint32_t gblocks[8] = {};
  uint32_t max_nplog = 0;
  uint32_t tile_log = 0;
  int max_c = 0;
};
// seg_bounds_dev: 2*nseg device words, [first, end) element indices of each segment.
bool seg_sort_plan(int dtype, const uint64_t* lo_host, const uint64_t* hi_host, uint32_t nseg,
                   SegSortPlan& p);
hipError_t seg_sort_run(const SegSortPlan& p, void* data, const uint64_t* seg_bounds_dev,
                        const uint32_t* tables_dev, hipStream_t stream);

// ---- element-wise (k_map.hip)
struct MapArgs {
  uint32_t count_lo;   // push block {count, pad, host_base, dev_base} (24 bytes)
  uint32_t pad;
  uint64_t host_base;  // arena host mirror base (relocation of captured pointers)
  uint64_t dev_base;   // arena device base
  uint64_t arena_bytes;
  uint8_t cap[64];     // capture bytes (first 64)
  uint32_t cap_size;
};
hipError_t transform(const px_desc& d, const void* in, void* out, uint64_t n, const MapArgs& a,
                     hipStream_t stream);
// Rewrite n stored host addresses inside [host_base, host_base+bytes) to device addresses.
hipError_t relocate_pointers(uint64_t* ptrs, uint64_t n, uint64_t host_base, uint64_t dev_base,
                             uint64_t bytes, hipStream_t stream);
// The reference's relocation probe: r = dev_base + (host_ptr - host_base); r[i] = r[i]*2+1.
hipError_t relocate_probe(uint64_t host_ptr, uint64_t host_base, uint64_t dev_base,
                          uint32_t count, hipStream_t stream);

// ---- search (k_search.hip): results written to *out_dev (uint64 index)
hipError_t find_first(const px_desc& pred, const void* in, uint64_t n, uint64_t* out_dev,
                      const Scratch& s, hipStream_t stream);
hipError_t argminmax(int dtype, const void* in, uint64_t n, bool want_max, bool want_last,
                     uint64_t* out_dev, const Scratch& s, hipStream_t stream);
hipError_t mismatch(int dtype, const void* a, const void* b, uint64_t n, uint64_t* out_dev,
                    const Scratch& s, hipStream_t stream);

// ---- streaming device-to-device copy (k_copy.hip): 16-byte non-temporal vectors when
// src and dst share their offset within a 16-byte line, else 4-byte words or bytes
hipError_t copy_bytes(void* dst, const void* src, uint64_t bytes, uint32_t max_blocks, hipStream_t stream);

// ---- fill kernel (graph-safe memset replacement)
hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream);

}  // namespace pxk
