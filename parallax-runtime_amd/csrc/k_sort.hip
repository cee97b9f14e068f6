// k_sort.hip -- bitonic sort for gfx950.
//
// The reference dispatches one global compare-exchange stage per (k, j) of the bitonic
// schedule with a host wait after each (KernelLauncher::launch_sort,
// src/kernel/kernel_launcher.cpp:1568-1606, bitonic.comp:15-30: 351 synchronous full
// passes at 2^26).  This engine sorts the same network with far fewer HBM passes:
//
//   * the network is written in its all-ascending form: merge level k starts with the
//     "flip" stage (partner i ^ (k-1)) followed by half-cleaners (partner i ^ j,
//     j = k/4 .. 1); every comparator puts the minimum at the lower index.  Positions
//     at or past n therefore always keep the +infinity value they start with, so they
//     are virtual: never loaded, never stored, and any n is accepted (the reference
//     needs a power of two and its funnel pads with numeric_limits::max in memory,
//     stdpar.hpp:159-176).
//   * every stage whose partner distance is below the LDS tile runs inside one kernel
//     with the tile resident in LDS (the tile sort, then one merge pass per level);
//   * the remaining (global) stages of a level run C at a time in registers: a thread
//     loads the 2^C elements that the C stages exchange among themselves, applies the
//     C stages, and stores them back -- one HBM pass for up to 6 stages (7 compiled for
//     4-byte keys; the default is set by the launch-shape configuration).
//
// Segmented form: the same kernels walk a block table, so thousands of segments of
// different lengths sort in the same few launches (each segment virtually padded to its
// own power of two).
#include <algorithm>
#include <type_traits>
#include <vector>

#include "dev_common.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

template <class T>
__device__ __forceinline__ T pos_inf();
template <>
__device__ __forceinline__ int32_t pos_inf<int32_t>() { return INT32_MAX; }
template <>
__device__ __forceinline__ uint32_t pos_inf<uint32_t>() { return UINT32_MAX; }
template <>
__device__ __forceinline__ int64_t pos_inf<int64_t>() { return INT64_MAX; }
template <>
__device__ __forceinline__ uint64_t pos_inf<uint64_t>() { return UINT64_MAX; }
template <>
__device__ __forceinline__ float pos_inf<float>() { return __builtin_huge_valf(); }
template <>
__device__ __forceinline__ double pos_inf<double>() { return __builtin_huge_val(); }

// min to a, max to b; a swap only when b < a (operator<, as std::sort).  Integers use
// min/max (2 instructions); floats keep compare + select, so NaN and -0.0 inputs still
// come out as a permutation of the input bits.
template <class T>
__device__ __forceinline__ void cmpx(T& a, T& b) {
  if constexpr (std::is_integral_v<T>) {
    const T lo = b < a ? b : a;
    const T hi = b < a ? a : b;
    a = lo;
    b = hi;
  } else {
    if (b < a) {
      const T t = a;
      a = b;
      b = t;
    }
  }
}

// Segment of the current block: plain (one segment) or from the block table.
struct SortSeg {
  uint64_t base, n;
  uint32_t nplog;   // log2 of the segment's padded length
  uint32_t lblock;  // block index inside the segment
};

__device__ __forceinline__ uint32_t ceil_log2(uint64_t n) {
  return n <= 1 ? 0u : 64u - static_cast<uint32_t>(__clzll(static_cast<long long>(n - 1)));
}

// Segment bounds travel as one array: seg[2*s] = first element, seg[2*s+1] = end.
__device__ __forceinline__ SortSeg sort_seg(const uint64_t* seg_off, const uint32_t* blk_seg,
                                            const uint32_t* blk_first, uint64_t n) {
  SortSeg s;
  if (seg_off == nullptr) {
    s.base = 0;
    s.n = n;
    s.lblock = blockIdx.x;
  } else {
    const uint32_t sg = blk_seg[blockIdx.x];
    s.base = seg_off[2 * sg];
    s.n = seg_off[2 * sg + 1] - s.base;
    s.lblock = blockIdx.x - blk_first[sg];
  }
  s.nplog = ceil_log2(s.n);
  return s;
}

// ---------------------------------------------------------------- tile kernels
// A tile of 2^tl elements lives in LDS; every stage below the tile size runs there.
// Stages are applied in register groups: a thread loads the 32 elements that five
// consecutive stages exchange among themselves (bits b .. b-4 of the index, the top one
// replaced by the flip mask on a level's first group), applies up to five stages, and
// writes them back -- one LDS round trip per five stages instead of one per stage.
// A full tile sort of 2^14 takes 23 round trips (105 stages); a merge takes 3 (14 stages).
//
// LDS layout: element i sits at i ^ ((i >> 5) & 63).  For every 32-element group (any b,
// with or without the flip mask) the 64 lanes of a wave then hit 64 distinct banks with
// 4-byte elements and 32 distinct 8-byte slots per half wave: no bank conflicts.
constexpr int kGroupLog = 5;
constexpr int kGroup = 1 << kGroupLog;

__device__ __forceinline__ uint32_t lds_slot(uint32_t i) { return i ^ ((i >> 5) & 63u); }

// Levels 1..5 of the network on 32 consecutive elements held in registers.
template <class T>
__device__ __forceinline__ void sort32(T (&v)[kGroup]) {
#pragma unroll
  for (int L = 1; L <= kGroupLog; ++L) {
    const int hk = 1 << (L - 1);
#pragma unroll
    for (int q = 0; q < kGroup; ++q)
      if (!(q & hk)) cmpx(v[q], v[q ^ ((hk << 1) - 1)]);
#pragma unroll
    for (int j = hk >> 1; j >= 1; j >>= 1) {
#pragma unroll
      for (int q = 0; q < kGroup; ++q)
        if (!(q & j)) cmpx(v[q], v[q | j]);
    }
  }
}

// Stages r0..4 of one register group (stage r acts on bit b - r; q's bit 4 - r selects
// that bit).  FLIP: stage 0 is the level's flip stage, so the elements with q's bit 4 set
// carry every lower bit inverted and are the upper partner in all later stages.
template <class T, bool FLIP>
__device__ __forceinline__ void group_stages(T (&v)[kGroup], int r0) {
#pragma unroll
  for (int r = 0; r < kGroupLog; ++r) {
    if (r < r0) continue;
#pragma unroll
    for (int q = 0; q < kGroup; ++q) {
      if ((q >> (kGroupLog - 1 - r)) & 1) continue;
      const int q2 = q | (1 << (kGroupLog - 1 - r));
      const bool q_high = FLIP && r >= 1 && ((q >> (kGroupLog - 1)) & 1);
      if (!q_high)
        cmpx(v[q], v[q2]);
      else
        cmpx(v[q2], v[q]);
    }
  }
}

// Group pass with the group's top bit B known at compile time.  The LDS slot map is
// linear over XOR (slot(a ^ b) = slot(a) ^ slot(b)), and every element of the group is
// irep ^ c(q) with a constant c(q), so each slot is slot(irep) ^ slot(c(q)): one XOR per
// element.  FLIP groups are a level's first group (kl = B + 1).
__host__ __device__ constexpr uint32_t group_const(int q, int B, bool flip) {
  return (static_cast<uint32_t>(q & (kGroup / 2 - 1)) << (B - (kGroupLog - 1))) ^
         ((q & (kGroup / 2)) ? (flip ? ((2u << B) - 1u) : (1u << B)) : 0u);
}

template <class T, int B, bool FLIP>
__device__ __attribute__((noinline)) void lds_group_pass(T* s, uint32_t gid, int r0) {
  constexpr uint32_t LOWB = B - (kGroupLog - 1);
  const uint32_t irep = ((gid >> LOWB) << (B + 1)) | (gid & ((1u << LOWB) - 1u));
  const uint32_t base = lds_slot(irep);
  T v[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) v[q] = s[base ^ lds_slot(group_const(q, B, FLIP))];
  group_stages<T, FLIP>(v, r0);
#pragma unroll
  for (int q = 0; q < kGroup; ++q) s[base ^ lds_slot(group_const(q, B, FLIP))] = v[q];
}

// Group pass with the geometry at run time (8-byte elements: the compile-time passes
// above cost more in call overhead and registers than they save in address math).
template <class T, bool FLIP>
__device__ __forceinline__ void lds_group_pass_rt(T* s, uint32_t gid, uint32_t b, int r0) {
  const uint32_t lowb = b - (kGroupLog - 1);
  const uint32_t irep = ((gid >> lowb) << (b + 1)) | (gid & ((1u << lowb) - 1u));
  const uint32_t m_top = FLIP ? ((2u << b) - 1u) : (1u << b);
  T v[kGroup];
#pragma unroll
  for (int q = 0; q < kGroup; ++q) {
    const uint32_t e = irep ^ (static_cast<uint32_t>(q & (kGroup / 2 - 1)) << lowb) ^
                       ((q & (kGroup / 2)) ? m_top : 0u);
    v[q] = s[lds_slot(e)];
  }
  group_stages<T, FLIP>(v, r0);
#pragma unroll
  for (int q = 0; q < kGroup; ++q) {
    const uint32_t e = irep ^ (static_cast<uint32_t>(q & (kGroup / 2 - 1)) << lowb) ^
                       ((q & (kGroup / 2)) ? m_top : 0u);
    s[lds_slot(e)] = v[q];
  }
}

template <class T, bool FLIP>
__device__ __forceinline__ void lds_group_dispatch(T* s, uint32_t gid, int b, int r0) {
  if constexpr (sizeof(T) == 8) {
    lds_group_pass_rt<T, FLIP>(s, gid, static_cast<uint32_t>(b), r0);
    return;
  }
  switch (b) {
#define PXK_GB(BB) \
  case BB: lds_group_pass<T, BB, FLIP>(s, gid, r0); break;
    PXK_GB(4) PXK_GB(5) PXK_GB(6) PXK_GB(7) PXK_GB(8)
    PXK_GB(9) PXK_GB(10) PXK_GB(11) PXK_GB(12) PXK_GB(13)
#undef PXK_GB
    default: break;
  }
}

// Stages on bits top .. 0 of one level (the first one the flip stage when flip), in
// groups of five; the last group is widened down to bit 0 and starts at a later stage.
template <class T>
__device__ __forceinline__ void lds_stages(T* s, uint32_t groups, int top, bool flip) {
  int t = top;
  while (t >= 0) {
    const int b = t >= kGroupLog - 1 ? t : kGroupLog - 1;
    const int r0 = b - t;
    if (threadIdx.x < groups) {
      if (flip)
        lds_group_dispatch<T, true>(s, threadIdx.x, b, r0);
      else
        lds_group_dispatch<T, false>(s, threadIdx.x, b, r0);
    }
    __syncthreads();
    t = b - kGroupLog;
    flip = false;
  }
}

template <class T>
struct SortCfg {
  static constexpr int kTileLog = (sizeof(T) == 4) ? 14 : 13;  // 64 KiB of LDS
  static constexpr int kThreads = 1 << (kTileLog - kGroupLog);  // one group per thread
  static constexpr int kMaxC = (sizeof(T) == 4) ? 7 : 6;  // register pass: 2^C elements per thread
  static int max_c() {
    const int c = sizeof(T) == 4 ? tune().sort_c4 : tune().sort_c8;
    return c < 1 ? 1 : (c > kMaxC ? kMaxC : c);
  }
};

// MERGE == false: full sort of each tile (levels 1 .. tl).  MERGE == true: the
// half-cleaners on bits tl-1 .. 0 of level kl (> tile).
template <class T, bool MERGE>
__global__ __launch_bounds__(SortCfg<T>::kThreads) void k_sort_local(T* data, uint64_t n,
                                                                     const uint64_t* seg_off,
                                                                     const uint32_t* blk_seg,
                                                                     const uint32_t* blk_first,
                                                                     uint32_t kl) {
  constexpr int TL = SortCfg<T>::kTileLog;
  constexpr uint32_t NT = SortCfg<T>::kThreads;
  extern __shared__ __align__(16) unsigned char s_raw[];
  T* s = reinterpret_cast<T*>(s_raw);
  const SortSeg sg = sort_seg(seg_off, blk_seg, blk_first, n);
  const uint32_t tl = sg.nplog < TL ? sg.nplog : TL;
  const uint64_t tbase = static_cast<uint64_t>(sg.lblock) << tl;
  if (tbase >= sg.n || tl == 0) return;
  if (MERGE && sg.nplog < kl) return;  // segment already sorted below this level
  // tiles shorter than one register group are padded to 32 with +inf in LDS
  const uint32_t te = tl < kGroupLog ? kGroup : (1u << tl);
  T* d = data + sg.base + tbase;
  const uint64_t lim = sg.n - tbase;  // valid elements of this tile (may exceed te)
  const uint32_t groups = te >> kGroupLog;

  if (te == (1u << TL) && lim >= te) {
#pragma unroll
    for (uint32_t k = 0; k < (1u << TL) / NT; ++k) {
      const uint32_t i = k * NT + threadIdx.x;
      s[lds_slot(i)] = d[i];
    }
  } else {
    for (uint32_t i = threadIdx.x; i < te; i += NT) s[lds_slot(i)] = i < lim ? d[i] : pos_inf<T>();
  }
  __syncthreads();
  if (!MERGE) {
    if (threadIdx.x < groups) {
      T v[kGroup];
      const uint32_t i0 = threadIdx.x << kGroupLog;
#pragma unroll
      for (int q = 0; q < kGroup; ++q) v[q] = s[lds_slot(i0 + q)];
      sort32<T>(v);
#pragma unroll
      for (int q = 0; q < kGroup; ++q) s[lds_slot(i0 + q)] = v[q];
    }
    __syncthreads();
    for (uint32_t L = kGroupLog + 1; L <= tl; ++L) lds_stages<T>(s, groups, static_cast<int>(L) - 1, true);
  } else {
    lds_stages<T>(s, groups, static_cast<int>(tl) - 1, false);
  }
  if (te == (1u << TL) && lim >= te) {
#pragma unroll
    for (uint32_t k = 0; k < (1u << TL) / NT; ++k) {
      const uint32_t i = k * NT + threadIdx.x;
      d[i] = s[lds_slot(i)];
    }
  } else {
    for (uint32_t i = threadIdx.x; i < te; i += NT)
      if (i < lim) d[i] = s[lds_slot(i)];
  }
}

// Register pass: C consecutive stages of level kl whose top partner bit is b (the
// group's bits are b-C+1 .. b).  FLIP: the first stage is the level's flip stage.
template <class T, int C, bool FLIP>
__global__ __launch_bounds__(256) void k_sort_global(T* data, uint64_t n, const uint64_t* seg_off,
                                                     const uint32_t* blk_seg,
                                                     const uint32_t* blk_first, uint32_t b,
                                                     uint32_t kl) {
  constexpr int Q = 1 << C;
  const SortSeg sg = sort_seg(seg_off, blk_seg, blk_first, n);
  if (sg.nplog < kl) return;  // segment too short for this level
  const uint64_t groups = (1ull << sg.nplog) >> C;
  const uint64_t gid = static_cast<uint64_t>(sg.lblock) * 256 + threadIdx.x;
  if (gid >= groups) return;
  const uint32_t lowb = b - C + 1;
  const uint64_t irep = ((gid >> lowb) << (b + 1)) | (gid & ((1ull << lowb) - 1));
  T* d = data + sg.base;
  const uint64_t lim = sg.n;

  // Element q: group bits b-C+1 .. b-1 are q's low C-1 bits (contiguous), the top one
  // selects irep or irep ^ m_top (the flip mask on a level's first pass).  Indices are
  // recomputed at the store instead of held in registers (C = 7 holds 128 elements).
  const uint64_t m_top = FLIP ? ((1ull << kl) - 1) : (1ull << b);
  auto elem = [&](int q) -> uint64_t {
    return irep ^ (static_cast<uint64_t>(q & (Q / 2 - 1)) << lowb) ^ ((q & (Q / 2)) ? m_top : 0ull);
  };
  T v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const uint64_t e = elem(q);
    v[q] = e < lim ? d[e] : pos_inf<T>();
  }
#pragma unroll
  for (int r = 0; r < C; ++r) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if ((q >> (C - 1 - r)) & 1) continue;
      const int q2 = q | (1 << (C - 1 - r));
      // The element of q sits at the lower index unless the flip mask (which sets every
      // bit below b) was applied to it and this stage's bit is one of those bits.
      const bool q_high = FLIP && r >= 1 && ((q >> (C - 1)) & 1);
      if (!q_high)
        cmpx(v[q], v[q2]);
      else
        cmpx(v[q2], v[q]);
    }
  }
  // A copy of irep the compiler cannot see through (an empty asm that only claims the
  // register): the store indices are recomputed from it instead of kept alive from the
  // loads, which would cost two registers per element.
  uint64_t irep_s = irep;
  asm volatile("" : "+v"(irep_s));
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const uint64_t e =
        irep_s ^ (static_cast<uint64_t>(q & (Q / 2 - 1)) << lowb) ^ ((q & (Q / 2)) ? m_top : 0ull);
    if (e < lim) d[e] = v[q];
  }
}


template <class T>
hipError_t launch_local(bool merge, uint32_t blocks, T* data, uint64_t n, const uint64_t* seg_off,
                        const uint32_t* blk_seg, const uint32_t* blk_first, uint32_t kl,
                        hipStream_t stream) {
  constexpr int TL = SortCfg<T>::kTileLog;
  const size_t lds = sizeof(T) << TL;
  constexpr int NT = SortCfg<T>::kThreads;
  if (merge)
    k_sort_local<T, true><<<blocks, NT, lds, stream>>>(data, n, seg_off, blk_seg, blk_first, kl);
  else
    k_sort_local<T, false><<<blocks, NT, lds, stream>>>(data, n, seg_off, blk_seg, blk_first, kl);
  return hipGetLastError();
}

template <class T>
hipError_t launch_global(int c, bool flip, uint32_t blocks, T* data, uint64_t n,
                         const uint64_t* seg_off, const uint32_t* blk_seg,
                         const uint32_t* blk_first, uint32_t b, uint32_t kl, hipStream_t stream) {
#define PXK_G(CC)                                                                                \
  case CC:                                                                                       \
    if constexpr (CC <= SortCfg<T>::kMaxC) {                                                     \
      if (flip)                                                                                  \
        k_sort_global<T, CC, true><<<blocks, 256, 0, stream>>>(data, n, seg_off, blk_seg, blk_first, b, kl); \
      else                                                                                       \
        k_sort_global<T, CC, false><<<blocks, 256, 0, stream>>>(data, n, seg_off, blk_seg, blk_first, b, kl); \
      break;                                                                                     \
    } else {                                                                                     \
      return hipErrorInvalidValue;                                                               \
    }
  switch (c) {
    PXK_G(1)
    PXK_G(2)
    PXK_G(3)
    PXK_G(4)
    PXK_G(5)
    PXK_G(6)
    PXK_G(7)
    default: return hipErrorInvalidValue;
  }
#undef PXK_G
  return hipGetLastError();
}

inline uint32_t host_ceil_log2(uint64_t n) {
  uint32_t l = 0;
  while ((1ull << l) < n) ++l;
  return l;
}

template <class T>
hipError_t sort_t(void* data_v, uint64_t n, hipStream_t stream, const LaunchHook& hook) {
  T* data = reinterpret_cast<T*>(data_v);
  if (n <= 1) return hipSuccess;
  constexpr int TL = SortCfg<T>::kTileLog;
  const uint32_t nplog = host_ceil_log2(n);
  const uint32_t tl = nplog < TL ? nplog : TL;
  // tiles that hold at least one real element
  const uint64_t ntiles = (n + (1ull << tl) - 1) >> tl;
  hook("sort_tile", 0);
  hipError_t e = launch_local<T>(false, static_cast<uint32_t>(ntiles), data, n, nullptr, nullptr,
                                 nullptr, 0, stream);
  hook("sort_tile", 1);
  if (e != hipSuccess) return e;
  for (uint32_t kl = tl + 1; kl <= nplog; ++kl) {
    int remaining = static_cast<int>(kl - tl);
    uint32_t b = kl - 1;
    bool first = true;
    while (remaining > 0) {
      const int c = std::min(remaining, SortCfg<T>::max_c());
      const uint64_t groups = (1ull << nplog) >> c;
      const uint64_t blocks = (groups + 255) / 256;
      hook("sort_global", 0);
      e = launch_global<T>(c, first, static_cast<uint32_t>(blocks), data, n, nullptr, nullptr,
                           nullptr, b, kl, stream);
      hook("sort_global", 1);
      if (e != hipSuccess) return e;
      b -= c;
      remaining -= c;
      first = false;
    }
    hook("sort_merge", 0);
    e = launch_local<T>(true, static_cast<uint32_t>(ntiles), data, n, nullptr, nullptr, nullptr, kl,
                        stream);
    hook("sort_merge", 1);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---- segmented: block tables per launch kind, built on the host.
//   local (tile sort and merges): blocks(s) = ceil(n_s / tile_s)
//   global pass with c stages:    blocks(s) = ceil((np_s >> c) / 256), only for segments
//                                 longer than one tile (others never reach a global stage)
template <class T>
bool build_plan(const uint64_t* lo, const uint64_t* hi, uint32_t nseg, SegSortPlan& p) {
  constexpr int TL = SortCfg<T>::kTileLog;
  const int MC = SortCfg<T>::max_c();
  p.nseg = nseg;
  p.tile_log = TL;
  p.max_c = MC;
  p.max_nplog = 0;
  std::vector<uint32_t> local_seg, local_first(nseg, 0);
  std::vector<std::vector<uint32_t>> gseg(MC + 1), gfirst(MC + 1, std::vector<uint32_t>(nseg, 0));
  for (uint32_t s = 0; s < nseg; ++s) {
    if (hi[s] < lo[s]) return false;
    const uint64_t len = hi[s] - lo[s];
    const uint32_t nl = host_ceil_log2(len);
    p.max_nplog = std::max(p.max_nplog, nl);
    const uint32_t tl = nl < static_cast<uint32_t>(TL) ? nl : static_cast<uint32_t>(TL);
    const uint64_t blocks = len <= 1 ? 0 : (len + (1ull << tl) - 1) >> tl;
    local_first[s] = static_cast<uint32_t>(local_seg.size());
    for (uint64_t k = 0; k < blocks; ++k) local_seg.push_back(s);
    for (int c = 1; c <= MC; ++c) {
      gfirst[c][s] = static_cast<uint32_t>(gseg[c].size());
      if (nl <= static_cast<uint32_t>(TL)) continue;
      const uint64_t gb = (((1ull << nl) >> c) + 255) / 256;
      for (uint64_t k = 0; k < gb; ++k) gseg[c].push_back(s);
    }
  }
  p.tables.clear();
  p.tables.insert(p.tables.end(), local_seg.begin(), local_seg.end());
  p.nlocal = static_cast<uint32_t>(local_seg.size());
  p.o_local_first = p.tables.size();
  p.tables.insert(p.tables.end(), local_first.begin(), local_first.end());
  for (int c = 1; c <= MC; ++c) {
    p.o_gseg[c] = p.tables.size();
    p.gblocks[c] = static_cast<uint32_t>(gseg[c].size());
    p.tables.insert(p.tables.end(), gseg[c].begin(), gseg[c].end());
    p.o_gfirst[c] = p.tables.size();
    p.tables.insert(p.tables.end(), gfirst[c].begin(), gfirst[c].end());
  }
  return true;
}

template <class T>
hipError_t seg_sort_run_t(const SegSortPlan& p, void* data_v, const uint64_t* off_d,
                          const uint32_t* w, hipStream_t stream) {
  T* data = reinterpret_cast<T*>(data_v);
  constexpr int TL = SortCfg<T>::kTileLog;
  if (p.nlocal == 0) return hipSuccess;
  hipError_t e = launch_local<T>(false, p.nlocal, data, 0, off_d, w, w + p.o_local_first, 0, stream);
  if (e != hipSuccess) return e;
  for (uint32_t kl = TL + 1; kl <= p.max_nplog; ++kl) {
    int remaining = static_cast<int>(kl - TL);
    uint32_t b = kl - 1;
    bool first = true;
    while (remaining > 0) {
      const int c = std::min(remaining, p.max_c);
      if (p.gblocks[c]) {
        e = launch_global<T>(c, first, p.gblocks[c], data, 0, off_d, w + p.o_gseg[c],
                             w + p.o_gfirst[c], b, kl, stream);
        if (e != hipSuccess) return e;
      }
      b -= c;
      remaining -= c;
      first = false;
    }
    // merge pass: blocks of segments shorter than this level return at once
    e = launch_local<T>(true, p.nlocal, data, 0, off_d, w, w + p.o_local_first, kl, stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

bool seg_sort_plan(int dtype, const uint64_t* lo_host, const uint64_t* hi_host, uint32_t nseg,
                   SegSortPlan& p) {
  p.dtype = dtype;
  PXK_DISPATCH(dtype, return build_plan<T>(lo_host, hi_host, nseg, p));
  return false;
}

hipError_t seg_sort_run(const SegSortPlan& p, void* data, const uint64_t* seg_off_dev,
                        const uint32_t* tables_dev, hipStream_t stream) {
  PXK_DISPATCH(p.dtype, return seg_sort_run_t<T>(p, data, seg_off_dev, tables_dev, stream));
  return hipErrorInvalidValue;
}

hipError_t sort(int dtype, void* data, uint64_t n, hipStream_t stream, const LaunchHook& hook) {
  PXK_DISPATCH(dtype, return sort_t<T>(data, n, stream, hook));
  return hipErrorInvalidValue;
}

}  // namespace pxk
