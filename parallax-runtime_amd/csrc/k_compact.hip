// k_compact.hip -- stable stream compaction for gfx950: copy_if, remove_if (negated
// predicate), unique (adjacent-difference flags) and partition, in one pass.
//
// Replaces the reference's flags.comp -> recursive scan of the flags in the element
// type -> host read of the kept count -> scatter.comp schedule
// (KernelLauncher::launch_compact, src/kernel/kernel_launcher.cpp:1701-1791; about
// 9 host-synchronised dispatches and 7-9 n*s bytes, with float positions that lose
// exactness above 2^24) with ONE launch: each tile evaluates the predicate on its
// elements, ranks the kept ones with DPP wave scans of integer counts, stages them in
// LDS in order, looks back for the number kept before the tile, and writes them as one
// contiguous, coalesced run.  Bytes moved: n*s read + kept*s written.
//
// Partitions need the total kept count first: a counting pass writes num_true to device
// memory and the next pass reads it there (no host round trip).
//  * stable partition (kept first, the rest after num_true, both in input order): the
//    compaction pass places rejected elements too;
//  * std::partition's order (the fallback of the partition funnel, reference
//    include/parallax/stdpar.hpp:369, libstdc++'s two-pointer form: the i-th predicate-
//    false element of [0, num_true) swaps with the i-th predicate-true element of
//    [num_true, n) counted from the end): the compaction pass writes the INDICES of the
//    misplaced elements (false below num_true, true at or above it) in order, 2m of them,
//    and a swap pass exchanges index i with index 2m-1-i for i < m.  In place.
#include "dev_common.hpp"
#include "dev_pred.hpp"
#include "dev_tiles.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

// Writes the staged run src[0, count) to dst (any element alignment): a scalar head up
// to the next 16-byte boundary, 16-byte non-temporal stores for the body, a scalar tail.
// `src` is 16-byte aligned and readable one vector past `count` (the staging area's
// spare vector): a body vector is read as the two aligned LDS vectors around it (16-byte
// reads, no bank conflicts) and shifted by the run's alignment, uniform per run.
template <class T, int BLOCK, class C>
__device__ __forceinline__ void write_run(T* dst, const T* src, C count) {
  constexpr int VEC = 16 / sizeof(T);
  using V = Vec<T, VEC>;
  const uint32_t mis = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(dst) & 15u) / sizeof(T));
  uint32_t head = (VEC - mis) % VEC;
  if (head > count) head = static_cast<uint32_t>(count);
  if (threadIdx.x < head) dst[threadIdx.x] = src[threadIdx.x];
  const uint32_t nvec = static_cast<uint32_t>((count - head) / VEC);
  V* vd = reinterpret_cast<V*>(dst + head);
  const V* vs = reinterpret_cast<const V*>(src);
  // the body starts `head` (< VEC, uniform) elements past an LDS vector
  auto body = [&](auto shc) {
    constexpr int SH = decltype(shc)::value;
#pragma unroll 1
    for (uint32_t v = threadIdx.x; v < nvec; v += BLOCK) {
      if constexpr (SH == 0) {
        st_stream(vd + v, vs[v]);
      } else {
        const V lo = vs[v], hi = vs[v + 1];
        V y;
#pragma unroll
        for (int j = 0; j < VEC; ++j) y.v[j] = (j + SH < VEC) ? lo.v[j + SH] : hi.v[j + SH - VEC];
        st_stream(vd + v, y);
      }
    }
  };
  switch (head) {
    case 0: body(std::integral_constant<int, 0>{}); break;
    case 1: body(std::integral_constant<int, 1>{}); break;
    case 2: if constexpr (VEC > 2) body(std::integral_constant<int, 2>{}); break;
    default: if constexpr (VEC > 3) body(std::integral_constant<int, 3>{}); break;
  }
  const uint32_t done = head + nvec * VEC;
  if (threadIdx.x < count - done) dst[done + threadIdx.x] = src[done + threadIdx.x];
}

// MODE kModeKept: kept elements; kModeStable: kept first, then the rest after
// *num_true_dev; kModeMisplaced: the indices (type C) of the elements on the wrong side of
// *num_true_dev (predicate false below it, true at or above it), in order.  The output
// and staging element type S is T, or C for kModeMisplaced.
template <class T, int BLOCK, int K, class C, int MODE, bool UNIQUE>
__global__ __launch_bounds__(BLOCK) void k_compact(
    const T* in_al, int64_t in_shift, std::conditional_t<MODE == kModeMisplaced, C, T>* out,
    SegGeom g, px_desc pd, uint64_t* counts, const uint64_t* num_true_dev, uint64_t* status,
    Ctl* ctl) {
  using S = std::conditional_t<MODE == kModeMisplaced, C, T>;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int TILE = BLOCK * K * VEC;
  constexpr int NW = BLOCK / kWave;
  constexpr int E = K * VEC;  // elements per thread (<= 32: one flag bit each)
  static_assert(E <= 32, "flag mask holds 32 elements");
  static_assert(TILE * sizeof(S) <= 128 * 1024, "staging fits the LDS");
  using V = Vec<T, VEC>;
  const V* vin = reinterpret_cast<const V*>(in_al);

  // the staged run, one spare vector (write_run reads one past the run) and, for the
  // kept-only mode, one slot per thread that rejected elements are written to
  constexpr int kSpare = 16 / sizeof(S) + (MODE == kModeKept ? BLOCK : 0);
  __shared__ __attribute__((aligned(16))) S s_stage[TILE + kSpare];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_next;
  __shared__ uint32_t s_epoch;
  __shared__ C s_w[NW];
  __shared__ C s_prefix;

  const Pred<T> pred(pd);
  const int lane = lane_id();
  // Wave index, tickets and tile geometry are wave-uniform and read as such (scalar
  // registers, scalar branches): a branch the compiler takes for divergent runs both arms
  // under lane masks, and its waits then cover both arms' loads.
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);

  // Positions inside the tile window are handled as 32-bit offsets rel in [0, TILE):
  // thread element e = k*VEC + j sits at rel = base_rel(k) + j; the segment covers
  // [lo_r, hi_r) of the window (lo_r = -1 when the segment starts before it; an empty
  // range when it ends before the window or starts after it).
  // `tid_o` is threadIdx.x made opaque once per tile: position arithmetic is redone per
  // tile rather than hoisted out of the tile loop, where its 32 invariants would be
  // live (and spilled) across the whole loop.
  int tid_o = static_cast<int>(threadIdx.x);
  auto base_rel = [&](int k) { return ((tid_o / kWave * K + k) * kWave + tid_o % kWave) * VEC; };

  struct Tile {
    T x[E];
    uint32_t flags, valid;
    C tex[K];
    C wave_off, kept;
    int lo_r, hi_r, vstart_r;
    int64_t w0, lo;
  };
  // the split of a partition (read once: the counting pass ran before on this stream)
  const uint64_t split = (MODE == kModeMisplaced) ? *num_true_dev : 0;

  // Raw loads of the tile and its valid bits (no LDS, no barrier: may run ahead).
  auto load = [&](Tile& u, const TileInfo& ti) {
    u.w0 = ti.w0;
    u.lo = ti.lo;
    // window-relative bounds in 32 bits (uniform 64-bit differences split into halves)
    const int64_t dl = ti.lo - u.w0, dh = ti.hi - u.w0;
    u.lo_r = static_cast<int32_t>(static_cast<uint64_t>(dl) >> 32) < 0 ? -1 : clamp_span(dl, TILE);
    u.hi_r = clamp_span(dh, TILE);
    u.vstart_r = u.lo_r > 0 ? u.lo_r : 0;
    const bool full = u.lo_r <= 0 && u.hi_r == TILE;

    // ---- raw loads and valid bits.  K buffer loads per thread for every tile, over the
    // window's vectors up to the segment's end (past them they read zero without touching
    // memory), so a wait for this tile counts out the loads issued after it; a partial
    // tile differs only in its valid bits (vector arithmetic, no memory branch).
    const uint32_t nvec = static_cast<uint32_t>(u.hi_r + VEC - 1) / VEC;
    const __amdgpu_buffer_rsrc_t rs = uniform_rsrc(vin + u.w0 / VEC, nvec * 16u);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const V v = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(
                                            rs, base_rel(k) / VEC * 16, 0, 2));
#pragma unroll
      for (int j = 0; j < VEC; ++j) u.x[k * VEC + j] = v.v[j];
    }
    if (full) {
      u.valid = (E == 32) ? 0xffffffffu : ((1u << E) - 1u);
    } else {
      u.valid = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r0 = base_rel(k);
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          if (r0 + j >= u.lo_r && r0 + j < u.hi_r) u.valid |= 1u << (k * VEC + j);
      }
    }
  };

  // Flags and ranks the loaded tile (ends with the block's counts in s_w read out);
  // thread 0 takes the block's next ticket into s_next on the way.
  auto rank = [&](Tile& u) {
    // ---- flags: the predicate (or adjacent difference) on valid positions
    if constexpr (UNIQUE) {
      u.flags = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r0 = base_rel(k);
        // previous element of lane's first item: lane-1's last item, or memory for lane 0
        T prev = dpp_mov<kWaveShr1, 0xf, 0xf>(u.x[k * VEC + VEC - 1]);
        if (lane == 0 && r0 - 1 >= u.lo_r && r0 - 1 < u.hi_r) prev = in_al[u.w0 + r0 - 1];
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const T pv = (j == 0) ? prev : u.x[k * VEC + j - 1];
          if ((r0 + j == u.lo_r) || !(u.x[k * VEC + j] == pv)) u.flags |= 1u << (k * VEC + j);
        }
      }
      u.flags &= u.valid;
    } else {
      u.flags = pred.template mask<E>(u.x) & u.valid;
    }
    if constexpr (MODE == kModeMisplaced) {
      // position pos = w0 + rel - lo is misplaced when pred(x) == (pos >= split):
      // flip the predicate bits of the positions below the split
      const int64_t pr64 = static_cast<int64_t>(split) + u.lo - u.w0;
      const int pr = pr64 < 0 ? 0 : (pr64 > TILE ? TILE : static_cast<int>(pr64));
      uint32_t left = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r0 = base_rel(k);
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          if (r0 + j < pr) left |= 1u << (k * VEC + j);
      }
      u.flags = (u.flags ^ left) & u.valid;
    }

    // ---- rank kept elements: per-chunk wave scans of counts
    C run = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const C c = static_cast<C>(__builtin_popcount((u.flags >> (k * VEC)) & ((1u << VEC) - 1u)));
      const C incl = wave_incl_scan(c);
      u.tex[k] = run + wave_shift_right1(incl);
      run += read_lane(incl, kWave - 1);
    }
    if (lane == 0) s_w[wid] = run;
    if (threadIdx.x == 0)
      s_next = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    u.wave_off = 0;
    for (int j = 0; j < wid; ++j) u.wave_off += s_w[j];
    u.kept = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) u.kept += s_w[j];
  };

  // Stages the kept (and, for partition, rejected) elements in LDS, in order.
  auto stage = [&](const Tile& u) {
    if constexpr (MODE == kModeKept) {
      // branch-free: every element is written, a rejected one to this thread's slot
      const C dummy = static_cast<C>(TILE + 16 / sizeof(S) + threadIdx.x);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        C r = u.wave_off + u.tex[k];
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          const int e = k * VEC + j;
          const C f = static_cast<C>((u.flags >> e) & 1u);
          s_stage[f ? r : dummy] = u.x[e];
          r += f;
        }
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      C r = u.wave_off + u.tex[k];
      const int r0 = base_rel(k);
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        const int e = k * VEC + j;
        if ((u.flags >> e) & 1u) {
          if constexpr (MODE == kModeMisplaced)
            s_stage[r] = static_cast<S>(u.w0 + r0 + j - u.lo);
          else
            s_stage[r] = u.x[e];
          ++r;
        } else if (MODE == kModeStable) {
          if ((u.valid >> e) & 1u) {
            const C vi = static_cast<C>(r0 + j - u.vstart_r);  // valid index inside the tile
            s_stage[u.kept + (vi - r)] = u.x[e];
          }
        }
      }
    }
  };

  // Writes the staged run(s); kept_before = kept in the segment before the tile.
  auto write = [&](const Tile& u, const TileInfo& ti, uint64_t seg_out, uint64_t kept_before) {
    write_run<S, BLOCK>(out + seg_out + kept_before, s_stage, u.kept);
    if constexpr (MODE == kModeStable) {
      const uint64_t num_true = *num_true_dev;
      const C sub_valid = static_cast<C>(u.hi_r > u.vstart_r ? u.hi_r - u.vstart_r : 0);
      if (sub_valid > 0) {
        const uint64_t rej_before = static_cast<uint64_t>(u.w0 + u.vstart_r - ti.lo) - kept_before;
        write_run<S, BLOCK>(out + seg_out + num_true + rej_before, s_stage + u.kept, sub_valid - u.kept);
      }
    }
  };

  if (threadIdx.x == 0) {
    s_epoch = ld_agent32(&ctl->epoch);
    s_tile = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const uint32_t epoch = s_epoch;

  // Persistent loop over tickets.  The block takes its next ticket while ranking the
  // current tile and issues that tile's loads once the current one is staged in LDS,
  // before the look-back and the writes, so their latency overlaps.  A block holds at
  // most two tickets and finishes the lower first; the lowest unfinished ticket waits
  // only on finished ones, so every wait ends.  In place (out == in) stays safe: a
  // tile's outputs land below its own end, and its look-back ends only after every
  // earlier tile has loaded.
  uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
  if (tile < g.ntiles) {
    TileInfo ti = tile_info<VEC, TILE>(g, tile, in_shift);
    Tile a;
    load(a, ti);
    while (true) {
      asm volatile("" : "+v"(tid_o));
      rank(a);
      const C tile_kept = a.kept;
      // publish the tile's count before staging, so later tiles' look-backs need not wait
      if (threadIdx.x == 0 && tile != ti.first_tile) publish<C>(status, tile, epoch, kStAgg, tile_kept);
      stage(a);
      // the tile's values now sit in LDS: their registers take the next tile's loads
      __builtin_amdgcn_sched_barrier(0);
      const uint32_t next = __builtin_amdgcn_readfirstlane(s_next);
      const bool more = next < g.ntiles;
      TileInfo tn = ti;
      Tile b;
      // wave 0 looks back: it issues its own loads of the next tile once the look-back
      // is done (a poll's wait for its load would also wait for every load before it)
      if (more) {
        tn = tile_info<VEC, TILE>(g, next, in_shift);
        asm volatile("" : "+v"(tid_o));
        if (wid != 0) load(b, tn);
      }

      // ---- kept before this tile (look-back), then the tile's count
      if (wid == 0) {
        C prefix;
        if (tile == ti.first_tile) {
          prefix = 0;
        } else {
          prefix = lookback<C>(status, tile, ti.first_tile, epoch);
        }
        if (tile != ti.last_tile && lane == 0)
          publish<C>(status, tile, epoch, kStIncl, prefix + tile_kept);
        if (lane == 0) {
          s_prefix = prefix;
          if (tile == ti.last_tile) {
            const uint64_t total = static_cast<uint64_t>(prefix) + tile_kept;
            if (counts) counts[ti.seg] = total;
            else ctl->result = total;
          }
        }
        if (more) load(b, tn);
      }
      __syncthreads();

      const uint64_t P = s_prefix;
      const uint64_t seg_out =
          (g.seg_lo != nullptr) ? (g.seg_out ? ld_table(g.seg_out, ti.seg) : ld_table(g.seg_lo, ti.seg)) : 0;
      // ---- coalesced writes of the staged run(s)
      write(a, ti, seg_out, P);
      __syncthreads();
      if (!more) break;
      a = b;
      tile = next;
      ti = tn;
    }
  }
  if (threadIdx.x == 0) block_end(ctl, s_epoch);
}

// Predicate count over one range (partition's first pass): grid-stride, last block
// writes the total.
template <class T>
__global__ __launch_bounds__(256) void k_count(const T* in, uint64_t n, px_desc pd,
                                               uint64_t* partials, Ctl* ctl, uint64_t* out) {
  const Pred<T> pred(pd);
  __shared__ uint64_t s_w[4];
  __shared__ int s_last;
  uint64_t c = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * 256)
    c += pred(in[i]) ? 1u : 0u;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  uint64_t w = wave_sum(c);
  if (lane == 0) s_w[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    st_carry(&partials[blockIdx.x], s_w[0] + s_w[1] + s_w[2] + s_w[3]);
    const uint32_t d =
        __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
    if (s_last) st_agent32(&ctl->done, 0u);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  uint64_t t = 0;
  for (uint32_t j = threadIdx.x; j < gridDim.x; j += 256) t += ld_carry(&partials[j]);
  w = wave_sum(t);
  if (lane == 0) s_w[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) *out = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

template <class T, class C, int MODE, bool UNIQUE>
hipError_t launch_compact_t(hipStream_t stream, const T* in_al, int64_t in_shift, void* out,
                            const SegGeom& g, const px_desc& pd, uint64_t* counts,
                            const uint64_t* num_true_dev, const Scratch& s) {
  using S = std::conditional_t<MODE == kModeMisplaced, C, T>;
  S* o = static_cast<S*>(out);
  const TileCfg c{g.tile_block, g.tile_k};
  const uint32_t grid = lookback_grid(g.ntiles, s.max_blocks, c.block);
  if (grid == 0) return hipSuccess;
#define PXK_CMP(B, KK)                                                                        \
  if constexpr (B * KK * 16 / sizeof(T) * sizeof(S) <= 128 * 1024) {                          \
    if (c.block == B && c.k == KK) {                                                          \
      k_compact<T, B, KK, C, MODE, UNIQUE><<<grid, B, 0, stream>>>(                           \
          in_al, in_shift, o, g, pd, counts, num_true_dev, s.status, s.ctl);                  \
      return hipGetLastError();                                                               \
    }                                                                                         \
  }
  PXK_CMP(256, 4)
  PXK_CMP(256, 8)
  PXK_CMP(512, 8)
  PXK_CMP(1024, 4)
  PXK_CMP(1024, 8)
#undef PXK_CMP
  return hipErrorInvalidValue;
}

template <class T>
hipError_t compact_t(const px_desc& pd, int mode, const void* in, void* out, const SegGeom& g,
                     uint64_t* counts, const uint64_t* num_true_dev, const Scratch& s,
                     hipStream_t stream) {
  const T* in_al = reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(in) & ~uintptr_t(15));
  const int64_t in_shift = elem_shift(in, sizeof(T));
  if (g.ntiles == 0) return hipSuccess;
  // the total element count decides the count width (a segment stays below 2^32)
  const bool wide = (g.seg_lo != nullptr) ? false : (g.n >= (1ull << 32));
  const size_t words = wide ? 2 : 1;
  if (static_cast<size_t>(g.ntiles) * words > s.status_words) return hipErrorInvalidValue;
  const bool uniq = (pd.fn == PX_PRED_UNIQUE);
  if (mode == kModeStable || mode == kModeMisplaced) {
    if (g.seg_lo != nullptr || num_true_dev == nullptr || uniq) return hipErrorInvalidValue;
    if (mode == kModeMisplaced) {
      const TileCfg want = misplaced_tile(dtype_of<T>(), g.n);
      if (g.tile_block != want.block || g.tile_k != want.k) return hipErrorInvalidValue;
      if (wide)
        return launch_compact_t<T, uint64_t, kModeMisplaced, false>(stream, in_al, in_shift, out, g, pd, counts, num_true_dev, s);
      return launch_compact_t<T, uint32_t, kModeMisplaced, false>(stream, in_al, in_shift, out, g, pd, counts, num_true_dev, s);
    }
    if (wide)
      return launch_compact_t<T, uint64_t, kModeStable, false>(stream, in_al, in_shift, out, g, pd, counts, num_true_dev, s);
    return launch_compact_t<T, uint32_t, kModeStable, false>(stream, in_al, in_shift, out, g, pd, counts, num_true_dev, s);
  }
  if (mode != kModeKept) return hipErrorInvalidValue;
  if (uniq) {
    if (wide)
      return launch_compact_t<T, uint64_t, kModeKept, true>(stream, in_al, in_shift, out, g, pd, counts, nullptr, s);
    return launch_compact_t<T, uint32_t, kModeKept, true>(stream, in_al, in_shift, out, g, pd, counts, nullptr, s);
  }
  if (wide)
    return launch_compact_t<T, uint64_t, kModeKept, false>(stream, in_al, in_shift, out, g, pd, counts, nullptr, s);
  return launch_compact_t<T, uint32_t, kModeKept, false>(stream, in_al, in_shift, out, g, pd, counts, nullptr, s);
}

// Swap pass of std::partition's order: idx[0, 2m) holds the misplaced positions in
// increasing order (m below the split, m at or above it); position idx[i] exchanges with
// idx[2m-1-i] for i < m.  The pairs are disjoint, so the swaps need no ordering.
template <class T, class C>
__global__ __launch_bounds__(256) void k_swap_pairs(T* data, const C* idx, const uint64_t* count2) {
  const uint64_t two_m = *count2;
  const uint64_t m = two_m / 2;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < m;
       i += static_cast<uint64_t>(gridDim.x) * 256) {
    const uint64_t a = idx[i];
    const uint64_t b = idx[two_m - 1 - i];
    const T va = data[a];
    const T vb = data[b];
    data[a] = vb;
    data[b] = va;
  }
}

template <class T>
hipError_t swap_pairs_t(void* data, const void* idx, bool wide, const uint64_t* count2_dev,
                        uint64_t n, const Scratch& s, hipStream_t stream) {
  uint64_t grid = (n / 2 + 255) / 256;
  if (grid < 1) grid = 1;
  if (grid > s.max_blocks) grid = s.max_blocks;
  T* d = static_cast<T*>(data);
  if (wide)
    k_swap_pairs<T, uint64_t><<<static_cast<uint32_t>(grid), 256, 0, stream>>>(d, static_cast<const uint64_t*>(idx), count2_dev);
  else
    k_swap_pairs<T, uint32_t><<<static_cast<uint32_t>(grid), 256, 0, stream>>>(d, static_cast<const uint32_t*>(idx), count2_dev);
  return hipGetLastError();
}

template <class T>
hipError_t count_t(const px_desc& pd, const void* in, uint64_t n, uint64_t* count_dev,
                   const Scratch& s, hipStream_t stream) {
  uint64_t grid = (n + 256 * 16 - 1) / (256 * 16);
  if (grid < 1) grid = 1;
  if (grid > s.max_blocks) grid = s.max_blocks;
  if (grid * sizeof(uint64_t) > s.partial_bytes) return hipErrorInvalidValue;
  k_count<T><<<static_cast<uint32_t>(grid), 256, 0, stream>>>(
      reinterpret_cast<const T*>(in), n, pd, reinterpret_cast<uint64_t*>(s.partials), s.ctl,
      count_dev);
  return hipGetLastError();
}

}  // namespace

hipError_t compact(const px_desc& pred, int mode, const void* in, void* out, const SegGeom& g,
                   uint64_t* counts, const uint64_t* num_true_dev, const Scratch& s,
                   hipStream_t stream) {
  PXK_DISPATCH(pred.dtype, return compact_t<T>(pred, mode, in, out, g, counts, num_true_dev, s, stream));
  return hipErrorInvalidValue;
}

hipError_t count_if(const px_desc& pred, const void* in, uint64_t n, uint64_t* count_dev,
                    const Scratch& s, hipStream_t stream) {
  PXK_DISPATCH(pred.dtype, return count_t<T>(pred, in, n, count_dev, s, stream));
  return hipErrorInvalidValue;
}

TileCfg misplaced_tile(int dtype, uint64_t n) {
  // 1024 x 8 vectors, or 1024 x 4 when 64-bit indices of 4-byte elements would not fit
  // the 128 KiB staging area
  return (dtype_size(dtype) == 4 && n >= (1ull << 32)) ? TileCfg{1024, 4} : TileCfg{1024, 8};
}

hipError_t swap_pairs(int dtype, void* data, const void* idx, bool wide_idx,
                      const uint64_t* count2_dev, uint64_t n, const Scratch& s,
                      hipStream_t stream) {
  PXK_DISPATCH(dtype, return swap_pairs_t<T>(data, idx, wide_idx, count2_dev, n, s, stream));
  return hipErrorInvalidValue;
}

}  // namespace pxk
