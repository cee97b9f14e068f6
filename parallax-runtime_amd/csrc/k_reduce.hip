// k_reduce.hip -- sum reduction for gfx950, plain, fused with an element map, and
// segmented.
//
// Plain form (reference: KernelLauncher::launch_reduce, src/kernel/kernel_launcher.cpp:908-985,
// which runs reduce.comp level after level with a host wait between levels): here ONE
// launch.  A grid of one 256-thread block per CU streams the input with non-temporal
// 16-byte loads, eight in flight per thread over two accumulator chains, in the carry
// type; each wave folds with DPP, waves fold through LDS, every block leaves one
// partial, and the last block to finish (agent-scope counter) folds the partials in a
// fixed order and writes the result.  The counter resets itself, so back-to-back calls
// need no reset launch.
//
// Fused map-reduce (transform_reduce / count_if): the same kernel with an element map
// applied between the load and the add, so Σ f(x) costs one read of x (the reference
// path, stdpar.hpp:246-293, writes f(x) to an arena array and reduces it: three passes).
//
// Segmented form: pass 1 walks the tiles of all segments (static assignment: no tile
// waits on another); single-tile segments write their sum directly, larger ones leave
// one partial per tile; pass 2 folds each segment's partials in tile order.  Segment s's
// sum goes to out[seg_out[s]] (out[s] without an output table).
#include "dev_common.hpp"
#include "dev_map.hpp"
#include "dev_tiles.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {


// 16-byte load, optionally non-temporal (streamed data is read once: keep it out of the
// caches the look-back and partials use).
template <class V, bool NT>
__device__ __forceinline__ V load16(const V* p) {
  static_assert(sizeof(V) == 16, "16-byte vectors only");
  if constexpr (NT) {
    const u32x4 r = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return __builtin_bit_cast(V, r);
  } else {
    return *p;
  }
}

// Identity map (plain reduce).
template <class T>
struct MapNone {
  __device__ __forceinline__ T operator()(T x) const { return x; }
};

// Adds map(v) for the VEC elements of one vector to acc.
template <class TI, class TO, int VEC, class M>
__device__ __forceinline__ void add_mapped(typename Traits<TO>::A& acc, const Vec<TI, VEC>& v,
                                           const M& m) {
  if constexpr (is_map_pred<M>::value) {
    acc = acc_add(acc, to_acc(static_cast<TO>(__builtin_popcount(m.p.template mask<VEC>(v.v)))));
  } else {
#pragma unroll
    for (int j = 0; j < VEC; ++j) acc = acc_add(acc, to_acc(m(v.v[j])));
  }
}

template <class TI, class TO, class M>
__device__ __forceinline__ TO map_one(TI x, const M& m) {
  if constexpr (is_map_pred<M>::value) {
    return m.p(x) ? TO(1) : TO(0);
  } else {
    return m(x);
  }
}

template <class TI, class TO, int VEC, int UNROLL, bool NT, class M>
__global__ __launch_bounds__(256) void k_reduce(const TI* __restrict__ in_al, int64_t lo,
                                                int64_t hi,
                                                typename Traits<TO>::A* __restrict__ partials,
                                                Ctl* ctl, TO* out, M m) {
  using A = typename Traits<TO>::A;
  using V = Vec<TI, VEC>;
  const V* vin = reinterpret_cast<const V*>(in_al);
  __shared__ A s_w[4];
  __shared__ int s_last;

  const int64_t a = (lo + VEC - 1) / VEC;  // interior vectors [a, b)
  const int64_t b = hi / VEC;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  int64_t i = a + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  // four independent accumulator chains (a double-double add is a long dependent
  // sequence), combined in a fixed order
  constexpr int NCH = 4;
  A ch[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) ch[c] = acc_zero<A>();
  for (; i + (UNROLL - 1) * stride < b; i += UNROLL * stride) {
    V v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) v[u] = load16<V, NT>(vin + i + u * stride);
    if constexpr (is_map_pred<M>::value) {
      // one predicate evaluation (one branch on its kind) for the whole group
      TI x[UNROLL * VEC];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
#pragma unroll
        for (int j = 0; j < VEC; ++j) x[u * VEC + j] = v[u].v[j];
      ch[0] = acc_add(ch[0], to_acc(static_cast<TO>(__builtin_popcount(m.p.template mask<UNROLL * VEC>(x)))));
    } else {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) add_mapped<TI, TO, VEC>(ch[u % NCH], v[u], m);
    }
  }
  for (; i < b; i += stride) add_mapped<TI, TO, VEC>(ch[0], load16<V, NT>(vin + i), m);
  A acc = acc_add(acc_add(ch[0], ch[1]), acc_add(ch[2], ch[3]));
  if (blockIdx.x == 0 && threadIdx.x < 2 * VEC) {
    // head [lo, min(a*VEC, hi)) and tail [max(a,b)*VEC, hi): fewer than VEC each
    int64_t pos;
    bool ok;
    if (threadIdx.x < VEC) {
      pos = lo + threadIdx.x;
      ok = pos < a * VEC && pos < hi;
    } else {
      pos = (a > b ? a : b) * VEC + (threadIdx.x - VEC);
      ok = pos < hi;
    }
    if (ok) acc = acc_add(acc, to_acc(map_one<TI, TO>(in_al[pos], m)));
  }

  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  A w = wave_sum(acc);
  if (lane == 0) s_w[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    const A tot = acc_add(acc_add(s_w[0], s_w[1]), acc_add(s_w[2], s_w[3]));
    st_carry(&partials[blockIdx.x], tot);
    const uint32_t d =
        __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
    if (s_last) st_agent32(&ctl->done, 0u);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  A c = acc_zero<A>();
  for (uint32_t j = threadIdx.x; j < gridDim.x; j += 256) c = acc_add(c, ld_carry(&partials[j]));
  w = wave_sum(c);
  if (lane == 0) s_w[wid] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    const A tot = acc_add(acc_add(s_w[0], s_w[1]), acc_add(s_w[2], s_w[3]));
    *out = from_acc<TO>(tot);
  }
}

// Segmented reduce, pass 1: tile partials.  Tiles are independent (static grid-stride
// assignment, no atomics); a segment held by one tile writes its result directly, the
// others leave one partial per tile for pass 2.  Software-pipelined: a block looks up its
// next tile and issues that tile's loads before it adds up the current one, so the
// table lookups and the memory latency of the next tile overlap this one's work.
template <class T, int K>
__global__ __launch_bounds__(256) void k_seg_reduce(const T* __restrict__ in_al, int64_t shift,
                                                    SegGeom g,
                                                    typename Traits<T>::A* __restrict__ partials,
                                                    T* __restrict__ out) {
  using A = typename Traits<T>::A;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int TILE = 256 * K * VEC;
  using V = Vec<T, VEC>;
  const V* vin = reinterpret_cast<const V*>(in_al);
  __shared__ A s_w[4];
  const int lane = lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);

  // Tile-relative bounds in 32 bits: the segment covers [lo_r, hi_r) of the tile window
  // (clamped to [0, TILE]); `nvec` of the window's vectors overlap it.  A tile window starts
  // at or after its segment's aligned start (w0 >= lo rounded down), so only the first
  // tile can start before `lo`.  Uniform 64-bit differences are split into 32-bit halves:
  // gfx9 compares signed 64-bit values only in vector registers.
  struct Rel {
    int lo_r, hi_r;
    uint32_t nvec;
  };
  auto rel = [&](const TileInfo& t) {
    Rel r;
    r.lo_r = clamp_span(t.lo - t.w0, TILE);
    r.hi_r = clamp_span(t.hi - t.w0, TILE);
    r.nvec = static_cast<uint32_t>(r.hi_r + VEC - 1) / VEC;
    return r;
  };
  // K vector loads per thread for every tile, so a wait for the current tile counts the
  // next tile's K loads out.  They are buffer loads over the tile window's vectors that
  // overlap its segment: a vector outside them reads as zero without touching memory (an
  // empty segment's tile reads nothing), and `add` masks the positions outside the
  // segment.
  auto load = [&](const TileInfo& t, const Rel& r, V (&v)[K]) {
    const __amdgpu_buffer_rsrc_t rs = uniform_rsrc(vin + t.w0 / VEC, r.nvec * 16u);
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(
                                       rs, static_cast<int>((k * 256u + threadIdx.x) * 16u), 0, 2));
  };
  // four accumulator chains (vector k adds into chain k % 4), folded in a fixed order
  auto add = [&](const Rel& r, const V (&v)[K]) {
    A acc[4] = {acc_zero<A>(), acc_zero<A>(), acc_zero<A>(), acc_zero<A>()};
    if (r.lo_r == 0 && r.hi_r == TILE) {  // whole tile inside the segment
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < VEC; ++j) acc[k & 3] = acc_add(acc[k & 3], to_acc(v[k].v[j]));
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int p0 = (k * 256 + static_cast<int>(threadIdx.x)) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          if (p0 + j >= r.lo_r && p0 + j < r.hi_r) acc[k & 3] = acc_add(acc[k & 3], to_acc(v[k].v[j]));
      }
    }
    return acc_add(acc_add(acc[0], acc[1]), acc_add(acc[2], acc[3]));
  };

  uint32_t tile = blockIdx.x;
  if (tile >= g.ntiles) return;
  TileInfo ti = tile_info<VEC, TILE>(g, tile, shift);
  Rel ri = rel(ti);
  // One tile: issue the next tile's loads into `nxt`, add up `cur`, write the result.
  // The two register sets alternate, so no copy waits for the next tile's loads.
  auto step = [&](const V (&cur)[K], V (&nxt)[K]) {
    const uint32_t next = tile + gridDim.x;
    const bool more = next < g.ntiles;
    // the last tile loads itself again: every step issues the same loads
    const TileInfo tn = more ? tile_info<VEC, TILE>(g, next, shift) : ti;
    const Rel rn = rel(tn);
    load(tn, rn, nxt);
    // the next tile's loads stay ahead of this tile's adds
    __builtin_amdgcn_sched_barrier(0);
    const A w = wave_sum(add(ri, cur));
    if (lane == 0) s_w[wid] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      const A tot = acc_add(acc_add(s_w[0], s_w[1]), acc_add(s_w[2], s_w[3]));
      if (ti.first_tile == ti.last_tile)
        out[g.seg_out ? ld_table(g.seg_out, ti.seg) : ti.seg] = from_acc<T>(tot);
      else
        partials[tile] = tot;
    }
    __syncthreads();
    ti = tn;
    ri = rn;
    tile = next;
    return more;
  };
  V va[K], vb[K];
  load(ti, ri, va);
  while (step(va, vb) && step(vb, va)) {
  }
}

// Segmented reduce, pass 2: one wave per segment of more than one tile sums its tile
// partials in a fixed order (lane-strided, then the wave tree), so results do not
// depend on scheduling.
template <class T>
__global__ __launch_bounds__(256) void k_seg_combine(SegGeom g,
                                                     const typename Traits<T>::A* __restrict__ partials,
                                                     T* __restrict__ out) {
  using A = typename Traits<T>::A;
  const uint32_t seg = blockIdx.x * 4 + threadIdx.x / kWave;
  if (seg >= g.nseg) return;
  const uint32_t first = g.tile_off[seg];
  const uint32_t last = g.tile_off[seg + 1] - 1;
  if (last == first) return;  // written by pass 1
  // four independent chains, so four partials per lane are in flight at once
  A c0 = acc_zero<A>(), c1 = acc_zero<A>(), c2 = acc_zero<A>(), c3 = acc_zero<A>();
  uint32_t j = first + lane_id();
  for (; j + 3 * kWave <= last; j += 4 * kWave) {
    c0 = acc_add(c0, partials[j]);
    c1 = acc_add(c1, partials[j + kWave]);
    c2 = acc_add(c2, partials[j + 2 * kWave]);
    c3 = acc_add(c3, partials[j + 3 * kWave]);
  }
  for (; j <= last; j += kWave) c0 = acc_add(c0, partials[j]);
  const A w = wave_sum(acc_add(acc_add(c0, c1), acc_add(c2, c3)));
  if (lane_id() == 0) out[g.seg_out ? g.seg_out[seg] : seg] = from_acc<T>(w);
}

// Launch of the (mapped) reduce over n elements of TI at `in` (any alignment).
template <class TI, class TO, bool TUNABLE, class M>
hipError_t reduce_launch(const void* in, uint64_t n, void* out_dev, const M& m, const Scratch& s,
                         hipStream_t stream) {
  using A = typename Traits<TO>::A;
  constexpr int VEC = 16 / sizeof(TI);
  const TI* in_al = reinterpret_cast<const TI*>(reinterpret_cast<uintptr_t>(in) & ~uintptr_t(15));
  const int64_t lo = elem_shift(in, sizeof(TI));
  const int64_t hi = lo + static_cast<int64_t>(n);
  const int64_t nvec = hi / VEC - (lo + VEC - 1) / VEC;
  const Tune& t = tune();
  const int unroll = (TUNABLE && t.red_unroll == 4) ? 4 : 8;
  int64_t grid = (nvec + 256 * unroll - 1) / (256 * unroll);
  if (grid < 1) grid = 1;
  // one block per CU (2^28: count_if even i32 at two per CU 0.190 ms, at one 0.164 ms)
  const int bpc = is_map_pred<M>::value ? 1 : (t.red_bpc > 0 ? t.red_bpc : 8);
  const int64_t cap = static_cast<int64_t>(s.max_blocks / 8) * bpc;
  if (grid > cap) grid = cap;
  if (static_cast<size_t>(grid) * sizeof(A) > s.partial_bytes) return hipErrorInvalidValue;
  const uint32_t gb = static_cast<uint32_t>(grid);
  A* part = reinterpret_cast<A*>(s.partials);
  TO* o = reinterpret_cast<TO*>(out_dev);
  if (!TUNABLE || (unroll == 8 && t.red_nt)) {
    k_reduce<TI, TO, VEC, 8, true><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o, m);
    return hipGetLastError();
  }
  if constexpr (TUNABLE) {
    if (unroll == 4 && t.red_nt)
      k_reduce<TI, TO, VEC, 4, true><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o, m);
    else if (unroll == 4)
      k_reduce<TI, TO, VEC, 4, false><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o, m);
    else
      k_reduce<TI, TO, VEC, 8, false><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o, m);
  }
  return hipGetLastError();
}

template <class T>
hipError_t reduce_t(const void* in, uint64_t n, void* out_dev, const Scratch& s,
                    hipStream_t stream) {
  return reduce_launch<T, T, true>(in, n, out_dev, MapNone<T>{}, s, stream);
}

// Fused map-reduce dispatch: same-type maps, converting identity maps, predicate counts.
template <class TI, class TO>
hipError_t map_reduce_tt(const px_desc& d, const void* in, uint64_t n, void* out_dev,
                         const Scratch& s, hipStream_t stream) {
  if (d.fn == PX_MAP_PRED) {
    if constexpr (std::is_integral_v<TO>) {
      MapPred<TI, TO> m{Pred<TI>(px_desc{d.magic, d.version, d.op, d.dtype, d.aux, d.fn2, d.flags, 0,
                                         d.c0, d.c1})};
      return reduce_launch<TI, TO, false>(in, n, out_dev, m, s, stream);
    }
    return hipErrorInvalidValue;
  }
  const TO a = bits_as<TO>(d.c0), b = bits_as<TO>(d.c1);
  if (d.fn == PX_MAP_IDENTITY)
    return reduce_launch<TI, TO, false>(in, n, out_dev, MapOp<TI, TO, PX_MAP_IDENTITY>{a, b}, s, stream);
  constexpr bool widening = (std::is_same_v<TI, int32_t> && std::is_same_v<TO, int64_t>) ||
                            (std::is_same_v<TI, uint32_t> && std::is_same_v<TO, uint64_t>) ||
                            (std::is_same_v<TI, float> && std::is_same_v<TO, double>);
  if constexpr (std::is_same_v<TI, TO> || widening) {
    switch (d.fn) {
#define PXK_MR(F) \
  case F: return reduce_launch<TI, TO, false>(in, n, out_dev, MapOp<TI, TO, F>{a, b}, s, stream);
      PXK_MR(PX_MAP_FILL)
      PXK_MR(PX_MAP_ADD)
      PXK_MR(PX_MAP_MUL)
      PXK_MR(PX_MAP_AXPB)
      PXK_MR(PX_MAP_NEG)
      PXK_MR(PX_MAP_ABS)
      PXK_MR(PX_MAP_SQUARE)
#undef PXK_MR
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

template <class TI>
hipError_t map_reduce_t(const px_desc& d, const void* in, uint64_t n, void* out_dev,
                        const Scratch& s, hipStream_t stream) {
  switch (d.aux) {
    case PX_I32: return map_reduce_tt<TI, int32_t>(d, in, n, out_dev, s, stream);
    case PX_U32: return map_reduce_tt<TI, uint32_t>(d, in, n, out_dev, s, stream);
    case PX_I64: return map_reduce_tt<TI, int64_t>(d, in, n, out_dev, s, stream);
    case PX_U64: return map_reduce_tt<TI, uint64_t>(d, in, n, out_dev, s, stream);
    case PX_F32: return map_reduce_tt<TI, float>(d, in, n, out_dev, s, stream);
    case PX_F64: return map_reduce_tt<TI, double>(d, in, n, out_dev, s, stream);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t seg_reduce_t(const void* in, const SegGeom& g, void* out, const Scratch& s,
                        hipStream_t stream) {
  using A = typename Traits<T>::A;
  const T* in_al = reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(in) & ~uintptr_t(15));
  const int64_t shift = elem_shift(in, sizeof(T));
  if (static_cast<size_t>(g.ntiles) * sizeof(A) > s.partial_bytes) return hipErrorInvalidValue;
  uint32_t grid = g.ntiles < s.max_blocks ? g.ntiles : s.max_blocks;
  if (grid == 0) return hipSuccess;
  A* part = reinterpret_cast<A*>(s.partials);
  T* o = reinterpret_cast<T*>(out);
  if (g.tile_block != 256) return hipErrorInvalidValue;
  if (g.tile_k == 4)
    k_seg_reduce<T, 4><<<grid, 256, 0, stream>>>(in_al, shift, g, part, o);
  else if (g.tile_k == 8)
    k_seg_reduce<T, 8><<<grid, 256, 0, stream>>>(in_al, shift, g, part, o);
  else if (g.tile_k == 16)
    k_seg_reduce<T, 16><<<grid, 256, 0, stream>>>(in_al, shift, g, part, o);
  else
    return hipErrorInvalidValue;
  if (g.ntiles > g.nseg)  // some segment spans more than one tile
    k_seg_combine<T><<<(g.nseg + 3) / 4, 256, 0, stream>>>(g, part, o);
  return hipGetLastError();
}

}  // namespace

hipError_t reduce(int dtype, const void* in, uint64_t n, void* out_dev, const Scratch& s,
                  hipStream_t stream) {
  PXK_DISPATCH(dtype, return reduce_t<T>(in, n, out_dev, s, stream));
  return hipErrorInvalidValue;
}

bool map_reduce_supported(const px_desc& d) {
  if (d.dtype >= PX_NUM_DTYPES || d.aux >= PX_NUM_DTYPES) return false;
  if (d.fn == PX_MAP_PRED) return d.aux <= PX_U64 && d.fn2 < PX_NUM_PREDS && d.fn2 != PX_PRED_UNIQUE;
  if (d.fn == PX_MAP_IDENTITY) return true;
  const bool widening = (d.dtype == PX_I32 && d.aux == PX_I64) || (d.dtype == PX_U32 && d.aux == PX_U64) ||
                        (d.dtype == PX_F32 && d.aux == PX_F64);
  return (d.dtype == d.aux || widening) && d.fn >= PX_MAP_FILL && d.fn <= PX_MAP_SQUARE;
}

hipError_t map_reduce(const px_desc& d, const void* in, uint64_t n, void* out_dev, const Scratch& s,
                      hipStream_t stream) {
  if (!map_reduce_supported(d)) return hipErrorInvalidValue;
  PXK_DISPATCH(d.dtype, return map_reduce_t<T>(d, in, n, out_dev, s, stream));
  return hipErrorInvalidValue;
}

hipError_t seg_reduce(int dtype, const void* in, const SegGeom& g, void* out, const Scratch& s,
                      hipStream_t stream) {
  PXK_DISPATCH(dtype, return seg_reduce_t<T>(in, g, out, s, stream));
  return hipErrorInvalidValue;
}

}  // namespace pxk
