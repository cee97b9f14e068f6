// k_reduce.hip -- sum reduction for gfx950.
//
// Plain form (reference: KernelLauncher::launch_reduce, src/kernel/kernel_launcher.cpp:908-985,
// which runs reduce.comp level after level with a host wait between levels): here ONE
// launch.  A grid sized to fill every CU streams the input with 16-byte loads, four
// in flight per thread, accumulating in registers in the carry type; each wave folds
// with DPP, waves fold through LDS, every block leaves one partial, and the last block
// to finish (agent-scope counter) folds the partials in a fixed order and writes the
// result.  The counter resets itself, so back-to-back calls need no reset launch.
//
// Segmented form: a persistent grid walks the tiles of all segments (static
// assignment: no tile waits on another); single-tile segments write their sum
// directly, larger ones leave per-tile partials and the last tile of the segment (a
// self-resetting per-segment counter) folds them in tile order.  Segment s's sum goes
// to out[seg_out[s]] (out[s] without an output table).
#include "dev_common.hpp"
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

template <class T, int VEC, int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_reduce(const T* __restrict__ in_al, int64_t lo,
                                                int64_t hi,
                                                typename Traits<T>::A* __restrict__ partials,
                                                Ctl* ctl, T* out) {
  using A = typename Traits<T>::A;
  using V = Vec<T, VEC>;
  const V* vin = reinterpret_cast<const V*>(in_al);
  __shared__ A s_w[4];
  __shared__ int s_last;

  const int64_t a = (lo + VEC - 1) / VEC;  // interior vectors [a, b)
  const int64_t b = hi / VEC;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  int64_t i = a + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  A acc = acc_zero<A>();
  A acc2 = acc_zero<A>();  // two chains for more add-level parallelism
  for (; i + (UNROLL - 1) * stride < b; i += UNROLL * stride) {
    V v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) v[u] = load16<V, NT>(vin + i + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        if (u & 1) acc2 = acc_add(acc2, to_acc(v[u].v[j]));
        else acc = acc_add(acc, to_acc(v[u].v[j]));
      }
    }
  }
  for (; i < b; i += stride) {
    const V v0 = load16<V, NT>(vin + i);
#pragma unroll
    for (int j = 0; j < VEC; ++j) acc = acc_add(acc, to_acc(v0.v[j]));
  }
  acc = acc_add(acc, acc2);
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
    if (ok) acc = acc_add(acc, to_acc(in_al[pos]));
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
    *out = from_acc<T>(tot);
  }
}

// Segmented reduce, pass 1: tile partials.  Tiles are independent (static grid-stride
// assignment, no atomics); a segment held by one tile writes its result directly, the
// others leave one partial per tile for pass 2.
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
  const int wid = threadIdx.x / kWave;

  for (uint32_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
    const TileInfo ti = tile_info<VEC, TILE>(g, tile, shift);
    A acc = acc_zero<A>();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t p0 = ti.w0 + (static_cast<int64_t>(k) * 256 + threadIdx.x) * VEC;
      if (p0 < ti.hi && p0 + VEC > ti.lo) {
        const V v = vin[p0 / VEC];
        if (p0 >= ti.lo && p0 + VEC <= ti.hi) {
#pragma unroll
          for (int j = 0; j < VEC; ++j) acc = acc_add(acc, to_acc(v.v[j]));
        } else {
#pragma unroll
          for (int j = 0; j < VEC; ++j)
            if (p0 + j >= ti.lo && p0 + j < ti.hi) acc = acc_add(acc, to_acc(v.v[j]));
        }
      }
    }
    const A w = wave_sum(acc);
    if (lane == 0) s_w[wid] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      const A tot = acc_add(acc_add(s_w[0], s_w[1]), acc_add(s_w[2], s_w[3]));
      if (ti.first_tile == ti.last_tile)
        out[g.seg_out ? g.seg_out[ti.seg] : ti.seg] = from_acc<T>(tot);
      else
        partials[tile] = tot;
    }
    __syncthreads();
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
  A c = acc_zero<A>();
  for (uint32_t j = first + lane_id(); j <= last; j += kWave) c = acc_add(c, partials[j]);
  const A w = wave_sum(c);
  if (lane_id() == 0) out[g.seg_out ? g.seg_out[seg] : seg] = from_acc<T>(w);
}

template <class T>
hipError_t reduce_t(const void* in, uint64_t n, void* out_dev, const Scratch& s,
                    hipStream_t stream) {
  using A = typename Traits<T>::A;
  constexpr int VEC = 16 / sizeof(T);
  const T* in_al = reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(in) & ~uintptr_t(15));
  const int64_t lo = elem_shift(in, sizeof(T));
  const int64_t hi = lo + static_cast<int64_t>(n);
  const int64_t nvec = hi / VEC - (lo + VEC - 1) / VEC;
  const Tune& t = tune();
  const int unroll = t.red_unroll == 4 ? 4 : 8;
  int64_t grid = (nvec + 256 * unroll - 1) / (256 * unroll);
  if (grid < 1) grid = 1;
  const int64_t cap = static_cast<int64_t>(s.max_blocks / 8) * (t.red_bpc > 0 ? t.red_bpc : 8);
  if (grid > cap) grid = cap;
  if (static_cast<size_t>(grid) * sizeof(A) > s.partial_bytes) return hipErrorInvalidValue;
  const uint32_t gb = static_cast<uint32_t>(grid);
  A* part = reinterpret_cast<A*>(s.partials);
  T* o = reinterpret_cast<T*>(out_dev);
  if (unroll == 4 && t.red_nt)
    k_reduce<T, VEC, 4, true><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o);
  else if (unroll == 4)
    k_reduce<T, VEC, 4, false><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o);
  else if (t.red_nt)
    k_reduce<T, VEC, 8, true><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o);
  else
    k_reduce<T, VEC, 8, false><<<gb, 256, 0, stream>>>(in_al, lo, hi, part, s.ctl, o);
  return hipGetLastError();
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

hipError_t seg_reduce(int dtype, const void* in, const SegGeom& g, void* out, const Scratch& s,
                      hipStream_t stream) {
  PXK_DISPATCH(dtype, return seg_reduce_t<T>(in, g, out, s, stream));
  return hipErrorInvalidValue;
}

}  // namespace pxk
