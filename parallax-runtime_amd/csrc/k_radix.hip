// k_radix.hip -- LSD radix sort of large arrays for gfx950.
//
// Replaces the reference's bitonic network for arrays above the radix threshold
// (KernelLauncher::launch_sort, src/kernel/kernel_launcher.cpp:1568-1606, and
// dispatch_sort_stage :1481-1566 over shaders/bitonic.comp:15-30: one host-synchronised
// dispatch per (k, j) stage, 351 full read + write passes at 2^26).  The result is the
// same ascending order by operator<: the input permuted into the order of an
// order-preserving integer key.
//
//   k_radix_hist  one read of the array: 256-bin histograms of every 8-bit digit of the
//                 keys (LDS counters, then agent-scope adds); the last block turns them
//                 into exclusive bin starts and clears the counters for the next call.
//   k_radix_pass  one launch per digit, least significant first (4 for 4-byte keys, 8
//                 for 8-byte keys), between the array and a scratch copy (an even
//                 number of passes, so the result lands in the array).  Tiles of 1024
//                 threads x KPT keys in ticket order.  Lanes own keys element-strided
//                 (key j of a lane sits at wave base + 64 j + lane), so ranking in
//                 (j, lane) order is input order and the pass is stable.  A key's rank
//                 among equal digits of its wave comes from eight ballots (the lanes
//                 whose digit matches) and per-wave LDS counters.  The tile's 256 digit
//                 counts are published together; each digit's count before the tile
//                 comes from a decoupled look-back that reads four predecessor tiles per
//                 round trip.  Keys are then reordered in LDS by (digit, rank), so each
//                 digit's keys leave the tile as one contiguous run.
//
// Keys: unsigned as is; signed with the sign bit flipped; IEEE floats with the sign bit
// flipped when positive and every bit flipped when negative (-0.0 before +0.0; NaNs by
// sign, after +inf or before -inf).  Counts are 32-bit, so n < 2^32.
// Bytes moved: n*s (histogram) + 2*n*s per digit.
#include <algorithm>

#include "dev_common.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

constexpr int kRadixBlock = 1024;
constexpr int kBins = 256;
constexpr int kHistBlock = 256;
constexpr int kLookWindow = 4;  // predecessor tiles read per look-back round trip

template <class T>
struct RadixKey;
template <>
struct RadixKey<uint32_t> {
  using U = uint32_t;
  __device__ static U enc(uint32_t x) { return x; }
};
template <>
struct RadixKey<int32_t> {
  using U = uint32_t;
  __device__ static U enc(int32_t x) { return static_cast<uint32_t>(x) ^ 0x80000000u; }
};
template <>
struct RadixKey<float> {
  using U = uint32_t;
  __device__ static U enc(float x) {
    const uint32_t u = __float_as_uint(x);
    return u ^ (static_cast<uint32_t>(static_cast<int32_t>(u) >> 31) | 0x80000000u);
  }
};
template <>
struct RadixKey<uint64_t> {
  using U = uint64_t;
  __device__ static U enc(uint64_t x) { return x; }
};
template <>
struct RadixKey<int64_t> {
  using U = uint64_t;
  __device__ static U enc(int64_t x) { return static_cast<uint64_t>(x) ^ 0x8000000000000000ull; }
};
template <>
struct RadixKey<double> {
  using U = uint64_t;
  __device__ static U enc(double x) {
    const uint64_t u = static_cast<uint64_t>(__double_as_longlong(x));
    return u ^ (static_cast<uint64_t>(static_cast<int64_t>(u) >> 63) | 0x8000000000000000ull);
  }
};

template <class T>
__device__ __forceinline__ uint32_t digit_of(T x, int shift) {
  return static_cast<uint32_t>(RadixKey<T>::enc(x) >> shift) & 0xffu;
}

template <class T>
constexpr int kKpt = sizeof(T) == 4 ? 16 : 8;  // keys per lane
template <class T>
constexpr int kRadixTile = kRadixBlock * kKpt<T>;

__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint32_t state, uint32_t v) {
  return (static_cast<uint64_t>((epoch << 2) | state) << 32) | v;
}

// ---------------------------------------------------------------- histograms
template <class T>
__global__ __launch_bounds__(kHistBlock) void k_radix_hist(const T* data, uint64_t n, Ctl* ctl) {
  constexpr int D = sizeof(T);  // 8-bit digits per key
  constexpr int VEC = 16 / sizeof(T);
  using V = Vec<T, VEC>;
  __shared__ uint32_t h[D * kBins];
  __shared__ uint32_t s_w[kHistBlock / kWave];
  __shared__ int s_last;

  for (int i = threadIdx.x; i < D * kBins; i += kHistBlock) h[i] = 0;
  __syncthreads();
  auto count = [&](T x) {
    const auto k = RadixKey<T>::enc(x);
#pragma unroll
    for (int d = 0; d < D; ++d) atomicAdd(&h[d * kBins + (static_cast<uint32_t>(k >> (8 * d)) & 0xffu)], 1u);
  };
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kHistBlock + threadIdx.x;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kHistBlock;
  const uint32_t mis = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(data) & 15u) / sizeof(T));
  const uint32_t head_v = mis ? static_cast<uint32_t>(VEC) - mis : 0u;
  const uint64_t head = head_v < static_cast<uint32_t>(n) ? head_v : static_cast<uint32_t>(n);  // n < 2^32
  if (gtid < head) count(data[gtid]);
  const V* vp = reinterpret_cast<const V*>(data + head);
  const uint64_t nv = (n - head) / VEC;
  for (uint64_t v = gtid; v < nv; v += gstride) {
    const V x = ld_stream(vp + v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) count(x.v[j]);
  }
  const uint64_t body_end = head + nv * VEC;
  if (gtid < n - body_end) count(data[body_end + gtid]);
  __syncthreads();
  for (int i = threadIdx.x; i < D * kBins; i += kHistBlock)
    if (h[i]) __hip_atomic_fetch_add(&ctl->radix_hist[i], h[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t d = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();

  // Last block: exclusive bin starts of every digit position; counters cleared.
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  for (int d = 0; d < D; ++d) {
    const uint32_t c = ld_agent32(&ctl->radix_hist[d * kBins + threadIdx.x]);
    const uint32_t incl = wave_incl_scan(c);
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    uint32_t off = incl - c;
    for (int w = 0; w < wid; ++w) off += s_w[w];
    st_agent32(&ctl->radix_bins[d * kBins + threadIdx.x], off);
    st_agent32(&ctl->radix_hist[d * kBins + threadIdx.x], 0u);
    __syncthreads();
  }
  if (threadIdx.x == 0) st_agent32(&ctl->done, 0u);
}

// ---------------------------------------------------------------- one digit pass
template <class T>
__global__ __launch_bounds__(kRadixBlock) void k_radix_pass(const T* __restrict__ in,
                                                          T* __restrict__ out, uint64_t n,
                                                          int shift, uint64_t* status, Ctl* ctl) {
  constexpr int KPT = kKpt<T>;
  constexpr int NW = kRadixBlock / kWave;
  constexpr int WSPAN = KPT * kWave;
  constexpr int TILE = kRadixTile<T>;
  constexpr int R = kLookWindow;
  const uint32_t ntiles = static_cast<uint32_t>((n + TILE - 1) / TILE);

  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_epoch;
  __shared__ uint32_t s_wh[NW][kBins];  // wave digit counts, then tile-local starts
  __shared__ uint32_t s_goff[kBins];    // output index of tile-local index 0, per digit
  __shared__ uint32_t s_wsum[kBins / kWave];
  __shared__ T s_keys[TILE];

  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const uint32_t* bins = ctl->radix_bins + (shift / 8) * kBins;
  const uint64_t lanes_below = (1ull << lane) - 1ull;

  if (threadIdx.x == 0) s_epoch = ld_agent32(&ctl->epoch);

  // Persistent loop over tickets; a tile waits only on lower tickets, all of which are
  // held by running blocks, so every wait ends.
  while (true) {
    if (threadIdx.x == 0)
      s_tile = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int b = lane; b < kBins; b += kWave) s_wh[wid][b] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t epoch = s_epoch;
    if (tile >= ntiles) break;
    const uint64_t tbase = static_cast<uint64_t>(tile) * TILE;
    const uint64_t wbase = tbase + static_cast<uint64_t>(wid) * WSPAN;

    // ---- loads (a wave's key j of all lanes is 64 consecutive keys)
    T key[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const uint64_t p = wbase + static_cast<uint64_t>(j) * kWave + lane;
      key[j] = p < n ? __builtin_nontemporal_load(in + p) : T(0);
    }

    // ---- rank among equal digits of the wave, in input order
    uint32_t rank[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const bool valid = wbase + static_cast<uint64_t>(j) * kWave + lane < n;
      const uint32_t d = digit_of(key[j], shift);
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t below = static_cast<uint32_t>(__popcll(peers & lanes_below));
      const uint32_t c = s_wh[wid][d];
      rank[j] = c + below;
      if (valid && below == 0) s_wh[wid][d] = c + static_cast<uint32_t>(__popcll(peers));
    }
    __syncthreads();

    // ---- per digit: wave offsets, tile count (published), block-wide exclusive scan
    uint32_t tot = 0, incl = 0;
    if (threadIdx.x < kBins) {
      const int b = threadIdx.x;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const uint32_t c = s_wh[w][b];
        s_wh[w][b] = tot;
        tot += c;
      }
      st_agent(&status[static_cast<uint64_t>(tile) * kBins + b],
               status_word(epoch, tile == 0 ? kStIncl : kStAgg, tot));
      incl = wave_incl_scan(tot);
      if (lane == kWave - 1) s_wsum[wid] = incl;
    }
    __syncthreads();

    // ---- digit counts before this tile (windowed look-back), output offsets
    if (threadIdx.x < kBins) {
      const int b = threadIdx.x;
      uint32_t lstart = incl - tot;
      for (int w = 0; w < wid; ++w) lstart += s_wsum[w];
      uint32_t prefix = 0;
      if (tile > 0) {
        int64_t t = static_cast<int64_t>(tile) - 1;
        while (true) {
          uint64_t wv[R];
#pragma unroll
          for (int r = 0; r < R; ++r)
            wv[r] = (t - r >= 0) ? ld_agent(&status[static_cast<uint64_t>(t - r) * kBins + b]) : 0ull;
          bool done = false;
          int r = 0;
          for (; r < R; ++r) {
            if (t - r < 0) {
              done = true;
              break;
            }
            const uint32_t f = static_cast<uint32_t>(wv[r] >> 32);
            if ((f >> 2) != epoch || (f & 3u) == 0u) break;  // not published yet: poll again
            prefix += static_cast<uint32_t>(wv[r]);
            if ((f & 3u) == kStIncl) {
              done = true;
              break;
            }
          }
          if (done) break;
          if (r == 0) __builtin_amdgcn_s_sleep(1);
          t -= r;
        }
        st_agent(&status[static_cast<uint64_t>(tile) * kBins + b],
                 status_word(epoch, kStIncl, prefix + tot));
      }
      s_goff[b] = bins[b] + prefix - lstart;  // modulo 2^32; n < 2^32
#pragma unroll
      for (int w = 0; w < NW; ++w) s_wh[w][b] += lstart;
    }
    __syncthreads();

    // ---- reorder the tile in LDS by (digit, rank)
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const bool valid = wbase + static_cast<uint64_t>(j) * kWave + lane < n;
      if (valid) s_keys[s_wh[wid][digit_of(key[j], shift)] + rank[j]] = key[j];
    }
    __syncthreads();

    // ---- write out: consecutive tile-local indices of one digit are consecutive outputs
    // Keys of this tile, in 32-bit arithmetic (n < 2^32).  hipcc lowered the 64-bit form
    // of this select (a vector compare, then a scalar select on SCC) to a select on the
    // borrow of the subtraction, which gave the full tile length for a partial last tile.
    const uint32_t left = static_cast<uint32_t>(n) - tile * static_cast<uint32_t>(TILE);
    const uint32_t tn = left < static_cast<uint32_t>(TILE) ? left : static_cast<uint32_t>(TILE);
    if (tn == TILE) {
#pragma unroll
      for (int i0 = 0; i0 < KPT; ++i0) {
        const uint32_t i = static_cast<uint32_t>(i0) * kRadixBlock + threadIdx.x;
        const T k = s_keys[i];
        out[static_cast<uint32_t>(s_goff[digit_of(k, shift)] + i)] = k;
      }
    } else {
      for (uint32_t i = threadIdx.x; i < tn; i += kRadixBlock) {
        const T k = s_keys[i];
        out[static_cast<uint32_t>(s_goff[digit_of(k, shift)] + i)] = k;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) block_end(ctl, s_epoch);
}

template <class T>
hipError_t radix_sort_t(void* data_v, uint64_t n, const Scratch& s, hipStream_t stream,
                        const LaunchHook& hook) {
  T* data = static_cast<T*>(data_v);
  T* tmp = static_cast<T*>(s.partials);
  constexpr uint64_t TILE = kRadixTile<T>;
  const uint64_t ntiles = (n + TILE - 1) / TILE;
  if (n >= (1ull << 32) || ntiles * kBins > s.status_words || s.partial_bytes < n * sizeof(T) || !tmp)
    return hipErrorInvalidValue;
  const uint32_t cus = s.max_blocks / 8 ? s.max_blocks / 8 : 1;
  const uint64_t hist_work = (n + kHistBlock * 64 - 1) / (kHistBlock * 64);
  const uint32_t hist_grid = static_cast<uint32_t>(std::min<uint64_t>(2ull * cus, hist_work ? hist_work : 1));
  hook("sort_hist", 0);
  k_radix_hist<T><<<hist_grid, kHistBlock, 0, stream>>>(data, n, s.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hook("sort_hist", 1);
  // One block per CU: every block of the grid is resident (a look-back waits only on
  // running blocks).
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(ntiles, cus));
  T* a = data;
  T* b = tmp;
  for (int d = 0; d < static_cast<int>(sizeof(T)); ++d) {
    hook("sort_radix", 0);
    k_radix_pass<T><<<grid, kRadixBlock, 0, stream>>>(a, b, n, 8 * d, s.status, s.ctl);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hook("sort_radix", 1);
    std::swap(a, b);
  }
  return hipSuccess;
}

}  // namespace

uint64_t radix_min() {
  return tune().sort_radix_min;
}

bool radix_sort_ok(int dtype, uint64_t n) {
  const uint64_t lo = radix_min();
  return dtype_size(dtype) > 0 && lo > 0 && n >= lo && n < (1ull << 32) - (1ull << 20);
}

uint64_t radix_status_words(int dtype, uint64_t n) {
  const uint64_t tile = dtype_size(dtype) == 4 ? kRadixTile<uint32_t> : kRadixTile<uint64_t>;
  return ((n + tile - 1) / tile) * kBins;
}

int radix_passes(int dtype) { return static_cast<int>(dtype_size(dtype)); }

hipError_t radix_sort(int dtype, void* data, uint64_t n, const Scratch& s, hipStream_t stream,
                      const LaunchHook& hook) {
  PXK_DISPATCH(dtype, return radix_sort_t<T>(data, n, s, stream, hook));
  return hipErrorInvalidValue;
}

}  // namespace pxk
