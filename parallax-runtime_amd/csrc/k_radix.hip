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

constexpr int kBins = 256;
// Predecessor tiles read per look-back round trip (with 24- / 16-key lanes: 2, 3, 4 tiles
// sort 2^26 u32 in 0.841 / 0.852 / 0.842 ms and 2^25 f64 in 1.281 / 1.281 / 1.295 ms).
constexpr int kLookWindow = 2;

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
constexpr bool kSignedKey = std::is_integral_v<T> && std::is_signed_v<T>;

// The key a sort orders by.  Signed integers key on (U)x - base with base = (U)min(x)
// (k_radix_min): order-preserving on [min, max], and small values of either sign then
// share their upper digits, whose passes are skipped (int32 in [-1000, 1000]: 2 of 4
// passes); base = (U)INT_MIN would be the sign-bit flip, at the same one VALU op.  Other
// types key on their order-preserving encoding (`base` unused).
template <class T>
__device__ __forceinline__ typename RadixKey<T>::U key_of(T x, typename RadixKey<T>::U base) {
  using U = typename RadixKey<T>::U;
  if constexpr (kSignedKey<T>) return static_cast<U>(static_cast<U>(x) - base);
  else return RadixKey<T>::enc(x);
}

// The element whose key is all ones (every digit 255): pads positions past n, which
// then sort after every real key of their tile.
template <class T>
__device__ __forceinline__ T radix_pad(typename RadixKey<T>::U base) {
  if constexpr (kSignedKey<T>) {
    return static_cast<T>(base - 1u);
  } else if constexpr (sizeof(T) == 4) {
    const uint32_t u = std::is_same_v<T, uint32_t> ? 0xffffffffu : 0x7fffffffu;
    return __builtin_bit_cast(T, u);
  } else {
    const uint64_t u = std::is_same_v<T, uint64_t> ? ~0ull : 0x7fffffffffffffffull;
    return __builtin_bit_cast(T, u);
  }
}

template <class T>
__device__ __forceinline__ uint32_t digit_of(T x, int shift, typename RadixKey<T>::U base) {
  return static_cast<uint32_t>(key_of(x, base) >> shift) & 0xffu;
}

// The key base the counting kernels read for a signed sort: (U)min(x) from ctl->radix_nmin
// (the complement of the smallest encoded key); 0 for other types.
template <class T>
__device__ __forceinline__ typename RadixKey<T>::U read_key_base(const Ctl* ctl) {
  using U = typename RadixKey<T>::U;
  if constexpr (kSignedKey<T>) {
    constexpr U sign = static_cast<U>(U(1) << (8 * sizeof(U) - 1));
    return static_cast<U>(static_cast<U>(~ld_agent(&ctl->radix_nmin)) ^ sign);
  } else {
    return U(0);
  }
}

// Keys per lane of a whole-array digit pass: 16 for 8-byte keys (1024-thread tiles of 16K
// keys, 128 KiB of LDS; 18 would need 161 KiB), 24 for 4-byte keys (the most without
// spills once the ticket and wave index sit in scalar registers: 26 spills).  Measured:
// 8-byte keys at 8 per lane 1.85 ms, at 16 per lane 1.51 ms (2^25 f64); 4-byte keys at
// 16 / 18 / 20 per lane 1.08 / 1.06 / 1.02 ms (2^26 u32: fewer tiles, fewer
// look-backs), with the LDS peer-mask ranking 20 / 22 per lane 0.886 / 0.869 ms, and
// with scalar ticket and wave index 22 / 24 per lane 0.857 / 0.840 ms (26 spills).  The
// segmented pass: 24 for 4-byte keys (served mixed passes 4.13 -> 4.06 ms at 20 against
// 16; served sort batches 2.81 -> 2.78 ms at 24 against 20), 16 for 8-byte keys.
template <class T>
constexpr int kKpt = sizeof(T) == 4 ? 24 : 16;
template <class T>
constexpr int kKptSeg = sizeof(T) == 4 ? 24 : 16;

__device__ __forceinline__ uint64_t status_word(uint32_t epoch, uint32_t state, uint32_t v) {
  return (static_cast<uint64_t>((epoch << 2) | state) << 32) | v;
}

// ---------------------------------------------------------------- histograms
// Lanes that must share the first lane's digit before hist_add counts a digit position
// by ballots (random 8-bit digits: 1 in 4 lanes on average).
constexpr int kSkewLanes = 8;

// Adds N encoded keys per lane to a wave's LDS histograms hw[D * 256].  Lanes adding to
// one LDS word serialise, so:
//  * a digit position that all N keys of every active lane share is added once, by the
//    first active lane (narrow keys have such digits in most waves); the test costs one
//    OR of differences per key and one ballot per digit position;
//  * a skewed position (the first lane's digit of key 0 shared by kSkewLanes lanes or
//    more: small integers' upper bytes, float exponents) counts its two most common
//    digits (the first lane's, and the first other lane's) by ballots, and only the
//    lanes holding another digit add to LDS;
//  * other positions add every key's digit.
template <int D, int N, class U>
__device__ __forceinline__ void hist_add(uint32_t* hw, const U (&k)[N], int lane) {
  U ref;  // the first active lane's first key (readfirstlane returns int: each half is
          // taken as uint32_t before widening)
  if constexpr (sizeof(U) == 4) {
    ref = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(k[0]));
  } else {
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(k[0] >> 32)));
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(k[0])));
    ref = (static_cast<U>(hi) << 32) | lo;
  }
  U diff = k[0] ^ ref;
#pragma unroll
  for (int j = 1; j < N; ++j) diff |= k[j] ^ ref;
  const uint64_t active = __ballot(1);
  const bool leader = (active & ((1ull << lane) - 1ull)) == 0ull;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const uint32_t r1 = static_cast<uint32_t>(ref >> (8 * d)) & 0xffu;
    if (__ballot(((diff >> (8 * d)) & 0xffu) != 0u) == 0ull) {
      if (leader) atomicAdd(&hw[d * kBins + r1], static_cast<uint32_t>(N * __popcll(active)));
      continue;
    }
    const uint32_t d0 = static_cast<uint32_t>(k[0] >> (8 * d)) & 0xffu;
    const uint64_t m1 = __ballot(d0 == r1);
    if (__popcll(m1) >= kSkewLanes) {
      const uint64_t rest = active & ~m1;
      const uint32_t r2 = rest ? static_cast<uint32_t>(__builtin_amdgcn_readlane(
                                     static_cast<int>(d0), __builtin_ctzll(rest)))
                               : r1;
      uint32_t c1 = 0, c2 = 0;  // wave-uniform counts of r1 and r2
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const uint32_t dj = static_cast<uint32_t>(k[j] >> (8 * d)) & 0xffu;
        const bool e1 = dj == r1;
        const bool e2 = !e1 && dj == r2;
        c1 += static_cast<uint32_t>(__popcll(__ballot(e1)));
        c2 += static_cast<uint32_t>(__popcll(__ballot(e2)));
        if (!e1 && !e2) atomicAdd(&hw[d * kBins + dj], 1u);
      }
      if (leader) {
        atomicAdd(&hw[d * kBins + r1], c1);
        if (c2) atomicAdd(&hw[d * kBins + r2], c2);
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) atomicAdd(&hw[d * kBins + (static_cast<uint32_t>(k[j] >> (8 * d)) & 0xffu)], 1u);
    }
  }
}

// Each wave counts into its own LDS histograms (no contention between waves); the
// block merges them and adds the totals to the control block's counters.
template <class T, int HB>
__global__ __launch_bounds__(HB) void k_radix_hist(const T* data, uint64_t n, Ctl* ctl) {
  constexpr int D = sizeof(T);  // 8-bit digits per key
  constexpr int VEC = 16 / sizeof(T);
  constexpr int NWH = HB / kWave;
  using V = Vec<T, VEC>;
  __shared__ uint32_t h[NWH][D * kBins];
  __shared__ uint32_t s_w[kBins / kWave];
  __shared__ int s_last;

  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  uint32_t* hw = h[wid];
  for (int i = lane; i < D * kBins; i += kWave) hw[i] = 0;  // the wave's own copy
  using U = typename RadixKey<T>::U;
  const U base = read_key_base<T>(ctl);
  auto count = [&](T x) {
    const U k = key_of(x, base);
#pragma unroll
    for (int d = 0; d < D; ++d) atomicAdd(&hw[d * kBins + (static_cast<uint32_t>(k >> (8 * d)) & 0xffu)], 1u);
  };
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * HB + threadIdx.x;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * HB;
  const uint32_t mis = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(data) & 15u) / sizeof(T));
  const uint32_t head_v = mis ? static_cast<uint32_t>(VEC) - mis : 0u;
  const uint64_t head = head_v < static_cast<uint32_t>(n) ? head_v : static_cast<uint32_t>(n);  // n < 2^32
  if (gtid < head) count(data[gtid]);
  const V* vp = reinterpret_cast<const V*>(data + head);
  const uint64_t nv = (n - head) / VEC;
  // UL vector loads per group, one digit test per group; the next group's loads are
  // issued before this group's LDS adds, so the adds overlap the memory latency
  constexpr int UL = 4;
  uint64_t v = gtid;
  V cur[UL], nxt[UL];
  bool have = v + (UL - 1) * gstride < nv;
  if (have) {
#pragma unroll
    for (int u = 0; u < UL; ++u) cur[u] = ld_stream(vp + v + u * gstride);
  }
  while (have) {
    const uint64_t vn = v + UL * gstride;
    const bool have_n = vn + (UL - 1) * gstride < nv;
    if (have_n) {
#pragma unroll
      for (int u = 0; u < UL; ++u) nxt[u] = ld_stream(vp + vn + u * gstride);
    }
    U k[UL * VEC];
#pragma unroll
    for (int u = 0; u < UL; ++u)
#pragma unroll
      for (int j = 0; j < VEC; ++j) k[u * VEC + j] = key_of(cur[u].v[j], base);
    hist_add<D>(hw, k, lane);
#pragma unroll
    for (int u = 0; u < UL; ++u) cur[u] = nxt[u];
    v = vn;
    have = have_n;
  }
  for (; v < nv; v += gstride) {
    const V x = ld_stream(vp + v);
    U k[VEC];
#pragma unroll
    for (int j = 0; j < VEC; ++j) k[j] = key_of(x.v[j], base);
    hist_add<D>(hw, k, lane);
  }
  const uint64_t body_end = head + nv * VEC;
  if (gtid < n - body_end) count(data[body_end + gtid]);
  __syncthreads();
  for (int i = threadIdx.x; i < D * kBins; i += HB) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NWH; ++w) c += h[w][i];
    if (c) __hip_atomic_fetch_add(&ctl->radix_hist[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t d = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();

  // Last block: exclusive bin starts of every digit position (one thread per bin);
  // counters cleared; a position where one bin holds all n keys is marked skipped.
  uint32_t skip = 0;
  for (int d = 0; d < D; ++d) {
    uint32_t c = 0, incl = 0;
    if (threadIdx.x < kBins) {
      c = ld_agent32(&ctl->radix_hist[d * kBins + threadIdx.x]);
      incl = wave_incl_scan(c);
      if (lane == kWave - 1) s_w[wid] = incl;
    }
    __syncthreads();
    if (threadIdx.x < kBins) {
      uint32_t off = incl - c;
      for (int w = 0; w < wid; ++w) off += s_w[w];
      st_agent32(&ctl->radix_bins[d * kBins + threadIdx.x], off);
      st_agent32(&ctl->radix_hist[d * kBins + threadIdx.x], 0u);
    }
    if (__syncthreads_or(threadIdx.x < kBins && c == static_cast<uint32_t>(n))) skip |= 1u << d;
  }
  if (threadIdx.x == 0) {
    st_agent32(&ctl->radix_skip, skip);
    if constexpr (kSignedKey<T>) {  // every block has read the minimum
      st_agent(&ctl->radix_base, static_cast<uint64_t>(base));
      st_agent(&ctl->radix_nmin, 0ull);
    }
    st_agent32(&ctl->done, 0u);
  }
}

// Signed integer sorts: the minimum of the array (its key base).  Every block adds the
// complement of its smallest encoded key to ctl->radix_nmin with an agent-scope max (the
// slot is zero between calls: the histogram kernel that reads it resets it).
template <class T>
__device__ __forceinline__ void radix_min_publish(uint64_t nm, Ctl* ctl) {
  __shared__ uint64_t s_nm[4];
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(static_cast<unsigned long long>(nm), off);
    nm = o > nm ? o : nm;
  }
  if (lane_id() == 0) s_nm[threadIdx.x / kWave] = nm;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) nm = s_nm[w] > nm ? s_nm[w] : nm;
    __hip_atomic_fetch_max(&ctl->radix_nmin, nm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_radix_min(const T* data, uint64_t n, Ctl* ctl) {
  using U = typename RadixKey<T>::U;
  constexpr int VEC = 16 / sizeof(T);
  using V = Vec<T, VEC>;
  uint64_t nm = 0;  // max of ~enc(x)
  auto add = [&](T x) {
    const uint64_t c = static_cast<U>(~RadixKey<T>::enc(x));
    nm = c > nm ? c : nm;
  };
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * 256;
  const uint32_t mis = static_cast<uint32_t>((reinterpret_cast<uintptr_t>(data) & 15u) / sizeof(T));
  const uint64_t head = (mis ? VEC - mis : 0u) < n ? (mis ? VEC - mis : 0u) : n;
  if (gtid < head) add(data[gtid]);
  const V* vp = reinterpret_cast<const V*>(data + head);
  const uint64_t nv = (n - head) / VEC;
  for (uint64_t v = gtid; v < nv; v += gstride) {
    const V x = ld_stream(vp + v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) add(x.v[j]);
  }
  const uint64_t body_end = head + nv * VEC;
  if (gtid < n - body_end) add(data[body_end + gtid]);
  radix_min_publish<T>(nm, ctl);
}

// ---------------------------------------------------------------- segment geometry
// Digit-pass blocks of the segmented sort: 1024 threads (one per CU, 16K-key tiles) or
// 512 (two per CU, 8K-key tiles), chosen per call by seg_radix_tile().  A segment just
// above the radix threshold fills less of a smaller tile; a long segment pays a larger
// tile's look-back less often.  Measured (us): 4,096 segments, lengths log-uniform
// [16, 2^18], narrow / full-range keys: u64 1024 1923 / 4673, 512 1890 / 4627, 256
// 2183 / 5562; u32 1024 1499 / 1917, 512 1403 / 1799, 256 1589 / 2049.  Equal segments,
// u32 / u64: 4 x 2^22 1024 410 / 871, 512 494 / 1051; 16 x 2^20 1024 402 / 856, 512
// 467 / 976; 64 x 2^16 1024 165 / 330, 512 234 / 483.
constexpr int kSegBlockSmall = 512;
constexpr int kSegBlockLarge = 1024;

// Tile t of the long segments covers [tbase, tend) of its segment (positions relative
// to the segment's first key: a segment holds fewer than 2^32 keys, the batch may not).
// `tile` is wave-uniform; the tables are read by scalar loads.
template <class T>
__device__ __forceinline__ void seg_tile(const RadixSegs& g, uint32_t tile, uint32_t& seg,
                                         uint32_t& tbase, uint32_t& tend) {
  const uint32_t TILE = g.tile;
  seg = ld_table(g.tile_seg, tile);
  const uint32_t len = static_cast<uint32_t>(ld_table(g.seg_hi, seg) - ld_table(g.seg_lo, seg));
  tbase = (tile - ld_table(g.seg_first, seg)) * TILE;
  tend = len - tbase < TILE ? len : tbase + TILE;
}

// ---------------------------------------------------------------- one digit pass
// Four waves per SIMD (128 VGPRs): 1024 threads per CU whatever the block size.
//
// SEG = false: one array [0, n) in `a`, `b` a scratch array; bin starts from the control
// block.  A digit that every key shares (marked by the histogram kernel) is skipped; the
// k digits that run alternate a -> b -> a ..; when k is odd, the launch of the first
// skipped digit copies the keys to the other array (an extra step in the alternation),
// so the last step always writes `a`.  SEG = true: the tiles of the long segments of a batch (RadixSegs tables; a
// tile never spans two segments); bin starts from the segment's own digit counts; a
// tile looks back only over its segment's tiles; a digit that every key shares is
// skipped (no block does anything).  The k digits that run move the keys between `a`
// (the array), `b` and g.tmp2 so that the last one writes the array and none reads and
// writes the same buffer: k even a -> b -> a ..; k odd (>= 3) a -> b -> tmp2 -> b ..
// -> a; k = 1: that digit's pass and the first skipped digit's launch, which copies the
// long segments' keys, go a -> b -> a in launch order.  A segment's keys sit at
// a + seg_lo[s] in the array and at seg_pack[s] in `b` and tmp2 (long segments packed
// one after the other), positions inside a segment are 32-bit.
template <class T, int BLOCK, bool SEG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_radix_pass(
    T* __restrict__ a, T* __restrict__ b, uint64_t n, int shift, RadixSegs g,
    uint64_t* status, Ctl* ctl) {
  static_assert(BLOCK >= kBins, "one thread per digit in the count phases");
  constexpr int KPT = SEG ? kKptSeg<T> : kKpt<T>;
  static_assert(BLOCK * KPT <= 65536, "ranks are kept in 16 bits");
  static_assert(KPT % 2 == 0, "ranks are packed two per register");
  constexpr int NW = BLOCK / kWave;
  constexpr int WSPAN = KPT * kWave;
  constexpr int TILE = BLOCK * KPT;
  constexpr int R = kLookWindow;
  constexpr int D = sizeof(T);
  const uint32_t ntiles = SEG ? g.ntiles : static_cast<uint32_t>((n + TILE - 1) / TILE);
  const T* __restrict__ in = a;
  T* __restrict__ out = b;
  bool in_a = true, out_a = false;  // SEG: which side is the array (segment bases differ)
  if constexpr (SEG) {
    const uint32_t skip = ld_agent32(&ctl->radix_skip);
    const int dig = shift / 8;
    const uint32_t runs = ~skip & ((1u << D) - 1u);
    const int k = __builtin_popcount(runs);
    if (k == 1) {
      const int r = __builtin_ctz(runs);
      const int s0 = __builtin_ctz(skip);  // D > 1, so some digit is skipped
      if (dig != r && dig != s0) return;
      const bool first = dig == (r < s0 ? r : s0);
      in = first ? a : b;
      out = first ? b : a;
      in_a = first;
      out_a = !first;
      if (dig == s0) {
        for (uint32_t t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
          uint32_t sg, tb, te;
          seg_tile<T>(g, t, sg, tb, te);
          const T* ti = in + (in_a ? ld_table(g.seg_lo, sg) : ld_table(g.seg_pack, sg));
          T* to = out + (out_a ? ld_table(g.seg_lo, sg) : ld_table(g.seg_pack, sg));
          for (uint32_t p = tb + threadIdx.x; p < te; p += BLOCK) to[p] = ti[p];
        }
        return;
      }
    } else {
      if ((skip >> dig) & 1u) return;  // every key shares this digit
      const int i = __builtin_popcount(runs & ((1u << dig) - 1u));  // this pass's run index
      // buffer 0 = a, 1 = b, 2 = tmp2, picked by selects (an indexed array of the
      // pointers would lose that they are global memory: flat loads and stores)
      auto buf = [&](int j) { return j == 0 ? a : (j == 1 ? b : static_cast<T*>(g.tmp2)); };
      auto dst = [&](int j) { return j == k - 1 ? 0 : ((k & 1) ? 1 + (j & 1) : 1 - (j & 1)); };
      const int bi = i == 0 ? 0 : dst(i - 1), bo = dst(i);
      in = buf(bi);
      out = buf(bo);
      in_a = bi == 0;
      out_a = bo == 0;
    }
  } else {
    const uint32_t skip = ld_agent32(&ctl->radix_skip);
    const int dig = shift / 8;
    const uint32_t runs = ~skip & ((1u << D) - 1u);
    const int k = __builtin_popcount(runs);  // odd k < D (D is even): skip != 0
    const bool copy = (k & 1) && dig == __builtin_ctz(skip | (1u << D));
    if (((skip >> dig) & 1u) && !copy) return;
    const uint32_t below = (1u << dig) - 1u;
    const int i = __builtin_popcount(runs & below) + (((k & 1) && (skip & below)) ? 1 : 0);
    in = (i & 1) ? b : a;
    out = (i & 1) ? a : b;
    if (copy) {
      constexpr int U = 8;  // loads in flight per thread
      const uint64_t stride = static_cast<uint64_t>(gridDim.x) * BLOCK;
      uint64_t x = static_cast<uint64_t>(blockIdx.x) * BLOCK + threadIdx.x;
      for (; x + (U - 1) * stride < n; x += U * stride) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = in[x + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) out[x + u * stride] = v[u];
      }
      for (; x < n; x += stride) out[x] = in[x];
      return;
    }
  }

  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_epoch;
  __shared__ uint32_t s_wh[NW][kBins];  // wave digit counts, then tile-local starts
  __shared__ uint32_t s_goff[kBins];    // output index of tile-local index 0, per digit
  __shared__ uint32_t s_wsum[kBins / kWave];
  __shared__ uint32_t s_bsum[kBins / kWave];  // segment bin-start scan (SEG)
  __shared__ __attribute__((aligned(16))) T s_keys[TILE];

  const int lane = lane_id();
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint32_t* bins = ctl->radix_bins + (shift / 8) * kBins;
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  // SEG: the bits of this digit that vary across the batch's long segments; a full tile
  // (no pad keys, which have every bit set) ranks on those bits only
  const uint32_t vary = SEG ? static_cast<uint32_t>(ld_agent(&ctl->radix_vary) >> shift) & 0xffu : 0xffu;
  using UK = typename RadixKey<T>::U;
  const UK base = kSignedKey<T> ? static_cast<UK>(ld_agent(&ctl->radix_base)) : UK(0);

  if (threadIdx.x == 0) s_epoch = ld_agent32(&ctl->epoch);
  __syncthreads();
  const uint32_t epoch = s_epoch;

  // Persistent loop over tickets; a tile waits only on lower tickets, all of which are
  // held by running blocks, so every wait ends.  Thread 0 takes the next ticket before
  // the tile's write-out, so the wait for it at the top of the loop does not include
  // the write-out's stores (a wait on a return waits for every memory operation issued
  // before it).  The address is taken as a per-lane value, so the fetch stays one lane's
  // atomic (a uniform one is folded into a wave-wide form that waits for its return at
  // once).
  auto take_ticket = [&]() {
    uint32_t zero = 0;
    asm volatile("" : "+v"(zero));
    return __hip_atomic_fetch_add(&ctl->ticket + zero, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  if (threadIdx.x == 0) s_tile = take_ticket();
  __syncthreads();
  while (true) {
    // the previous tile's last reads of s_wh came before the end-of-tile barrier, and a
    // wave's digit counts are its own until the ranking barrier
#pragma unroll
    for (int q = 0; q < kBins / kWave; ++q) s_wh[wid][q * kWave + lane] = 0u;
    // the wave's peer-mask table (256 64-bit lane masks, zero between keys), in the tile
    // buffer, which holds nothing until the reorder
    uint64_t* const mt = reinterpret_cast<uint64_t*>(s_keys) + wid * kBins;
#pragma unroll
    for (int q = 0; q < kBins / kWave; ++q) mt[q * kWave + lane] = 0ull;
    const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
    if (tile >= ntiles) break;
    // Positions in 32-bit arithmetic (n < 2^32; hipcc lowered a 64-bit form of the
    // tile-length select -- a vector compare, then a scalar select on SCC -- to a select
    // on the subtraction's borrow, giving a partial last tile its full length): key j of
    // this lane is real when j < nv.
    // Positions past n hold the pad key: digit 255 in every pass and last in input
    // order, so they take the tile's last local indices [tn, TILE), which are never
    // written out (only the last tile has them, and nothing reads its counts).
    // The tile covers positions [tbase, tend) of the array.
    // SEG: positions relative to the segment, through the segment's base pointers
    uint32_t tbase, tend, first_tile = 0, seg = 0;
    const T* __restrict__ tin = in;
    T* __restrict__ tout = out;
    if constexpr (SEG) {
      seg = ld_table(g.tile_seg, tile);
      first_tile = ld_table(g.seg_first, seg);
      const uint64_t lo64 = ld_table(g.seg_lo, seg), pk = ld_table(g.seg_pack, seg);
      const uint32_t len = static_cast<uint32_t>(ld_table(g.seg_hi, seg) - lo64);
      tbase = (tile - first_tile) * static_cast<uint32_t>(TILE);
      tend = len - tbase < static_cast<uint32_t>(TILE) ? len : tbase + static_cast<uint32_t>(TILE);
      tin = in + (in_a ? lo64 : pk);
      tout = out + (out_a ? lo64 : pk);
    } else {
      const uint32_t n32 = static_cast<uint32_t>(n);
      tbase = tile * static_cast<uint32_t>(TILE);
      tend = n32 - tbase < static_cast<uint32_t>(TILE) ? n32 : tbase + static_cast<uint32_t>(TILE);
    }
    const uint32_t wbase = tbase + static_cast<uint32_t>(wid * WSPAN);
    const uint32_t first = wbase + static_cast<uint32_t>(lane);
    const uint32_t left = first < tend ? (tend - first + kWave - 1) / kWave : 0u;
    const int nv = left < static_cast<uint32_t>(KPT) ? static_cast<int>(left) : KPT;

    // ---- loads (a wave's key j of all lanes is 64 consecutive keys)
    T key[KPT];
    const T* lp = tin + first;
    const bool full = tend - tbase == static_cast<uint32_t>(TILE);

    // ---- rank among equal digits of the wave, in input order (two 16-bit ranks per
    // register: a tile holds fewer than 2^16 keys)
    uint32_t rank2[(KPT + 1) / 2] = {};
    auto rank_keys = [&](auto masked) {
#pragma unroll
      for (int j = 0; j < KPT; ++j) {
        const uint32_t d = digit_of(key[j], shift, base);
        // Lanes whose digit differs from this lane's: some bit's ballot disagrees with
        // this lane's bit (sign-extended to a full word, the same for both halves).
        // Masked: a bit every key of the batch shares (all 64 lanes hold real keys)
        // agrees in every lane and is skipped (a uniform branch).
        uint32_t diff_lo = 0, diff_hi = 0;
        // diff |= m ^ own in one three-input op per half (truth table 0xF6 for
        // S0 | (S1 ^ S2) with S0, S1, S2 = 0xF0, 0xCC, 0xAA); 0x09 = ~(S0 | (S1 ^ S2))
        // gives the complement, the peers, at once.
        auto add_bit = [&](int b, auto table) {
          uint32_t own = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int32_t>(d), b, 1));
          asm volatile("" : "+v"(own));  // ballot from the extracted bit: one compare
          const uint64_t m = __ballot(own != 0u);
          diff_lo = __builtin_amdgcn_bitop3_b32(diff_lo, static_cast<uint32_t>(m), own, decltype(table)::value);
          diff_hi = __builtin_amdgcn_bitop3_b32(diff_hi, static_cast<uint32_t>(m >> 32), own, decltype(table)::value);
        };
        uint64_t peers;
        if constexpr (decltype(masked)::value) {
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if ((vary >> b) & 1u) add_bit(b, std::integral_constant<int, 0xF6>{});
          peers = ~((static_cast<uint64_t>(diff_hi) << 32) | diff_lo);
        } else {
          // Some keys find their peers through LDS instead: every lane ORs its bit into
          // its digit's mask, reads the mask back and clears it (a wave's LDS operations
          // run in order, so the read sees all 64 ORs and the next key a cleared word):
          // three LDS operations and no ballots, so ranking alternates between the VALU
          // and LDS units (every second 4-byte key, every third 8-byte key: 2^26 u32
          // 0.94 -> 0.88 ms, 2^25 f64 1.29 -> 1.28 ms; all keys through LDS is no faster
          // than none)
          if (sizeof(T) == 4 ? (j & 1) != 0 : j % 3 == 2) {
            __hip_atomic_fetch_or(&mt[d], 1ull << lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            peers = __hip_atomic_load(&mt[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&mt[d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
#pragma unroll
            for (int b = 0; b < 7; ++b) add_bit(b, std::integral_constant<int, 0xF6>{});
            add_bit(7, std::integral_constant<int, 0x09>{});
            peers = (static_cast<uint64_t>(diff_hi) << 32) | diff_lo;
          }
        }
        const uint32_t below = static_cast<uint32_t>(__popcll(peers & lanes_below));
        const uint32_t c = s_wh[wid][d];
        rank2[j / 2] |= (c + below) << (16 * (j & 1));
        if (below == 0) s_wh[wid][d] = c + static_cast<uint32_t>(__popcll(peers));
      }
    };
    if (full) {
      // no conditions between the loads and the ranking: key j is ranked as soon as it
      // has arrived, while the later ones are still in flight
#pragma unroll
      for (int j = 0; j < KPT; ++j) key[j] = __builtin_nontemporal_load(lp + j * kWave);
      if (SEG && vary != 0xffu)
        rank_keys(std::true_type{});
      else
        rank_keys(std::false_type{});
    } else {
#pragma unroll
      for (int j = 0; j < KPT; ++j) key[j] = j < nv ? __builtin_nontemporal_load(lp + j * kWave) : radix_pad<T>(base);
      rank_keys(std::false_type{});
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
               status_word(epoch, tile == first_tile ? kStIncl : kStAgg, tot));
      incl = wave_incl_scan(tot);
      if (lane == kWave - 1) s_wsum[wid] = incl;
    }
    uint32_t scount = 0, sincl = 0;  // the segment's count of this digit (SEG)
    if constexpr (SEG) {
      if (threadIdx.x < kBins) {
        scount = g.seg_cnt[(static_cast<uint64_t>(seg) * D + shift / 8) * kBins + threadIdx.x];
        sincl = wave_incl_scan(scount);
        if (lane == kWave - 1) s_bsum[wid] = sincl;
      }
    }
    __syncthreads();

    // ---- tile-local starts of every (wave, digit)
    uint32_t lstart = 0;
    if (threadIdx.x < kBins) {
      const int b = threadIdx.x;
      lstart = incl - tot;
      for (int w = 0; w < wid; ++w) lstart += s_wsum[w];
#pragma unroll
      for (int w = 0; w < NW; ++w) s_wh[w][b] += lstart;
    }
    __syncthreads();

    // ---- reorder the tile in LDS by (digit, rank); it needs no global offsets, so it
    // runs before the look-back and gives the predecessors time to publish
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      s_keys[s_wh[wid][digit_of(key[j], shift, base)] + ((rank2[j / 2] >> (16 * (j & 1))) & 0xffffu)] = key[j];
    }

    // ---- digit counts before this tile (windowed look-back), output offsets
    if (threadIdx.x < kBins) {
      const int b = threadIdx.x;
      uint32_t prefix = 0;
      if (tile > first_tile) {
        const int64_t tstop = first_tile;  // the segment's (or array's) first tile
        int64_t t = static_cast<int64_t>(tile) - 1;
        while (true) {
          uint64_t wv[R];
#pragma unroll
          for (int r = 0; r < R; ++r)
            wv[r] = (t - r >= tstop) ? ld_agent(&status[static_cast<uint64_t>(t - r) * kBins + b]) : 0ull;
          bool done = false;
          int r = 0;
          for (; r < R; ++r) {
            if (t - r < tstop) {
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
      uint32_t bin;
      if constexpr (SEG) {
        bin = sincl - scount;
        for (int w = 0; w < wid; ++w) bin += s_bsum[w];
      } else {
        bin = bins[b];
      }
      s_goff[b] = bin + prefix - lstart;  // modulo 2^32; n < 2^32
    }
    __syncthreads();

    uint32_t next_tile = 0;
    if (threadIdx.x == 0) next_tile = take_ticket();

    // ---- write out: consecutive tile-local indices of one digit are consecutive outputs.
    // Every thread issues exactly KPT stores, so the wait for the ticket below counts
    // them out: in a partial tile, indices past its end store the tile's last key to its
    // place once more (the same value to the same address).
    const uint32_t tlast = tend - tbase - 1u;
#pragma unroll
    for (int i0 = 0; i0 < KPT; ++i0) {
      const uint32_t i = min(static_cast<uint32_t>(i0) * BLOCK + threadIdx.x, tlast);
      const T k = s_keys[i];
      tout[static_cast<uint32_t>(s_goff[digit_of(k, shift, base)] + i)] = k;
    }
    // every thread has read this tile's s_tile (before the ranking barrier)
    if (threadIdx.x == 0) s_tile = next_tile;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t d = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (d == gridDim.x - 1) {  // last block: reset for the next launch
      st_agent32(&ctl->ticket, 0u);
      st_agent32(&ctl->done, 0u);
      st_agent32(&ctl->epoch, next_epoch(epoch));
    }
  }
}

// Threads per digit-pass block (256, 512 or 1024); blocks per CU = 1024 / block.
// Smaller arrays take smaller blocks: more tiles, so more CUs share each pass (measured
// for 2^k u32 keys, 1024 / 512 / 256 threads: 2^17 77 / 68 / 65 us, 2^19 86 / 80 / 86,
// 2^20 98 / 98 / 113); 8-byte keys take the radix path only from 2^20, where 1024 is
// best.
int radix_block(size_t elem, uint64_t n) {
  const int b = tune().radix_block;
  if (b == 256 || b == 512 || b == 1024) return b;
  if (elem == 4 && n < (1ull << 19)) return 256;
  if (elem == 4 && n < (1ull << 20)) return 512;
  return 1024;
}

uint64_t radix_tile(size_t elem, uint64_t n) {
  return static_cast<uint64_t>(radix_block(elem, n)) * (elem == 4 ? kKpt<uint32_t> : kKpt<uint64_t>);
}

template <class T, int BLOCK>
hipError_t radix_passes_t(T* data, T* tmp, uint64_t n, uint32_t grid, const Scratch& s,
                          hipStream_t stream, const LaunchHook& hook) {
  // Every launch gets (array, scratch); each chooses its direction on the device.
  for (int d = 0; d < static_cast<int>(sizeof(T)); ++d) {
    hook("sort_radix", 0);
    k_radix_pass<T, BLOCK, false><<<grid, BLOCK, 0, stream>>>(data, tmp, n, 8 * d, RadixSegs{}, s.status, s.ctl);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hook("sort_radix", 1);
  }
  return hipSuccess;
}

template <class T>
hipError_t radix_sort_t(void* data_v, uint64_t n, const Scratch& s, hipStream_t stream,
                        const LaunchHook& hook) {
  T* data = static_cast<T*>(data_v);
  T* tmp = static_cast<T*>(s.partials);
  const int block = radix_block(sizeof(T), n);
  const uint64_t TILE = radix_tile(sizeof(T), n);
  const uint64_t ntiles = (n + TILE - 1) / TILE;
  if (n >= (1ull << 32) || ntiles * kBins > s.status_words || s.partial_bytes < n * sizeof(T) || !tmp)
    return hipErrorInvalidValue;
  const uint32_t cus = s.max_blocks / 8 ? s.max_blocks / 8 : 1;
  // Histograms: no more blocks than 16 vectors per thread need.  One 512-thread block per CU (measured, 2^26 u32 / 2^25 f64: 121 / 128 us; 256
  // threads x 1 or 2 per CU 132 / 137 and 154 / 161 us, 1024 x 1 145 / 151, 512 x 2
  // 193 / 207): every block adds its D * 256 counts to the control block with
  // device-scope atomics at the end, so more blocks cost more than they hide.
  const int HB = (tune().radix_hist_block == 256 || tune().radix_hist_block == 1024) ? tune().radix_hist_block : 512;
  const uint64_t hist_per_cu = tune().radix_hist_bpc > 0 ? static_cast<uint64_t>(tune().radix_hist_bpc) : 1;
  const uint64_t hist_work = (n + HB * 16ull * (16 / sizeof(T)) - 1) / (HB * 16ull * (16 / sizeof(T)));
  const uint32_t hist_grid = static_cast<uint32_t>(std::min<uint64_t>(hist_per_cu * cus, hist_work ? hist_work : 1));
  hook("sort_hist", 0);
  if constexpr (kSignedKey<T>) {
    k_radix_min<T><<<std::min<uint32_t>(cus * 4, static_cast<uint32_t>((n + 4095) / 4096)), 256, 0, stream>>>(data, n, s.ctl);
    const hipError_t e0 = hipGetLastError();
    if (e0 != hipSuccess) return e0;
  }
  if (HB == 1024)
    k_radix_hist<T, 1024><<<hist_grid, 1024, 0, stream>>>(data, n, s.ctl);
  else if (HB == 512)
    k_radix_hist<T, 512><<<hist_grid, 512, 0, stream>>>(data, n, s.ctl);
  else
    k_radix_hist<T, 256><<<hist_grid, 256, 0, stream>>>(data, n, s.ctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hook("sort_hist", 1);
  // 1024 threads per CU (1024 / block blocks fit one CU's registers and LDS).  Tiles
  // are taken by ticket when a block runs, so a look-back only ever waits on tiles held
  // by running blocks, whatever the grid.
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(ntiles, static_cast<uint64_t>(cus) * (1024 / block)));
  if (block == 256) return radix_passes_t<T, 256>(data, tmp, n, grid, s, stream, hook);
  if (block == 1024) return radix_passes_t<T, 1024>(data, tmp, n, grid, s, stream, hook);
  return radix_passes_t<T, 512>(data, tmp, n, grid, s, stream, hook);
}

// ---------------------------------------------------------------- segmented
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * 256)
    p[i] = 0u;
}

// Digit counts per long segment, and which digits every key shares.  Each block takes a
// contiguous range of tiles and keeps counting in LDS (a copy per wave) while
// consecutive tiles belong to the same segment; it adds the counts to the segment's
// totals when the segment changes and at the end (agent-scope adds, since a segment's
// tiles may span blocks).  Keys are read as 16-byte vectors (a tile's unaligned head
// and tail one key at a time); a digit position that a whole vector of every active
// lane shares is added once by one lane.  The same pass ORs the encoded keys and their
// complements; the last block marks the digits every key shares (their passes are
// skipped) and clears the ORs for the next call.
template <class T>
__global__ __launch_bounds__(256) void k_segradix_hist(const T* data_all, RadixSegs g, Ctl* ctl) {
  constexpr int D = sizeof(T);
  constexpr int HB = 256;
  constexpr int NWH = HB / kWave;
  constexpr int VEC = 16 / sizeof(T);
  constexpr int UL = 4;  // vector loads in flight per thread
  using U = typename RadixKey<T>::U;
  using V = Vec<T, VEC>;
  __shared__ uint32_t h[NWH][D * kBins];
  __shared__ unsigned long long s_or, s_nor;
  __shared__ int s_last;

  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  uint32_t* hw = h[wid];
  if (threadIdx.x == 0) {
    s_or = 0ull;
    s_nor = 0ull;
  }
  for (int i = lane; i < D * kBins; i += kWave) hw[i] = 0u;  // the wave's own copy
  __syncthreads();
  U acc_or = 0, acc_nor = 0;
  const U base = read_key_base<T>(ctl);
  auto count1 = [&](T x) {
    const U k = key_of(x, base);
    acc_or |= k;
    acc_nor |= static_cast<U>(~k);
#pragma unroll
    for (int d = 0; d < D; ++d) atomicAdd(&hw[d * kBins + (static_cast<uint32_t>(k >> (8 * d)) & 0xffu)], 1u);
  };
  auto count_keys = [&](auto& k) {  // encoded keys of this lane
#pragma unroll
    for (auto& x : k) {
      acc_or |= x;
      acc_nor |= static_cast<U>(~x);
    }
    hist_add<D>(hw, k, lane);
  };
  auto flush = [&](uint32_t seg) {
    __syncthreads();
    uint32_t* cnt = g.seg_cnt + static_cast<uint64_t>(seg) * D * kBins;
    for (int i = threadIdx.x; i < D * kBins; i += HB) {
      uint32_t c = 0;
#pragma unroll
      for (int w = 0; w < NWH; ++w) c += h[w][i];
      if (c) __hip_atomic_fetch_add(&cnt[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    for (int i = lane; i < D * kBins; i += kWave) hw[i] = 0u;
  };
  // gridDim.x <= g.ntiles: every block has at least one tile
  const uint32_t t0 = static_cast<uint32_t>(static_cast<uint64_t>(blockIdx.x) * g.ntiles / gridDim.x);
  const uint32_t t1 = static_cast<uint32_t>(static_cast<uint64_t>(blockIdx.x + 1) * g.ntiles / gridDim.x);
  uint32_t cur = 0xffffffffu;
  for (uint32_t tile = t0; tile < t1; ++tile) {
    uint32_t seg, tbase, tend;
    seg_tile<T>(g, tile, seg, tbase, tend);
    if (seg != cur) {
      if (cur != 0xffffffffu) flush(cur);
      cur = seg;
    }
    const T* data = data_all + ld_table(g.seg_lo, seg);  // positions are relative to the segment
    const uint32_t mis = static_cast<uint32_t>(((reinterpret_cast<uintptr_t>(data + tbase)) & 15u) / sizeof(T));
    uint32_t head = mis ? static_cast<uint32_t>(VEC) - mis : 0u;
    if (head > tend - tbase) head = tend - tbase;
    const uint32_t nvec = (tend - tbase - head) / VEC;
    const uint32_t body_end = tbase + head + nvec * VEC;
    if (threadIdx.x < head) count1(data[tbase + threadIdx.x]);
    if (threadIdx.x < tend - body_end) count1(data[body_end + threadIdx.x]);
    const V* vp = reinterpret_cast<const V*>(data + tbase + head);
    for (uint32_t v0 = threadIdx.x; v0 < nvec; v0 += HB * UL) {
      V x[UL];
#pragma unroll
      for (int u = 0; u < UL; ++u) {
        const uint32_t v = v0 + u * HB;
        if (v < nvec) x[u] = ld_stream(vp + v);
      }
      if (v0 + (UL - 1) * HB < nvec) {  // a whole group: one digit test for UL vectors
        U k[UL * VEC];
#pragma unroll
        for (int u = 0; u < UL; ++u)
#pragma unroll
          for (int j = 0; j < VEC; ++j) k[u * VEC + j] = key_of(x[u].v[j], base);
        count_keys(k);
      } else {
#pragma unroll
        for (int u = 0; u < UL; ++u) {
          if (v0 + u * HB < nvec) {
            U k[VEC];
#pragma unroll
            for (int j = 0; j < VEC; ++j) k[j] = key_of(x[u].v[j], base);
            count_keys(k);
          }
        }
      }
    }
  }
  if (cur != 0xffffffffu) flush(cur);

  atomicOr(&s_or, static_cast<unsigned long long>(acc_or));
  atomicOr(&s_nor, static_cast<unsigned long long>(acc_nor));
  __syncthreads();
  if (threadIdx.x == 0) {
    // 16 slots: the blocks ending together do not all update one word
    __hip_atomic_fetch_or(&ctl->radix_or[blockIdx.x & 15u], static_cast<uint64_t>(s_or), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(&ctl->radix_nor[blockIdx.x & 15u], static_cast<uint64_t>(s_nor), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const uint32_t d = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  __threadfence();
  uint64_t all_or = 0, all_nor = 0;
  for (int q = 0; q < 16; ++q) {
    all_or |= ld_agent(&ctl->radix_or[q]);
    all_nor |= ld_agent(&ctl->radix_nor[q]);
    st_agent(&ctl->radix_or[q], 0ull);
    st_agent(&ctl->radix_nor[q], 0ull);
  }
  const uint64_t varying = all_or & all_nor;  // set in some key, clear in another
  uint32_t skip = 0;
  for (int d = 0; d < D; ++d)
    if (((varying >> (8 * d)) & 0xffu) == 0u) skip |= 1u << d;
  st_agent32(&ctl->radix_skip, skip);
  st_agent(&ctl->radix_vary, varying);
  if constexpr (kSignedKey<T>) {  // every block has read the minimum
    st_agent(&ctl->radix_base, static_cast<uint64_t>(base));
    st_agent(&ctl->radix_nmin, 0ull);
  }
  st_agent32(&ctl->done, 0u);
}

// Signed integer keys of a batched sort: the minimum over every long segment (one key
// base for the batch), as k_radix_min.
template <class T>
__global__ __launch_bounds__(256) void k_segradix_min(const T* data_all, RadixSegs g, Ctl* ctl) {
  using U = typename RadixKey<T>::U;
  constexpr int VEC = 16 / sizeof(T);
  using V = Vec<T, VEC>;
  uint64_t nm = 0;
  auto add = [&](T x) {
    const uint64_t c = static_cast<U>(~RadixKey<T>::enc(x));
    nm = c > nm ? c : nm;
  };
  for (uint32_t t = blockIdx.x; t < g.ntiles; t += gridDim.x) {
    uint32_t sg, tb, te;
    seg_tile<T>(g, t, sg, tb, te);
    const T* data = data_all + ld_table(g.seg_lo, sg);
    // 16-byte loads for the tile's aligned body, single keys for its head and tail
    const uint32_t mis = static_cast<uint32_t>(((reinterpret_cast<uintptr_t>(data + tb)) & 15u) / sizeof(T));
    uint32_t head = mis ? static_cast<uint32_t>(VEC) - mis : 0u;
    if (head > te - tb) head = te - tb;
    const uint32_t nvec = (te - tb - head) / VEC;
    const uint32_t body_end = tb + head + nvec * VEC;
    if (threadIdx.x < head) add(data[tb + threadIdx.x]);
    if (threadIdx.x < te - body_end) add(data[body_end + threadIdx.x]);
    const V* vp = reinterpret_cast<const V*>(data + tb + head);
    for (uint32_t v = threadIdx.x; v < nvec; v += 256) {
      const V x = ld_stream(vp + v);
#pragma unroll
      for (int j = 0; j < VEC; ++j) add(x.v[j]);
    }
  }
  radix_min_publish<T>(nm, ctl);
}

template <class T>
hipError_t seg_radix_sort_t(void* data_v, const RadixSegs& g, const Scratch& s,
                            hipStream_t stream, hipStream_t pass_stream, hipEvent_t prep_done) {
  if (g.ntiles == 0) return hipSuccess;
  T* data = static_cast<T*>(data_v);
  T* tmp = static_cast<T*>(g.tmp_b);
  if (g.tile != static_cast<uint32_t>(kSegBlockLarge * kKptSeg<T>) &&
      g.tile != static_cast<uint32_t>(kSegBlockSmall * kKptSeg<T>))
    return hipErrorInvalidValue;
  if (static_cast<uint64_t>(g.ntiles) * kBins > s.status_words || !tmp || !g.tmp2 || !g.seg_cnt ||
      !g.seg_pack)
    return hipErrorInvalidValue;
  const uint32_t cus = s.max_blocks / 8 ? s.max_blocks / 8 : 1;
  const uint64_t ncnt = static_cast<uint64_t>(g.nseg) * sizeof(T) * kBins;
  k_fill_u32<<<static_cast<uint32_t>(std::min<uint64_t>((ncnt + 255) / 256, 1024)), 256, 0, stream>>>(g.seg_cnt, ncnt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // Counting blocks per CU: each adds its counts to a segment's totals with device-scope
  // atomics whenever its tiles change segment, so few long segments count best on 2 (a
  // served sort group of ~30 segments: 3.40 vs 3.48 ms for the group) and many on 4
  // (4,096 segments of 16..2^18 keys: 1.83 vs 1.88 ms).
  const uint64_t hist_per_cu = tune().seg_hist_bpc > 0 ? static_cast<uint64_t>(tune().seg_hist_bpc)
                                                       : (g.nseg <= 128 ? 2 : 4);
  const uint32_t hgrid = static_cast<uint32_t>(std::min<uint64_t>(g.ntiles, hist_per_cu * cus));
  if constexpr (kSignedKey<T>) {
    k_segradix_min<T><<<static_cast<uint32_t>(std::min<uint64_t>(g.ntiles, 4ull * cus)), 256, 0, stream>>>(data, g, s.ctl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  k_segradix_hist<T><<<hgrid, 256, 0, stream>>>(data, g, s.ctl);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (pass_stream && prep_done) {
    if ((e = hipEventRecord(prep_done, stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(pass_stream, prep_done, 0)) != hipSuccess) return e;
    stream = pass_stream;
  }
  const bool large = g.tile == static_cast<uint32_t>(kSegBlockLarge * kKptSeg<T>);
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(g.ntiles, large ? cus : 2 * cus));
  for (int d = 0; d < static_cast<int>(sizeof(T)); ++d) {
    if (large)
      k_radix_pass<T, kSegBlockLarge, true><<<grid, kSegBlockLarge, 0, stream>>>(data, tmp, 0, 8 * d, g, s.status, s.ctl);
    else
      k_radix_pass<T, kSegBlockSmall, true><<<grid, kSegBlockSmall, 0, stream>>>(data, tmp, 0, 8 * d, g, s.status, s.ctl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

// Radix below these sizes loses to the bitonic path: each digit pass is at least one
// tile's serial load, rank and write on one CU (about 10 us).  Measured (2^k keys,
// bitonic / radix with the block chosen by radix_block): u32 2^16 61 / 62 us, 2^17
// 72 / 65, 2^18 85 / 70, 2^19 103 / 80; f64 2^18 112 / 134, 2^19 144 / 157, 2^20
// 203 / 202.
uint64_t radix_min(int dtype) {
  const uint64_t t = tune().sort_radix_min;
  if (t != ~0ull) return t;
  return dtype_size(dtype) == 4 ? (1ull << 17) : (1ull << 20);
}

bool radix_sort_ok(int dtype, uint64_t n) {
  const uint64_t lo = radix_min(dtype);
  return dtype_size(dtype) > 0 && lo > 0 && n >= lo && n < (1ull << 32) - (1ull << 20);
}

uint64_t radix_status_words(int dtype, uint64_t n) {
  const uint64_t tile = radix_tile(dtype_size(dtype), n);
  return ((n + tile - 1) / tile) * kBins;
}

int radix_passes(int dtype) { return static_cast<int>(dtype_size(dtype)); }

uint64_t seg_radix_min_len(int dtype) {
  if (tune().seg_radix_min) return tune().seg_radix_min;
  // Segments longer than this take the segmented radix sort, even though such a segment
  // fills only part of its tile: the bitonic tile sort of mid-size segments is slower.
  // Measured on 4,096 segments, lengths log-uniform [16, 2^18] (us, narrow / full-range
  // keys): u64 with the threshold at 8192 2051 / 4820, 2048 1956 / 4757, 512 1913 / 4674,
  // 256 1913 / 4647, 64 1924 / 4707; u32 at 16384 1542 / 1958, 1024 1516 / 1920, 512
  // 1512 / 1927.
  return dtype_size(dtype) == 4 ? 1024 : 512;
}

uint32_t seg_radix_tile(int dtype, const uint64_t* lo, const uint64_t* hi, uint32_t n) {
  const uint64_t kpt = dtype_size(dtype) == 4 ? kKptSeg<uint32_t> : kKptSeg<uint64_t>;
  const uint64_t small = kSegBlockSmall * kpt, large = kSegBlockLarge * kpt;
  if (tune().seg_radix_block) return static_cast<uint32_t>(tune().seg_radix_block * kpt);
  uint64_t ps = 0, pl = 0;  // tile slots at each size
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t len = hi[i] - lo[i];
    ps += (len + small - 1) / small * small;
    pl += (len + large - 1) / large * large;
  }
  return static_cast<uint32_t>(pl * 20 <= ps * 21 ? large : small);
}

uint64_t seg_radix_cnt_bytes(int dtype, uint32_t nseg) {
  return static_cast<uint64_t>(nseg) * dtype_size(dtype) * kBins * sizeof(uint32_t);
}

hipError_t seg_radix_sort(int dtype, void* data, const RadixSegs& seg, const Scratch& s,
                          hipStream_t stream, hipStream_t pass_stream, hipEvent_t prep_done) {
  PXK_DISPATCH(dtype, return seg_radix_sort_t<T>(data, seg, s, stream, pass_stream, prep_done));
  return hipErrorInvalidValue;
}

hipError_t radix_sort(int dtype, void* data, uint64_t n, const Scratch& s, hipStream_t stream,
                      const LaunchHook& hook) {
  PXK_DISPATCH(dtype, return radix_sort_t<T>(data, n, s, stream, hook));
  return hipErrorInvalidValue;
}

}  // namespace pxk
