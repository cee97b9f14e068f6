// k_scan.hip -- single-pass inclusive / exclusive prefix scan for gfx950.
//
// Replaces the reference's scan.comp + scan_add.comp pair and its recursive launch
// schedule (KernelLauncher::launch_scan, src/kernel/kernel_launcher.cpp:1280-1342: per-
// block scan, recursive scan of block sums, add pass; about 4*n*s bytes in 7 host-
// synchronised dispatches at 2^28) and the exclusive shift pass (launch_exclusive_scan,
// :1437-1475) with ONE launch that reads every element once and writes it once (2*n*s).
//
// Tile = BLOCK threads x K chunks x one 16-byte vector per lane.  Wave w owns K
// consecutive 64-vector chunks; per chunk each lane sums its vector, the wave scans the
// 64 sums with DPP and carries the running total to the next chunk; the wave totals meet
// in LDS.  Tiles take tickets in order; each publishes its aggregate, wave 0 looks back
// (decoupled look-back: one predecessor status word per lane and poll, stopping at the
// nearest inclusive prefix), publishes its inclusive prefix, and the tile writes its
// outputs with non-temporal 16-byte stores.  Tiles of a batched (segmented) call look
// back only inside their own segment.
//
// Arithmetic: everything inside a wave's chunks runs in the local type L; the wave
// totals are combined into per-wave carries by one 16-lane scan in the carry type A, and
// the look-back runs in A.  For double (L = double, A = double-double) an output is
// hi + (lo + local), where local = the element's prefix inside its wave; it passes
// through at most 1 (vector pair) + 6 (wave scan) + KR - 1 (chunk carry) + 1 (vector
// element) + 2 (output) roundings, each bounded by eps times the sum of |x| over the
// prefix, and the carries add about eps^2 -- within 2 ceil(log2 n) eps sum|x| (n = the
// segment's length; a short segment fills fewer lanes and chunks, and sums of zeros are
// exact).
#include <cstring>

#include "dev_common.hpp"
#include "dev_tiles.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

template <class T, int BLOCK, int K, bool EXCL>
__global__ __launch_bounds__(BLOCK) void k_scan(const T* in_al, int64_t in_shift, T* out,
                                              int64_t out_shift, SegGeom g,
                                              typename Traits<T>::A init, uint64_t* status,
                                              Ctl* ctl) {
  using A = typename Traits<T>::A;
  using L = typename Traits<T>::L;
  constexpr int VEC = 16 / sizeof(T);
  // A ticket covers S sub-tiles of KR register chunks each; all but the last sub-tile
  // wait in LDS (lane-private slots) while the last is scanned in registers, so one
  // look-back serves S * KR chunks without raising register pressure.
  constexpr int KR = K > 8 ? 8 : K;
  constexpr int S = K / KR;
  static_assert(KR * S == K, "tile chunks must be a multiple of the register chunks");
  constexpr int SUB = BLOCK * KR * VEC;
  constexpr int TILE = BLOCK * K * VEC;
  constexpr int NW = BLOCK / kWave;
  static_assert(NW <= 16, "the wave totals are scanned in one 16-lane DPP row");
  using V = Vec<T, VEC>;
  const V* vin = reinterpret_cast<const V*>(in_al);
  // the 16-byte-aligned address at or below `out`, by pointer arithmetic (through an
  // integer the compiler loses that it is global memory and stores through flat
  // instructions, which the LDS wait counter also counts)
  V* vout = reinterpret_cast<V*>(out - out_shift);

  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_epoch;
  __shared__ L s_w[S][NW];
  __shared__ A s_prefix;
  __shared__ V s_stash[S > 1 ? (S - 1) * KR * BLOCK : 1];

  const int lane = lane_id();
  // wave index and tickets are read as wave-uniform values (scalar registers and
  // branches): a branch taken for divergent runs both arms under lane masks, and its
  // waits then cover both arms' loads
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);

  // Masked load of sub-tile `sub` (zeros outside the tile's segment window).
  auto load = [&](T (&dst)[KR][VEC], const TileInfo& t, int sub) {
    const int64_t b0 = t.w0 + static_cast<int64_t>(sub) * SUB;
    if (b0 >= t.lo && b0 + SUB <= t.hi) {  // whole sub-tile inside: no masks
      const V* tp = vin + b0 / VEC + wid * (KR * kWave) + lane;
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const V v = ld_stream(tp + k * kWave);
#pragma unroll
        for (int j = 0; j < VEC; ++j) dst[k][j] = v.v[j];
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int64_t p0 = b0 + ((static_cast<int64_t>(wid) * KR + k) * kWave + lane) * VEC;
      if (p0 >= t.lo && p0 + VEC <= t.hi) {
        const V v = vin[p0 / VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) dst[k][j] = v.v[j];
      } else if (p0 < t.hi && p0 + VEC > t.lo) {
        const V v = vin[p0 / VEC];
#pragma unroll
        for (int j = 0; j < VEC; ++j) dst[k][j] = (p0 + j >= t.lo && p0 + j < t.hi) ? v.v[j] : T(0);
      } else {
#pragma unroll
        for (int j = 0; j < VEC; ++j) dst[k][j] = T(0);
      }
    }
    // a partial sub-tile's loads complete here (vmcnt 0), so the whole sub-tiles' loads
    // into the same registers need no wait of their own (the arms join before them)
    __builtin_amdgcn_s_waitcnt(0x0F70);
  };

  // Per-chunk wave scans carried across the wave's chunks (local type); returns the
  // wave total, tex[k] = this lane's exclusive prefix inside the wave.
  auto local_scan = [&](const T (&x)[KR][VEC], L (&tex)[KR]) {
    L run = acc_zero<L>();
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      L c = to_loc(x[k][0]);
#pragma unroll
      for (int j = 1; j < VEC; ++j) c = acc_add(c, to_loc(x[k][j]));
      const L incl = wave_incl_scan(c);
      const L tot = read_lane(incl, kWave - 1);
      tex[k] = acc_add(run, wave_shift_right1(incl));
      run = acc_add(run, tot);
    }
    return run;
  };

  // Outputs of one sub-tile (16-byte stores when the output lines up with the input
  // vectors); `base` is the carry into this wave's first chunk.
  auto emit = [&](const T (&x)[KR][VEC], const L (&tex)[KR], A base, const TileInfo& ti, int sub) {
    const int64_t odelta = ti.out_delta - in_shift;  // output element = position + odelta
    const bool vec_ok = ((out_shift + odelta) & (VEC - 1)) == 0;
    const int64_t b0 = ti.w0 + static_cast<int64_t>(sub) * SUB;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      T xv[VEC];
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        xv[j] = x[k][j];
        // Widening carries: convert again here rather than keep the widened values of
        // the whole tile alive from the local scan (twice the registers).
        if constexpr (sizeof(L) > sizeof(T)) asm volatile("" : "+v"(xv[j]));
      }
      V y;
      if constexpr (sizeof(A) == sizeof(L)) {
        A r = acc_add(base, tex[k]);
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          if (EXCL) {
            y.v[j] = from_acc<T>(r);
            r = acc_add(r, to_acc(xv[j]));
          } else {
            r = acc_add(r, to_acc(xv[j]));
            y.v[j] = from_acc<T>(r);
          }
        }
      } else {
        L r = tex[k];
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          if (EXCL) {
            y.v[j] = out_of<T>(base, r);
            r = acc_add(r, to_loc(xv[j]));
          } else {
            r = acc_add(r, to_loc(xv[j]));
            y.v[j] = out_of<T>(base, r);
          }
        }
      }
      const int64_t p0 = b0 + ((static_cast<int64_t>(wid) * KR + k) * kWave + lane) * VEC;
      if (vec_ok && p0 >= ti.lo && p0 + VEC <= ti.hi) {
        st_stream(&vout[(p0 + odelta + out_shift) / VEC], y);
      } else if (p0 < ti.hi && p0 + VEC > ti.lo) {
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          if (p0 + j >= ti.lo && p0 + j < ti.hi) out[p0 + j + odelta] = y.v[j];
      }
    }
  };

  // Two tiles in registers (this one and the next) fit 128 VGPRs when the chunk carries
  // are 32-bit or the tile has 4 chunks; wider carries over 8 chunks keep one tile.
  constexpr bool kPrefetch = S == 1 && 8 * KR + KR * static_cast<int>(sizeof(L)) / 4 <= 72;
  if constexpr (kPrefetch) {
    // One sub-tile per ticket: the block takes its next ticket and issues that tile's
    // loads before the look-back of the current one, so the look-back's round trips
    // overlap the next tile's memory latency.  A block holds at most two tickets and
    // finishes the lower first; the lowest unfinished ticket waits only on finished
    // ones, so every wait still ends.  Wave totals are double-buffered by tile parity:
    // a wave may publish the next tile's total while another still reads this one's.
    __shared__ uint32_t s_next;
    __shared__ L s_w2[2][NW];
    __shared__ A s_off[NW];  // carry into each wave's first chunk
    if (threadIdx.x == 0) {
      s_epoch = ld_agent32(&ctl->epoch);
      s_tile = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t epoch = s_epoch;
    uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
    if (tile < g.ntiles) {
      TileInfo ti = tile_info<VEC, TILE>(g, tile, in_shift);
      T x[KR][VEC];
      load(x, ti, 0);
      for (uint32_t it = 0;; ++it) {
        const int par = static_cast<int>(it & 1u);
        // the wave total now; the chunk carries (the same scans) again after the
        // look-back, so they hold no registers while the next tile's loads are in flight
        L tex[KR];
        const L run = local_scan(x, tex);
        if (lane == 0) s_w2[par][wid] = run;
        if (threadIdx.x == 0)
          s_next = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const uint32_t next = __builtin_amdgcn_readfirstlane(s_next);
        const bool more = next < g.ntiles;
        TileInfo tn = ti;
        T xn[KR][VEC];
        // wave 0 looks back: it issues its own loads of the next tile once the look-back
        // is done (a poll's wait for its load would also wait for every load before it)
        if (more) {
          tn = tile_info<VEC, TILE>(g, next, in_shift);
          if (wid != 0) load(xn, tn, 0);
        }
        if (wid == 0) {
          // wave totals scanned across lanes 0..NW-1 in the carry type
          const A wt = lane < NW ? loc_to_acc<A>(s_w2[par][lane]) : acc_zero<A>();
          const A wincl = row_incl_scan(wt);
          const A tile_tot = read_lane(wincl, NW - 1);
          A prefix;
          if (tile == ti.first_tile) {
            prefix = EXCL ? init : acc_zero<A>();
          } else {
            if (lane == 0) publish<A>(status, tile, epoch, kStAgg, tile_tot);
            prefix = lookback<A>(status, tile, ti.first_tile, epoch);
          }
          if (tile != ti.last_tile && lane == 0)
            publish<A>(status, tile, epoch, kStIncl, acc_add(prefix, tile_tot));
          const A wexcl = dpp_mov<kRowShr1, 0xf, 0xf>(wincl);
          if (lane < NW) s_off[lane] = acc_add(prefix, wexcl);
          if (more) load(xn, tn, 0);
        }
        __syncthreads();
        const A w = s_off[wid];
        local_scan(x, tex);
        emit(x, tex, w, ti, 0);
        if (!more) break;
#pragma unroll
        for (int k = 0; k < KR; ++k)
#pragma unroll
          for (int j = 0; j < VEC; ++j) x[k][j] = xn[k][j];
        tile = next;
        ti = tn;
      }
    }
    if (threadIdx.x == 0) block_end(ctl, epoch);
    return;
  }

  if (threadIdx.x == 0) s_epoch = ld_agent32(&ctl->epoch);

  // Persistent loop over tickets; a tile waits only on lower tickets, all of which are
  // held by running blocks, so every wait ends.
  while (true) {
    if (threadIdx.x == 0)
      s_tile = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
    const uint32_t epoch = s_epoch;
    if (tile >= g.ntiles) break;
    const TileInfo ti = tile_info<VEC, TILE>(g, tile, in_shift);
    T x[KR][VEC];

    // ---- leading sub-tiles: wave totals only, values parked in LDS
#pragma unroll
    for (int sub = 0; sub < S - 1; ++sub) {
      load(x, ti, sub);
      L c = acc_zero<L>();
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        V v;
#pragma unroll
        for (int j = 0; j < VEC; ++j) {
          v.v[j] = x[k][j];
          c = acc_add(c, to_loc(x[k][j]));
        }
        s_stash[(sub * KR + k) * BLOCK + threadIdx.x] = v;
      }
      const L tot = read_lane(wave_incl_scan(c), kWave - 1);
      if (lane == 0) s_w[sub][wid] = tot;
    }

    // ---- last sub-tile scanned in registers
    load(x, ti, S - 1);
    L tex[KR];
    const L run = local_scan(x, tex);
    if (lane == 0) s_w[S - 1][wid] = run;
    __syncthreads();

    if (wid == 0) {
      A tile_tot = loc_to_acc<A>(s_w[0][0]);
#pragma unroll
      for (int sub = 0; sub < S; ++sub)
#pragma unroll
        for (int j = (sub == 0 ? 1 : 0); j < NW; ++j) tile_tot = acc_add_loc(tile_tot, s_w[sub][j]);
      A prefix;
      if (tile == ti.first_tile) {
        prefix = EXCL ? init : acc_zero<A>();
      } else {
        if (lane == 0) publish<A>(status, tile, epoch, kStAgg, tile_tot);
        prefix = lookback<A>(status, tile, ti.first_tile, epoch);
      }
      if (tile != ti.last_tile && lane == 0)
        publish<A>(status, tile, epoch, kStIncl, acc_add(prefix, tile_tot));
      if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();

    // ---- carries into each sub-tile and wave, in a fixed order
    A sub_base[S];
    A wave_off[S];
    {
      A b = s_prefix;
#pragma unroll
      for (int sub = 0; sub < S; ++sub) {
        sub_base[sub] = b;
        A w = acc_zero<A>();
        for (int j = 0; j < wid; ++j) w = acc_add_loc(w, s_w[sub][j]);
        wave_off[sub] = w;
        if (sub + 1 < S)
          for (int j = 0; j < NW; ++j) b = acc_add_loc(b, s_w[sub][j]);
      }
    }
    emit(x, tex, acc_add(sub_base[S - 1], wave_off[S - 1]), ti, S - 1);

    // ---- parked sub-tiles: rescan from LDS (each lane reads back its own slots)
#pragma unroll
    for (int sub = 0; sub < S - 1; ++sub) {
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const V v = s_stash[(sub * KR + k) * BLOCK + threadIdx.x];
#pragma unroll
        for (int j = 0; j < VEC; ++j) x[k][j] = v.v[j];
      }
      local_scan(x, tex);
      emit(x, tex, acc_add(sub_base[sub], wave_off[sub]), ti, sub);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) block_end(ctl, s_epoch);
}

template <class T>
hipError_t scan_t(const void* in, void* out, const SegGeom& g, bool exclusive, const void* init,
                  const Scratch& s, hipStream_t stream) {
  using A = typename Traits<T>::A;
  const T* in_al = reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(in) & ~uintptr_t(15));
  const int64_t in_shift = elem_shift(in, sizeof(T));
  const int64_t out_shift = elem_shift(out, sizeof(T));
  if (static_cast<size_t>(g.ntiles) * (sizeof(A) / 4) > s.status_words) return hipErrorInvalidValue;
  A iv{};
  if (exclusive && init) {
    T t;
    std::memcpy(&t, init, sizeof(T));
    if constexpr (sizeof(A) == 16) {
      double h = static_cast<double>(t);
      std::memcpy(&iv, &h, sizeof(double));  // dd{t, 0}
    } else if constexpr (Traits<T>::is_float) {
      iv = static_cast<A>(t);
    } else {
      iv = static_cast<A>(t);
    }
  }
  const TileCfg c{g.tile_block, g.tile_k};
  const uint32_t grid = lookback_grid(g.ntiles, s.max_blocks, c.block);
  if (grid == 0) return hipSuccess;
  T* o = reinterpret_cast<T*>(out);
#define PXK_SCAN(B, KK)                                                                            \
  if (c.block == B && c.k == KK) {                                                                 \
    if (exclusive)                                                                                 \
      k_scan<T, B, KK, true><<<grid, B, 0, stream>>>(in_al, in_shift, o, out_shift, g, iv, s.status, s.ctl); \
    else                                                                                           \
      k_scan<T, B, KK, false><<<grid, B, 0, stream>>>(in_al, in_shift, o, out_shift, g, iv, s.status, s.ctl); \
    return hipGetLastError();                                                                      \
  }
  PXK_SCAN(256, 4)
  PXK_SCAN(256, 8)
  PXK_SCAN(512, 8)
  PXK_SCAN(1024, 4)
  PXK_SCAN(1024, 8)
  PXK_SCAN(1024, 16)
#undef PXK_SCAN
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t scan(int dtype, const void* in, void* out, const SegGeom& g, bool exclusive,
                const void* init, const Scratch& s, hipStream_t stream) {
  PXK_DISPATCH(dtype, return scan_t<T>(in, out, g, exclusive, init, s, stream));
  return hipErrorInvalidValue;
}

namespace {
// Inclusive scan of short segments: one wave per segment, 64 elements per step (a wave
// scan in the accumulator type plus the running carry), segments taken grid-stride.  A
// look-back tile would spend a block's fixed work on a handful of elements.
template <class T>
__global__ __launch_bounds__(256) void k_seg_scan_short(const T* __restrict__ in, T* __restrict__ out,
                                                        const uint64_t* lo, const uint64_t* hi,
                                                        const uint64_t* out_lo, uint32_t nshort) {
  using A = typename Traits<T>::A;
  constexpr uint32_t WPB = 256 / kWave;
  const int lane = lane_id();
  const uint32_t nwaves = gridDim.x * WPB;
  for (uint32_t sg = blockIdx.x * WPB + threadIdx.x / kWave; sg < nshort; sg += nwaves) {  // wave-uniform
    const uint64_t b = lo[sg], e = hi[sg];
    const uint64_t ob = out_lo ? out_lo[sg] : b;
    A carry = acc_zero<A>();
    for (uint64_t p = b; p < e; p += kWave) {
      const uint64_t i = p + lane;
      const A x = i < e ? to_acc(in[i]) : acc_zero<A>();
      const A incl = acc_add(carry, wave_incl_scan(x));
      if (i < e) out[ob + (i - b)] = from_acc<T>(incl);
      carry = read_lane(incl, kWave - 1);
    }
  }
}
}  // namespace

hipError_t seg_scan_short(int dtype, const void* in, void* out, const uint64_t* lo, const uint64_t* hi,
                          const uint64_t* out_lo, uint32_t nshort, const Scratch& s, hipStream_t stream) {
  if (nshort == 0) return hipSuccess;
  const uint32_t want = (nshort + 3) / 4;  // four waves (segments) per block
  const uint32_t grid = want < s.max_blocks ? want : s.max_blocks;
  PXK_DISPATCH(dtype, {
    k_seg_scan_short<T><<<grid, 256, 0, stream>>>(static_cast<const T*>(in), static_cast<T*>(out), lo, hi,
                                                  out_lo, nshort);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

}  // namespace pxk
