// dev_common.hpp -- device-side building blocks shared by every gfx950 primitive kernel.
//
//  * element traits: per element type T, the carry type A the kernels accumulate in
//      int32/uint32 -> uint32 (wrapping), int64/uint64 -> uint64 (wrapping),
//      float -> double, double -> double-double (hi, lo),
//    and the local type L a scan uses inside one wave's chunks (A, except double for
//    double: only the carries between waves and tiles need the double-double).
//    Integer sums are exact modulo 2^w in any order; the float carries keep every
//    reduce / scan result within a few units of the last place of the exact sum.
//  * wave64 cross-lane steps on DPP (row_shr, row_bcast15/31, quad_perm, row mirrors,
//    wave_shr) -- no LDS inside a wave; LDS is used only between the waves of a block.
//  * agent-scope status words for single-pass decoupled look-back: each 32-bit word of
//    a carry travels in a 64-bit atomic word {flag:32 | payload:32}, so a reader never
//    needs a fence and never sees a torn value (all words must carry the same flag).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace pxk {

constexpr int kWave = 64;

// ---------------------------------------------------------------- double-double
struct dd {
  double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  // TwoSum of the high parts, then fold both low parts and renormalise.
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  const double h = s + e;
  const double l = e - (h - s);
  return dd{h, l};
}

// a + b for a double b: TwoSum of a.hi and b, a.lo folded in, renormalised.
__device__ __forceinline__ dd dd_add_d(dd a, double b) {
  const double s = a.hi + b;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b - bb);
  e += a.lo;
  const double h = s + e;
  return dd{h, e - (h - s)};
}

// ---------------------------------------------------------------- carry traits
template <class T>
struct Traits;
template <>
struct Traits<int32_t> {
  using A = uint32_t;
  using L = uint32_t;
  static constexpr bool is_float = false;
};
template <>
struct Traits<uint32_t> {
  using A = uint32_t;
  using L = uint32_t;
  static constexpr bool is_float = false;
};
template <>
struct Traits<int64_t> {
  using A = uint64_t;
  using L = uint64_t;
  static constexpr bool is_float = false;
};
template <>
struct Traits<uint64_t> {
  using A = uint64_t;
  using L = uint64_t;
  static constexpr bool is_float = false;
};
template <>
struct Traits<float> {
  using A = double;
  using L = double;
  static constexpr bool is_float = true;
};
template <>
struct Traits<double> {
  using A = dd;
  using L = double;
  static constexpr bool is_float = true;
};

__device__ __forceinline__ uint32_t acc_add(uint32_t a, uint32_t b) { return a + b; }
__device__ __forceinline__ uint64_t acc_add(uint64_t a, uint64_t b) { return a + b; }
__device__ __forceinline__ double acc_add(double a, double b) { return a + b; }
__device__ __forceinline__ dd acc_add(dd a, dd b) { return dd_add(a, b); }

template <class A>
__device__ __forceinline__ A acc_zero() {
  return A{};
}
template <>
__device__ __forceinline__ dd acc_zero<dd>() {
  return dd{0.0, 0.0};
}

__device__ __forceinline__ uint32_t to_acc(int32_t x) { return static_cast<uint32_t>(x); }
__device__ __forceinline__ uint32_t to_acc(uint32_t x) { return x; }
__device__ __forceinline__ uint64_t to_acc(int64_t x) { return static_cast<uint64_t>(x); }
__device__ __forceinline__ uint64_t to_acc(uint64_t x) { return x; }
__device__ __forceinline__ double to_acc(float x) { return static_cast<double>(x); }
__device__ __forceinline__ dd to_acc(double x) { return dd{x, 0.0}; }

// Local (in-wave) scan values: the carry type, except double for double.
template <class T>
__device__ __forceinline__ auto to_loc(T x) {
  return to_acc(x);
}
template <>
__device__ __forceinline__ auto to_loc<double>(double x) {
  return x;
}
// A local value as a carry, and a carry plus a local value.
template <class A, class L>
__device__ __forceinline__ A loc_to_acc(L v) {
  if constexpr (sizeof(A) == sizeof(L)) return v;
  else return dd{v, 0.0};
}
__device__ __forceinline__ uint32_t acc_add_loc(uint32_t a, uint32_t b) { return a + b; }
__device__ __forceinline__ uint64_t acc_add_loc(uint64_t a, uint64_t b) { return a + b; }
__device__ __forceinline__ double acc_add_loc(double a, double b) { return a + b; }
__device__ __forceinline__ dd acc_add_loc(dd a, double b) { return dd_add_d(a, b); }
// The output element of a carry plus a local value (double-double: hi + (lo + r)).
template <class T, class A, class L>
__device__ __forceinline__ T out_of(A base, L r) {
  if constexpr (sizeof(A) == sizeof(L)) return from_acc<T>(acc_add(base, r));
  else return base.hi + (base.lo + r);
}

template <class T>
__device__ __forceinline__ T from_acc(typename Traits<T>::A a);
template <>
__device__ __forceinline__ int32_t from_acc<int32_t>(uint32_t a) {
  return static_cast<int32_t>(a);
}
template <>
__device__ __forceinline__ uint32_t from_acc<uint32_t>(uint32_t a) {
  return a;
}
template <>
__device__ __forceinline__ int64_t from_acc<int64_t>(uint64_t a) {
  return static_cast<int64_t>(a);
}
template <>
__device__ __forceinline__ uint64_t from_acc<uint64_t>(uint64_t a) {
  return a;
}
template <>
__device__ __forceinline__ float from_acc<float>(double a) {
  return static_cast<float>(a);  // round to nearest even
}
template <>
__device__ __forceinline__ double from_acc<double>(dd a) {
  return a.hi + a.lo;
}

// ---------------------------------------------------------------- DPP moves
// Move a value of any 4-byte-multiple type across lanes with one DPP mov per dword.
// Lanes whose source is outside the row / disabled by the masks receive 0.
template <int CTRL, int ROW_MASK, int BANK_MASK, class V>
__device__ __forceinline__ V dpp_mov(V v) {
  static_assert(sizeof(V) % 4 == 0, "dpp_mov needs dword-sized values");
  constexpr int W = sizeof(V) / 4;
  union U {
    V v;
    int w[W];
  } in, out;
  in.v = v;
#pragma unroll
  for (int i = 0; i < W; ++i)
    out.w[i] = __builtin_amdgcn_update_dpp(0, in.w[i], CTRL, ROW_MASK, BANK_MASK, false);
  return out.v;
}

template <class V>
__device__ __forceinline__ V read_lane(V v, int lane) {
  constexpr int W = sizeof(V) / 4;
  union U {
    V v;
    int w[W];
  } in, out;
  in.v = v;
#pragma unroll
  for (int i = 0; i < W; ++i) out.w[i] = __builtin_amdgcn_readlane(in.w[i], lane);
  return out.v;
}

// DPP control codes (GFX9 encoding).
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowShr1 = 0x111;
constexpr int kRowShr2 = 0x112;
constexpr int kRowShr4 = 0x114;
constexpr int kRowShr8 = 0x118;
constexpr int kWaveShr1 = 0x138;
constexpr int kRowMirror = 0x140;
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;

// Inclusive scan across the 64 lanes of a wave: Hillis-Steele inside each 16-lane row
// (row_shr 1,2,4,8), then row_bcast15 (rows 1,3) and row_bcast31 (rows 2,3).
template <class A>
__device__ __forceinline__ A wave_incl_scan(A v) {
  v = acc_add(v, dpp_mov<kRowShr1, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr2, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr4, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr8, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowBcast15, 0xa, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowBcast31, 0xc, 0xf>(v));
  return v;
}

// Inclusive scan inside each 16-lane row (row_shr 1, 2, 4, 8).
template <class A>
__device__ __forceinline__ A row_incl_scan(A v) {
  v = acc_add(v, dpp_mov<kRowShr1, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr2, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr4, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowShr8, 0xf, 0xf>(v));
  return v;
}

// Exclusive counterpart of an inclusive wave scan: shift right by one lane (lane 0 -> 0).
template <class A>
__device__ __forceinline__ A wave_shift_right1(A v) {
  return dpp_mov<kWaveShr1, 0xf, 0xf>(v);
}

// Sum over the 64 lanes, identical in every lane: butterfly inside each row
// (quad xor 1, quad xor 2, half-row mirror, row mirror), then the four row totals in
// a fixed order through readlane.
template <class A>
__device__ __forceinline__ A wave_sum(A v) {
  v = acc_add(v, dpp_mov<kQuadXor1, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kQuadXor2, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowHalfMirror, 0xf, 0xf>(v));
  v = acc_add(v, dpp_mov<kRowMirror, 0xf, 0xf>(v));
  A r0 = read_lane(v, 0), r1 = read_lane(v, 16), r2 = read_lane(v, 32), r3 = read_lane(v, 48);
  return acc_add(acc_add(r0, r1), acc_add(r2, r3));
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---------------------------------------------------------------- read-only tables
// Entry i of a table that no kernel writes while it runs, at a wave-uniform index, as a
// scalar load (constant address space).  A vector load would count in the same in-order
// counter as the tile's stores, and a wait for it would also wait for every store issued
// before it.
template <class T>
__device__ __forceinline__ T ld_table(const T* p, uint32_t i) {
  using CP = const __attribute__((address_space(4))) T*;
  return ((CP)(p))[static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(i))];
}

// ---------------------------------------------------------------- buffer resources
// A raw buffer over [p, p + nbytes) for a wave-uniform p and nbytes: loads past nbytes
// return zero without touching memory, stores there are dropped.  readfirstlane returns
// int, so each half of the address is taken as uint32_t before widening (a sign-extended
// low half would set every bit of the high half).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_rsrc(const void* p, uint32_t nbytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32)));
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a)));
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>((static_cast<uint64_t>(hi) << 32) | lo), 0,
                                           static_cast<int>(__builtin_amdgcn_readfirstlane(nbytes)), 0x00020000);
}

// A wave-uniform 64-bit difference clamped to [0, m] (m < 2^31) with 32-bit compares on
// its halves: gfx9 compares signed 64-bit values only in vector registers, whose
// temporaries can collide with the destinations of loads in flight.
__device__ __forceinline__ int clamp_span(int64_t d, int m) {
  const int32_t dh = static_cast<int32_t>(static_cast<uint64_t>(d) >> 32);
  const uint32_t dl = static_cast<uint32_t>(d);
  return dh < 0 ? 0 : (dh > 0 ? m : (dl < static_cast<uint32_t>(m) ? static_cast<int>(dl) : m));
}

// ---------------------------------------------------------------- agent-scope words
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Store / load a carry value word by word through agent-scope atomics (used for the
// per-block partials that another block combines).
template <class A>
__device__ __forceinline__ void st_carry(A* p, A v) {
  constexpr int W = sizeof(A) / 4;
  union U {
    A v;
    uint32_t w[W];
  } u;
  u.v = v;
  uint32_t* q = reinterpret_cast<uint32_t*>(p);
#pragma unroll
  for (int i = 0; i < W; ++i) st_agent32(q + i, u.w[i]);
}
template <class A>
__device__ __forceinline__ A ld_carry(const A* p) {
  constexpr int W = sizeof(A) / 4;
  union U {
    A v;
    uint32_t w[W];
  } u;
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
#pragma unroll
  for (int i = 0; i < W; ++i) u.w[i] = ld_agent32(q + i);
  return u.v;
}

__device__ __forceinline__ uint32_t next_epoch(uint32_t e) { return e >= kEpochMax ? 1u : e + 1u; }

// Block start: take a ticket and read the epoch (thread 0), broadcast through LDS.
__device__ __forceinline__ void block_begin(Ctl* ctl, uint32_t* s_ticket, uint32_t* s_epoch) {
  if (threadIdx.x == 0) {
    *s_ticket = __hip_atomic_fetch_add(&ctl->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *s_epoch = ld_agent32(&ctl->epoch);
  }
  __syncthreads();
}

// Block end (thread 0 only): the last block of the launch resets ticket/done and
// advances the epoch for the next launch on this context.
__device__ __forceinline__ void block_end(Ctl* ctl, uint32_t epoch) {
  uint32_t d = __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (d == gridDim.x - 1) {
    st_agent32(&ctl->ticket, 0u);
    st_agent32(&ctl->done, 0u);
    st_agent32(&ctl->epoch, next_epoch(epoch));
  }
}

// ---------------------------------------------------------------- look-back status
// Tile status: W = sizeof(A)/4 words per tile, word = (flag << 32) | payload, with
// flag = (epoch << 2) | state, state 1 = aggregate, 2 = inclusive prefix.
constexpr uint32_t kStAgg = 1u;
constexpr uint32_t kStIncl = 2u;

template <class A>
__device__ __forceinline__ void publish(uint64_t* status, uint64_t tile, uint32_t epoch,
                                        uint32_t state, A v) {
  constexpr int W = sizeof(A) / 4;
  union U {
    A v;
    uint32_t w[W];
  } u;
  u.v = v;
  const uint64_t flag = static_cast<uint64_t>((epoch << 2) | state) << 32;
#pragma unroll
  for (int i = 0; i < W; ++i) st_agent(status + tile * W + i, flag | u.w[i]);
}

// Read tile `tile`'s status.  Returns the state (0 = not yet valid for this epoch).
template <class A>
__device__ __forceinline__ uint32_t peek(const uint64_t* status, uint64_t tile, uint32_t epoch,
                                         A& v) {
  constexpr int W = sizeof(A) / 4;
  union U {
    A v;
    uint32_t w[W];
  } u;
  uint32_t f0 = 0;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const uint64_t word = ld_agent(status + tile * W + i);
    const uint32_t f = static_cast<uint32_t>(word >> 32);
    if (i == 0) f0 = f;
    ok = ok && (f == f0);
    u.w[i] = static_cast<uint32_t>(word);
  }
  v = u.v;
  if (!ok || (f0 >> 2) != epoch) return 0u;
  return f0 & 3u;
}

// Decoupled look-back (Merrill & Garland 2016), run by one whole wave for tile `tile`
// (> first).  Each poll covers kLookbackR * 64 predecessors: lane L inspects tiles
// pred - (L*R + r), r = 0..R-1, nearest first, with independent loads, so the inclusive
// prefix frontier advances 64*R tiles per memory round trip.  Tiles before `first` (the
// first tile of the segment) count as an inclusive prefix of 0.  The wave stops at the
// closest inclusive prefix once every tile between it and `tile` has published; it
// waits on nothing beyond that.  Returns the exclusive prefix of `tile` in every lane.
// Every predecessor has already taken its ticket (tickets are handed out in order), so
// every wait ends.
constexpr int kLookbackR = 1;

template <class A>
__device__ __forceinline__ A lookback(const uint64_t* status, uint64_t tile, int64_t first,
                                      uint32_t epoch) {
  constexpr int R = kLookbackR;
  const int lane = lane_id();
  A run = acc_zero<A>();
  int64_t pred = static_cast<int64_t>(tile) - 1;
  while (true) {
    A v[R];
    uint32_t st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t t = pred - (lane * R + r);
      v[r] = acc_zero<A>();
      st[r] = (t >= first) ? peek<A>(status, static_cast<uint64_t>(t), epoch, v[r]) : kStIncl;
    }
    while (true) {
      // this lane: first inclusive entry ri (R: none) unless an earlier entry is unpublished
      int ri = R;
      bool blocked = false;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (ri == R && !blocked) {
          if (st[r] == 0u) blocked = true;
          else if (st[r] == kStIncl) ri = r;
        }
      }
      const uint64_t incl = __ballot(ri < R);
      const uint64_t blk = __ballot(blocked);
      if (incl != 0ull) {
        const int p = __ffsll(static_cast<unsigned long long>(incl)) - 1;
        const uint64_t below = (p == 0) ? 0ull : (~0ull >> (64 - p));
        if ((blk & below) == 0ull) {
          A c = acc_zero<A>();
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (lane < p || (lane == p && r <= ri)) c = acc_add(c, v[r]);
          return acc_add(run, wave_sum(c));
        }
      } else if (blk == 0ull) {
        A c = acc_zero<A>();
#pragma unroll
        for (int r = 0; r < R; ++r) c = acc_add(c, v[r]);
        run = acc_add(run, wave_sum(c));
        break;  // every entry is an aggregate: next window
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (st[r] == 0u)
          st[r] = peek<A>(status, static_cast<uint64_t>(pred - (lane * R + r)), epoch, v[r]);
    }
    pred -= static_cast<int64_t>(kWave) * R;
  }
}

// ---------------------------------------------------------------- vector IO
template <class T, int N>
struct alignas(sizeof(T) * N) Vec {
  T v[N];
};

using u32x4 = __attribute__((ext_vector_type(4))) unsigned int;

// 16-byte non-temporal load / store of a Vec (streamed data: read or written once).
template <class V>
__device__ __forceinline__ V ld_stream(const V* p) {
  static_assert(sizeof(V) == 16, "16-byte vectors only");
  return __builtin_bit_cast(V, __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p)));
}
template <class V>
__device__ __forceinline__ void st_stream(V* p, const V& v) {
  static_assert(sizeof(V) == 16, "16-byte vectors only");
  __builtin_nontemporal_store(__builtin_bit_cast(u32x4, v), reinterpret_cast<u32x4*>(p));
}

}  // namespace pxk
