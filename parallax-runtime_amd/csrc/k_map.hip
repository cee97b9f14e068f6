// k_map.hip -- element-wise skeletons, pointer relocation, search and fill kernels.
//
// launch / launch_with_captures / launch_transform (reference KernelLauncher,
// src/kernel/kernel_launcher.cpp:277-817) run compiler-emitted SPIR-V for a user lambda.
// The MI355X engine runs a descriptor functor (parallax/descriptor.h px_mapop) whose
// constants come from the capture bytes, exactly where the reference binds the closure
// (uniform@2).  Captured host pointers into the arena are relocated on the device with
// the reference's formula gpu = dev_base + (host_ptr - host_base) (relocate.comp:23-29).
//
// Search kernels (reference launch_find / launch_argminmax / launch_mismatch,
// kernel_launcher.cpp:987-1278) finish on the device instead of combining per-block
// winners on the host.
#include <algorithm>
#include <type_traits>

#include "dev_common.hpp"
#include "dev_map.hpp"
#include "dev_pred.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

template <class U>
__device__ __forceinline__ U cap_or_const(const MapArgs& a, int slot, uint64_t c) {
  if (a.cap_size >= (slot + 1) * sizeof(U) && (slot + 1) * sizeof(U) <= sizeof(a.cap)) {
    U v;
    __builtin_memcpy(&v, a.cap + slot * sizeof(U), sizeof(U));
    return v;
  }
  return bits_as<U>(c);
}

template <class TI, class TO>
__global__ __launch_bounds__(256) void k_transform(const TI* in, TO* out, uint64_t n, px_desc d,
                                                   MapArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (d.fn == PX_MAP_PRED) {
    const Pred<TI> p(px_desc{d.magic, d.version, d.op, d.dtype, d.aux, d.fn2, d.flags, 0, d.c0, d.c1});
    for (; i < n; i += stride) out[i] = p(in[i]) ? TO(1) : TO(0);
    return;
  }
  if (d.fn == PX_MAP_GATHER) {
    uint64_t tp = 0, size = 0;
    if (a.cap_size >= 16) {
      __builtin_memcpy(&tp, a.cap, 8);
      __builtin_memcpy(&size, a.cap + 8, 8);
    }
    // relocate a captured arena host pointer: gpu = dev_base + (host_ptr - host_base)
    if (tp >= a.host_base && tp < a.host_base + a.arena_bytes) tp = a.dev_base + (tp - a.host_base);
    const TO* table = reinterpret_cast<const TO*>(tp);
    for (; i < n; i += stride) {
      if constexpr (std::is_integral_v<TI>) {
        const uint64_t idx = static_cast<uint64_t>(in[i]);
        out[i] = (table != nullptr && idx < size) ? table[idx] : TO{};
      } else {
        out[i] = TO{};
      }
    }
    return;
  }
  const TO ca = cap_or_const<TO>(a, 0, d.c0);
  const TO cb = cap_or_const<TO>(a, 1, d.c1);
  const uint32_t fn = d.fn;
  for (; i < n; i += stride) out[i] = map_apply<TO>(fn, static_cast<TO>(in[i]), ca, cb);
}

__global__ void k_relocate_pointers(uint64_t* ptrs, uint64_t n, uint64_t host_base,
                                    uint64_t dev_base, uint64_t bytes) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = ptrs[i];
  if (p >= host_base && p < host_base + bytes) ptrs[i] = dev_base + (p - host_base);
}

__global__ void k_relocate_probe(uint64_t host_ptr, uint64_t host_base, uint64_t dev_base,
                                 uint32_t count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float* r = reinterpret_cast<float*>(dev_base + (host_ptr - host_base));
  r[i] = r[i] * 2.0f + 1.0f;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// First index (relative to lo) in [lo, hi) of in_al whose element passes the test.
// Vectors of 16 bytes from the aligned-down pointer; the head and tail of fewer than VEC
// elements are tested by block 0.  Blocks run in rounds of UL vectors per thread in
// increasing index order: a round's hits are reduced in LDS and published with one
// device-scope min, after which the block stops (its later rounds hold larger indices);
// a block also stops once a published hit lies below its next round.  The published
// minimum only decreases, so a stale read of it only delays a stop.
constexpr int kSearchUL = 4;

template <class T, class Test>
__device__ __forceinline__ void search_first(int64_t lo, int64_t hi, uint64_t* out, const Test& test) {
  constexpr int VEC = 16 / sizeof(T);
  constexpr int UL = kSearchUL;
  __shared__ unsigned long long s_min;
  __shared__ unsigned long long s_seen;
  const uint64_t n = static_cast<uint64_t>(hi - lo);
  const int64_t a = (lo + VEC - 1) / VEC;  // interior vectors [a, b)
  const int64_t b = hi / VEC;
  if (blockIdx.x == 0 && threadIdx.x < 2 * VEC) {
    int64_t pos;
    bool ok;
    if (threadIdx.x < VEC) {
      pos = lo + threadIdx.x;
      ok = pos < a * VEC && pos < hi;
    } else {
      pos = (a > b ? a : b) * VEC + (threadIdx.x - VEC);
      ok = pos < hi;
    }
    if (ok && test.one(pos))
      atomicMin(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(pos - lo));
  }
  if (threadIdx.x == 0) {
    s_min = n;
    s_seen = ld_agent(out);
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t r = a + static_cast<int64_t>(blockIdx.x) * 256; r < b; r += UL * stride) {
    if (s_seen + static_cast<uint64_t>(lo) <= static_cast<uint64_t>(r * VEC)) break;  // uniform
    typename Test::Group g;
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int64_t v = r + u * stride + threadIdx.x;
      if (v < b) test.load(g, u, v);
    }
    const uint64_t next_seen = threadIdx.x == 0 ? ld_agent(out) : 0ull;  // in flight with the loads
    uint64_t best = n;
#pragma unroll
    for (int u = 0; u < UL; ++u) {
      const int64_t v = r + u * stride + threadIdx.x;
      if (v < b && best == n) {
        const uint32_t m = test.mask(g, u);
        if (m) best = static_cast<uint64_t>(v * VEC + __builtin_ctz(m) - lo);
      }
    }
    if (best < n) atomicMin(&s_min, static_cast<unsigned long long>(best));
    __syncthreads();
    const uint64_t found = s_min;
    if (found < n) {  // uniform
      if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(found));
      break;
    }
    if (threadIdx.x == 0) s_seen = next_seen;
    __syncthreads();
  }
}

template <class T>
struct FindTest {
  using V = Vec<T, 16 / sizeof(T)>;
  struct Group {
    V x[kSearchUL];
  };
  const V* vin;
  const T* in_al;
  Pred<T> pred;
  __device__ void load(Group& g, int u, int64_t v) const { g.x[u] = ld_stream(vin + v); }
  __device__ uint32_t mask(const Group& g, int u) const { return pred.template mask<16 / sizeof(T)>(g.x[u].v); }
  __device__ bool one(int64_t pos) const { return pred(in_al[pos]); }
};

template <class T>
struct MismatchTest {
  static constexpr int VEC = 16 / sizeof(T);
  using V = Vec<T, VEC>;
  struct Group {
    V x[kSearchUL], y[kSearchUL];
  };
  const T* a_al;
  const T* b_al;
  __device__ void load(Group& g, int u, int64_t v) const {
    g.x[u] = ld_stream(reinterpret_cast<const V*>(a_al) + v);
    g.y[u] = ld_stream(reinterpret_cast<const V*>(b_al) + v);
  }
  __device__ uint32_t mask(const Group& g, int u) const {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < VEC; ++j) m |= (g.x[u].v[j] == g.y[u].v[j] ? 0u : 1u) << j;
    return m;
  }
  __device__ bool one(int64_t pos) const { return !(a_al[pos] == b_al[pos]); }
};

template <class T>
__global__ __launch_bounds__(256) void k_find(const T* in_al, int64_t lo, int64_t hi, px_desc pd, uint64_t* out) {
  const FindTest<T> t{reinterpret_cast<const typename FindTest<T>::V*>(in_al), in_al, Pred<T>(pd)};
  search_first<T>(lo, hi, out, t);
}

// a_al and b_al share their offset from a 16-byte boundary (else k_mismatch_scalar).
template <class T>
__global__ __launch_bounds__(256) void k_mismatch(const T* a_al, const T* b_al, int64_t lo, int64_t hi, uint64_t* out) {
  const MismatchTest<T> t{a_al, b_al};
  search_first<T>(lo, hi, out, t);
}

template <class T>
__global__ __launch_bounds__(256) void k_mismatch_scalar(const T* a, const T* b, uint64_t n, uint64_t* out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    if (!(a[i] == b[i])) {
      atomicMin(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(i));
      return;
    }
  }
}

// (value, index) winner: min (or max) by operator<, ties to the first (or last) index.
template <class T>
struct Best {
  T v;
  uint64_t i;
};
template <class T>
__device__ __forceinline__ bool better(const Best<T>& a, const Best<T>& b, bool want_max,
                                       bool want_last, uint64_t n) {
  if (a.i >= n) return false;
  if (b.i >= n) return true;
  const bool lt = want_max ? (b.v < a.v) : (a.v < b.v);
  const bool gt = want_max ? (a.v < b.v) : (b.v < a.v);
  if (lt) return true;
  if (gt) return false;
  return want_last ? (a.i > b.i) : (a.i < b.i);
}

// A thread keeps one winner per vector slot j (its elements at vector position j, in
// increasing index order), so one comparison per element keeps that slot's winner
// (strict for the first index among ties, non-strict for the last) and the VEC chains
// of comparisons are independent; a winner's index is its 32-bit vector ordinal within
// the thread.  The slots' winners, the head and tail elements and the winners of other
// threads combine by better().
template <class T, bool MAX, bool LAST>
__global__ __launch_bounds__(256) void k_argminmax(const T* in_al, int64_t lo, int64_t hi,
                                                   Best<T>* partials, Ctl* ctl, uint64_t* out) {
  constexpr int VEC = 16 / sizeof(T);
  constexpr int UL = 8;  // vector loads in flight per thread
  constexpr bool want_max = MAX, want_last = LAST;
  using V = Vec<T, VEC>;
  __shared__ Best<T> s_b[256];
  __shared__ int s_last;
  const uint64_t n = static_cast<uint64_t>(hi - lo);
  Best<T> b{T{}, n};  // index n: none yet
  auto take = [&](T x, int64_t pos) {
    const Best<T> c{x, static_cast<uint64_t>(pos - lo)};
    if (better(c, b, want_max, want_last, n)) b = c;
  };
  auto wins = [](T x, T cur) {
    if constexpr (MAX)
      return LAST ? !(x < cur) : (cur < x);
    else
      return LAST ? !(cur < x) : (x < cur);
  };
  const int64_t va = (lo + VEC - 1) / VEC;  // interior vectors [va, vb)
  const int64_t vb = hi / VEC;
  if (blockIdx.x == 0 && threadIdx.x < VEC) {  // the head (before the first whole vector)
    const int64_t pos = lo + threadIdx.x;
    if (pos < va * VEC && pos < hi) take(in_al[pos], pos);
  }
  const V* vin = reinterpret_cast<const V*>(in_al);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  const int64_t v0 = va + static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  int64_t v = v0;
  T bv[VEC];
  uint32_t bk[VEC];  // vector ordinal of slot j's winner; none when the thread has no vector
  constexpr uint32_t kNone = 0xffffffffu;
#pragma unroll
  for (int j = 0; j < VEC; ++j) bk[j] = kNone;
  uint32_t k = 0;
  if (v < vb) {  // the thread's first vector seeds every slot
    const V x = ld_stream(vin + v);
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      bv[j] = x.v[j];
      bk[j] = 0;
    }
    v += stride;
    k = 1;
  }
  auto step = [&](const V& x, uint32_t kx) {
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const bool w = wins(x.v[j], bv[j]);
      bv[j] = w ? x.v[j] : bv[j];
      bk[j] = w ? kx : bk[j];
    }
  };
  for (; v + (UL - 1) * stride < vb; v += UL * stride, k += UL) {
    V x[UL];
#pragma unroll
    for (int u = 0; u < UL; ++u) x[u] = ld_stream(vin + v + u * stride);
#pragma unroll
    for (int u = 0; u < UL; ++u) step(x[u], k + static_cast<uint32_t>(u));
  }
  for (; v < vb; v += stride, ++k) step(ld_stream(vin + v), k);
#pragma unroll
  for (int j = 0; j < VEC; ++j)
    if (bk[j] != kNone) take(bv[j], (v0 + static_cast<int64_t>(bk[j]) * stride) * VEC + j);
  if (blockIdx.x == 0 && threadIdx.x < VEC) {  // the tail (after the last whole vector)
    const int64_t pos = (va > vb ? va : vb) * VEC + threadIdx.x;
    if (pos < hi) take(in_al[pos], pos);
  }
  s_b[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w && better(s_b[threadIdx.x + w], s_b[threadIdx.x], want_max, want_last, n))
      s_b[threadIdx.x] = s_b[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = s_b[0];
    __threadfence();
    const uint32_t d =
        __hip_atomic_fetch_add(&ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (d == gridDim.x - 1) ? 1 : 0;
    if (s_last) st_agent32(&ctl->done, 0u);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  Best<T> c{T{}, n};
  for (uint32_t j = threadIdx.x; j < gridDim.x; j += 256) {
    const volatile Best<T>* vp = &partials[j];
    Best<T> p;
    p.v = vp->v;
    p.i = vp->i;
    if (better(p, c, want_max, want_last, n)) c = p;
  }
  s_b[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w && better(s_b[threadIdx.x + w], s_b[threadIdx.x], want_max, want_last, n))
      s_b[threadIdx.x] = s_b[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s_b[0].i;
}

inline uint32_t ew_grid(uint64_t n, uint32_t cap) {
  uint64_t g = (n + 256 * 4 - 1) / (256 * 4);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<uint32_t>(g);
}

template <class TI>
hipError_t transform_in(const px_desc& d, const void* in, void* out, uint64_t n, const MapArgs& a,
                        hipStream_t stream) {
  const uint32_t grid = ew_grid(n, 8192);
  PXK_DISPATCH(d.aux, {
    k_transform<TI, T><<<grid, 256, 0, stream>>>(reinterpret_cast<const TI*>(in),
                                                 reinterpret_cast<T*>(out), n, d, a);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t transform(const px_desc& d, const void* in, void* out, uint64_t n, const MapArgs& a,
                     hipStream_t stream) {
  if (n == 0) return hipSuccess;
  switch (d.dtype) {
    case PX_I32: return transform_in<int32_t>(d, in, out, n, a, stream);
    case PX_U32: return transform_in<uint32_t>(d, in, out, n, a, stream);
    case PX_I64: return transform_in<int64_t>(d, in, out, n, a, stream);
    case PX_U64: return transform_in<uint64_t>(d, in, out, n, a, stream);
    case PX_F32: return transform_in<float>(d, in, out, n, a, stream);
    case PX_F64: return transform_in<double>(d, in, out, n, a, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t relocate_pointers(uint64_t* ptrs, uint64_t n, uint64_t host_base, uint64_t dev_base,
                             uint64_t bytes, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_relocate_pointers<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(ptrs, n, host_base,
                                                                                dev_base, bytes);
  return hipGetLastError();
}

hipError_t relocate_probe(uint64_t host_ptr, uint64_t host_base, uint64_t dev_base, uint32_t count,
                          hipStream_t stream) {
  if (count == 0) return hipSuccess;
  k_relocate_probe<<<(count + 255) / 256, 256, 0, stream>>>(host_ptr, host_base, dev_base, count);
  return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  k_fill_u32<<<ew_grid(n, 4096), 256, 0, stream>>>(p, n, v);
  return hipGetLastError();
}

// Aligned-down pointer and element range [lo, lo + n) of an operand.
template <class T>
inline const T* aligned_down(const void* p, int64_t& lo) {
  const uintptr_t u = reinterpret_cast<uintptr_t>(p);
  lo = static_cast<int64_t>((u & 15u) / sizeof(T));
  return reinterpret_cast<const T*>(u & ~static_cast<uintptr_t>(15));
}

// Search grids: at most per_cu 256-thread blocks per CU (measured at 2^28 i32, 4 / 8 per
// CU: find with no hit 215 / 161 us, mismatch 323 / 361 us), no more blocks than rounds
// of kSearchUL vectors need.
inline uint32_t search_grid(uint64_t n, size_t elem, const Scratch& s, uint32_t per_cu) {
  const uint64_t per_block = 256ull * kSearchUL * (16 / elem);
  const uint64_t cap = std::max<uint64_t>(static_cast<uint64_t>(s.max_blocks / 8) * per_cu, 1);
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<uint32_t>(g);
}

hipError_t find_first(const px_desc& pred, const void* in, uint64_t n, uint64_t* out_dev,
                      const Scratch& s, hipStream_t stream) {
  k_set_u64<<<1, 1, 0, stream>>>(out_dev, n);
  if (n == 0) return hipGetLastError();
  PXK_DISPATCH(pred.dtype, {
    int64_t lo;
    const T* al = aligned_down<T>(in, lo);
    k_find<T><<<search_grid(n, sizeof(T), s, 8), 256, 0, stream>>>(al, lo, lo + static_cast<int64_t>(n), pred, out_dev);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

hipError_t mismatch(int dtype, const void* a, const void* b, uint64_t n, uint64_t* out_dev,
                    const Scratch& s, hipStream_t stream) {
  k_set_u64<<<1, 1, 0, stream>>>(out_dev, n);
  if (n == 0) return hipGetLastError();
  PXK_DISPATCH(dtype, {
    int64_t lo, lo_b;
    const T* aa = aligned_down<T>(a, lo);
    const T* ba = aligned_down<T>(b, lo_b);
    if (lo == lo_b && (reinterpret_cast<uintptr_t>(a) & (sizeof(T) - 1)) == 0 &&
        (reinterpret_cast<uintptr_t>(b) & (sizeof(T) - 1)) == 0) {
      k_mismatch<T><<<search_grid(n, sizeof(T), s, 4), 256, 0, stream>>>(aa, ba, lo, lo + static_cast<int64_t>(n), out_dev);
    } else {  // different offsets from a 16-byte boundary: element loads
      k_mismatch_scalar<T><<<ew_grid(n, s.max_blocks), 256, 0, stream>>>(reinterpret_cast<const T*>(a),
                                                                         reinterpret_cast<const T*>(b), n, out_dev);
    }
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

hipError_t argminmax(int dtype, const void* in, uint64_t n, bool want_max, bool want_last,
                     uint64_t* out_dev, const Scratch& s, hipStream_t stream) {
  if (n == 0) {
    k_set_u64<<<1, 1, 0, stream>>>(out_dev, 0);
    return hipGetLastError();
  }
  PXK_DISPATCH(dtype, {
    // one 256-thread block per CU at most, 8 vectors per thread per round (2^28 f32 /
    // i32: 1 / 2 / 4 blocks per CU 0.167 / 0.210 / 0.278 ms; with one comparison chain
    // per thread instead of one per vector slot, 4 per CU was best at 0.260 ms)
    const uint64_t per_block = 256ull * 8 * (16 / sizeof(T));
    const uint64_t cap = s.max_blocks / 8 ? s.max_blocks / 8 : 1;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>((n + per_block - 1) / per_block, 1), cap));
    if (grid * sizeof(Best<T>) > s.partial_bytes) return hipErrorInvalidValue;
    int64_t lo;
    const T* al = aligned_down<T>(in, lo);
    Best<T>* part = reinterpret_cast<Best<T>*>(s.partials);
    const int64_t hi = lo + static_cast<int64_t>(n);
    if (want_max && want_last)
      k_argminmax<T, true, true><<<grid, 256, 0, stream>>>(al, lo, hi, part, s.ctl, out_dev);
    else if (want_max)
      k_argminmax<T, true, false><<<grid, 256, 0, stream>>>(al, lo, hi, part, s.ctl, out_dev);
    else if (want_last)
      k_argminmax<T, false, true><<<grid, 256, 0, stream>>>(al, lo, hi, part, s.ctl, out_dev);
    else
      k_argminmax<T, false, false><<<grid, 256, 0, stream>>>(al, lo, hi, part, s.ctl, out_dev);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

}  // namespace pxk
