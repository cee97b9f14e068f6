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

template <class T>
__global__ __launch_bounds__(256) void k_find(const T* in, uint64_t n, px_desc pd, uint64_t* out) {
  const Pred<T> pred(pd);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  uint64_t best = n;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    if (pred(in[i])) {
      best = i;
      break;
    }
  }
  if (best < n) atomicMin(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(best));
}

template <class T>
__global__ __launch_bounds__(256) void k_mismatch(const T* a, const T* b, uint64_t n, uint64_t* out) {
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

template <class T>
__global__ __launch_bounds__(256) void k_argminmax(const T* in, uint64_t n, bool want_max,
                                                   bool want_last, Best<T>* partials, Ctl* ctl,
                                                   uint64_t* out) {
  __shared__ Best<T> s_b[256];
  __shared__ int s_last;
  Best<T> b{T{}, n};
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    const Best<T> c{in[i], i};
    if (better(c, b, want_max, want_last, n)) b = c;
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

hipError_t find_first(const px_desc& pred, const void* in, uint64_t n, uint64_t* out_dev,
                      const Scratch& s, hipStream_t stream) {
  (void)s;
  k_set_u64<<<1, 1, 0, stream>>>(out_dev, n);
  if (n == 0) return hipGetLastError();
  const uint32_t grid = ew_grid(n, s.max_blocks);
  PXK_DISPATCH(pred.dtype, {
    k_find<T><<<grid, 256, 0, stream>>>(reinterpret_cast<const T*>(in), n, pred, out_dev);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

hipError_t mismatch(int dtype, const void* a, const void* b, uint64_t n, uint64_t* out_dev,
                    const Scratch& s, hipStream_t stream) {
  k_set_u64<<<1, 1, 0, stream>>>(out_dev, n);
  if (n == 0) return hipGetLastError();
  const uint32_t grid = ew_grid(n, s.max_blocks);
  PXK_DISPATCH(dtype, {
    k_mismatch<T><<<grid, 256, 0, stream>>>(reinterpret_cast<const T*>(a),
                                            reinterpret_cast<const T*>(b), n, out_dev);
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
  const uint32_t grid = ew_grid(n, s.max_blocks);
  PXK_DISPATCH(dtype, {
    if (grid * sizeof(Best<T>) > s.partial_bytes) return hipErrorInvalidValue;
    k_argminmax<T><<<grid, 256, 0, stream>>>(reinterpret_cast<const T*>(in), n, want_max, want_last,
                                             reinterpret_cast<Best<T>*>(s.partials), s.ctl, out_dev);
    return hipGetLastError();
  });
  return hipErrorInvalidValue;
}

}  // namespace pxk
