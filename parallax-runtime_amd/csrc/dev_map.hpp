// dev_map.hpp -- element maps of the descriptor (px_mapop) on the device: the map of
// the transform kernels (k_map.hip) and the functors the fused map-reduce instantiates
// (k_reduce.hip).  Integer arithmetic wraps (two's complement, no undefined overflow);
// AXPB on floats is one fma (single rounding).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "dev_pred.hpp"
#include "parallax/descriptor.h"

namespace pxk {

template <class U>
__device__ __forceinline__ U wrap_add(U a, U b) {
  if constexpr (std::is_integral_v<U>) {
    using W = std::make_unsigned_t<U>;
    return static_cast<U>(static_cast<W>(static_cast<W>(a) + static_cast<W>(b)));
  } else {
    return a + b;
  }
}
template <class U>
__device__ __forceinline__ U wrap_mul(U a, U b) {
  if constexpr (std::is_integral_v<U>) {
    using W = std::make_unsigned_t<U>;
    return static_cast<U>(static_cast<W>(static_cast<W>(a) * static_cast<W>(b)));
  } else {
    return a * b;
  }
}

template <class U>
__device__ __forceinline__ U map_apply(uint32_t fn, U x, U a, U b) {
  switch (fn) {
    case PX_MAP_IDENTITY: return x;
    case PX_MAP_FILL: return a;
    case PX_MAP_ADD: return wrap_add(x, a);
    case PX_MAP_MUL: return wrap_mul(x, a);
    case PX_MAP_AXPB:
      if constexpr (std::is_floating_point_v<U>) {
        return __builtin_fma(a, x, b);  // single rounding
      } else {
        return wrap_add(wrap_mul(a, x), b);
      }
    case PX_MAP_NEG:
      if constexpr (std::is_floating_point_v<U>) {
        return -x;
      } else {
        return wrap_mul(x, static_cast<U>(-1));
      }
    case PX_MAP_ABS:
      if constexpr (std::is_floating_point_v<U>) {
        return __builtin_fabs(x);
      } else if constexpr (std::is_signed_v<U>) {
        return x < 0 ? wrap_mul(x, static_cast<U>(-1)) : x;
      } else {
        return x;
      }
    case PX_MAP_SQUARE: return wrap_mul(x, x);
    default: return x;
  }
}

// Map functors with the op fixed at compile time (fused map-reduce).  Each converts the
// input to the output type first, then maps, as the transform kernel does.
template <class TI, class TO, int FN>
struct MapOp {
  TO a, b;
  __device__ __forceinline__ TO operator()(TI x) const {
    return map_apply<TO>(FN, static_cast<TO>(x), a, b);
  }
};

// Predicate map (count_if): 1 when pred(x), else 0; evaluated a vector at a time.
template <class TI, class TO>
struct MapPred {
  Pred<TI> p;
};

template <class M>
struct is_map_pred : std::false_type {};
template <class TI, class TO>
struct is_map_pred<MapPred<TI, TO>> : std::true_type {};

}  // namespace pxk
