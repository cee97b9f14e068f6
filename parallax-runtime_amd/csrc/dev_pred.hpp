// dev_pred.hpp -- predicate semantics on the device (descriptor predicates of
// parallax/descriptor.h).  Integer remainder follows C++ truncation; x % 0 is false and
// x % -1 is 0 (both undefined in C++; pinned here).  Floats use IEEE comparisons, so a
// NaN fails every ordered predicate; MOD_EQ on floats uses fmod.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "parallax/descriptor.h"

namespace pxk {

template <class T>
__host__ __device__ __forceinline__ T bits_as(uint64_t b) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, static_cast<uint32_t>(b));
  } else {
    return __builtin_bit_cast(T, b);
  }
}

template <class T>
__device__ __forceinline__ bool pred_eval(uint32_t fn, uint32_t flags, T c0, T c1, T x) {
  bool r = false;
  switch (fn) {
    case PX_PRED_TRUE: r = true; break;
    case PX_PRED_EQ: r = (x == c0); break;
    case PX_PRED_NE: r = (x != c0); break;
    case PX_PRED_LT: r = (x < c0); break;
    case PX_PRED_LE: r = (x <= c0); break;
    case PX_PRED_GT: r = (x > c0); break;
    case PX_PRED_GE: r = (x >= c0); break;
    case PX_PRED_IN_RANGE: r = (c0 <= x) && (x < c1); break;
    case PX_PRED_MOD_EQ:
      if constexpr (std::is_floating_point_v<T>) {
        r = (fmod(x, c0) == c1);
      } else {
        if (c0 == T(0)) {
          r = false;
        } else if constexpr (std::is_signed_v<T>) {
          const T rem = (c0 == T(-1)) ? T(0) : static_cast<T>(x % c0);
          r = (rem == c1);
        } else {
          r = (static_cast<T>(x % c0) == c1);
        }
      }
      break;
    case PX_PRED_NONZERO: r = (x != T(0)); break;
    case PX_PRED_BITS_ANY:
      if constexpr (std::is_integral_v<T>) r = (x & c0) != 0;
      break;
    default: r = false; break;
  }
  return (flags & PX_FLAG_NEGATE) ? !r : r;
}

// Predicate with its constants unpacked once per kernel.  MOD_EQ by a positive power of
// two m is one AND and one compare, (x & am) == ac: the truncated remainder of x is
// x & (m-1) for x >= 0 and (x & (m-1)) - m for x < 0 when nonzero, so
//   c == 0:          am = m-1,            ac = 0 (any sign);
//   0 < c < m:       am = (m-1) | sign,   ac = c (x >= 0);
//   -m < c < 0:      am = (m-1) | sign,   ac = sign | (c + m) (x < 0);
//   |c| >= m:        never (am = 0, ac = 1).
template <class T>
struct Pred {
  uint32_t fn, flags;
  T c0, c1;
  T mod_mask;  // m - 1 when MOD_EQ's divisor m is a positive power of two, else 0
  T am, ac;    // the AND mask and compare value of that case
  __host__ __device__ __forceinline__ Pred(const px_desc& d)
      : fn(d.fn), flags(d.flags), c0(bits_as<T>(d.c0)), c1(bits_as<T>(d.c1)), mod_mask(T(0)), am(T(0)), ac(T(1)) {
    if constexpr (std::is_integral_v<T>) {
      if (fn == PX_PRED_MOD_EQ && c0 > T(0) && (c0 & (c0 - T(1))) == T(0)) {
        mod_mask = c0 - T(1);
        using U = std::make_unsigned_t<T>;
        const U mm = static_cast<U>(mod_mask), sign = static_cast<U>(U(1) << (8 * sizeof(T) - 1));
        if (c1 == T(0)) {
          am = static_cast<T>(mm);
          ac = T(0);
        } else if (c1 > T(0) && c1 <= mod_mask) {
          am = static_cast<T>(std::is_signed_v<T> ? (mm | sign) : mm);
          ac = c1;
        } else if (std::is_signed_v<T> && c1 < T(0) && c1 >= static_cast<T>(-mod_mask)) {
          am = static_cast<T>(mm | sign);
          ac = static_cast<T>(sign | static_cast<U>(c1 + c0));
        }
      }
    }
  }
  __device__ __forceinline__ bool operator()(T x) const { return pred_eval<T>(fn, flags, c0, c1, x); }

  // Bit i of the result: predicate of x[i], for i < N (N <= 32), with one branch on the
  // predicate kind for the whole array.
  template <int N>
  __device__ __forceinline__ uint32_t mask(const T (&x)[N]) const {
    static_assert(N <= 32, "at most 32 elements");
    uint32_t m = 0;
#define PXK_PMASK(EXPR)                                     \
  _Pragma("unroll") for (int i = 0; i < N; ++i) {           \
    const T v = x[i];                                       \
    if (EXPR) m |= 1u << i;                                 \
  }
    switch (fn) {
      case PX_PRED_TRUE: m = (N == 32) ? 0xffffffffu : ((1u << N) - 1u); break;
      case PX_PRED_EQ: PXK_PMASK(v == c0) break;
      case PX_PRED_NE: PXK_PMASK(v != c0) break;
      case PX_PRED_LT: PXK_PMASK(v < c0) break;
      case PX_PRED_LE: PXK_PMASK(v <= c0) break;
      case PX_PRED_GT: PXK_PMASK(v > c0) break;
      case PX_PRED_GE: PXK_PMASK(v >= c0) break;
      case PX_PRED_IN_RANGE: PXK_PMASK((c0 <= v) && (v < c1)) break;
      case PX_PRED_NONZERO: PXK_PMASK(v != T(0)) break;
      case PX_PRED_MOD_EQ:
        if constexpr (std::is_integral_v<T>) {
          if (mod_mask != T(0)) {
            const T a = am, c = ac;
            PXK_PMASK(T(v & a) == c)
            break;
          }
        }
        PXK_PMASK(pred_eval<T>(PX_PRED_MOD_EQ, 0u, c0, c1, v))
        break;
      case PX_PRED_BITS_ANY:
        if constexpr (std::is_integral_v<T>) {
          PXK_PMASK((v & c0) != T(0))
        }
        break;
      default: break;
    }
#undef PXK_PMASK
    if (flags & PX_FLAG_NEGATE) m = ~m & ((N == 32) ? 0xffffffffu : ((1u << N) - 1u));
    return m;
  }
};

}  // namespace pxk
