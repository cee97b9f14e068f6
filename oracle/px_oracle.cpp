// px_oracle.cpp -- CPU restatement of the reference runtime's fallback algorithms.
// TEST INFRASTRUCTURE ONLY (see px_oracle.h): the checker for the HIP path, never the
// thing measured or shipped.
//
// The reference's funnels in include/parallax/stdpar.hpp run an ISO host loop whenever a
// kernel lookup misses.  Those loops are the semantics the GPU engine must reproduce:
//   reduce        stdpar.hpp:146-148   T acc = init; acc = acc + data[i]   (init = E{}, :460)
//   incl. scan    stdpar.hpp:209-210   acc = acc + in[i]; out[i] = acc
//   excl. scan    stdpar.hpp:237-238   out[i] = acc; acc = acc + t          (acc = init)
//   sort          stdpar.hpp:178       std::sort(data, data + n)
//   copy_if       stdpar.hpp:319-321   stable filter
//   remove_if     stdpar.hpp:343-345   stable in-place removal of pred-true elements
//   partition     stdpar.hpp:369       std::partition
//   unique        stdpar.hpp:392       std::unique
//   count_if      stdpar.hpp:290-292
//   for_each map  stdpar.hpp:81        f(data[i]) in place
//   transform     stdpar.hpp:119       out[i] = f(in[i])
// The search entry points have no host loop in stdpar.hpp; their semantics are the
// reference launcher's host combine over per-block winners:
//   find          kernel_launcher.cpp:1180-1189  first index whose element passes; n if none
//   mismatch      kernel_launcher.cpp:1272-1277  first index with a[i] != b[i]; n if none
//   argminmax     kernel_launcher.cpp:1078-1098  min (max) by <, ties to the smaller index
//                 (the larger with want_last); 0 for n == 0.  Values compare in the
//                 element type here (the reference widens them to double, which merges
//                 int64 values that differ beyond 2^53).
// Integer arithmetic is done in the unsigned type of the same width, which is the
// two's-complement wrap the reference's signed loops produce in practice, but defined.
#include "px_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <type_traits>

namespace {

template <class T>
struct Wrap {  // arithmetic type with defined wraparound for integers
  using type = T;
};
template <>
struct Wrap<int32_t> {
  using type = uint32_t;
};
template <>
struct Wrap<int64_t> {
  using type = uint64_t;
};

template <class T>
inline T add(T a, T b) {
  if constexpr (std::is_integral_v<T>) {
    using W = typename Wrap<T>::type;
    return static_cast<T>(static_cast<W>(static_cast<W>(a) + static_cast<W>(b)));
  } else {
    return a + b;
  }
}
template <class T>
inline T mul(T a, T b) {
  if constexpr (std::is_integral_v<T>) {
    using W = typename Wrap<T>::type;
    return static_cast<T>(static_cast<W>(static_cast<W>(a) * static_cast<W>(b)));
  } else {
    return a * b;
  }
}

// Error-free transformation (Knuth TwoSum) and cascaded double-double accumulation
// (Ogita-Rump-Oishi Sum2).  Result error <= u|S| + gamma_{n-1}^2 sum|x|.
inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
struct DD {
  double hi = 0.0, lo = 0.0;
  void add(double x) {
    double s, e;
    two_sum(hi, x, s, e);
    hi = s;
    lo += e;
  }
  double value() const { return hi + lo; }
};

template <class T>
T load_const(uint64_t bits) {
  T v;
  std::memcpy(&v, &bits, sizeof(T));
  return v;
}

// Predicate semantics shared by every compaction / count / find descriptor.
template <class T>
bool pred(const px_desc& d, T x) {
  const T c0 = load_const<T>(d.c0), c1 = load_const<T>(d.c1);
  bool r = false;
  switch (d.fn) {
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
        r = (std::fmod(x, c0) == c1);
      } else {
        // C++ truncated remainder; x % 0 is defined here as "false" and x % -1 as 0
        // (both are undefined behaviour in C++, the engine pins them the same way).
        if (c0 == T(0)) {
          r = false;
        } else if constexpr (std::is_signed_v<T>) {
          T rem = (c0 == T(-1)) ? T(0) : static_cast<T>(x % c0);
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
  return (d.flags & PX_FLAG_NEGATE) ? !r : r;
}

#define PXO_DISPATCH(dtype, ...)                             \
  switch (dtype) {                                           \
    case PX_I32: { using T = int32_t; __VA_ARGS__; break; }  \
    case PX_U32: { using T = uint32_t; __VA_ARGS__; break; } \
    case PX_I64: { using T = int64_t; __VA_ARGS__; break; }  \
    case PX_U64: { using T = uint64_t; __VA_ARGS__; break; } \
    case PX_F32: { using T = float; __VA_ARGS__; break; }    \
    case PX_F64: { using T = double; __VA_ARGS__; break; }   \
    default: break;                                          \
  }

template <class T>
void reduce_t(const T* in, size_t n, T* out) {
  T acc{};  // init = E{} (stdpar.hpp:460); device_reduce folds from init (:146-148)
  for (size_t i = 0; i < n; ++i) acc = add(acc, in[i]);
  *out = acc;
}

template <class T>
void incl_scan_t(const T* in, T* out, size_t n) {
  T acc{};
  for (size_t i = 0; i < n; ++i) {
    acc = add(acc, in[i]);
    out[i] = acc;
  }
}

template <class T>
void excl_scan_t(const T* in, T* out, size_t n, T init) {
  T acc = init;
  for (size_t i = 0; i < n; ++i) {
    T t = in[i];  // read before write: in may alias out
    out[i] = acc;
    acc = add(acc, t);
  }
}

// Map functor semantics; arithmetic in the output type U.
template <class U>
U map_apply(const px_desc& d, uint32_t fn, U x, U a, U b) {
  switch (fn) {
    case PX_MAP_IDENTITY: return x;
    case PX_MAP_FILL: return a;
    case PX_MAP_ADD: return add(x, a);
    case PX_MAP_MUL: return mul(x, a);
    case PX_MAP_AXPB:
      if constexpr (std::is_floating_point_v<U>) {
        return std::fma(a, x, b);  // one rounding, as on the device
      } else {
        return add(mul(a, x), b);
      }
    case PX_MAP_NEG:
      if constexpr (std::is_floating_point_v<U>) {
        return -x;
      } else {
        return add(U(0), mul(x, static_cast<U>(-1)));
      }
    case PX_MAP_ABS:
      if constexpr (std::is_floating_point_v<U>) {
        return std::fabs(x);
      } else if constexpr (std::is_signed_v<U>) {
        return x < 0 ? mul(x, static_cast<U>(-1)) : x;
      } else {
        return x;
      }
    case PX_MAP_SQUARE: return mul(x, x);
    default: (void)d; return x;
  }
}

template <class U>
void map_consts(const px_desc& d, const void* cap, size_t cap_size, U& a, U& b) {
  a = load_const<U>(d.c0);
  b = load_const<U>(d.c1);
  if (cap && cap_size >= sizeof(U)) std::memcpy(&a, cap, sizeof(U));
  if (cap && cap_size >= 2 * sizeof(U))
    std::memcpy(&b, static_cast<const char*>(cap) + sizeof(U), sizeof(U));
}

template <class T, class U>
void transform_tu(const px_desc& d, const T* in, U* out, size_t n, const void* cap,
                  size_t cap_size) {
  if (d.fn == PX_MAP_PRED) {
    px_desc pd = d;
    pd.fn = d.fn2;
    for (size_t i = 0; i < n; ++i) out[i] = pred<T>(pd, in[i]) ? U(1) : U(0);
    return;
  }
  if (d.fn == PX_MAP_GATHER) {
    const U* table = nullptr;
    uint64_t size = 0;
    if (cap && cap_size >= 16) {
      std::memcpy(&table, cap, sizeof(table));
      std::memcpy(&size, static_cast<const char*>(cap) + 8, 8);
    }
    if constexpr (std::is_integral_v<T>) {
      for (size_t i = 0; i < n; ++i) {
        uint64_t idx = static_cast<uint64_t>(in[i]);
        out[i] = (table && idx < size) ? table[idx] : U{};
      }
    } else {
      for (size_t i = 0; i < n; ++i) out[i] = U{};
    }
    return;
  }
  U a, b;
  map_consts<U>(d, cap, cap_size, a, b);
  for (size_t i = 0; i < n; ++i) out[i] = map_apply<U>(d, d.fn, static_cast<U>(in[i]), a, b);
}

}  // namespace

extern "C" {

int pxo_abi_version(void) { return 1; }

void pxo_reduce(int dtype, const void* in, size_t n, void* out) {
  PXO_DISPATCH(dtype, reduce_t<T>(static_cast<const T*>(in), n, static_cast<T*>(out)))
}

void pxo_sum_exact(int dtype, const void* in, size_t n, double* hi, double* lo, double* abs_sum) {
  DD s;
  double a = 0.0;
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    if constexpr (std::is_floating_point_v<T>) {
      DD aa;
      for (size_t i = 0; i < n; ++i) {
        s.add(static_cast<double>(p[i]));
        aa.add(std::fabs(static_cast<double>(p[i])));
      }
      a = aa.value();
    } else {
      T acc{};
      for (size_t i = 0; i < n; ++i) acc = add(acc, p[i]);
      s.hi = static_cast<double>(acc);
      for (size_t i = 0; i < n; ++i) a += std::fabs(static_cast<double>(p[i]));
    }
  })
  double h, l;
  two_sum(s.hi, s.lo, h, l);
  if (hi) *hi = h;
  if (lo) *lo = l;
  if (abs_sum) *abs_sum = a;
}

void pxo_inclusive_scan(int dtype, const void* in, void* out, size_t n) {
  PXO_DISPATCH(dtype, incl_scan_t<T>(static_cast<const T*>(in), static_cast<T*>(out), n))
}

void pxo_exclusive_scan(int dtype, const void* in, void* out, size_t n, const void* init) {
  PXO_DISPATCH(dtype, {
    T iv{};
    if (init) std::memcpy(&iv, init, sizeof(T));
    excl_scan_t<T>(static_cast<const T*>(in), static_cast<T*>(out), n, iv);
  })
}

void pxo_scan_exact(int dtype, const void* in, size_t n, int exclusive, const void* init,
                    double* out_val, double* out_abs) {
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    T iv{};
    if (init) std::memcpy(&iv, init, sizeof(T));
    DD s, a;
    if (exclusive) {
      s.add(static_cast<double>(iv));
      a.add(std::fabs(static_cast<double>(iv)));
    }
    for (size_t i = 0; i < n; ++i) {
      const double x = static_cast<double>(p[i]);
      if (exclusive) {
        out_val[i] = s.value();
        if (out_abs) out_abs[i] = a.value();
        s.add(x);
        a.add(std::fabs(x));
      } else {
        s.add(x);
        a.add(std::fabs(x));
        out_val[i] = s.value();
        if (out_abs) out_abs[i] = a.value();
      }
    }
  })
}

void pxo_sort(int dtype, void* data, size_t n) {
  PXO_DISPATCH(dtype, { T* p = static_cast<T*>(data); std::sort(p, p + n); })
}

int pxo_pred_eval(const px_desc* d, const void* x) {
  int r = 0;
  PXO_DISPATCH(d->dtype, { T v; std::memcpy(&v, x, sizeof(T)); r = pred<T>(*d, v) ? 1 : 0; })
  return r;
}

size_t pxo_copy_if(const px_desc* d, const void* in, void* out, size_t n) {
  size_t k = 0;
  if (d->fn == PX_PRED_UNIQUE) {
    // adjacent-unique flags: keep i == 0 or in[i] != in[i-1] (std::unique keeps the first)
    PXO_DISPATCH(d->dtype, {
      const T* p = static_cast<const T*>(in);
      T* o = static_cast<T*>(out);
      for (size_t i = 0; i < n; ++i)
        if (i == 0 || !(p[i] == p[i - 1])) o[k++] = p[i];
    })
    return k;
  }
  PXO_DISPATCH(d->dtype, {
    const T* p = static_cast<const T*>(in);
    T* o = static_cast<T*>(out);
    for (size_t i = 0; i < n; ++i)
      if (pred<T>(*d, p[i])) o[k++] = p[i];
  })
  return k;
}

size_t pxo_remove_if(const px_desc* d, void* data, size_t n) {
  size_t k = 0;
  PXO_DISPATCH(d->dtype, {
    T* p = static_cast<T*>(data);
    for (size_t i = 0; i < n; ++i)
      if (!pred<T>(*d, p[i])) p[k++] = p[i];
  })
  return k;
}

size_t pxo_partition(const px_desc* d, void* data, size_t n) {
  size_t k = 0;
  PXO_DISPATCH(d->dtype, {
    T* p = static_cast<T*>(data);
    const px_desc dd = *d;
    T* mid = std::partition(p, p + n, [&dd](const T& x) { return pred<T>(dd, x); });
    k = static_cast<size_t>(mid - p);
  })
  return k;
}

size_t pxo_stable_partition(const px_desc* d, void* data, size_t n) {
  size_t k = 0;
  PXO_DISPATCH(d->dtype, {
    T* p = static_cast<T*>(data);
    const px_desc dd = *d;
    T* mid = std::stable_partition(p, p + n, [&dd](const T& x) { return pred<T>(dd, x); });
    k = static_cast<size_t>(mid - p);
  })
  return k;
}

size_t pxo_unique(int dtype, void* data, size_t n) {
  size_t k = 0;
  PXO_DISPATCH(dtype, {
    T* p = static_cast<T*>(data);
    k = static_cast<size_t>(std::unique(p, p + n) - p);
  })
  return k;
}

long pxo_count_if(const px_desc* d, const void* in, size_t n) {
  long c = 0;
  PXO_DISPATCH(d->dtype, {
    const T* p = static_cast<const T*>(in);
    for (size_t i = 0; i < n; ++i)
      if (pred<T>(*d, p[i])) ++c;
  })
  return c;
}

void pxo_map(const px_desc* d, void* data, size_t n, const void* captures, size_t capture_size) {
  px_desc dd = *d;
  dd.aux = d->dtype;  // in-place: output type is the element type
  PXO_DISPATCH(d->dtype, transform_tu<T, T>(dd, static_cast<const T*>(data), static_cast<T*>(data),
                                            n, captures, capture_size))
}

void pxo_transform(const px_desc* d, const void* in, void* out, size_t n, const void* captures,
                   size_t capture_size) {
  // output type d->aux; arithmetic in the output type after a static_cast
  switch (d->dtype) {
#define PXO_TIN(CODE, TIN)                                                              \
  case CODE: {                                                                          \
    const TIN* p = static_cast<const TIN*>(in);                                         \
    PXO_DISPATCH(d->aux, transform_tu<TIN, T>(*d, p, static_cast<T*>(out), n, captures, \
                                              capture_size))                            \
    break;                                                                              \
  }
    PXO_TIN(PX_I32, int32_t)
    PXO_TIN(PX_U32, uint32_t)
    PXO_TIN(PX_I64, int64_t)
    PXO_TIN(PX_U64, uint64_t)
    PXO_TIN(PX_F32, float)
    PXO_TIN(PX_F64, double)
#undef PXO_TIN
    default: break;
  }
}

void pxo_seg_reduce(int dtype, const void* in, const uint64_t* offsets, size_t nseg, void* out) {
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    T* o = static_cast<T*>(out);
    for (size_t s = 0; s < nseg; ++s)
      reduce_t<T>(p + offsets[s], offsets[s + 1] - offsets[s], o + s);
  })
}

void pxo_seg_sum_exact(int dtype, const void* in, const uint64_t* offsets, size_t nseg,
                       double* hi, double* abs_sum) {
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    for (size_t s = 0; s < nseg; ++s) {
      double h, l, a;
      pxo_sum_exact(dtype, p + offsets[s], offsets[s + 1] - offsets[s], &h, &l, &a);
      hi[s] = h;
      abs_sum[s] = a;
    }
  })
}

void pxo_seg_inclusive_scan(int dtype, const void* in, void* out, const uint64_t* offsets,
                            size_t nseg) {
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    T* o = static_cast<T*>(out);
    for (size_t s = 0; s < nseg; ++s)
      incl_scan_t<T>(p + offsets[s], o + offsets[s], offsets[s + 1] - offsets[s]);
  })
}

void pxo_seg_scan_exact(int dtype, const void* in, const uint64_t* offsets, size_t nseg,
                        double* out_val, double* out_abs) {
  const size_t es = (dtype == PX_I32 || dtype == PX_U32 || dtype == PX_F32) ? 4 : 8;
  const char* p = static_cast<const char*>(in);
  for (size_t s = 0; s < nseg; ++s)
    pxo_scan_exact(dtype, p + offsets[s] * es, offsets[s + 1] - offsets[s], 0, nullptr,
                   out_val + offsets[s], out_abs ? out_abs + offsets[s] : nullptr);
}

void pxo_seg_sort(int dtype, void* data, const uint64_t* offsets, size_t nseg) {
  PXO_DISPATCH(dtype, {
    T* p = static_cast<T*>(data);
    for (size_t s = 0; s < nseg; ++s) std::sort(p + offsets[s], p + offsets[s + 1]);
  })
}

size_t pxo_find(const px_desc* d, const void* in, size_t n) {
  PXO_DISPATCH(d->dtype, {
    const T* p = static_cast<const T*>(in);
    for (size_t i = 0; i < n; ++i)
      if (pred<T>(*d, p[i])) return i;
  })
  return n;
}

size_t pxo_mismatch(int dtype, const void* a, const void* b, size_t n) {
  PXO_DISPATCH(dtype, {
    const T* x = static_cast<const T*>(a);
    const T* y = static_cast<const T*>(b);
    for (size_t i = 0; i < n; ++i)
      if (!(x[i] == y[i])) return i;
  })
  return n;
}

size_t pxo_argminmax(int dtype, const void* in, size_t n, int want_max, int want_last) {
  if (n == 0) return 0;
  size_t best = 0;
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    for (size_t i = 1; i < n; ++i) {
      const bool strictly = want_max ? (p[best] < p[i]) : (p[i] < p[best]);
      const bool tie = !(p[i] < p[best]) && !(p[best] < p[i]);
      if (strictly || (tie && want_last)) best = i;
    }
  })
  return best;
}

size_t pxo_seg_unique(int dtype, const void* in, void* out, const uint64_t* offsets, size_t nseg,
                      uint64_t* out_offsets) {
  size_t total = 0;
  PXO_DISPATCH(dtype, {
    const T* p = static_cast<const T*>(in);
    T* o = static_cast<T*>(out);
    out_offsets[0] = 0;
    for (size_t s = 0; s < nseg; ++s) {
      const size_t len = offsets[s + 1] - offsets[s];
      std::memcpy(o + total, p + offsets[s], len * sizeof(T));
      total += static_cast<size_t>(std::unique(o + total, o + total + len) - (o + total));
      out_offsets[s + 1] = total;
    }
  })
  return total;
}

}  // extern "C"
