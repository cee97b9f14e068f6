// parallax/stdpar.hpp -- C++17 parallel-algorithm front end for the MI355X engine.
//
// The algorithm surface (parallax::reduce, sort, inclusive_scan, ... with the
// parallax::par / par_unseq policies) and the detail::device_* funnel templates keep the
// reference's names, template parameters and parameter types (reference
// include/parallax/stdpar.hpp:46-394, 423-631): a funnel's __PRETTY_FUNCTION__ is the
// registry key a compiler plugin registers under, so the keys must be the same strings.
// The bodies are this engine's: every funnel hands the caller's pointer straight to the
// C ABI (the runtime keeps arena and device pointers in place and stages ordinary host
// memory itself), and runs the ISO loop when the lookup misses.
//
// Keys without a functor (device_reduce, device_sort, device_scan, device_exclusive_scan,
// device_unique) are resolved by the runtime's built-in table; funnels with a user
// functor need a descriptor registered under their key (parallax_kernel_register).
#ifndef PARALLAX_STDPAR_HPP
#define PARALLAX_STDPAR_HPP

#include <parallax/px_device.h>
#include <parallax/runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <execution>
#include <iterator>
#include <numeric>
#include <string>
#include <type_traits>

namespace parallax {

struct parallel_policy {};
struct parallel_unsequenced_policy {};
inline constexpr parallel_policy par{};
inline constexpr parallel_unsequenced_policy par_unseq{};

namespace detail {

// Looks up `base` + `suffix` once per funnel instantiation.
inline parallax_kernel_t lookup_suffixed(const char* base, const char* suffix) {
  const std::string key = std::string(base) + suffix;
  return parallax_kernel_lookup(key.c_str());
}

template <class T>
inline constexpr int is_fp = std::is_floating_point_v<T> ? 1 : 0;

// ---- map funnels
template <class T, class F>
__attribute__((noinline)) void device_invoke(T* data, std::size_t n, F f) {
  static parallax_kernel_t k = parallax_kernel_lookup(__PRETTY_FUNCTION__);
  if (k) {
    if constexpr (std::is_empty_v<F>)
      parallax_kernel_launch(k, static_cast<void*>(data), n, sizeof(T));
    else
      parallax_kernel_launch_with_captures(k, data, n, static_cast<void*>(&f), sizeof(F), sizeof(T));
    return;
  }
  for (std::size_t i = 0; i < n; ++i) f(data[i]);
}

template <class Tin, class Tout, class F>
__attribute__((noinline)) void device_transform(const Tin* in, Tout* out, std::size_t n, F f) {
  static parallax_kernel_t k = parallax_kernel_lookup(__PRETTY_FUNCTION__);
  if (k) {
    void* ip = const_cast<Tin*>(in);
    if constexpr (std::is_empty_v<F>)
      parallax_kernel_launch_transform2(k, ip, out, n, sizeof(Tin), sizeof(Tout));
    else
      parallax_kernel_launch_transform2_captures(k, ip, out, n, sizeof(Tin), sizeof(Tout),
                                                 static_cast<void*>(&f), sizeof(F));
    return;
  }
  for (std::size_t i = 0; i < n; ++i) out[i] = f(in[i]);
}

// ---- reduction: the device sums in the element type (identity 0), init is combined here
template <class T>
__attribute__((noinline)) T device_reduce(const T* data, std::size_t n, T init) {
  static parallax_kernel_t k = parallax_kernel_lookup(__PRETTY_FUNCTION__);
  if (k) {
    T sum{};
    parallax_reduce(k, const_cast<T*>(data), n, sizeof(T), &sum);
    return init + sum;
  }
  T acc = init;
  for (std::size_t i = 0; i < n; ++i) acc = acc + data[i];
  return acc;
}

// ---- sort: any n (positions past the next power of two are virtual +inf on the device)
template <class T>
__attribute__((noinline)) void device_sort(T* data, std::size_t n) {
  static parallax_kernel_t k = parallax_kernel_lookup(__PRETTY_FUNCTION__);
  if (k) {
    parallax_sort(k, data, n, sizeof(T));
    return;
  }
  std::sort(data, data + n);
}

// ---- scans
template <class T>
__attribute__((noinline)) void device_scan(const T* in, T* out, std::size_t n) {
  static parallax_kernel_t ks = lookup_suffixed(__PRETTY_FUNCTION__, ":scan");
  static parallax_kernel_t ka = lookup_suffixed(__PRETTY_FUNCTION__, ":add");
  if (ks && ka) {
    if (in != out) std::memcpy(out, in, n * sizeof(T));  // the ABI scans in place
    parallax_scan(ks, ka, out, n, sizeof(T));
    return;
  }
  T acc{};
  for (std::size_t i = 0; i < n; ++i) {
    acc = acc + in[i];
    out[i] = acc;
  }
}

template <class T>
__attribute__((noinline)) void device_exclusive_scan(const T* in, T* out, std::size_t n, T init) {
  static parallax_kernel_t ks = lookup_suffixed(__PRETTY_FUNCTION__, ":scan");
  static parallax_kernel_t ka = lookup_suffixed(__PRETTY_FUNCTION__, ":add");
  static parallax_kernel_t kh = lookup_suffixed(__PRETTY_FUNCTION__, ":shift");
  if (ks && ka && kh) {
    // this engine never writes the input, so no scratch copy is needed
    parallax_exclusive_scan(ks, ka, kh, const_cast<T*>(in), out, n, sizeof(T), &init);
    return;
  }
  T acc = init;
  for (std::size_t i = 0; i < n; ++i) {
    const T x = in[i];
    out[i] = acc;
    acc = acc + x;
  }
}

// ---- transform_reduce / count_if: the mapping descriptor (":xform" / ":pred") and the
// sum (":reduce") run as ONE fused device pass (parallax_transform_reduce); the reference
// wrote f(x) to an arena array and reduced it (stdpar.hpp:246-293 there).
template <class T, class U, class F>
__attribute__((noinline)) U device_transform_reduce(const T* in, std::size_t n, F f) {
  static parallax_kernel_t kx = lookup_suffixed(__PRETTY_FUNCTION__, ":xform");
  static parallax_kernel_t kr = lookup_suffixed(__PRETTY_FUNCTION__, ":reduce");
  if (kx && kr && std::is_empty_v<F>) {
    U sum{};
    parallax_transform_reduce(kx, kr, const_cast<T*>(in), n, sizeof(T), sizeof(U), &sum);
    return sum;
  }
  U acc{};
  for (std::size_t i = 0; i < n; ++i) acc = acc + static_cast<U>(f(in[i]));
  return acc;
}

template <class T, class Pred>
__attribute__((noinline)) long device_count_if(const T* in, std::size_t n, Pred pred) {
  static parallax_kernel_t kp = lookup_suffixed(__PRETTY_FUNCTION__, ":pred");
  static parallax_kernel_t kr = lookup_suffixed(__PRETTY_FUNCTION__, ":reduce");
  if (kp && kr && std::is_empty_v<Pred>) {
    int c = 0;  // the ":pred" / ":reduce" pair counts in int, as the reference registers it
    parallax_transform_reduce(kp, kr, const_cast<T*>(in), n, sizeof(T), sizeof(int), &c);
    return c;
  }
  long c = 0;
  for (std::size_t i = 0; i < n; ++i)
    if (pred(in[i])) ++c;
  return c;
}

// ---- compaction family: one fused device pass per call (":flags" carries the
// predicate, ":scatter" the mode; ":scan" / ":add" are validated for compatibility)
struct CompactKernels {
  parallax_kernel_t f, s, a, x;
  bool ok() const { return f && s && a && x; }
};
inline CompactKernels compact_kernels(const char* key) {
  return {lookup_suffixed(key, ":flags"), lookup_suffixed(key, ":scan"),
          lookup_suffixed(key, ":add"), lookup_suffixed(key, ":scatter")};
}

template <class T, class Pred>
__attribute__((noinline)) std::size_t device_copy_if(const T* in, T* out, std::size_t n, Pred pred) {
  static CompactKernels k = compact_kernels(__PRETTY_FUNCTION__);
  if (k.ok() && std::is_empty_v<Pred>)
    return parallax_copy_if(k.f, k.s, k.a, k.x, const_cast<T*>(in), out, n, sizeof(T), is_fp<T>);
  std::size_t m = 0;
  for (std::size_t i = 0; i < n; ++i)
    if (pred(in[i])) out[m++] = in[i];
  return m;
}

// remove_if / partition / unique compact in place: the runtime stages host memory itself
// and its compaction kernels are safe with input == output (every tile writes only below
// its own end, after every earlier tile has read its input), so no temporary is needed
// and no size limit applies.  partition comes back in std::partition's order, so the
// device and the host loop give the same bytes.
template <class T, class Pred>
__attribute__((noinline)) std::size_t device_remove_if(T* data, std::size_t n, Pred pred) {
  static CompactKernels k = compact_kernels(__PRETTY_FUNCTION__);
  if (k.ok() && std::is_empty_v<Pred>)
    return parallax_copy_if(k.f, k.s, k.a, k.x, data, data, n, sizeof(T), is_fp<T>);
  std::size_t m = 0;
  for (std::size_t i = 0; i < n; ++i)
    if (!pred(data[i])) data[m++] = data[i];
  return m;
}

template <class T, class Pred>
__attribute__((noinline)) std::size_t device_partition(T* data, std::size_t n, Pred pred) {
  static CompactKernels k = compact_kernels(__PRETTY_FUNCTION__);
  if (k.ok() && std::is_empty_v<Pred>)
    return parallax_copy_if(k.f, k.s, k.a, k.x, data, data, n, sizeof(T), is_fp<T>);
  return static_cast<std::size_t>(std::partition(data, data + n, pred) - data);
}

template <class T>
__attribute__((noinline)) std::size_t device_unique(T* data, std::size_t n) {
  static CompactKernels k = compact_kernels(__PRETTY_FUNCTION__);
  if (k.ok()) return parallax_copy_if(k.f, k.s, k.a, k.x, data, data, n, sizeof(T), is_fp<T>);
  return static_cast<std::size_t>(std::unique(data, data + n) - data);
}

// ---- offload gates: parallel policies on contiguous ranges
template <class P>
inline constexpr bool parallel_v =
    std::is_same_v<std::remove_cvref_t<P>, parallel_policy> ||
    std::is_same_v<std::remove_cvref_t<P>, parallel_unsequenced_policy>
#if defined(__cpp_lib_execution)
    || std::is_same_v<std::remove_cvref_t<P>, std::execution::parallel_policy> ||
    std::is_same_v<std::remove_cvref_t<P>, std::execution::parallel_unsequenced_policy>
#endif
    ;

template <class Policy, class... Its>
inline constexpr bool offload_v = parallel_v<Policy> && (parallax::is_contiguous_iterator_v<Its> && ...);

template <class It>
inline std::size_t count(It first, It last) {
  return static_cast<std::size_t>(std::distance(first, last));
}

}  // namespace detail

// ---------------------------------------------------------------- algorithm surface
template <class Policy, class It, class F>
void for_each(Policy&&, It first, It last, F f) {
  if constexpr (detail::offload_v<Policy, It>) {
    const std::size_t n = detail::count(first, last);
    if (n) detail::device_invoke(&*first, n, f);
  } else {
    std::for_each(first, last, f);
  }
}

template <class Policy, class InIt, class OutIt, class F>
OutIt transform(Policy&&, InIt first, InIt last, OutIt d_first, F f) {
  const std::size_t n = detail::count(first, last);
  if constexpr (detail::offload_v<Policy, InIt, OutIt>) {
    if (n) detail::device_transform(&*first, &*d_first, n, f);
    return std::next(d_first, static_cast<std::ptrdiff_t>(n));
  } else {
    return std::transform(first, last, d_first, f);
  }
}

template <class Policy, class It, class T>
T reduce(Policy&&, It first, It last, T init) {
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    if (n == 0) return init;
    return init + static_cast<T>(detail::device_reduce<E>(&*first, n, E{}));
  } else {
    return std::reduce(first, last, init);
  }
}

template <class Policy, class It, class V>
void fill(Policy&& p, It first, It last, const V& value) {
  using E = typename std::iterator_traits<It>::value_type;
  const E v = static_cast<E>(value);
  if constexpr (detail::offload_v<Policy, It>)
    parallax::for_each(std::forward<Policy>(p), first, last, [v](E& x) { x = v; });
  else
    std::fill(first, last, value);
}

template <class Policy, class It, class Gen>
void generate(Policy&& p, It first, It last, Gen gen) {
  using E = typename std::iterator_traits<It>::value_type;
  if constexpr (detail::offload_v<Policy, It>)
    parallax::for_each(std::forward<Policy>(p), first, last, [gen](E& x) { x = gen(); });
  else
    std::generate(first, last, gen);
}

template <class Policy, class It>
void sort(Policy&&, It first, It last) {
  if constexpr (detail::offload_v<Policy, It>) {
    const std::size_t n = detail::count(first, last);
    if (n) detail::device_sort(&*first, n);
  } else {
    std::sort(first, last);
  }
}

template <class Policy, class It, class OutIt>
OutIt inclusive_scan(Policy&&, It first, It last, OutIt d_first) {
  const std::size_t n = detail::count(first, last);
  if constexpr (detail::offload_v<Policy, It, OutIt>) {
    if (n) detail::device_scan(&*first, &*d_first, n);
    return std::next(d_first, static_cast<std::ptrdiff_t>(n));
  } else {
    return std::inclusive_scan(first, last, d_first);
  }
}

template <class Policy, class It, class OutIt, class T>
OutIt exclusive_scan(Policy&&, It first, It last, OutIt d_first, T init) {
  const std::size_t n = detail::count(first, last);
  if constexpr (detail::offload_v<Policy, It, OutIt>) {
    using E = typename std::iterator_traits<It>::value_type;
    if (n) detail::device_exclusive_scan(&*first, &*d_first, n, static_cast<E>(init));
    return std::next(d_first, static_cast<std::ptrdiff_t>(n));
  } else {
    return std::exclusive_scan(first, last, d_first, init);
  }
}

template <class Policy, class It, class T, class BinOp, class UnOp>
T transform_reduce(Policy&&, It first, It last, T init, BinOp bop, UnOp uop) {
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    using U = std::decay_t<decltype(uop(std::declval<E>()))>;
    const std::size_t n = detail::count(first, last);
    if (n == 0) return init;
    return init + static_cast<T>(detail::device_transform_reduce<E, U, UnOp>(&*first, n, uop));
  } else {
    return std::transform_reduce(first, last, init, bop, uop);
  }
}

template <class Policy, class It, class Pred>
typename std::iterator_traits<It>::difference_type count_if(Policy&&, It first, It last, Pred pred) {
  using D = typename std::iterator_traits<It>::difference_type;
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    return n ? static_cast<D>(detail::device_count_if<E, Pred>(&*first, n, pred)) : D{};
  } else {
    return std::count_if(first, last, pred);
  }
}

template <class Policy, class It, class Pred>
bool all_of(Policy&& p, It first, It last, Pred pred) {
  return parallax::count_if(std::forward<Policy>(p), first, last, pred) == std::distance(first, last);
}
template <class Policy, class It, class Pred>
bool any_of(Policy&& p, It first, It last, Pred pred) {
  return parallax::count_if(std::forward<Policy>(p), first, last, pred) > 0;
}
template <class Policy, class It, class Pred>
bool none_of(Policy&& p, It first, It last, Pred pred) {
  return parallax::count_if(std::forward<Policy>(p), first, last, pred) == 0;
}

template <class Policy, class It, class OutIt, class Pred>
OutIt copy_if(Policy&&, It first, It last, OutIt d_first, Pred pred) {
  if constexpr (detail::offload_v<Policy, It, OutIt>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    const std::size_t m = n ? detail::device_copy_if<E, Pred>(&*first, &*d_first, n, pred) : 0;
    return std::next(d_first, static_cast<std::ptrdiff_t>(m));
  } else {
    return std::copy_if(first, last, d_first, pred);
  }
}

template <class Policy, class It, class Pred>
It remove_if(Policy&&, It first, It last, Pred pred) {
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    const std::size_t m = n ? detail::device_remove_if<E, Pred>(&*first, n, pred) : 0;
    return std::next(first, static_cast<std::ptrdiff_t>(m));
  } else {
    return std::remove_if(first, last, pred);
  }
}

template <class Policy, class It, class Pred>
It partition(Policy&&, It first, It last, Pred pred) {
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    const std::size_t m = n ? detail::device_partition<E, Pred>(&*first, n, pred) : 0;
    return std::next(first, static_cast<std::ptrdiff_t>(m));
  } else {
    return std::partition(first, last, pred);
  }
}

template <class Policy, class It>
It unique(Policy&&, It first, It last) {
  if constexpr (detail::offload_v<Policy, It>) {
    using E = typename std::iterator_traits<It>::value_type;
    const std::size_t n = detail::count(first, last);
    const std::size_t m = n ? detail::device_unique<E>(&*first, n) : 0;
    return std::next(first, static_cast<std::ptrdiff_t>(m));
  } else {
    return std::unique(first, last);
  }
}

}  // namespace parallax

#endif  // PARALLAX_STDPAR_HPP
