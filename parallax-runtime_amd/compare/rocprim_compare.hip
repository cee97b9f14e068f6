// rocprim_compare.hip -- out-of-library comparator, not part of the engine: the same
// primitives as the plain benchmark suite through rocPRIM (ROCm's own device library,
// header-only in /opt/rocm/include/rocprim), on inputs of the same sizes and
// distributions, timed the same way (hipEvents around each call, median of 10 after 3
// warm-ups, temp storage allocated beforehand).  Build: make -C parallax-runtime_amd
// rocprim-compare; run: parallax-runtime_amd/lib/rocprim_compare.
//
//   reduce          2^28 float32, sum
//   inclusive_scan  2^28 uint32 in [0, 16)
//   select          2^28 int32, x % 2 == 0 (copy_if)
//   radix_sort_keys 2^26 uint32 uniform, 2^25 float64 uniform in [0, 1)
#include <hip/hip_runtime.h>

#include <cstring>  // rocPRIM's iterator headers use memset

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#define CHECK(x)                                                                       \
  do {                                                                                 \
    const hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                            \
      std::printf("hip error %s at line %d\n", hipGetErrorString(e_), __LINE__);       \
      return 1;                                                                        \
    }                                                                                  \
  } while (0)

namespace {

struct IsEven {
  __device__ bool operator()(int x) const { return (x & 1) == 0; }
};

// Device time of f() (median of 10 after 3 warm-ups); prep() runs before each call,
// outside the timed region.
template <class F, class P>
float median_ms(F f, P prep) {
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  std::vector<float> ms;
  for (int i = 0; i < 13; ++i) {
    prep();
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e0, nullptr);
    f();
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (i >= 3) ms.push_back(t);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

int streams() {
  const size_t n = size_t(1) << 28;
  std::mt19937_64 g(1);
  std::vector<float> hf(n);
  std::vector<uint32_t> hu(n);
  std::vector<int> hi(n);
  std::normal_distribution<float> nd;
  for (size_t i = 0; i < n; ++i) {
    hf[i] = nd(g);
    hu[i] = static_cast<uint32_t>(g() >> 60);  // [0, 16)
    hi[i] = static_cast<int>(g());
  }
  float *f, *fr;
  uint32_t *u, *uo;
  int *ii, *io;
  size_t* cnt;
  CHECK(hipMalloc(&f, n * 4));
  CHECK(hipMalloc(&u, n * 4));
  CHECK(hipMalloc(&uo, n * 4));
  CHECK(hipMalloc(&ii, n * 4));
  CHECK(hipMalloc(&io, n * 4));
  CHECK(hipMalloc(&cnt, 64));
  CHECK(hipMalloc(&fr, 64));
  CHECK(hipMemcpy(f, hf.data(), n * 4, hipMemcpyHostToDevice));
  CHECK(hipMemcpy(u, hu.data(), n * 4, hipMemcpyHostToDevice));
  CHECK(hipMemcpy(ii, hi.data(), n * 4, hipMemcpyHostToDevice));
  size_t tr = 0, ts = 0, tc = 0;
  CHECK(rocprim::reduce(nullptr, tr, f, fr, 0.0f, n, rocprim::plus<float>()));
  CHECK(rocprim::inclusive_scan(nullptr, ts, u, uo, n, rocprim::plus<uint32_t>()));
  CHECK(rocprim::select(nullptr, tc, ii, io, cnt, n, IsEven()));
  void* tmp = nullptr;
  CHECK(hipMalloc(&tmp, std::max(tr, std::max(ts, tc))));
  auto none = [] {};
  const float r = median_ms([&] { (void)rocprim::reduce(tmp, tr, f, fr, 0.0f, n, rocprim::plus<float>()); }, none);
  std::printf("rocprim reduce f32 2^28: %.4f ms, %.0f GB/s\n", r, n * 4.0 / r / 1e6);
  const float s = median_ms([&] { (void)rocprim::inclusive_scan(tmp, ts, u, uo, n, rocprim::plus<uint32_t>()); }, none);
  std::printf("rocprim inclusive_scan u32 2^28: %.4f ms, %.0f GB/s\n", s, 2.0 * n * 4 / s / 1e6);
  const float c = median_ms([&] { (void)rocprim::select(tmp, tc, ii, io, cnt, n, IsEven()); }, none);
  size_t kept = 0;
  CHECK(hipMemcpy(&kept, cnt, sizeof(kept), hipMemcpyDeviceToHost));
  std::printf("rocprim select (copy_if x %% 2 == 0) i32 2^28: %.4f ms, %.0f GB/s (kept %zu)\n", c,
              (n * 4.0 + kept * 4.0) / c / 1e6, kept);
  for (void* p : {static_cast<void*>(f), static_cast<void*>(u), static_cast<void*>(uo), static_cast<void*>(ii),
                  static_cast<void*>(io), static_cast<void*>(cnt), static_cast<void*>(fr), tmp})
    CHECK(hipFree(p));
  return 0;
}

template <class T>
int sort(const char* name, std::vector<T>& h) {
  const size_t n = h.size();
  T *a, *b, *src;
  CHECK(hipMalloc(&a, n * sizeof(T)));
  CHECK(hipMalloc(&b, n * sizeof(T)));
  CHECK(hipMalloc(&src, n * sizeof(T)));
  CHECK(hipMemcpy(src, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  size_t tb = 0;
  CHECK(rocprim::radix_sort_keys(nullptr, tb, a, b, n));
  void* tmp = nullptr;
  CHECK(hipMalloc(&tmp, tb));
  const float t = median_ms([&] { (void)rocprim::radix_sort_keys(tmp, tb, a, b, n, 0, 8 * sizeof(T)); },
                            [&] { (void)hipMemcpy(a, src, n * sizeof(T), hipMemcpyDeviceToDevice); });
  std::vector<T> out(n);
  CHECK(hipMemcpy(out.data(), b, n * sizeof(T), hipMemcpyDeviceToHost));
  std::sort(h.begin(), h.end());
  std::printf("rocprim radix_sort_keys %s: %.4f ms (sorted: %s)\n", name, t, out == h ? "yes" : "NO");
  for (void* p : {static_cast<void*>(a), static_cast<void*>(b), static_cast<void*>(src), tmp}) CHECK(hipFree(p));
  return 0;
}

}  // namespace

int main() {
  if (streams()) return 1;
  std::mt19937_64 g(2);
  std::vector<uint32_t> u(size_t(1) << 26);
  for (auto& x : u) x = static_cast<uint32_t>(g());
  std::vector<double> d(size_t(1) << 25);
  for (auto& x : d) x = static_cast<double>(g() >> 11) * (1.0 / 9007199254740992.0);
  if (sort("u32 2^26", u)) return 1;
  if (sort("f64 2^25", d)) return 1;
  return 0;
}
