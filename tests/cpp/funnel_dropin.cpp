// funnel_dropin.cpp -- a user program written against parallax/stdpar.hpp, as the
// reference's funnels are used (std::vector data, parallax::par policy).  Without a GPU
// (or with PARALLAX_FORCE_CPU=1) every lookup misses and the ISO loops run; on the GPU
// the fixed-function funnels resolve through the runtime's built-in table and copy_if
// through a descriptor registered under its funnel key.  Prints one PASS/FAIL line per
// check and the device launch counter; exit status 0 iff every check passed.
#include <parallax/descriptor.h>
#include <parallax/px_device.h>
#include <parallax/stdpar.hpp>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

struct IsEven {
  bool operator()(int x) const { return x % 2 == 0; }
};
struct Sq {
  long operator()(int x) const { return static_cast<long>(x) * x; }
};

static int failures = 0;
static void check(bool ok, const char* what) {
  std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++failures;
}

// Large in-place unique through the built-in device_unique key: a sorted 2^28 uint64
// vector (2 GiB, twice the default arena), checked against std::unique on a copy.
static int big_unique() {
  const size_t n = size_t(1) << 28;
  std::vector<uint64_t> v(n);
  std::mt19937_64 rng(11);
  uint64_t x = 0;
  for (size_t i = 0; i < n; ++i) {
    x += (rng() % 3 == 0) ? 1 + (rng() & 7) : 0;
    v[i] = x;
  }
  std::vector<uint64_t> ref = v;
  uint64_t st0[7] = {0}, st1[7] = {0};
  pxr_stats(st0);
  auto e = parallax::unique(parallax::par, v.begin(), v.end());
  pxr_stats(st1);
  auto e_ref = std::unique(ref.begin(), ref.end());
  const size_t k = static_cast<size_t>(e - v.begin());
  const size_t k_ref = static_cast<size_t>(e_ref - ref.begin());
  check(k == k_ref && std::memcmp(v.data(), ref.data(), k * sizeof(uint64_t)) == 0,
        "unique 2^28 uint64 in place");
  std::printf("kept %zu of %zu\n", k, n);
  std::printf("device_launches %llu\n", static_cast<unsigned long long>(st1[0] - st0[0]));
  return failures == 0 ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "big") == 0) return big_unique();
  // the descriptor a compiler registrar would emit for device_copy_if<int, IsEven>
  static px_desc flags{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_FLAGS, PX_I32, 0, PX_PRED_MOD_EQ, 0, 0, 2, 0};
  static px_desc scan{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_SCAN, PX_I32, 0, 0, 0, 0, 0, 0};
  static px_desc add{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_SCAN_ADD, PX_I32, 0, 0, 0, 0, 0, 0};
  static px_desc scatter{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_SCATTER, PX_I32, PX_SCATTER_COMPACT, 0, 0, 0, 0, 0};
  const char* key =
      "std::size_t parallax::detail::device_copy_if(const T*, T*, std::size_t, Pred) "
      "[with T = int; Pred = IsEven; std::size_t = long unsigned int]";
  const std::string k(key);
  parallax_kernel_register((k + ":flags").c_str(), reinterpret_cast<unsigned*>(&flags), PX_DESC_WORDS);
  parallax_kernel_register((k + ":scan").c_str(), reinterpret_cast<unsigned*>(&scan), PX_DESC_WORDS);
  parallax_kernel_register((k + ":add").c_str(), reinterpret_cast<unsigned*>(&add), PX_DESC_WORDS);
  parallax_kernel_register((k + ":scatter").c_str(), reinterpret_cast<unsigned*>(&scatter), PX_DESC_WORDS);

  // partition and remove_if with registered functors, in place on the caller's vector
  static px_desc part{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_SCATTER, PX_I32, PX_SCATTER_PARTITION, 0, 0, 0, 0, 0};
  static px_desc not_even{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_FLAGS, PX_I32, 0, PX_PRED_MOD_EQ, PX_FLAG_NEGATE, 0, 2, 0};
  const std::string kp(
      "std::size_t parallax::detail::device_partition(T*, std::size_t, Pred) "
      "[with T = int; Pred = IsEven; std::size_t = long unsigned int]");
  const std::string kr(
      "std::size_t parallax::detail::device_remove_if(T*, std::size_t, Pred) "
      "[with T = int; Pred = IsEven; std::size_t = long unsigned int]");
  parallax_kernel_register((kp + ":flags").c_str(), reinterpret_cast<unsigned*>(&flags), PX_DESC_WORDS);
  parallax_kernel_register((kp + ":scan").c_str(), reinterpret_cast<unsigned*>(&scan), PX_DESC_WORDS);
  parallax_kernel_register((kp + ":add").c_str(), reinterpret_cast<unsigned*>(&add), PX_DESC_WORDS);
  parallax_kernel_register((kp + ":scatter").c_str(), reinterpret_cast<unsigned*>(&part), PX_DESC_WORDS);
  parallax_kernel_register((kr + ":flags").c_str(), reinterpret_cast<unsigned*>(&not_even), PX_DESC_WORDS);
  parallax_kernel_register((kr + ":scan").c_str(), reinterpret_cast<unsigned*>(&scan), PX_DESC_WORDS);
  parallax_kernel_register((kr + ":add").c_str(), reinterpret_cast<unsigned*>(&add), PX_DESC_WORDS);
  parallax_kernel_register((kr + ":scatter").c_str(), reinterpret_cast<unsigned*>(&scatter), PX_DESC_WORDS);

  // count_if and transform_reduce with registered functors: one fused device pass each
  static px_desc pred{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_TRANSFORM, PX_I32, PX_I32, PX_MAP_PRED, 0, PX_PRED_MOD_EQ, 2, 0};
  static px_desc sum_i32{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_REDUCE, PX_I32, 0, 0, 0, 0, 0, 0};
  static px_desc sq{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_TRANSFORM, PX_I32, PX_I64, PX_MAP_SQUARE, 0, 0, 0, 0};
  static px_desc sum_i64{PX_DESC_MAGIC, PX_DESC_VERSION, PX_OP_REDUCE, PX_I64, 0, 0, 0, 0, 0, 0};
  const std::string kc(
      "long int parallax::detail::device_count_if(const T*, std::size_t, Pred) "
      "[with T = int; Pred = IsEven; std::size_t = long unsigned int]");
  const std::string kt(
      "U parallax::detail::device_transform_reduce(const T*, std::size_t, F) "
      "[with T = int; U = long int; F = Sq; std::size_t = long unsigned int]");
  parallax_kernel_register((kc + ":pred").c_str(), reinterpret_cast<unsigned*>(&pred), PX_DESC_WORDS);
  parallax_kernel_register((kc + ":reduce").c_str(), reinterpret_cast<unsigned*>(&sum_i32), PX_DESC_WORDS);
  parallax_kernel_register((kt + ":xform").c_str(), reinterpret_cast<unsigned*>(&sq), PX_DESC_WORDS);
  parallax_kernel_register((kt + ":reduce").c_str(), reinterpret_cast<unsigned*>(&sum_i64), PX_DESC_WORDS);

  std::mt19937_64 rng(7);
  const size_t n = 1000003;
  std::vector<int> vi(n);
  for (auto& x : vi) x = static_cast<int>(rng() % 2001) - 1000;
  std::vector<double> vd(n);
  for (auto& x : vd) x = static_cast<double>(rng() % 100000) / 7.0;

  // reduce
  long long want = 0;
  for (int x : vi) want += x;
  check(parallax::reduce(parallax::par, vi.begin(), vi.end(), 0) == static_cast<int>(want), "reduce int");
  const double sd = parallax::reduce(parallax::par, vd.begin(), vd.end(), 0.0);
  double exact = 0;
  for (double x : vd) exact += x;
  check(std::fabs(sd - exact) <= 1e-9 * std::fabs(exact), "reduce double");
  // scans
  std::vector<int> inc(n), exc(n);
  parallax::inclusive_scan(parallax::par, vi.begin(), vi.end(), inc.begin());
  std::vector<int> inc_ref(n);
  std::inclusive_scan(vi.begin(), vi.end(), inc_ref.begin());
  check(inc == inc_ref, "inclusive_scan");
  parallax::exclusive_scan(parallax::par, vi.begin(), vi.end(), exc.begin(), 5);
  std::vector<int> exc_ref(n);
  std::exclusive_scan(vi.begin(), vi.end(), exc_ref.begin(), 5);
  check(exc == exc_ref, "exclusive_scan");
  // sort (non power of two) then unique
  std::vector<int> s = vi, s_ref = vi;
  parallax::sort(parallax::par, s.begin(), s.end());
  std::sort(s_ref.begin(), s_ref.end());
  check(s == s_ref, "sort");
  auto e = parallax::unique(parallax::par, s.begin(), s.end());
  auto e_ref = std::unique(s_ref.begin(), s_ref.end());
  check((e - s.begin()) == (e_ref - s_ref.begin()) && std::equal(s.begin(), e, s_ref.begin()), "unique");
  // copy_if with a registered functor
  std::vector<int> out(n), out_ref;
  auto oe = parallax::copy_if(parallax::par, vi.begin(), vi.end(), out.begin(), IsEven{});
  std::copy_if(vi.begin(), vi.end(), std::back_inserter(out_ref), IsEven{});
  check(static_cast<size_t>(oe - out.begin()) == out_ref.size() &&
            std::equal(out_ref.begin(), out_ref.end(), out.begin()), "copy_if IsEven");
  // partition: the same bytes as the host loop's std::partition (GPU or not)
  std::vector<int> pv = vi, pv_ref = vi;
  auto pm = parallax::partition(parallax::par, pv.begin(), pv.end(), IsEven{});
  auto pm_ref = std::partition(pv_ref.begin(), pv_ref.end(), IsEven{});
  check((pm - pv.begin()) == (pm_ref - pv_ref.begin()) && pv == pv_ref, "partition IsEven (std order)");
  // remove_if in place
  std::vector<int> rv = vi, rv_ref = vi;
  auto rm = parallax::remove_if(parallax::par, rv.begin(), rv.end(), IsEven{});
  auto rm_ref = std::remove_if(rv_ref.begin(), rv_ref.end(), IsEven{});
  check((rm - rv.begin()) == (rm_ref - rv_ref.begin()) && std::equal(rv.begin(), rm, rv_ref.begin()),
        "remove_if IsEven");
  // registered count_if / transform_reduce (fused map + sum on the device)
  const long ce = parallax::count_if(parallax::par, vi.begin(), vi.end(), IsEven{});
  check(ce == std::count_if(vi.begin(), vi.end(), IsEven{}), "count_if IsEven");
  const long sq2 = parallax::transform_reduce(parallax::par, vi.begin(), vi.end(), 5L, std::plus<>{}, Sq{});
  long sq_ref = 5;
  for (int x : vi) sq_ref += static_cast<long>(x) * x;
  check(sq2 == sq_ref, "transform_reduce Sq");
  // a capturing lambda has no descriptor: it stays on the CPU and is still correct
  const int t = 17;
  const long c = parallax::count_if(parallax::par, vi.begin(), vi.end(), [t](int x) { return x > t; });
  check(c == std::count_if(vi.begin(), vi.end(), [t](int x) { return x > t; }), "count_if (host loop)");

  uint64_t st[7] = {0};
  pxr_stats(st);
  std::printf("device_launches %llu\n", static_cast<unsigned long long>(st[0]));
  return failures == 0 ? 0 : 1;
}
