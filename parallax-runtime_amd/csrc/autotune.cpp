// autotune.cpp -- tile-shape autotuner of the batched (segmented) plans.
//
// A plan is built once and replayed many times, so its tile shape is picked by
// measurement: every candidate of seg_tile_candidates() runs once to warm up, then
// three times between HIP events on the plan's stream; the fastest (minimum of its
// three) wins.  Choices are cached per (op, dtype, log2 segment count, log2 element
// count) for the life of the process, and kept across processes in the text file named
// by PARALLAX_TUNE_CACHE when that is set (one "op dtype nseg_log2 n_log2 block k ms"
// line per choice).
#include "autotune.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace parallax {

namespace {

using Key = std::tuple<int, int, int, int>;

uint32_t log2_bucket(uint64_t v) {
  uint32_t l = 0;
  while (v > 1) {
    v >>= 1;
    ++l;
  }
  return l;
}

struct Cache {
  std::mutex mu;
  std::map<Key, TuneChoice> map;
  bool loaded = false;
};

Cache& cache() {
  static Cache c;
  return c;
}

const char* cache_file() {
  const char* f = std::getenv("PARALLAX_TUNE_CACHE");
  return (f && *f) ? f : nullptr;
}

void load_file(Cache& c) {
  c.loaded = true;
  const char* f = cache_file();
  if (!f) return;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return;
  int op, dt, a, b, blk, k;
  double ms;
  while (std::fscanf(fp, "%d %d %d %d %d %d %lf", &op, &dt, &a, &b, &blk, &k, &ms) == 7) {
    const pxk::TileCfg t{blk, k};
    const bool ok = op == PX_OP_REDUCE ? (blk == 256 && (k == 4 || k == 8 || k == 16))
                                       : pxk::tile_cfg_supported(t, op);
    if (ok) c.map[Key{op, dt, a, b}] = TuneChoice{t, ms, true};
  }
  std::fclose(fp);
}

void append_file(const Key& k, const TuneChoice& t) {
  const char* f = cache_file();
  if (!f) return;
  FILE* fp = std::fopen(f, "a");
  if (!fp) return;
  std::fprintf(fp, "%d %d %d %d %d %d %.6f\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
               std::get<3>(k), t.tile.block, t.tile.k, t.ms);
  std::fclose(fp);
}

}  // namespace

TuneChoice autotune_tile(StreamContext& c, int op, int dtype, const SegSpec& sp,
                         const std::function<int(pxk::TileCfg)>& run) {
  uint64_t total = 0;
  for (uint32_t s = 0; s < sp.nseg; ++s) total += sp.hi[s] - sp.lo[s];
  const Key key{op, dtype, static_cast<int>(log2_bucket(sp.nseg)), static_cast<int>(log2_bucket(total))};
  Cache& ch = cache();
  {
    std::lock_guard<std::mutex> lk(ch.mu);
    if (!ch.loaded) load_file(ch);
    auto it = ch.map.find(key);
    if (it != ch.map.end()) return it->second;
  }
  TuneChoice best{seg_tile_default(op, dtype, sp), 0.0, false};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return best;
  }
  double best_ms = 1e30;
  for (const pxk::TileCfg& t : seg_tile_candidates(op, dtype)) {
    if (run(t) != PX_OK) continue;  // warm-up (tables, scratch)
    double m = 1e30;
    bool ok = true;
    for (int r = 0; r < 3 && ok; ++r) {
      ok = hipEventRecord(e0, c.stream()) == hipSuccess && run(t) == PX_OK &&
           hipEventRecord(e1, c.stream()) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.f;
      if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < m) m = ms;
    }
    if (ok && m < best_ms) {
      best_ms = m;
      best = TuneChoice{t, m, true};
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best.tuned) {
    std::lock_guard<std::mutex> lk(ch.mu);
    ch.map[key] = best;
    append_file(key, best);
  }
  return best;
}

}  // namespace parallax
