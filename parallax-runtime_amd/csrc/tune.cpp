// tune.cpp -- launch-shape configuration (defaults, PARALLAX_TUNE overrides).
#include <cstdlib>
#include <cstring>
#include <string>

#include "kernels.hpp"

namespace pxk {

bool tile_cfg_supported(TileCfg c, int op) {
  return (c.block == 256 && (c.k == 4 || c.k == 8)) || (c.block == 512 && c.k == 8) ||
         (c.block == 1024 && (c.k == 4 || c.k == 8 || (op == PX_OP_SCAN && c.k == 16)));
}

namespace {
TileCfg g_override[3][PX_NUM_DTYPES] = {};
int override_slot(int op) {
  return op == PX_OP_REDUCE ? 0 : op == PX_OP_SCAN ? 1 : op == PX_OP_FLAGS ? 2 : -1;
}
}  // namespace

TileCfg tile_override(int op, int dtype) {
  const int i = override_slot(op);
  if (i < 0 || dtype < 0 || dtype >= PX_NUM_DTYPES) return TileCfg{0, 0};
  return g_override[i][dtype];
}

bool set_tile_override(int op, int dtype, TileCfg c) {
  const int i = override_slot(op);
  if (i < 0 || dtype < 0 || dtype >= PX_NUM_DTYPES) return false;
  const bool clear = c.block == 0 && c.k == 0;
  const bool ok = clear || (op == PX_OP_REDUCE ? (c.block == 256 && (c.k == 4 || c.k == 8 || c.k == 16))
                                               : tile_cfg_supported(c, op));
  if (!ok) return false;
  g_override[i][dtype] = c;
  return true;
}

uint32_t lookback_grid(uint32_t ntiles, uint32_t max_blocks, int block) {
  const uint32_t cus = max_blocks / 8;
  const int bpc = tune().bpc > 0 ? tune().bpc : (1024 / block > 0 ? 1024 / block : 1);
  const uint32_t cap = cus * static_cast<uint32_t>(bpc);
  return ntiles < cap ? ntiles : cap;
}

namespace {
Tune defaults() {
  Tune t;
  for (int d = 0; d < PX_NUM_DTYPES; ++d) {
    // 1024 x 8 for every type: with the in-wave arithmetic of f64 in double (only the
    // carries double-double) its scan fits 128 VGPRs there too (2^27 f64: 0.44 ms on
    // 1024 x 8, 0.94 ms on 1024 x 4).
    t.scan[d] = TileCfg{1024, 8};
    t.compact[d] = TileCfg{1024, 8};
  }
  return t;
}

Tune load() {
  Tune t = defaults();
  const char* env = std::getenv("PARALLAX_TUNE");
  if (!env) return t;
  std::string s(env);
  size_t pos = 0;
  TileCfg scan{0, 0}, compact{0, 0};
  while (pos < s.size()) {
    size_t comma = s.find(',', pos);
    if (comma == std::string::npos) comma = s.size();
    const std::string kv = s.substr(pos, comma - pos);
    pos = comma + 1;
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) continue;
    const std::string k = kv.substr(0, eq);
    const int v = std::atoi(kv.c_str() + eq + 1);
    if (k == "scan_block") scan.block = v;
    else if (k == "scan_k") scan.k = v;
    else if (k == "compact_block") compact.block = v;
    else if (k == "compact_k") compact.k = v;
    else if (k == "bpc" || k == "scan_bpc") t.bpc = v;
    else if (k == "red_unroll") t.red_unroll = v;
    else if (k == "red_bpc") t.red_bpc = v;
    else if (k == "red_nt") t.red_nt = v;
    else if (k == "sort_c4" && v >= 1 && v <= 7) t.sort_c4 = v;
    else if (k == "sort_c8" && v >= 1 && v <= 6) t.sort_c8 = v;
    else if (k == "sort_radix_min" && v >= 0) t.sort_radix_min = static_cast<uint64_t>(v);
    else if (k == "radix_block" && (v == 256 || v == 512 || v == 1024)) t.radix_block = v;
    else if (k == "radix_hist_bpc" && v >= 1 && v <= 16) t.radix_hist_bpc = v;
    else if (k == "radix_hist_block" && (v == 256 || v == 512 || v == 1024)) t.radix_hist_block = v;
    else if (k == "seg_hist_bpc" && v >= 1 && v <= 16) t.seg_hist_bpc = v;
    else if (k == "seg_radix_min" && v >= 1) t.seg_radix_min = static_cast<uint64_t>(v);
    else if (k == "seg_radix_block" && (v == 512 || v == 1024)) t.seg_radix_block = v;
    else if (k == "copy_bpc" && v >= 1 && v <= 32) t.copy_bpc = v;
    else if (k == "copy_unroll" && (v == 1 || v == 2 || v == 4 || v == 8)) t.copy_unroll = v;
  }
  for (int d = 0; d < PX_NUM_DTYPES; ++d) {
    TileCfg a{scan.block ? scan.block : t.scan[d].block, scan.k ? scan.k : t.scan[d].k};
    if (tile_cfg_supported(a, PX_OP_SCAN)) t.scan[d] = a;
    TileCfg b{compact.block ? compact.block : t.compact[d].block, compact.k ? compact.k : t.compact[d].k};
    if (tile_cfg_supported(b, PX_OP_FLAGS)) t.compact[d] = b;
  }
  return t;
}
}  // namespace

Tune& tune() {
  static Tune t = load();
  return t;
}

}  // namespace pxk
