// tune.cpp -- launch-shape configuration (defaults, PARALLAX_TUNE overrides).
#include <cstdlib>
#include <cstring>
#include <string>

#include "kernels.hpp"

namespace pxk {

bool tile_cfg_supported(TileCfg c) {
  return (c.block == 256 || c.block == 512) && (c.k == 4 || c.k == 8);
}

namespace {
Tune load() {
  Tune t;
  const char* env = std::getenv("PARALLAX_TUNE");
  if (!env) return t;
  std::string s(env);
  size_t pos = 0;
  while (pos < s.size()) {
    size_t comma = s.find(',', pos);
    if (comma == std::string::npos) comma = s.size();
    const std::string kv = s.substr(pos, comma - pos);
    pos = comma + 1;
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) continue;
    const std::string k = kv.substr(0, eq);
    const int v = std::atoi(kv.c_str() + eq + 1);
    TileCfg scan = t.scan;
    if (k == "scan_block") scan.block = v;
    else if (k == "scan_k") scan.k = v;
    else if (k == "scan_bpc") t.scan_bpc = v;
    else if (k == "red_unroll") t.red_unroll = v;
    else if (k == "red_bpc") t.red_bpc = v;
    else if (k == "red_nt") t.red_nt = v;
    if (tile_cfg_supported(scan)) t.scan = scan;
  }
  return t;
}
}  // namespace

Tune& tune() {
  static Tune t = load();
  return t;
}

}  // namespace pxk
