// autotune.hpp -- tile-shape autotuner of the batched plans (see autotune.cpp).
#pragma once

#include <functional>

#include "engine.hpp"

namespace parallax {

struct TuneChoice {
  pxk::TileCfg tile;
  double ms;    // best measured duration of one call (0 when not measured)
  bool tuned;   // false: the default was kept (nothing could be timed)
};

// Picks the fastest tile shape for a batched `op` on `dtype` over segments `sp` by timing
// `run(tile)` (which must enqueue the whole call on c's stream) for every candidate.
TuneChoice autotune_tile(StreamContext& c, int op, int dtype, const SegSpec& sp,
                         const std::function<int(pxk::TileCfg)>& run);

}  // namespace parallax
