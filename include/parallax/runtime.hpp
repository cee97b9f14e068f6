// parallax/runtime.hpp -- process-wide runtime objects (reference runtime.hpp:11-16).
// All are created lazily on first use and are nullptr when no GPU is present (or
// PARALLAX_FORCE_CPU is set), in which case every funnel takes its CPU loop.
#pragma once

#include "parallax/arena.hpp"
#include "parallax/hip_backend.hpp"
#include "parallax/kernel_launcher.hpp"

namespace parallax {

HipBackend* get_global_backend();
UnifiedArena* get_global_arena();
KernelLauncher* get_global_launcher();

}  // namespace parallax
