// api_device.cpp -- the device-resident (pxd_*), batched (pxb_*) and runtime (pxr_*)
// entry points of parallax/px_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "autotune.hpp"
#include "engine.hpp"
#include "parallax/px_device.h"

using parallax::StreamContext;

struct pxd_ctx {
  std::unique_ptr<StreamContext> c;
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

struct pxb_plan {
  std::unique_ptr<StreamContext> c;  // the plan owns its scratch and tables
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  pxk::TileCfg tile{0, 0};  // tile shape of the captured call (0: fixed-shape op)
  double tuned_ms = 0.0;
};

namespace {

int nodev() {
  parallax::set_last_error(PX_ENODEV, "no GPU device (or PARALLAX_FORCE_CPU is set)");
  return PX_ENODEV;
}

int hip_rc(hipError_t e, const char* what) {
  if (e == hipSuccess) return PX_EOK;
  parallax::set_last_error(PX_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return PX_EDEVICE;
}

parallax::SegSpec contiguous(const uint64_t* off, size_t nseg) {
  parallax::SegSpec sp;
  sp.lo = off;
  sp.hi = off + 1;
  sp.nseg = static_cast<uint32_t>(nseg);
  return sp;
}

bool ok_ctx(pxd_ctx_t ctx) { return ctx && ctx->c && ctx->c->ok(); }

}  // namespace

extern "C" {

const char* pxr_version(void) { return "parallax-mi355x 0.1.0 (gfx950)"; }

int pxr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int pxr_device_info(char* name, size_t name_len, int* compute_units, size_t* hbm_bytes,
                    size_t* lds_per_block) {
  parallax::HipBackend* be = parallax::global_backend();
  if (!be) return nodev();
  const auto& caps = be->capabilities();
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", caps.name.c_str(), caps.arch.c_str());
  }
  if (compute_units) *compute_units = caps.compute_units;
  if (hbm_bytes) *hbm_bytes = caps.hbm_bytes;
  if (lds_per_block) *lds_per_block = caps.lds_per_block;
  return PX_EOK;
}

int pxr_last_error(char* buf, size_t len) {
  if (buf && len) std::snprintf(buf, len, "%s", parallax::last_error_text());
  return parallax::last_error_code();
}

int pxr_stats(uint64_t* out7) {
  if (!out7) return PX_EARG;
  const auto& s = parallax::stats();
  out7[0] = s.launches;
  out7[1] = s.reduce;
  out7[2] = s.scan;
  out7[3] = s.sort;
  out7[4] = s.compact;
  out7[5] = s.transform;
  out7[6] = s.seg;
  return PX_EOK;
}

void* pxd_malloc(size_t bytes) {
  if (!parallax::global_backend()) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

void pxd_free(void* p) {
  if (p) (void)hipFree(p);
}

// The copy has completed when these return (hipMemcpy alone may return before a
// device-side or staged copy lands, and work on non-blocking streams is not ordered after
// the null stream).
static int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, const char* what) {
  if (!parallax::global_backend()) return nodev();
  const int rc = hip_rc(hipMemcpy(dst, src, bytes, kind), what);
  if (rc != PX_EOK) return rc;
  return hip_rc(hipStreamSynchronize(nullptr), what);
}
int pxd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return copy_sync(dst, src, bytes, hipMemcpyHostToDevice, "memcpy h2d");
}
int pxd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost, "memcpy d2h");
}
int pxd_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  // the streaming-copy kernel on the null stream, waited for
  parallax::HipBackend* be = parallax::global_backend();
  if (!be) return nodev();
  const uint32_t mb = static_cast<uint32_t>(std::max(1, be->capabilities().compute_units)) * 8u;
  const int rc = hip_rc(pxk::copy_bytes(dst, src, bytes, mb, nullptr), "memcpy d2d");
  if (rc != PX_EOK) return rc;
  return hip_rc(hipStreamSynchronize(nullptr), "memcpy d2d");
}
int pxd_memset(void* dst, int value, size_t bytes) {
  if (!parallax::global_backend()) return nodev();
  const int rc = hip_rc(hipMemset(dst, value, bytes), "memset");
  if (rc != PX_EOK) return rc;
  return hip_rc(hipStreamSynchronize(nullptr), "memset");
}
int pxd_device_sync(void) {
  if (!parallax::global_backend()) return nodev();
  return hip_rc(hipDeviceSynchronize(), "device sync");
}

pxd_ctx_t pxd_ctx_create(void* hip_stream) {
  parallax::HipBackend* be = parallax::global_backend();
  if (!be) {
    nodev();
    return nullptr;
  }
  auto* ctx = new pxd_ctx;
  ctx->c = std::make_unique<StreamContext>(be, static_cast<hipStream_t>(hip_stream));
  if (!ctx->c->ok()) {
    parallax::set_last_error(PX_ENOMEM, "context creation failed");
    delete ctx;
    return nullptr;
  }
  (void)hipEventCreate(&ctx->t0);
  (void)hipEventCreate(&ctx->t1);
  return ctx;
}

void pxd_ctx_destroy(pxd_ctx_t ctx) {
  if (!ctx) return;
  if (ctx->t0) (void)hipEventDestroy(ctx->t0);
  if (ctx->t1) (void)hipEventDestroy(ctx->t1);
  delete ctx;
}

void* pxd_ctx_stream(pxd_ctx_t ctx) { return ok_ctx(ctx) ? ctx->c->stream() : nullptr; }

int pxd_sync(pxd_ctx_t ctx) {
  if (!ok_ctx(ctx)) return PX_EARG;
  return hip_rc(parallax::wait_stream(ctx->c->stream()), "stream sync");
}

int pxd_timer_begin(pxd_ctx_t ctx) {
  if (!ok_ctx(ctx)) return PX_EARG;
  return hip_rc(hipEventRecord(ctx->t0, ctx->c->stream()), "event record");
}

int pxd_timer_end(pxd_ctx_t ctx, double* ms) {
  if (!ok_ctx(ctx) || !ms) return PX_EARG;
  int rc = hip_rc(hipEventRecord(ctx->t1, ctx->c->stream()), "event record");
  if (rc) return rc;
  rc = hip_rc(hipEventSynchronize(ctx->t1), "event sync");
  if (rc) return rc;
  float f = 0.f;
  rc = hip_rc(hipEventElapsedTime(&f, ctx->t0, ctx->t1), "elapsed");
  *ms = f;
  return rc;
}

#define PXD_GUARD()                                  \
  if (!ok_ctx(ctx)) return PX_EARG;                  \
  std::lock_guard<std::mutex> lk(ctx->c->mu)

int pxd_reduce(pxd_ctx_t ctx, int dtype, const void* d_in, size_t n, void* d_out) {
  PXD_GUARD();
  return parallax::run_reduce(*ctx->c, dtype, d_in, n, d_out);
}

int pxd_copy(pxd_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes) {
  PXD_GUARD();
  if (bytes && (!d_dst || !d_src)) return PX_EARG;
  return hip_rc(pxk::copy_bytes(d_dst, d_src, bytes, static_cast<uint32_t>(ctx->c->cu_count()) * 8u,
                                ctx->c->stream()),
                "copy");
}

int pxd_transform_reduce(pxd_ctx_t ctx, const px_desc* map, const void* d_in, size_t n,
                         void* d_out) {
  PXD_GUARD();
  if (!map) return PX_EARG;
  px_desc m = *map;
  if (m.op == PX_OP_MAP) m.aux = m.dtype;
  return parallax::run_map_reduce(*ctx->c, m, d_in, n, d_out);
}

int pxd_inclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out, size_t n) {
  PXD_GUARD();
  return parallax::run_scan(*ctx->c, dtype, d_in, d_out, n, false, nullptr);
}

int pxd_exclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out, size_t n,
                       const void* h_init) {
  PXD_GUARD();
  return parallax::run_scan(*ctx->c, dtype, d_in, d_out, n, true, h_init);
}

int pxd_sort(pxd_ctx_t ctx, int dtype, void* d_data, size_t n) {
  PXD_GUARD();
  return parallax::run_sort(*ctx->c, dtype, d_data, n);
}

int pxd_copy_if(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out, size_t n,
                uint64_t* d_count) {
  PXD_GUARD();
  if (!flags) return PX_EARG;
  return parallax::run_compact(*ctx->c, *flags, d_in, d_out, n, d_count);
}

int pxd_partition(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out, size_t n,
                  uint64_t* d_num_true) {
  PXD_GUARD();
  if (!flags || flags->fn == PX_PRED_UNIQUE) return PX_EARG;
  return parallax::run_partition(*ctx->c, *flags, d_in, d_out, n, d_num_true);
}

int pxd_transform(pxd_ctx_t ctx, const px_desc* map, const void* d_in, void* d_out, size_t n,
                  const void* captures, size_t capture_size) {
  PXD_GUARD();
  if (!map) return PX_EARG;
  return parallax::run_transform(*ctx->c, *map, d_in, d_out, n, captures, capture_size, 0, 0, 0);
}

int pxd_find(pxd_ctx_t ctx, const px_desc* pred, const void* d_in, size_t n, uint64_t* d_index) {
  PXD_GUARD();
  if (!pred) return PX_EARG;
  return parallax::run_find(*ctx->c, *pred, d_in, n, d_index);
}

int pxd_argminmax(pxd_ctx_t ctx, int dtype, const void* d_in, size_t n, int want_max,
                  int want_last, uint64_t* d_index) {
  PXD_GUARD();
  return parallax::run_argminmax(*ctx->c, dtype, d_in, n, want_max != 0, want_last != 0, d_index);
}

int pxd_mismatch(pxd_ctx_t ctx, int dtype, const void* d_a, const void* d_b, size_t n,
                 uint64_t* d_index) {
  PXD_GUARD();
  return parallax::run_mismatch(*ctx->c, dtype, d_a, d_b, n, d_index);
}

int pxd_relocate_pointers(pxd_ctx_t ctx, uint64_t* d_ptrs, size_t n, uint64_t host_base,
                          uint64_t dev_base, uint64_t bytes) {
  PXD_GUARD();
  return hip_rc(pxk::relocate_pointers(d_ptrs, n, host_base, dev_base, bytes, ctx->c->stream()),
                "relocate");
}

int pxb_seg_reduce(pxd_ctx_t ctx, int dtype, const void* d_in, const uint64_t* h_offsets,
                   size_t nseg, void* d_out) {
  PXD_GUARD();
  if (!h_offsets) return PX_EARG;
  return parallax::run_seg_reduce(*ctx->c, dtype, d_in, contiguous(h_offsets, nseg), d_out);
}

int pxb_seg_inclusive_scan(pxd_ctx_t ctx, int dtype, const void* d_in, void* d_out,
                           const uint64_t* h_offsets, size_t nseg) {
  PXD_GUARD();
  if (!h_offsets) return PX_EARG;
  return parallax::run_seg_scan(*ctx->c, dtype, d_in, d_out, contiguous(h_offsets, nseg));
}

int pxb_seg_sort(pxd_ctx_t ctx, int dtype, void* d_data, const uint64_t* h_offsets, size_t nseg) {
  PXD_GUARD();
  if (!h_offsets) return PX_EARG;
  return parallax::run_seg_sort(*ctx->c, dtype, d_data, contiguous(h_offsets, nseg));
}

int pxb_seg_copy_if(pxd_ctx_t ctx, const px_desc* flags, const void* d_in, void* d_out,
                    const uint64_t* h_offsets, size_t nseg, uint64_t* d_counts) {
  PXD_GUARD();
  if (!flags || !h_offsets) return PX_EARG;
  return parallax::run_seg_compact(*ctx->c, *flags, d_in, d_out, contiguous(h_offsets, nseg),
                                   d_counts);
}

// ---- graph plans
namespace {
int plan_call(StreamContext& c, int op, int dtype, const px_desc* flags, const void* d_in,
              void* d_out, const parallax::SegSpec& sp, uint64_t* d_counts, pxk::TileCfg tile) {
  switch (op) {
    case PX_OP_REDUCE: return parallax::run_seg_reduce(c, dtype, d_in, sp, d_out, tile);
    case PX_OP_SCAN: return parallax::run_seg_scan(c, dtype, d_in, d_out, sp, tile);
    case PX_OP_SORT: return parallax::run_seg_sort(c, dtype, d_out, sp);
    case PX_OP_FLAGS:
      if (!flags) return PX_EARG;
      return parallax::run_seg_compact(c, *flags, d_in, d_out, sp, d_counts, tile);
    default: return PX_EUNSUPPORTED;
  }
}

bool autotune_enabled() {
  const char* e = std::getenv("PARALLAX_AUTOTUNE");
  return !(e && e[0] == '0');
}
}  // namespace

pxb_plan_t pxb_plan_create(pxd_ctx_t ctx, int op, int dtype, const px_desc* flags,
                           const void* d_in, void* d_out, const uint64_t* h_offsets, size_t nseg,
                           uint64_t* d_counts) {
  if (!ok_ctx(ctx) || !h_offsets) return nullptr;
  auto* p = new pxb_plan;
  p->c = std::make_unique<StreamContext>(ctx->c->backend(), ctx->c->stream());
  if (!p->c->ok()) {
    delete p;
    return nullptr;
  }
  const parallax::SegSpec sp = contiguous(h_offsets, nseg);
  int rc = PX_EOK;
  // An in-place scan / compaction runs its tuning and warm-up calls into a scratch
  // output, and a sort warms up on a scratch copy of its keys (at the same offset within
  // a 16-byte line), so the caller's data is intact when the plan is returned.
  const int dt = (op == PX_OP_FLAGS && flags) ? static_cast<int>(flags->dtype) : dtype;
  void* warm_out = d_out;
  void* tmp = nullptr;
  const bool warm_copy = op == PX_OP_SORT || ((op == PX_OP_SCAN || op == PX_OP_FLAGS) && d_in == d_out);
  if (warm_copy && nseg > 0 && dt >= 0 && dt < PX_NUM_DTYPES && d_out) {
    const size_t data_bytes = static_cast<size_t>(h_offsets[nseg]) * pxk::dtype_size(dt);
    const size_t mis = reinterpret_cast<uintptr_t>(d_out) & 15u;
    if (hipMalloc(&tmp, data_bytes + 16) != hipSuccess) {
      (void)hipGetLastError();
      parallax::set_last_error(PX_ENOMEM, "plan scratch output");
      pxb_plan_destroy(p);
      return nullptr;
    }
    warm_out = op == PX_OP_SORT ? static_cast<char*>(tmp) + mis : tmp;
    if (op == PX_OP_SORT && data_bytes &&
        hipMemcpyAsync(warm_out, d_out, data_bytes, hipMemcpyDeviceToDevice, p->c->stream()) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(tmp);
      parallax::set_last_error(PX_EDEVICE, "plan warm-up copy");
      pxb_plan_destroy(p);
      return nullptr;
    }
  }
  // 0) tile shape: timed on this batch (scan / compaction / reduce)
  if (op == PX_OP_REDUCE || op == PX_OP_SCAN || op == PX_OP_FLAGS) {
    p->tile = parallax::seg_tile_default(op, dt, sp);
    if (autotune_enabled() && nseg > 0 && dt >= 0 && dt < PX_NUM_DTYPES &&
        pxk::tile_override(op, dt).block == 0) {
      const parallax::TuneChoice ch = parallax::autotune_tile(
          *p->c, op, dt, sp, [&](pxk::TileCfg t) {
            return plan_call(*p->c, op, dtype, flags, d_in, warm_out, sp, d_counts, t);
          });
      p->tile = ch.tile;
      p->tuned_ms = ch.ms;
    }
  }
  // 1) warm-up run: uploads the tables and sizes the scratch
  rc = plan_call(*p->c, op, dtype, flags, d_in, warm_out, sp, d_counts, p->tile);
  if (rc == PX_EOK) rc = hip_rc(hipStreamSynchronize(p->c->stream()), "plan warm-up");
  if (tmp) (void)hipFree(tmp);
  // 2) capture the same call with uploads frozen
  if (rc == PX_EOK) {
    hipStream_t s = p->c->stream();
    rc = hip_rc(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc == PX_EOK) {
      p->c->freeze(true);
      const int crc = plan_call(*p->c, op, dtype, flags, d_in, d_out, sp, d_counts, p->tile);
      p->c->freeze(false);
      const int erc = hip_rc(hipStreamEndCapture(s, &p->graph), "end capture");
      rc = crc ? crc : erc;
    }
    if (rc == PX_EOK)
      rc = hip_rc(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0), "instantiate");
  }
  if (rc != PX_EOK) {
    pxb_plan_destroy(p);
    return nullptr;
  }
  return p;
}

int pxr_tile_override(int op, int dtype, int block, int items) {
  if (!pxk::set_tile_override(op, dtype, pxk::TileCfg{block, items})) {
    parallax::set_last_error(PX_EARG, "pxr_tile_override: unsupported op, dtype or tile shape");
    return PX_EARG;
  }
  return PX_EOK;
}

int pxb_plan_tile(pxb_plan_t plan, int* block, int* items, double* tuned_ms) {
  if (!plan) return PX_EARG;
  if (block) *block = plan->tile.block;
  if (items) *items = plan->tile.k;
  if (tuned_ms) *tuned_ms = plan->tuned_ms;
  return PX_EOK;
}

int pxb_plan_launch(pxb_plan_t plan) {
  if (!plan || !plan->exec) return PX_EARG;
  return hip_rc(hipGraphLaunch(plan->exec, plan->c->stream()), "graph launch");
}

void pxb_plan_destroy(pxb_plan_t plan) {
  if (!plan) return;
  if (plan->c) (void)hipStreamSynchronize(plan->c->stream());
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  delete plan;
}

}  // extern "C"
