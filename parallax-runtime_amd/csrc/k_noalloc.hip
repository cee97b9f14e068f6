// k_noalloc.hip -- pxd_inclusive_scan_noalloc: the plain inclusive scan for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation).  Same kernel and tiling as pxd_inclusive_scan; its scratch -- the
// control block and the look-back status words -- is the file-scope device array below
// instead of a stream context's allocations.
//
// Each GPU holds its own copy of `scratch` until the process ends.  Host code never uses
// the array's name as an address: at a device's first call its address is taken with
// hipGetSymbolAddress (with that device current) and kept per device.  Setup clears the
// array with fill kernels (epoch 1, everything else 0); every scan leaves the control
// block clean and advances the epoch, and status words are valid only for their epoch,
// so later calls clear nothing until the 30-bit epoch is about to wrap.  All calls on one
// GPU share its array: calls that overlap on one GPU (other streams, other threads)
// race, and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.hpp"
#include "kernels.hpp"
#include "parallax/px_device.h"

namespace {

constexpr size_t kCtlBytes = (sizeof(pxk::Ctl) + 255) & ~size_t(255);
// 2^31 elements of any type at the default tile shapes: f64 tiles of 8,192 elements with
// 4 status words each need 262,144 tiles x 32 bytes = 8 MiB; one more tile for an
// unaligned start, plus margin.
constexpr uint64_t kMaxElems = 1ull << 31;
constexpr size_t kStatusBytes = (8u << 20) + (64u << 10);
constexpr size_t kScratchBytes = kCtlBytes + kStatusBytes;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
  unsigned char* ptr = nullptr;  // device address of this GPU's copy of `scratch`
  int cus = 0;
  uint64_t launches = 0;         // scans since the last clear (epoch wrap guard)
};
std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

int fail(int code, const char* what) {
  parallax::set_last_error(code, what);
  return code;
}

// Clears the control block and the status words on `stream` and sets the epoch to 1
// (fill kernels: no memset node if the caller captures the stream).
hipError_t clear(const DeviceScratch& d, hipStream_t stream) {
  hipError_t e = pxk::fill_u32(reinterpret_cast<uint32_t*>(d.ptr), kScratchBytes / 4, 0u, stream);
  if (e != hipSuccess) return e;
  return pxk::fill_u32(reinterpret_cast<uint32_t*>(d.ptr + offsetof(pxk::Ctl, epoch)), 1, 1u, stream);
}

}  // namespace

extern "C" int pxd_inclusive_scan_noalloc(int dtype, const void* d_in, void* d_out, size_t n,
                                          void* hip_stream) {
  if (dtype < 0 || dtype >= PX_NUM_DTYPES) return fail(PX_EARG, "scan_noalloc: bad dtype");
  if (n == 0) return PX_EOK;
  if (!d_in || !d_out) return fail(PX_EARG, "scan_noalloc: null pointer");
  if (n > kMaxElems)
    return fail(PX_EARG, "scan_noalloc: more than 2^31 elements (the no-alloc scratch limit)");
  const hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return fail(PX_ENODEV, "scan_noalloc: no current device");
  }
  const pxk::SegGeom g = parallax::scan_geometry(dtype, d_in, n);
  const size_t words = static_cast<size_t>(g.ntiles) * (pxk::carry_bytes(dtype) / 4);
  if (words * 8 > kStatusBytes)
    return fail(PX_EARG, "scan_noalloc: the tile shape needs more status words than the no-alloc scratch holds");
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceScratch& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(PX_EDEVICE, "scan_noalloc: hipGetSymbolAddress of the scratch array failed");
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
      (void)hipGetLastError();
      return fail(PX_EDEVICE, "scan_noalloc: compute-unit count unavailable");
    }
    DeviceScratch nd;
    nd.ptr = static_cast<unsigned char*>(p);
    nd.cus = cus;
    if (clear(nd, stream) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PX_EDEVICE, "scan_noalloc: scratch clear failed");
    }
    d = nd;
  }
  if (++d.launches >= pxk::kEpochMax - 2) {
    if (clear(d, stream) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PX_EDEVICE, "scan_noalloc: scratch clear failed");
    }
    d.launches = 1;
  }
  pxk::Scratch s;
  s.ctl = reinterpret_cast<pxk::Ctl*>(d.ptr);
  s.status = reinterpret_cast<uint64_t*>(d.ptr + kCtlBytes);
  s.status_words = kStatusBytes / 8;
  s.partials = nullptr;
  s.partial_bytes = 0;
  s.max_blocks = static_cast<uint32_t>(d.cus) * 8u;
  const hipError_t e = pxk::scan(dtype, d_in, d_out, g, false, nullptr, s, stream);
  if (e != hipSuccess) return fail(PX_EDEVICE, hipGetErrorString(e));
  ++parallax::stats().scan;
  ++parallax::stats().launches;
  return PX_EOK;
}
