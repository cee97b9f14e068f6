// k_copy.hip -- device-to-device streaming copy for gfx950: the read-and-write ceiling
// every look-back kernel (scan, compaction, radix pass) is measured against, and the
// copy behind pxd_copy / pxd_memcpy_d2d.
//
// Grid-stride over 16-byte vectors, U non-temporal loads in flight per thread before the
// matching non-temporal stores (streamed data is touched once: it should not displace
// the L2 / MALL lines other work uses).  A head and a tail of fewer than 16 bytes (when
// source and destination share their offset within a 16-byte line) are copied byte by
// byte by block 0; other misalignments copy 4-byte words or bytes.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace pxk {

namespace {

template <int U>
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                uint64_t nvec, const unsigned char* __restrict__ hsrc,
                                                unsigned char* __restrict__ hdst, uint32_t head,
                                                const unsigned char* __restrict__ tsrc,
                                                unsigned char* __restrict__ tdst, uint32_t tail) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  for (; i + (U - 1) * stride < nvec; i += U * stride) {
    u32x4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < nvec; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
  if (blockIdx.x == 0) {
    if (threadIdx.x < head) hdst[threadIdx.x] = hsrc[threadIdx.x];
    if (threadIdx.x < tail) tdst[threadIdx.x] = tsrc[threadIdx.x];
  }
}

template <class W>
__global__ __launch_bounds__(256) void k_copy_words(const W* __restrict__ src, W* __restrict__ dst, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * 256;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

hipError_t copy_bytes(void* dst, const void* src, uint64_t bytes, uint32_t max_blocks, hipStream_t stream) {
  if (bytes == 0 || dst == src) return hipSuccess;
  const uintptr_t s = reinterpret_cast<uintptr_t>(src), d = reinterpret_cast<uintptr_t>(dst);
  // one 256-thread block per CU with 4 loads in flight per thread streams best (1 GiB,
  // measured: 5.95 TB/s; 2 / 4 / 8 blocks per CU 5.66 / 4.94 / 4.66 at 4 loads, 1 block
  // with 1 / 2 / 8 loads 3.35 / 5.37 / 5.88 TB/s)
  const TileCfg cfg{tune().copy_bpc > 0 ? tune().copy_bpc : 1, tune().copy_unroll > 0 ? tune().copy_unroll : 4};
  const uint32_t cus = max_blocks / 8 ? max_blocks / 8 : 1;
  if ((s & 15u) == (d & 15u)) {
    const uint32_t head = static_cast<uint32_t>(std::min<uint64_t>((16u - (s & 15u)) & 15u, bytes));
    const uint64_t nvec = (bytes - head) / 16;
    const uint32_t tail = static_cast<uint32_t>(bytes - head - nvec * 16);
    const auto* vs = reinterpret_cast<const u32x4*>(s + head);
    auto* vd = reinterpret_cast<u32x4*>(d + head);
    const auto* ts = reinterpret_cast<const unsigned char*>(s + head + nvec * 16);
    auto* td = reinterpret_cast<unsigned char*>(d + head + nvec * 16);
    const uint64_t want = (nvec + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, static_cast<uint64_t>(cus) * cfg.block)));
    const auto* hs = reinterpret_cast<const unsigned char*>(src);
    auto* hd = reinterpret_cast<unsigned char*>(dst);
    switch (cfg.k) {
      case 1: k_copy16<1><<<grid, 256, 0, stream>>>(vs, vd, nvec, hs, hd, head, ts, td, tail); break;
      case 2: k_copy16<2><<<grid, 256, 0, stream>>>(vs, vd, nvec, hs, hd, head, ts, td, tail); break;
      case 8: k_copy16<8><<<grid, 256, 0, stream>>>(vs, vd, nvec, hs, hd, head, ts, td, tail); break;
      default: k_copy16<4><<<grid, 256, 0, stream>>>(vs, vd, nvec, hs, hd, head, ts, td, tail); break;
    }
    return hipGetLastError();
  }
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((bytes + 1023) / 1024, static_cast<uint64_t>(cus) * 8)));
  if (((s | d | bytes) & 3u) == 0)
    k_copy_words<uint32_t><<<grid, 256, 0, stream>>>(reinterpret_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), bytes / 4);
  else
    k_copy_words<unsigned char><<<grid, 256, 0, stream>>>(static_cast<const unsigned char*>(src), static_cast<unsigned char*>(dst), bytes);
  return hipGetLastError();
}

}  // namespace pxk
