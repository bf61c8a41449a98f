// Transpose of a CSR graph built on the device: the GPU counterpart of
// hb_graph_transpose (host counting sort), producing the same arrays.  It
// replaces the host construction of the transpose that the systolic mode
// needs (hyperball_impl.rs:795-796 reads it; utils/mod.rs:84-118 defines it:
// (u, v) in G  <=>  (v, u) in G^T).
//
//   1. in-degrees: one atomic increment per arc;
//   2. transpose offsets: exclusive scan of the in-degrees (rocPRIM);
//   3. the source node of every arc, expanded from the offsets;
//   4. a stable radix sort of (successor, source) pairs by successor
//      (rocPRIM).  Arcs enter in CSR order, i.e. by increasing source, so
//      stability leaves every transpose row sorted by source.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hb_kernels.hpp"

namespace hb {
namespace {

// In-degrees in 64 bits: with 2^32 arcs or more one node may have 2^32
// predecessors.
__global__ __launch_bounds__(kBlock) void indegree_kernel(const uint32_t *succ, uint64_t a,
                                                          unsigned long long *cnt) {
  for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < a;
       k += uint64_t(gridDim.x) * blockDim.x)
    atomicAdd(cnt + succ[k], 1ull);
}

// Eight lanes per node write the node id over its row.
__global__ __launch_bounds__(kBlock) void sources_kernel(const uint64_t *off, uint64_t n,
                                                         uint32_t *src) {
  constexpr uint32_t L = 8;
  const uint64_t groups = uint64_t(gridDim.x) * (blockDim.x / L);
  const uint32_t gl = threadIdx.x % L;
  for (uint64_t u = uint64_t(blockIdx.x) * (blockDim.x / L) + threadIdx.x / L; u < n;
       u += groups) {
    const uint64_t e = off[u + 1];
    for (uint64_t k = off[u] + gl; k < e; k += L) src[k] = uint32_t(u);
  }
}

uint32_t grid_of(uint64_t items, uint32_t per_block) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return uint32_t(g);
}

struct DeviceScratch {
  void *p = nullptr;
  ~DeviceScratch() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
};

}  // namespace

hipError_t launch_transpose(const uint64_t *off, const uint32_t *succ, uint64_t n, uint64_t a,
                            uint64_t *toff, uint32_t *tsucc, hipStream_t s) {
  hipError_t e;
  DeviceScratch cnt, src, keys, tmp;
  if ((e = cnt.alloc((n + 1) * sizeof(uint64_t))) != hipSuccess) return e;
  unsigned long long *d_cnt = static_cast<unsigned long long *>(cnt.p);
  if ((e = launch_fill_u32(static_cast<uint32_t *>(cnt.p), 2 * (n + 1), 0u, s)) != hipSuccess)
    return e;
  if (a) {
    indegree_kernel<<<grid_of(a, kBlock), kBlock, 0, s>>>(succ, a, d_cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  size_t scan_bytes = 0;
  if ((e = rocprim::exclusive_scan(nullptr, scan_bytes, d_cnt, toff, uint64_t(0), size_t(n + 1),
                                   rocprim::plus<uint64_t>(), s)) != hipSuccess)
    return e;
  size_t sort_bytes = 0;
  int bits = 1;
  while (bits < 32 && (uint64_t(1) << bits) < n) bits++;
  if (a) {
    if ((e = rocprim::radix_sort_pairs(nullptr, sort_bytes, succ, static_cast<uint32_t *>(nullptr),
                                       static_cast<const uint32_t *>(nullptr), tsucc, size_t(a), 0,
                                       bits, s)) != hipSuccess)
      return e;
  }
  if ((e = tmp.alloc(std::max(scan_bytes, sort_bytes))) != hipSuccess) return e;
  if ((e = rocprim::exclusive_scan(tmp.p, scan_bytes, d_cnt, toff, uint64_t(0), size_t(n + 1),
                                   rocprim::plus<uint64_t>(), s)) != hipSuccess)
    return e;
  if (!a) return hipSuccess;
  if ((e = src.alloc(a * sizeof(uint32_t))) != hipSuccess) return e;
  if ((e = keys.alloc(a * sizeof(uint32_t))) != hipSuccess) return e;
  sources_kernel<<<grid_of(n, kBlock / 8), kBlock, 0, s>>>(off, n, static_cast<uint32_t *>(src.p));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = rocprim::radix_sort_pairs(tmp.p, sort_bytes, succ, static_cast<uint32_t *>(keys.p),
                                     static_cast<const uint32_t *>(src.p), tsucc, size_t(a), 0,
                                     bits, s)) != hipSuccess)
    return e;
  // the scratch buffers are freed on return: finish the work that reads them
  return hipStreamSynchronize(s);
}

}  // namespace hb
