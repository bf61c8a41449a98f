// CPU emulation of the part of HIP that the HyperBall library uses, for tests
// only: it lets the library's own kernel and engine sources run on the host
// under AddressSanitizer, so an out-of-bounds access or a divergent
// cross-lane operation shows up on the CPU with a stack trace.
//
// Execution model.  A launch runs its blocks one after another; the threads
// of a block are OS threads (one per lane).  Cross-lane operations (__shfl*,
// the lane-group ballot and barrier) exchange values through per-segment
// slots guarded by barriers among the lanes of the segment (width G or 64),
// __syncthreads is a barrier over the block.  A barrier that does not fill
// within a minute (a lane that skipped a cross-lane operation its segment
// performs) aborts with a message.  Streams are synchronous; graph capture
// records launches and copies and replays them in order.
//
// Not emulated: lockstep execution outside cross-lane operations (kernels
// must order lane-to-lane LDS traffic with a lane-group barrier, as the
// library does), concurrency between blocks, and device memory ordering.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
// device arrays are host arrays here: a symbol's address is its own
#define HIP_SYMBOL(x) (&(x))
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
// callers pass wavefront-uniform values
#define __builtin_amdgcn_readfirstlane(x) (x)
// register allocation hints mean nothing here
#define amdgpu_waves_per_eu(...) unused

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};

struct float2 {
  float x, y;
};
struct uint2 {
  uint32_t x, y;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
struct uint4 {
  uint32_t x, y, z, w;
};

namespace emu {
struct Idx {
  uint32_t x = 0, y = 0, z = 0;
};
}  // namespace emu
extern thread_local emu::Idx threadIdx, blockIdx, blockDim, gridDim;

// ---- runtime types -------------------------------------------------------
enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNoDevice = 100,
  hipErrorNotSupported = 801,
};
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
};
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
};
#define hipStreamNonBlocking 1u
#define hipHostMallocDefault 0u
#define hipEventDisableTiming 2u

struct EmuStream;
struct EmuEvent;
struct EmuGraph;
typedef EmuStream *hipStream_t;
typedef EmuEvent *hipEvent_t;
typedef EmuGraph *hipGraph_t;
typedef EmuGraph *hipGraphExec_t;
typedef struct {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  size_t totalGlobalMem;
} hipDeviceProp_t;

hipError_t hipMalloc(void **p, size_t bytes);
inline hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t attr, int d);
// Two workgroups per CU for every kernel.
template <class K>
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, K, int, size_t) {
  *n = 2;
  return hipSuccess;
}
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t *e);
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
// streams run each operation when it is issued, so an event has always
// completed by the time a stream could wait for it
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);

// ---- device side -----------------------------------------------------------
namespace emu {
// Value of `v` at lane src(lane-in-segment) of the caller's width-W segment.
uint64_t exchange(uint64_t v, int width, const std::function<int(int)> &src);
// Bits of pred over the caller's width-W segment, at their wave lane positions.
uint64_t segment_ballot(bool pred, int width);
void segment_barrier(int width);
void block_barrier();

template <int G>
inline uint64_t ballot(bool pred) {
  return segment_ballot(pred, G);
}
template <int G>
inline void group_barrier() {
  segment_barrier(G);
}
template <class T>
inline uint64_t to_bits(T v) {
  static_assert(sizeof(T) <= 8, "exchange of at most 8 bytes");
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T from_bits(uint64_t b) {
  T v;
  std::memcpy(&v, &b, sizeof(T));
  return v;
}

// Launch: blocks run in order on the thread pool, or the launch is recorded
// when the stream is capturing.
void run_grid(dim3 grid, dim3 block, const std::function<void()> &body);
bool record(hipStream_t s, std::function<void()> op);

inline dim3 to_dim3(dim3 d) { return d; }
template <class I, class = std::enable_if_t<std::is_integral<I>::value>>
inline dim3 to_dim3(I n) {
  return dim3(uint32_t(n));
}

template <class... P, class G, class B, class... A>
void launch(void (*kernel)(P...), G grid, B block, size_t, hipStream_t s, A &&...args) {
  std::tuple<std::decay_t<P>...> t(std::forward<A>(args)...);
  const dim3 g = to_dim3(grid), b = to_dim3(block);
  auto op = [kernel, t, g, b]() { run_grid(g, b, [&]() { std::apply(kernel, t); }); };
  if (!record(s, op)) op();
}
}  // namespace emu

inline uint32_t __lane_id() { return threadIdx.x & 63u; }
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline void __syncthreads() { emu::block_barrier(); }

template <class T>
inline T __shfl(T v, int src, int width = 64) {
  return emu::from_bits<T>(emu::exchange(emu::to_bits(v), width, [&](int) { return src % width; }));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  return emu::from_bits<T>(
      emu::exchange(emu::to_bits(v), width, [&](int l) { return (l ^ mask) % width; }));
}
template <class T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
  return emu::from_bits<T>(emu::exchange(emu::to_bits(v), width, [&](int l) {
    return l >= int(delta) ? l - int(delta) : l;
  }));
}

template <class T, class U>
inline T atomicAdd(T *p, U v) {
  return __atomic_fetch_add(p, T(v), __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicOr(T *p, U v) {
  return __atomic_fetch_or(p, T(v), __ATOMIC_SEQ_CST);
}
