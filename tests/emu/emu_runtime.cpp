// Host implementation of the HIP emulation declared in hip/hip_runtime.h
// (test infrastructure: see that header for the execution model).
#include "hip/hip_runtime.h"

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <semaphore>
#include <thread>

thread_local emu::Idx threadIdx, blockIdx, blockDim, gridDim;

namespace emu {
namespace {

constexpr int kMaxThreads = 1024;

std::atomic<uint64_t> g_phases{0};       // completed barrier phases (progress)
std::atomic<int> g_running{0};            // launches in progress

// Barrier among n lanes: arrive, spin briefly, then sleep on the generation
// word.  A watchdog aborts the process when no barrier phase completes for a
// minute while a launch runs (a lane skipped a cross-lane operation).
struct Barrier {
  int n = 1;
  std::atomic<int> count{0};
  std::atomic<uint32_t> gen{0};
  void wait() {
    if (n == 1) return;
    const uint32_t g = gen.load(std::memory_order_acquire);
    if (count.fetch_add(1, std::memory_order_acq_rel) == n - 1) {
      count.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
      g_phases.fetch_add(1, std::memory_order_relaxed);
      gen.notify_all();
      return;
    }
    for (int i = 0; i < 200; i++) {
      if (gen.load(std::memory_order_acquire) != g) return;
      std::this_thread::yield();
    }
    while (gen.load(std::memory_order_acquire) == g) gen.wait(g, std::memory_order_acquire);
  }
};

void watchdog() {
  uint64_t last = 0;
  int idle = 0;
  for (;;) {
    std::this_thread::sleep_for(std::chrono::seconds(5));
    const uint64_t now = g_phases.load();
    if (g_running.load() && now == last) {
      if (++idle >= 12) {
        std::fprintf(stderr, "emu: no barrier completed for 60 s: a lane skipped a cross-lane "
                             "operation (or __syncthreads) that the rest of its segment performs\n");
        std::abort();
      }
    } else {
      idle = 0;
    }
    last = now;
  }
}

// Exchange state of one wave at one width: a barrier per segment and a slot
// per lane.
struct WidthState {
  std::vector<std::unique_ptr<Barrier>> bars;
  uint64_t slots[64];
};

struct BlockCtx {
  int nthreads = 0;
  Barrier block;
  std::vector<std::array<WidthState, 7>> waves;
  explicit BlockCtx(int n) : nthreads(n), waves((n + 63) / 64) {
    block.n = n;
    for (size_t w = 0; w < waves.size(); w++) {
      const int lanes = std::min(64, n - int(w) * 64);
      for (int k = 0; k < 7; k++) {
        const int width = 1 << k;
        auto &ws = waves[w][k];
        for (int s = 0; s < 64 / width; s++) {
          auto b = std::make_unique<Barrier>();
          b->n = std::max(1, std::min(width, lanes - s * width));
          ws.bars.push_back(std::move(b));
        }
      }
    }
  }
};

BlockCtx *g_block = nullptr;

int log2i(int w) {
  int k = 0;
  while ((1 << k) < w) k++;
  if ((1 << k) != w || k > 6) {
    std::fprintf(stderr, "emu: unsupported cross-lane width %d\n", w);
    std::abort();
  }
  return k;
}

// Worker pool: thread i runs lane i of the current block; only the lanes of
// the block are woken.
struct Pool {
  std::mutex m;
  std::condition_variable done;
  int active = 0, finished = 0;
  const std::function<void()> *body = nullptr;
  dim3 grid, block, bidx;
  std::vector<std::unique_ptr<std::binary_semaphore>> go;
  std::vector<std::thread> threads;
  bool stop = false;

  Pool() {
    for (int i = 0; i < kMaxThreads; i++) go.push_back(std::make_unique<std::binary_semaphore>(0));
    for (int i = 0; i < kMaxThreads; i++) threads.emplace_back([this, i] { loop(i); });
  }
  ~Pool() {
    stop = true;
    for (auto &g : go) g->release();
    for (auto &t : threads) t.join();
  }
  void loop(int i) {
    for (;;) {
      go[i]->acquire();
      if (stop) return;
      threadIdx = Idx{uint32_t(i), 0, 0};
      blockIdx = Idx{bidx.x, bidx.y, bidx.z};
      blockDim = Idx{block.x, block.y, block.z};
      gridDim = Idx{grid.x, grid.y, grid.z};
      (*body)();
      std::lock_guard<std::mutex> l(m);
      if (++finished == active) done.notify_one();
    }
  }
  void run_block(dim3 g, dim3 b, dim3 idx, const std::function<void()> &f) {
    grid = g;
    block = b;
    bidx = idx;
    body = &f;
    active = int(b.x);
    finished = 0;
    for (int i = 0; i < active; i++) go[i]->release();
    std::unique_lock<std::mutex> l(m);
    done.wait(l, [&] { return finished == active; });
  }
};

Pool &pool() {
  static Pool p;
  return p;
}

std::mutex g_launch;  // one grid at a time

}  // namespace

uint64_t exchange(uint64_t v, int width, const std::function<int(int)> &src) {
  BlockCtx &b = *g_block;
  const int t = int(threadIdx.x), lane = t & 63, seg = lane / width;
  WidthState &ws = b.waves[t / 64][log2i(width)];
  ws.slots[lane] = v;
  ws.bars[seg]->wait();
  const uint64_t r = ws.slots[seg * width + src(lane % width)];
  ws.bars[seg]->wait();
  return r;
}

uint64_t segment_ballot(bool pred, int width) {
  BlockCtx &b = *g_block;
  const int t = int(threadIdx.x), lane = t & 63, seg = lane / width;
  WidthState &ws = b.waves[t / 64][log2i(width)];
  ws.slots[lane] = pred ? 1 : 0;
  ws.bars[seg]->wait();
  uint64_t r = 0;
  for (int i = seg * width; i < (seg + 1) * width; i++)
    if (ws.slots[i]) r |= uint64_t(1) << i;
  ws.bars[seg]->wait();
  return r;
}

void segment_barrier(int width) {
  BlockCtx &b = *g_block;
  const int t = int(threadIdx.x), lane = t & 63;
  b.waves[t / 64][log2i(width)].bars[lane / width]->wait();
}

void block_barrier() { g_block->block.wait(); }

void run_grid(dim3 grid, dim3 block, const std::function<void()> &body) {
  if (block.y != 1 || block.z != 1 || block.x == 0 || block.x > uint32_t(kMaxThreads)) {
    std::fprintf(stderr, "emu: unsupported block shape %u x %u x %u\n", block.x, block.y, block.z);
    std::abort();
  }
  std::lock_guard<std::mutex> l(g_launch);
  static std::once_flag dog;
  std::call_once(dog, [] { std::thread(watchdog).detach(); });
  g_running++;
  BlockCtx ctx(int(block.x));
  g_block = &ctx;
  for (uint32_t z = 0; z < grid.z; z++)
    for (uint32_t y = 0; y < grid.y; y++)
      for (uint32_t x = 0; x < grid.x; x++) pool().run_block(grid, block, dim3(x, y, z), body);
  g_block = nullptr;
  g_running--;
}

}  // namespace emu

// ---- runtime ----------------------------------------------------------------
struct EmuStream {
  bool capturing = false;
  std::vector<std::function<void()>> ops;
};
struct EmuEvent {
  std::chrono::steady_clock::time_point t;
};
struct EmuGraph {
  std::vector<std::function<void()>> ops;
};

namespace emu {
bool record(hipStream_t s, std::function<void()> op) {
  if (!s || !s->capturing) return false;
  s->ops.push_back(std::move(op));
  return true;
}
}  // namespace emu

hipError_t hipMalloc(void **p, size_t bytes) {
  *p = std::malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  auto op = [dst, src, bytes] { std::memcpy(dst, src, bytes); };
  if (!emu::record(s, op)) op();
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  return (s && s->capturing) ? hipErrorInvalidValue : hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->name, "CPU emulation");
  std::strcpy(p->gcnArchName, "gfx950:emulated");
  p->multiProcessorCount = 4;
  p->totalGlobalMem = size_t(16) << 30;
  return hipSuccess;
}

hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t attr, int) {
  if (attr != hipDeviceAttributeMultiprocessorCount) return hipErrorInvalidValue;
  hipDeviceProp_t p;
  hipGetDeviceProperties(&p, 0);
  *v = p.multiProcessorCount;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  *s = new EmuStream();
  return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new EmuEvent();
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (s && s->capturing) return hipErrorNotSupported;
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s || s->capturing) return hipErrorInvalidValue;
  s->capturing = true;
  s->ops.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->capturing) return hipErrorInvalidValue;
  s->capturing = false;
  *g = new EmuGraph{std::move(s->ops)};
  s->ops.clear();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t) {
  *e = new EmuGraph{g->ops};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  if (s && s->capturing) return hipErrorNotSupported;
  for (auto &op : e->ops) op();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
