// HyperBall engine: host control of the device state and the C ABI of
// include/hyperball.h.
//
// The per-iteration control logic restates HyperBall::iterate
// (src/algo/hyperball/hyperball_impl.rs:775-949): the systolic / local /
// pre-local mode choice, the neighbourhood-function bookkeeping (last value
// kept unclamped, pushed value clamped to be monotone, relative increment),
// and the swaps of the double-buffered state.  Work distribution differs by
// design: instead of the reference's arc-balanced chunks and its local list,
// a systolic iteration compacts the bitmap (must-be-checked | modified) into
// an ordered worklist on the device and launches the union kernel over it.
// Registers and modified bits do not depend on the schedule (the reference's
// own test asserts its parallel run equals a sequential Jacobi run,
// hyperball_impl.rs:1302-1319).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hb_common.hpp"
#include "hb_hll.hpp"
#include "hb_kernels.hpp"

// Nodes up to which iteration 0 reads seeded pairs instead of hashing (4 bytes
// a node: 128 MiB, half the last-level cache).  The CPU emulation lowers it so
// that its small graphs take both paths.
#ifndef HB_SEED_ITEMS_MAX_NODES
#define HB_SEED_ITEMS_MAX_NODES (uint64_t(1) << 25)
#endif

namespace hb {
namespace {

#define HB_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      fail(e_ == hipErrorOutOfMemory ? HB_E_OOM : HB_E_HIP,                                 \
           std::string(#call) + " failed: " + hipGetErrorString(e_));                      \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t count = 0;
  bool host = false;  // pinned host memory that kernels access over the host link
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t n, bool in_host = false) {
    release();
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(1, n) * sizeof(T);
    hipError_t e = in_host ? hipHostMalloc(&q, bytes, hipHostMallocDefault) : hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(HB_E_OOM, std::string(in_host ? "hipHostMalloc" : "hipMalloc") + " of " +
                         std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    p = static_cast<T *>(q);
    count = n;
    host = in_host;
  }
  void release() {
    if (p) (void)(host ? hipHostFree(p) : hipFree(p));
    p = nullptr;
    count = 0;
    host = false;
  }
};

struct Discount {
  hb_discount_fn f;
  void *ctx;
};

const char *kNotRun =
    "HyperBall was not run. Please call HyperBall::run before accessing computed fields";

}  // namespace

struct Instance {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  hb_progress_fn progress = nullptr;  // called after every iteration
  void *progress_ctx = nullptr;
  bool in_batch = false;  // owned by a batch: hb_destroy refuses it
  uint32_t cus = 256;

  uint64_t n = 0, a = 0;
  uint32_t log2m = 0, m = 0, r = 0, W = 0, cdw = 0;
  uint32_t lanes = 1;  // lanes per node in the union kernel
  uint32_t nseg = 1;   // counter segments: above 4096 registers a group works segment by
                       // segment, and no row takes the heavy-row phase
  uint32_t cstride = 0;  // dwords between counters: counter_stride(cdw)
  uint64_t sentinel = 0;
  double amm = 0.0;
  uint32_t hash = 0;
  uint64_t seed = 0, weight_seed = 0;
  bool weighted = false, has_t = false, do_sd = false, do_hc = false;
  uint32_t acc_mode = ACC_F64;
  std::vector<Discount> discounts;
  // Rows longer than heavy_degree take a whole workgroup each (the union
  // kernel's heavy-row phase).
  // heavy_nodes holds every node with more than kHeavyFloor successors,
  // sorted by decreasing degree, so a threshold selects a prefix.
  static constexpr uint64_t kHeavyFloor = 64;
  static constexpr uint64_t kDefaultHeavy = 512;
  uint64_t heavy_degree = kDefaultHeavy;
  std::vector<uint32_t> heavy_nodes;
  std::vector<uint64_t> heavy_degs;
  DevBuf<uint32_t> d_heavy;
  DevBuf<uint32_t> d_heavy_sel;  // worklist iterations: the heavy rows with work
  uint64_t heavy_count = 0;

  // Dense iterations visit nodes in `order`: ids sorted by decreasing
  // out-degree inside consecutive windows of kOrderWindow nodes, so the lane
  // groups of a wavefront get rows of similar length while successor reads
  // keep the locality of the id order.
  static constexpr uint64_t kOrderWindow = 256;
  DevBuf<uint32_t> d_order;
  DevBuf<uint64_t> off, toff, d_weights, d_wprefix;
  DevBuf<uint32_t> succ, tsucc;
  DevBuf<uint32_t> cnt[2], mod[2], must[2];
  DevBuf<uint32_t> chg[2];  // change masks, swapped with the modified bitmaps
  // unweighted, n <= HB_SEED_ITEMS_MAX_NODES: each node's seeded register
  // (F_SEEDED reads it)
  DevBuf<uint32_t> seed_items;
  uint32_t chg_shift = 0;   // log2 slabs per change-mask bit
  // Delta gathering (read only the changed slabs of successors) pays when
  // counters are large and few of their slabs changed: it costs a 4-byte mask
  // read per modified successor and splits each lane's load by slab.
  // Measured on the web graph at log2m 10, r 6: a dense iteration takes about
  // 14.2 + 6.6 f ms with it against 19.1 ms without, f being the fraction of
  // slabs gathered, and nothing is gained below log2m 10.
  static constexpr uint32_t kDeltaMinLog2m = 10;
  static constexpr double kDeltaMaxFraction = 0.7;
  uint64_t last_changed = 0;  // change-mask bits set by the last iteration
  bool delta = false;         // the last planned iteration gathers deltas
  int delta_mode = 0;         // 0: by the rule above, 1: always, 2: never
  // Window kernel (launch_window_union): the union step of iterations that
  // read the counters (not the seeded first one, no delta gathering), for
  // 64-register counters in device memory: those that check every node, and
  // worklist iterations (systolic or local) whose worklist is expected to
  // hold at least kWindowListShare of the nodes (they need no compacted
  // list: the kernel reads the bitmaps).  indeg[s] counts the arcs into s
  // from other nodes (iterations that check every node count visited arcs
  // from it).  Rows longer than kWindowHeavy (or the heavy-row threshold, if
  // larger) take a workgroup each; shorter ones a wavefront or a lane group
  // inside their block.  Successors outside the tile are read untested while
  // at least kUntestedShare of the nodes changed in the last iteration.
  static constexpr uint64_t kWindowHeavy = 8192;
  static constexpr double kUntestedShare = 0.5;
  static constexpr double kWindowListShare = 0.125;
  DevBuf<uint32_t> indeg;
  int window_mode = 0;  // 0: by the rule above, 1: whenever the kernel can run, 2: never
  uint32_t win_block = 0, win_half = 0;
  uint64_t win_heavy = 0;  // 0: max(heavy_degree, kWindowHeavy)
  bool window_now = false;  // the iteration under way runs the window kernel
  DevBuf<double> cache, lc;
  DevBuf<uint8_t> sd, hc;
  std::vector<std::unique_ptr<DevBuf<uint8_t>>> dc;
  DevBuf<uint32_t> list, bsum, list_len;
  DevBuf<Stats> stats;
  DevBuf<uint64_t> stat_slots;  // kStatSlots lines of per-wavefront partial statistics
  DevBuf<uint32_t> work;
  Stats *h_stats = nullptr;
  uint32_t max_grid = 0;
  uint64_t bitmap_words = 0;  // u32 words, even

  int cur = 0, mcur = 0, mustcur = 0;
  bool inited = false;
  uint64_t iteration = 0;
  bool systolic = false, local = false, pre_local = false;
  bool last_full = false;  // the last iteration checked every node (not a worklist)
  double last = 0.0, rel_inc = 0.0;
  uint64_t modified = 0;
  std::vector<double> nf;
  std::vector<hb_iteration_stats> history;
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  // Streamed counters (HB_COUNTERS_STREAM): both arrays in pinned host
  // memory; an iteration walks tiles of tile_nodes nodes.  Per tile: upload
  // its own counters (s_own[2], alternately per tile, so the next tile's
  // upload need not wait for this tile's passes), merge them and then every
  // uploaded block (block_nodes nodes, s_src[2] alternately) that holds
  // modified nodes into the partial next counters (s_part[2], alternately per
  // tile), finish, download.  Uploads run on s_up, downloads on s_down,
  // kernels on `stream`; events order the buffer reuse.
  uint32_t counter_memory = HB_COUNTERS_DEVICE;
  uint64_t tile_nodes = 0, block_nodes = 0, tiles = 0, blocks = 0;
  DevBuf<uint32_t> s_own[2], s_part[2], s_src[2], s_counts;
  uint32_t *h_counts = nullptr;  // pinned: per tile the nodes to work on, per block the modified
  hipStream_t s_up = nullptr, s_down = nullptr;
  enum { kOwnReady0, kOwnReady1, kOwnFree0, kOwnFree1, kSrcReady0, kSrcReady1, kSrcFree0,
         kSrcFree1, kPartDone0, kPartDone1, kPartFree0, kPartFree1, kPreReady, kOtherRead,
         kDownDone, kStreamEvents };
  hipEvent_t sev[kStreamEvents] = {};

  ~Instance() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamSynchronize(stream);
    }
    for (auto q : {s_up, s_down})
      if (q) (void)hipStreamSynchronize(q);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : sev)
      if (e) (void)hipEventDestroy(e);
    for (auto q : {s_up, s_down})
      if (q) (void)hipStreamDestroy(q);
    if (h_counts) (void)hipHostFree(h_counts);
    if (h_stats) (void)hipHostFree(h_stats);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }

  size_t acc_bytes() const { return acc_mode == ACC_F32 ? 4 : 8; }
  bool do_centrality() const { return do_sd || do_hc || !discounts.empty(); }
  void use_device() const { HB_HIP(hipSetDevice(device)); }

  void set_heavy_degree(uint64_t h) {
    heavy_degree = std::max(h, kHeavyFloor);
    uint64_t k = 0;
    while (k < heavy_degs.size() && heavy_degs[k] > heavy_degree) k++;
    heavy_count = k;
  }

  // Light/heavy split autotuner: times the first iteration that gathers
  // whole successor counters (iteration 1, every node checked, delta
  // gathering off; iteration 0 reads seeded registers only, unless weighted)
  // under each candidate threshold that changes the heavy-row set and keeps
  // the fastest.  Results do not depend on the threshold; the run state is
  // reset as by init().  Candidates: every 64 * 2^k that leaves a different,
  // non-empty set of heavy rows, then "no heavy rows".
  std::vector<std::pair<uint64_t, double>> autotune(uint32_t reps) {
    if (reps == 0) reps = 1;
    std::vector<uint64_t> cand;
    uint64_t last_count = UINT64_MAX;
    for (uint64_t t = kHeavyFloor;; t *= 2) {
      set_heavy_degree(t);
      if (heavy_count == 0) break;
      if (heavy_count != last_count) cand.push_back(t);
      last_count = heavy_count;
    }
    cand.push_back(UINT64_MAX);
    const bool was_timing = timing;
    const int was_delta = delta_mode;
    timing = true;
    delta_mode = 2;
    std::vector<std::pair<uint64_t, double>> table;
    try {
      for (uint64_t t : cand) {
        set_heavy_degree(t);
        double best = 1e300;
        for (uint32_t r = 0; r < reps; r++) {
          init();
          iterate();
          iterate();
          best = std::min(best, history.back().union_ms);
        }
        table.push_back({t, best});
      }
    } catch (...) {
      timing = was_timing;
      delta_mode = was_delta;
      throw;
    }
    timing = was_timing;
    delta_mode = was_delta;
    auto it = std::min_element(table.begin(), table.end(),
                               [](const auto &a, const auto &b) { return a.second < b.second; });
    set_heavy_degree(it->first);
    init();
    return table;
  }

  void fill_zero(void *p, size_t bytes) {
    HB_HIP(launch_fill_u32(static_cast<uint32_t *>(p), bytes / 4, 0u, stream));
  }

  void init() {
    use_device();
    // cnt[0] is zeroed for seed_kernel (weights) or written whole by the
    // fused seeding and count (no weights).  cnt[1] is not initialised:
    // iteration 0 writes every node's counter there (every node starts
    // modified), and hb_get_counters reports it as zero until then.
    if (weighted) fill_zero(cnt[0].p, size_t(n) * cstride * 4);
    SeedArgs s{};
    s.counters = cnt[0].p;
    s.n = n;
    s.cdw = cstride;
    s.r = r;
    s.log2m = log2m;
    s.sentinel = sentinel;
    s.hash = hash;
    s.seed = seed;
    s.weights = weighted ? d_weights.p : nullptr;
    s.weight_prefix = weighted ? d_wprefix.p : nullptr;
    s.weight_seed = weight_seed;
    s.chg = chg[0].p;
    s.chg_shift = chg_shift;
    s.items = seed_items.p;
    if (weighted) {
      HB_HIP(launch_seed(s, stream));
      HB_HIP(launch_count_all(int(r), int(log2m), cnt[0].p, n, amm, m, lc.p, cache.p, stream));
    } else {
      HB_HIP(launch_seed_count(s, m, lc.p, cache.p, stream));
    }
    if (do_sd) fill_zero(sd.p, n * acc_bytes());
    if (do_hc) fill_zero(hc.p, n * acc_bytes());
    for (auto &d : dc) fill_zero(d->p, n * acc_bytes());
    HB_HIP(launch_fill_bits(mod[0].p, n, stream));
    fill_zero(mod[1].p, bitmap_words * 4);
    fill_zero(must[0].p, bitmap_words * 4);
    fill_zero(must[1].p, bitmap_words * 4);
    HB_HIP(hipStreamSynchronize(stream));
    cur = mcur = mustcur = 0;
    last_changed = 0;
    iteration = 0;
    systolic = local = pre_local = last_full = false;
    modified = 0;
    last = double(n);
    rel_inc = 0.0;
    nf.assign(1, last);
    history.clear();
    inited = true;
  }

  // Host part of HyperBall::iterate before the device work: mode choice
  // (hyperball_impl.rs:786-861) and the arguments of the union launch.
  struct Plan {
    UnionArgs ua;
    bool list_mode;
    bool mark_first;  // mark the predecessors of the modified nodes before the union
    double base_nf;
    uint32_t grid;
  };

  Plan plan_iteration() {
    if (!inited) fail(HB_E_INVALID_ARG, "hb_iterate called before hb_init");
    Plan p{};
    const uint64_t modified_prev = modified;
    const bool prev_systolic = systolic;
    systolic = has_t && iteration > 0 && modified_prev < n / 4;
    p.base_nf = systolic ? last : 0.0;
    local = pre_local;
    pre_local = systolic && a > 0 && modified_prev < (n * n) / (a * 10);

    uint32_t flags = 0;
    p.list_mode = false;
    if (systolic) flags |= F_SYSTOLIC;
    if (do_centrality()) flags |= F_CENTRALITY;
    if (local) {
      p.list_mode = true;
      flags |= F_CHECK_ALL;
    } else if (systolic && prev_systolic) {
      p.list_mode = true;
    } else if (systolic && last_full) {
      // The first systolic iteration after one that checked every node: no
      // predecessors were marked, so the reference checks every node
      // again.  A node none of whose successors changed keeps its counter
      // (the full check already merged them all), so marking the
      // predecessors of the modified nodes first and running over the
      // worklist gives the same counters.
      p.list_mode = true;
      p.mark_first = true;
    } else {
      flags |= F_CHECK_ALL;  // standard, or the first systolic iteration (all nodes checked)
    }
    if (p.list_mode) flags |= F_LIST;
    last_full = !p.list_mode;
    {
      // expected fraction of successor slabs to gather: one non-zero slab per
      // counter after unweighted seeding; then the mean share of changed
      // slabs over the nodes the last iteration modified
      const double bits = double(std::min<uint32_t>(32, m / 32));
      const double est = iteration == 0 ? (weighted ? 1.0 : 1.0 / bits)
                         : modified_prev ? double(last_changed) / (double(modified_prev) * bits)
                                         : 1.0;
      delta = delta_mode == 1 ||
              (delta_mode == 0 && log2m >= kDeltaMinLog2m && est < kDeltaMaxFraction);
      if (delta) flags |= F_DELTA;
    }
    if (iteration == 0 && !weighted) flags |= F_SEEDED;

    UnionArgs &ua = p.ua;
    ua.off = off.p;
    ua.succ = succ.p;
    ua.toff = toff.p;
    ua.tsucc = tsucc.p;
    ua.curr = cnt[cur].p;
    ua.next = cnt[1 - cur].p;
    ua.curr_mod = mod[mcur].p;
    ua.next_mod = mod[1 - mcur].p;
    ua.must = must[mustcur].p;
    ua.next_must = must[1 - mustcur].p;
    ua.cache = cache.p;
    ua.list = list.p;
    ua.list_len = list_len.p;
    ua.order = d_order.p;
    ua.count = n;
    ua.flags = flags;
    ua.acc_mode = acc_mode;
    ua.sd = do_sd ? static_cast<void *>(sd.p) : nullptr;
    ua.hc = do_hc ? static_cast<void *>(hc.p) : nullptr;
    ua.ndisc = int(discounts.size());
    for (size_t k = 0; k < discounts.size(); k++) {
      ua.dc[k] = dc[k]->p;
      ua.disc[k] = discounts[k].f(iteration + 1, discounts[k].ctx);
    }
    ua.dist = double(iteration + 1);
    ua.alpha_m_m = amm;
    ua.m = m;
    ua.lc_table = lc.p;
    ua.seed_items = seed_items.p;
    ua.log2m = log2m;
    ua.hash = hash;
    ua.seed = seed;
    ua.sentinel = sentinel;
    ua.stats = stats.p;
    ua.stat_slots = stat_slots.p;
    ua.work = work.p;
    ua.heavy_degree = heavy_count ? heavy_degree : UINT64_MAX;
    ua.heavy_list = d_heavy.p;
    ua.heavy_count = heavy_count;
    ua.heavy_len = nullptr;
    ua.chg_cur = chg[mcur].p;
    ua.chg_next = chg[1 - mcur].p;
    ua.chg_shift = chg_shift;

    const uint64_t npb = kBlock / lanes;
    p.grid = max_grid;
    if (!p.list_mode)
      p.grid = uint32_t(std::min<uint64_t>(max_grid, std::max<uint64_t>(8, (n + npb - 1) / npb)));
    p.grid = (p.grid + 7) / 8 * 8;
    return p;
  }

  // Host part of HyperBall::iterate after the device work
  // (hyperball_impl.rs:909-946): swaps, neighbourhood function, counters.
  void finish_iteration(const Plan &p, const Stats &st, float union_ms, float iteration_ms) {
    hb_iteration_stats hs{};
    hs.iteration = iteration;
    hs.systolic = systolic;
    hs.local = local;
    hs.pre_local = pre_local;
    hs.checked = st.checked;
    hs.visited_arcs = st.visited;
    hs.scanned_arcs = st.scanned;
    hs.transpose_arcs = st.tscanned;
    hs.copied = st.copied;
    hs.modified = st.modified;
    hs.gathered_slabs = st.gathered;
    hs.delta_gathering = delta;
    hs.streamed = streamed_now;
    hs.window = window_now;
    last_changed = st.changed;
    hs.union_ms = union_ms;
    hs.iteration_ms = iteration_ms;

    cur = 1 - cur;
    mcur = 1 - mcur;
    if (systolic) mustcur = 1 - mustcur;
    double current = p.base_nf + stats_nf(st);
    last = current;
    const double last_output = nf.back();
    if (current < last_output) current = last_output;
    rel_inc = current / last_output;
    nf.push_back(current);
    hs.nf = current;
    history.push_back(hs);
    modified = st.modified;
    iteration++;
    if (progress) progress(&history.back(), progress_ctx);
  }

  // HyperBall::iterate with streamed counters: the same plan and per-node
  // body as iterate(), worked tile by tile (see the streamed-counter fields).
  // A tile with no node to check and no node modified last iteration is
  // skipped: its next counters already equal its current ones in the other
  // host array.  Source blocks without modified nodes are not uploaded.
  void iterate_streamed() {
    use_device();
    Plan p = plan_iteration();
    p.ua.flags &= ~(F_SEEDED | F_DELTA);  // counters are read, whole, from the uploaded ranges
    delta = false;
    if (timing) HB_HIP(hipEventRecord(ev[0], stream));
    HB_HIP(launch_prepare_iteration(p.ua.next_mod, systolic ? p.ua.next_must : nullptr,
                                    bitmap_words, stats.p, work.p, stat_slots.p, stream));
    if (p.mark_first) {
      fill_zero(must[mustcur].p, bitmap_words * 4);
      HB_HIP(launch_compact(mod[mcur].p, nullptr, n, bsum.p, list.p, list_len.p, stream));
      HB_HIP(launch_mark_predecessors(list.p, list_len.p, toff.p, tsucc.p, must[mustcur].p,
                                      max_grid, stream));
    }
    fill_zero(s_counts.p, (tiles + blocks) * 4);
    HB_HIP(launch_stream_summary(p.ua, n, tile_nodes, block_nodes, tiles, s_counts.p, stream));
    HB_HIP(hipMemcpyAsync(h_counts, s_counts.p, (tiles + blocks) * 4, hipMemcpyDeviceToHost,
                          stream));
    HB_HIP(hipStreamSynchronize(stream));
    if (timing) HB_HIP(hipEventRecord(ev[1], stream));
    const uint64_t bytes_per_node = uint64_t(cstride) * 4;
    const uint32_t *host_cur = cnt[cur].p;
    uint32_t *host_next = cnt[1 - cur].p;
    const uint64_t bpt = tile_nodes / block_nodes;  // blocks per tile
    std::vector<uint64_t> work_tiles;
    for (uint64_t t = 0; t < tiles; t++)
      if (h_counts[t]) work_tiles.push_back(t);
    // Which tile each own buffer holds this iteration, and which of its
    // blocks have arrived.  A tile's blocks that the previous tile needed as
    // sources were uploaded straight into its own buffer; blocks of the
    // previous tile are read from the other own buffer while it still holds
    // them.
    int64_t held[2] = {-1, -1};
    std::vector<char> loaded[2] = {std::vector<char>(bpt, 0), std::vector<char>(bpt, 0)};
    auto upload = [&](uint32_t *dst, uint64_t lo, uint64_t hi) {
      HB_HIP(hipMemcpyAsync(dst, host_cur + lo * cstride, (hi - lo) * bytes_per_node,
                            hipMemcpyHostToDevice, s_up));
    };
    int pi = 0, k = 0;
    for (size_t w = 0; w < work_tiles.size(); w++) {
      const uint64_t t = work_tiles[w];
      const int64_t next_tile = w + 1 < work_tiles.size() ? int64_t(work_tiles[w + 1]) : -1;
      const uint64_t tl = t * tile_nodes, th = std::min(n, tl + tile_nodes);
      const int other = 1 - pi;
      if (held[pi] != int64_t(t)) {
        HB_HIP(hipStreamWaitEvent(s_up, sev[kOwnFree0 + pi], 0));
        held[pi] = int64_t(t);
        std::fill(loaded[pi].begin(), loaded[pi].end(), 0);
      }
      for (uint64_t j = 0; j < bpt && tl + j * block_nodes < th; j++) {
        if (loaded[pi][j]) continue;
        const uint64_t bl = tl + j * block_nodes, bh = std::min(th, bl + block_nodes);
        upload(s_own[pi].p + (bl - tl) * cstride, bl, bh);
        loaded[pi][j] = 1;
      }
      HB_HIP(hipEventRecord(sev[kOwnReady0 + pi], s_up));
      HB_HIP(hipStreamWaitEvent(stream, sev[kOwnReady0 + pi], 0));
      HB_HIP(hipStreamWaitEvent(stream, sev[kPartFree0 + pi], 0));
      StreamPass sp{s_own[pi].p, tl, th, s_own[pi].p, s_part[pi].p, tl, th, 1};
      HB_HIP(launch_stream_merge(int(r), int(log2m), p.ua, sp, stream));
      sp.first = 0;
      bool other_read = false;  // passes of this tile read the other own buffer
      for (uint64_t b = 0; b < blocks; b++) {
        const uint64_t bl = b * block_nodes, bh = std::min(n, bl + block_nodes);
        if ((bl >= tl && bh <= th) || h_counts[tiles + b] == 0) continue;
        const int64_t bt = int64_t(b / bpt);
        const uint64_t j = b % bpt, off = j * block_nodes * cstride;
        sp.lo = bl;
        sp.hi = bh;
        if (held[other] == bt && loaded[other][j]) {
          sp.src = s_own[other].p + off;
          other_read = true;
        } else if (bt == next_tile) {
          if (held[other] != bt) {
            if (other_read) {
              HB_HIP(hipEventRecord(sev[kOtherRead], stream));
              HB_HIP(hipStreamWaitEvent(s_up, sev[kOtherRead], 0));
              other_read = false;
            }
            HB_HIP(hipStreamWaitEvent(s_up, sev[kOwnFree0 + other], 0));
            held[other] = bt;
            std::fill(loaded[other].begin(), loaded[other].end(), 0);
          }
          upload(s_own[other].p + off, bl, bh);
          loaded[other][j] = 1;
          HB_HIP(hipEventRecord(sev[kPreReady], s_up));
          HB_HIP(hipStreamWaitEvent(stream, sev[kPreReady], 0));
          sp.src = s_own[other].p + off;
        } else {
          HB_HIP(hipStreamWaitEvent(s_up, sev[kSrcFree0 + k], 0));
          upload(s_src[k].p, bl, bh);
          HB_HIP(hipEventRecord(sev[kSrcReady0 + k], s_up));
          HB_HIP(hipStreamWaitEvent(stream, sev[kSrcReady0 + k], 0));
          sp.src = s_src[k].p;
          HB_HIP(launch_stream_merge(int(r), int(log2m), p.ua, sp, stream));
          HB_HIP(hipEventRecord(sev[kSrcFree0 + k], stream));
          k ^= 1;
          continue;
        }
        HB_HIP(launch_stream_merge(int(r), int(log2m), p.ua, sp, stream));
      }
      HB_HIP(launch_stream_finish(int(r), int(log2m), p.ua, sp, stream));
      // nothing issued so far reads either own buffer after this point
      HB_HIP(hipEventRecord(sev[kOwnFree0], stream));
      HB_HIP(hipEventRecord(sev[kOwnFree1], stream));
      HB_HIP(hipEventRecord(sev[kPartDone0 + pi], stream));
      HB_HIP(hipStreamWaitEvent(s_down, sev[kPartDone0 + pi], 0));
      HB_HIP(hipMemcpyAsync(host_next + tl * cstride, s_part[pi].p, (th - tl) * bytes_per_node,
                            hipMemcpyDeviceToHost, s_down));
      HB_HIP(hipEventRecord(sev[kPartFree0 + pi], s_down));
      pi = other;
    }
    HB_HIP(hipEventRecord(sev[kDownDone], s_down));
    HB_HIP(hipStreamWaitEvent(stream, sev[kDownDone], 0));
    if (timing) HB_HIP(hipEventRecord(ev[2], stream));
    HB_HIP(launch_reduce_stats(stats.p, stat_slots.p, h_stats, stream));
    if (timing) HB_HIP(hipEventRecord(ev[3], stream));
    HB_HIP(hipStreamSynchronize(stream));
    float ums = 0, ims = 0;
    if (timing) {
      HB_HIP(hipEventElapsedTime(&ums, ev[1], ev[2]));
      HB_HIP(hipEventElapsedTime(&ims, ev[0], ev[3]));
    }
    finish_iteration(p, *h_stats, ums, ims);
  }

  // Streamed counters: whether this iteration streams them (hb_set_streaming).
  int stream_mode = 0;
  bool streamed_now = false;  // the iteration under way streams them
  bool stream_this_iteration() const {
    if (counter_memory != HB_COUNTERS_STREAM || stream_mode == 2) return false;
    if (stream_mode == 1) return true;
    if (iteration == 0) return weighted;  // the seeded iteration reads no counters
    const double visited = double(modified) * double(a) / double(std::max<uint64_t>(n, 1));
    const double streamed = double(n) * double(tiles > 3 ? tiles - 2 : 1);
    return visited >= streamed;
  }

  // Turns a planned dense iteration into a window-kernel launch: untested
  // reads outside the tile while most nodes changed, and the kernel's own
  // heavy-row threshold (a prefix of heavy_nodes, sorted by decreasing degree).
  bool plan_window(Plan &p, uint64_t modified_prev) {
    const uint32_t f = p.ua.flags;
    if (window_mode == 2 || !indeg.p || (f & (F_SEEDED | F_DELTA))) return false;
    if (p.list_mode || (f & F_SYSTOLIC)) {
      // A worklist iteration runs as a window iteration when its worklist is
      // expected to hold a large share of the nodes: the modified nodes and
      // their predecessors (about the mean in-degree each).
      const double expected = double(modified_prev) * (1.0 + double(a) / double(n));
      if (window_mode == 0 && expected < kWindowListShare * double(n)) return false;
      if (p.mark_first) {
        // The first systolic iteration checks every node, as the reference
        // does (no predecessors were marked), instead of marking the
        // predecessors of the modified nodes first: a node with no modified
        // successor merges nothing, so the counters are the same, and the
        // kernel tests each successor's modified bit anyway.
        p.mark_first = false;
        p.list_mode = false;
        p.ua.flags = (p.ua.flags & ~F_LIST) | F_CHECK_ALL;
      }
    } else if (!(f & F_CHECK_ALL)) {
      return false;
    } else if (double(modified_prev) >= kUntestedShare * double(n)) {
      p.ua.flags |= F_UNTESTED;
    }
    const uint64_t h = win_heavy ? win_heavy : std::max(heavy_degree, kWindowHeavy);
    uint64_t k = 0;
    while (k < heavy_degs.size() && heavy_degs[k] > h) k++;
    p.ua.heavy_degree = k ? h : UINT64_MAX;
    p.ua.heavy_count = k;
    p.ua.indeg = indeg.p;
    p.ua.win_block = win_block;
    p.ua.win_half = win_half;
    return true;
  }

  void iterate() {
    streamed_now = stream_this_iteration();
    if (streamed_now) {
      iterate_streamed();
      return;
    }
    use_device();
    const uint64_t modified_prev = modified;
    Plan p = plan_iteration();
    window_now = plan_window(p, modified_prev);
    if (timing) HB_HIP(hipEventRecord(ev[0], stream));
    HB_HIP(launch_prepare_iteration(p.ua.next_mod, systolic ? p.ua.next_must : nullptr,
                                    bitmap_words, stats.p, work.p, stat_slots.p, stream));
    if (p.mark_first) {
      fill_zero(must[mustcur].p, bitmap_words * 4);
      HB_HIP(launch_compact(mod[mcur].p, nullptr, n, bsum.p, list.p, list_len.p, stream));
      HB_HIP(launch_mark_predecessors(list.p, list_len.p, toff.p, tsucc.p, must[mustcur].p,
                                      max_grid, stream));
    }
    if (p.list_mode && !window_now)
      HB_HIP(launch_compact(p.ua.must, p.ua.curr_mod, n, bsum.p, list.p, list_len.p, stream));
    if (!window_now && p.ua.heavy_count && !(p.ua.flags & F_CHECK_ALL)) {
      // worklist iteration: only the heavy rows to check or to copy
      uint32_t *len = work.p + size_t(kHeavySelCounter) * kWorkStride;
      HB_HIP(launch_select_heavy(d_heavy.p, p.ua.heavy_count, p.ua.must, p.ua.curr_mod,
                                 d_heavy_sel.p, len, stream));
      p.ua.heavy_list = d_heavy_sel.p;
      p.ua.heavy_len = len;
    }
    if (timing) HB_HIP(hipEventRecord(ev[1], stream));
    if (window_now)
      HB_HIP(launch_window_union(int(r), int(log2m), p.ua, stream));
    else
      HB_HIP(launch_union(int(r), int(log2m), p.ua, p.grid, stream));
    if (timing) HB_HIP(hipEventRecord(ev[2], stream));
    HB_HIP(launch_reduce_stats(stats.p, stat_slots.p, h_stats, stream));  // h_stats: pinned
    if (timing) HB_HIP(hipEventRecord(ev[3], stream));
    HB_HIP(hipStreamSynchronize(stream));
    float ums = 0, ims = 0;
    if (timing) {
      HB_HIP(hipEventElapsedTime(&ums, ev[1], ev[2]));
      HB_HIP(hipEventElapsedTime(&ims, ev[0], ev[3]));
    }
    finish_iteration(p, *h_stats, ums, ims);
  }

  bool stop_after(uint64_t i, uint64_t upper_bound, double threshold) const {
    if (modified == 0) return true;
    if (!std::isnan(threshold) && i > 3 && rel_inc < 1.0 + threshold) return true;
    return i + 1 >= upper_bound;
  }

  // HyperBall::run (hyperball_impl.rs:534-583).
  void run(uint64_t upper_bound, double threshold) {
    upper_bound = std::min(upper_bound, n);
    init();
    for (uint64_t i = 0; i < upper_bound; i++) {
      iterate();
      if (stop_after(i, upper_bound, threshold)) break;
    }
  }

  void ensure_iteration() const {
    if (!inited || iteration == 0) fail(HB_E_NOT_RUN, kNotRun);
  }

  std::vector<double> read_acc(const DevBuf<uint8_t> &buf) {
    use_device();
    std::vector<double> out(n);
    if (n == 0) return out;
    if (acc_mode == ACC_F64) {
      HB_HIP(hipMemcpyAsync(out.data(), buf.p, n * 8, hipMemcpyDeviceToHost, stream));
      HB_HIP(hipStreamSynchronize(stream));
    } else if (acc_mode == ACC_F32_KAHAN) {
      std::vector<float> tmp(2 * n);
      HB_HIP(hipMemcpyAsync(tmp.data(), buf.p, n * 8, hipMemcpyDeviceToHost, stream));
      HB_HIP(hipStreamSynchronize(stream));
      for (uint64_t v = 0; v < n; v++) out[v] = double(tmp[2 * v]);
    } else {
      std::vector<float> tmp(n);
      HB_HIP(hipMemcpyAsync(tmp.data(), buf.p, n * 4, hipMemcpyDeviceToHost, stream));
      HB_HIP(hipStreamSynchronize(stream));
      for (uint64_t v = 0; v < n; v++) out[v] = double(tmp[v]);
    }
    return out;
  }

  std::vector<double> read_counts() {
    use_device();
    std::vector<double> out(n);
    if (n) {
      HB_HIP(hipMemcpyAsync(out.data(), cache.p, n * 8, hipMemcpyDeviceToHost, stream));
      HB_HIP(hipStreamSynchronize(stream));
    }
    return out;
  }
};

namespace {

void check_csr(const hb_csr *g, const char *what) {
  if (!g) fail(HB_E_INVALID_ARG, std::string(what) + " is null");
  if (g->num_nodes > 0xffffffffull)
    fail(HB_E_INVALID_ARG, std::string(what) + ": more than 2^32 - 1 nodes");
  if (!g->offsets) fail(HB_E_INVALID_ARG, std::string(what) + ": null offsets");
  if (g->num_arcs && !g->successors) fail(HB_E_INVALID_ARG, std::string(what) + ": null successors");
  const uint64_t n = g->num_nodes;
  if (g->offsets[0] != 0 || g->offsets[n] != g->num_arcs)
    fail(HB_E_INVALID_ARG, std::string(what) + ": offsets must run from 0 to num_arcs");
  parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t v = b; v < e; v++)
      if (g->offsets[v + 1] < g->offsets[v])
        fail(HB_E_INVALID_ARG, std::string(what) + ": offsets are not monotone");
  });
  parallel_for(g->num_arcs, [&](uint64_t b, uint64_t e, unsigned) {
    for (uint64_t k = b; k < e; k++)
      if (g->successors[k] >= n)
        fail(HB_E_INVALID_ARG, std::string(what) + ": successor out of range");
  });
}

std::string min_alignment(uint64_t bits) {
  if (bits % 128 == 0) return "u128";
  if (bits % 64 == 0) return "u64";
  if (bits % 32 == 0) return "u32";
  if (bits % 16 == 0) return "u16";
  return "u8";
}

template <class T>
void upload(DevBuf<T> &buf, const T *src, size_t count, hipStream_t s) {
  buf.alloc(count);
  if (count) HB_HIP(hipMemcpyAsync(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}

Instance *checked(hb_instance *hb) {
  if (!hb) fail(HB_E_INVALID_ARG, "null instance");
  return reinterpret_cast<Instance *>(hb);
}

void copy_out(const std::vector<double> &v, double *out, size_t cap) {
  if (!out) fail(HB_E_INVALID_ARG, "null output buffer");
  if (cap < v.size())
    fail(HB_E_INVALID_ARG, "output buffer holds " + std::to_string(cap) + " values, " +
                               std::to_string(v.size()) + " needed");
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}

}  // namespace

// Many independent instances advanced together: every batched iteration is
// one fixed sequence (descriptor upload, fused kernels, statistics download)
// that is captured once in a HIP graph and replayed.  All per-iteration
// choices reach the device through the descriptor table, so the captured
// kernel arguments never change.
struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<std::unique_ptr<Instance>> inst;
  DevBuf<BatchDesc> d_desc;
  BatchDesc *h_desc = nullptr;  // pinned
  DevBuf<Stats> d_stats;
  DevBuf<uint64_t> slots;  // per instance: kStatSlots statistics slots
  Stats *h_stats = nullptr;     // pinned
  DevBuf<uint32_t> work;
  DevBuf<uint64_t> total_chunks, total_window;
  DevBuf<uint32_t> shape_inst, shape_off;  // instances grouped by shape
  BatchLaunch L{};
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<double> step_ms;

  ~Batch() {
    if (stream) {
      (void)hipSetDevice(device);
      (void)hipStreamSynchronize(stream);
    }
    drop_graph();
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    inst.clear();
    if (h_desc) (void)hipHostFree(h_desc);
    if (h_stats) (void)hipHostFree(h_stats);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  std::vector<uint32_t> h_shape_inst;

  // Heavy-row ranges, shape by shape (each shape's total is a kernel argument
  // captured in the graph: a change drops it).
  void refresh_heavy() {
    for (int sh = 0; sh < kBatchShapes; sh++) {
      uint64_t acc = 0;
      for (uint32_t j = L.first[sh]; j < L.first[sh] + L.count[sh]; j++) {
        const uint32_t i = h_shape_inst[j];
        h_desc[i].heavy_base = acc;
        acc += inst[i]->heavy_count;
        h_desc[i].heavy_end = acc;
      }
      if (acc != L.total_heavy[sh]) drop_graph();
      L.total_heavy[sh] = acc;
    }
  }

  void enqueue_step() {
    const int k = int(inst.size());
    HB_HIP(hipMemcpyAsync(d_desc.p, h_desc, k * sizeof(BatchDesc), hipMemcpyHostToDevice, stream));
    HB_HIP(launch_batch_iteration(L, stream));
    HB_HIP(hipMemcpyAsync(h_stats, d_stats.p, k * sizeof(Stats), hipMemcpyDeviceToHost, stream));
  }

  void capture() {
    HB_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_step();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HB_HIP(hipStreamEndCapture(stream, &graph));
    HB_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }

  void step() {
    HB_HIP(hipEventRecord(ev[0], stream));
    if (use_graph) {
      if (!exec) capture();
      HB_HIP(hipGraphLaunch(exec, stream));
    } else {
      enqueue_step();
    }
    HB_HIP(hipEventRecord(ev[1], stream));
    HB_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    HB_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    step_ms.push_back(ms);
  }

  // Per-instance run state of a batched run.
  std::vector<uint64_t> ub, iters;
  std::vector<char> running;
  std::vector<Instance::Plan> plans;
  double stop_threshold = NAN;

  void begin(uint64_t upper_bound, double threshold) {
    HB_HIP(hipSetDevice(device));
    const size_t k = inst.size();
    ub.assign(k, 0);
    iters.assign(k, 0);
    running.assign(k, 0);
    plans.resize(k);
    stop_threshold = threshold;
    for (size_t i = 0; i < k; i++) {
      inst[i]->init();
      ub[i] = std::min(upper_bound, inst[i]->n);
      running[i] = ub[i] > 0;
    }
    refresh_heavy();
    step_ms.clear();
  }

  // One batched iteration of every running instance; false when none runs.
  bool advance() {
    const size_t k = inst.size();
    bool any = false;
    for (size_t i = 0; i < k; i++) {
      BatchDesc &d = h_desc[i];
      if (!running[i]) {
        d.active = 0;
        continue;
      }
      any = true;
      const uint64_t modified_prev = inst[i]->modified;
      plans[i] = inst[i]->plan_iteration();
      inst[i]->window_now = inst[i]->plan_window(plans[i], modified_prev);
      d.window = inst[i]->window_now;
      d.u = plans[i].ua;
      d.u.stats = d_stats.p + i;
      d.u.work = work.p;
      d.u.stat_slots = slots.p + i * size_t(kStatSlots) * kStatSlotWords;
      d.active = 1;
      d.list_mode = plans[i].list_mode && !d.window;  // a window iteration reads the bitmaps
      d.mark_first = plans[i].mark_first;
      d.clear_must = inst[i]->systolic;
    }
    if (!any) return false;
    step();
    for (size_t i = 0; i < k; i++) {
      if (!running[i]) continue;
      inst[i]->finish_iteration(plans[i], h_stats[i], 0.0f, float(step_ms.back()));
      if (inst[i]->stop_after(iters[i], ub[i], stop_threshold)) running[i] = 0;
      iters[i]++;
    }
    return true;
  }

  // HyperBall::run (hyperball_impl.rs:534-583) for every instance; each one
  // stops by its own rule, the batch ends when all have stopped.
  void run(uint64_t upper_bound, double threshold) {
    begin(upper_bound, threshold);
    while (advance()) {
    }
  }

  // Light/heavy split autotuner for the batch: one threshold for every
  // instance, chosen by timing the second batched iteration (all instances
  // dense and gathering whole counters: delta gathering off; the first reads
  // seeded registers only) under each 64 * 2^k below the longest row, and
  // "no heavy rows".
  std::vector<std::pair<uint64_t, double>> autotune(uint32_t reps) {
    if (reps == 0) reps = 1;
    uint64_t top = 0;
    for (auto &I : inst)
      if (!I->heavy_degs.empty()) top = std::max(top, I->heavy_degs.front());
    std::vector<uint64_t> cand;
    for (uint64_t t = Instance::kHeavyFloor; t < top; t *= 2) cand.push_back(t);
    cand.push_back(UINT64_MAX);
    std::vector<std::pair<uint64_t, double>> table;
    std::vector<int> was_delta;
    for (auto &I : inst) {
      was_delta.push_back(I->delta_mode);
      I->delta_mode = 2;
    }
    auto restore = [&] {
      for (size_t i = 0; i < inst.size(); i++) inst[i]->delta_mode = was_delta[i];
    };
    try {
      for (uint64_t t : cand) {
        for (auto &I : inst) I->set_heavy_degree(t);
        double best = 1e300;
        for (uint32_t r = 0; r < reps; r++) {
          begin(UINT64_MAX, NAN);
          advance();
          advance();  // does nothing when every instance stopped after the first
          if (!step_ms.empty()) best = std::min(best, step_ms.back());
        }
        table.push_back({t, best});
      }
    } catch (...) {
      restore();
      throw;
    }
    restore();
    auto it = std::min_element(table.begin(), table.end(),
                               [](const auto &a, const auto &b) { return a.second < b.second; });
    for (auto &I : inst) {
      I->set_heavy_degree(it->first);
      I->init();
    }
    refresh_heavy();
    step_ms.clear();
    return table;
  }
};

// Builds an instance (HyperBallBuilder::with_hyper_log_log + build,
// hyperball_impl.rs:91-133, :321-403).  With shared_stream the instance works
// on that stream instead of its own.
std::unique_ptr<Instance> make_instance(const hb_csr *graph, const hb_csr *transpose,
                                        const hb_config *cfg, hipStream_t shared_stream) {
    if (!cfg) fail(HB_E_INVALID_ARG, "null configuration");
    check_csr(graph, "graph");
    const uint64_t n = graph->num_nodes;
    if (transpose) {
      check_csr(transpose, "transpose");
      if (transpose->num_nodes != n)
        fail(HB_E_INVALID_ARG, "the transpose should have same number of nodes of the graph (" +
                                   std::to_string(n) + "). Got " +
                                   std::to_string(transpose->num_nodes) + ".");
      if (transpose->num_arcs != graph->num_arcs)
        fail(HB_E_INVALID_ARG, "the transpose should have same number of nodes of the graph (" +
                                   std::to_string(graph->num_arcs) + "). Got " +
                                   std::to_string(transpose->num_arcs) + ".");
    }
    if (cfg->acc_precision > HB_ACC_F32) fail(HB_E_INVALID_ARG, "unknown accumulator precision");
    if (cfg->hash > HB_HASH_JENKINS) fail(HB_E_INVALID_ARG, "unknown hash");
    if (cfg->counter_memory > HB_COUNTERS_STREAM) fail(HB_E_INVALID_ARG, "unknown counter memory");
    uint64_t num_elements = cfg->num_elements ? cfg->num_elements : n;
    if (cfg->weights && !cfg->num_elements) {
      num_elements = 0;
      for (uint64_t v = 0; v < n; v++) num_elements += cfg->weights[v];
    }
    if (cfg->log2m < 4)
      fail(HB_E_INVALID_ARG,
           "the logarithm of the number of registers per counter should be at least 4; got " +
               std::to_string(cfg->log2m));
    const uint32_t r = cfg->reg_bits ? cfg->reg_bits : register_size(num_elements);
    if (cfg->log2m > 30) fail(HB_E_INVALID_ARG, "log2m out of range");
    const uint64_t m = uint64_t(1) << cfg->log2m;
    if ((m * r) % 64 != 0)
      fail(HB_E_INVALID_ARG,
           "W should allow counter backends to be aligned. Use " + min_alignment(m * r) +
               " or smaller unsigned integer types");
    int lanes, spl, nseg;
    if (!union_shape(int(r), int(cfg->log2m), &lanes, &spl, &nseg))
      fail(HB_E_INVALID_ARG, "no kernel for register width " + std::to_string(r) + " and log2m " +
                                 std::to_string(cfg->log2m) +
                                 " (this library holds widths 5 and 6, log2m 5..16)");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      fail(HB_E_NO_DEVICE, "no HIP device is visible");
    }
    if (cfg->device < 0 || cfg->device >= ndev)
      fail(HB_E_NO_DEVICE, "device " + std::to_string(cfg->device) + " does not exist");
    hipDeviceProp_t prop;
    HB_HIP(hipGetDeviceProperties(&prop, cfg->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(HB_E_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                               "; this library is built for gfx950 only");

    auto inst = std::make_unique<Instance>();
    Instance &I = *inst;
    I.device = cfg->device;
    I.lanes = uint32_t(lanes);
    I.nseg = uint32_t(nseg);
    I.use_device();
    I.cus = uint32_t(prop.multiProcessorCount);
    if (shared_stream) {
      I.stream = shared_stream;
      I.own_stream = false;
    } else {
      HB_HIP(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
    }
    for (auto &e : I.ev) HB_HIP(hipEventCreate(&e));
    void *hs = nullptr;
    HB_HIP(hipHostMalloc(&hs, sizeof(Stats), hipHostMallocDefault));
    I.h_stats = static_cast<Stats *>(hs);

    I.n = n;
    I.a = graph->num_arcs;
    I.log2m = cfg->log2m;
    I.m = uint32_t(m);
    I.r = r;
    I.W = uint32_t(m * r / 64);
    I.cdw = uint32_t(m * r / 32);
    I.cstride = counter_stride(I.cdw);
    while ((m / 32) >> I.chg_shift > 32) I.chg_shift++;
    I.sentinel = uint64_t(1) << ((1u << r) - 2);
    I.amm = alpha_m_m(cfg->log2m);
    I.hash = cfg->hash;
    I.seed = cfg->seed;
    I.weight_seed = cfg->weight_seed;
    I.weighted = cfg->weights != nullptr;
    I.has_t = transpose != nullptr || cfg->device_transpose != 0;
    I.do_sd = cfg->sum_of_distances != 0;
    I.do_hc = cfg->sum_of_inverse_distances != 0;
    I.acc_mode = cfg->acc_precision;

    hipStream_t s = I.stream;
    upload(I.off, graph->offsets, n + 1, s);
    upload(I.succ, graph->successors, graph->num_arcs, s);
    if (transpose) {
      upload(I.toff, transpose->offsets, n + 1, s);
      upload(I.tsucc, transpose->successors, transpose->num_arcs, s);
    } else if (cfg->device_transpose) {
      I.toff.alloc(n + 1);
      I.tsucc.alloc(graph->num_arcs);
      HB_HIP(launch_transpose(I.off.p, I.succ.p, n, graph->num_arcs, I.toff.p, I.tsucc.p, s));
    }
    std::vector<uint64_t> prefix;
    if (I.weighted) {
      prefix.resize(n + 1, 0);
      for (uint64_t v = 0; v < n; v++) prefix[v + 1] = prefix[v] + cfg->weights[v];
      upload(I.d_weights, cfg->weights, n, s);
      upload(I.d_wprefix, prefix.data(), n + 1, s);
    }
    {
      std::vector<uint32_t> order(n);
      parallel_for((n + Instance::kOrderWindow - 1) / Instance::kOrderWindow,
                   [&](uint64_t b, uint64_t e, unsigned) {
                     for (uint64_t w = b; w < e; w++) {
                       const uint64_t lo = w * Instance::kOrderWindow;
                       const uint64_t hi = std::min(n, lo + Instance::kOrderWindow);
                       for (uint64_t v = lo; v < hi; v++) order[v] = uint32_t(v);
                       std::stable_sort(order.begin() + lo, order.begin() + hi,
                                        [&](uint32_t x, uint32_t y) {
                                          return graph->offsets[x + 1] - graph->offsets[x] >
                                                 graph->offsets[y + 1] - graph->offsets[y];
                                        });
                     }
                   });
      upload(I.d_order, order.data(), n, s);
      HB_HIP(hipStreamSynchronize(s));
    }
    std::vector<double> lc(m + 1, 0.0);
    for (uint64_t z = 1; z <= m; z++) lc[z] = double(m) * std::log(double(m) / double(z));
    upload(I.lc, lc.data(), m + 1, s);

    I.bitmap_words = std::max<uint64_t>(2, ((n + 63) / 64) * 2);
    for (int k = 0; k < 2; k++) {
      I.cnt[k].alloc(size_t(n) * I.cstride, cfg->counter_memory != HB_COUNTERS_DEVICE);
      I.mod[k].alloc(I.bitmap_words);
      I.must[k].alloc(I.bitmap_words);
      I.chg[k].alloc(n);
    }
    // seeded pairs for iteration 0 only while they stay in the last-level
    // cache (random reads of a larger array cost more than hashing again)
    if (!I.weighted && n <= HB_SEED_ITEMS_MAX_NODES) I.seed_items.alloc(n);
    I.cache.alloc(n);
    if (I.do_sd) I.sd.alloc(n * I.acc_bytes());
    if (I.do_hc) I.hc.alloc(n * I.acc_bytes());
    I.list.alloc(n);
    I.bsum.alloc(compact_scratch_words(n));
    I.list_len.alloc(1);
    I.max_grid = std::max<uint32_t>(64, I.cus * 8);
    I.max_grid = (I.max_grid + 7) / 8 * 8;
    {
      std::vector<std::pair<uint64_t, uint32_t>> heavy;
      for (uint64_t v = 0; v < n; v++) {
        const uint64_t d = graph->offsets[v + 1] - graph->offsets[v];
        if (nseg == 1 && d > Instance::kHeavyFloor) heavy.push_back({d, uint32_t(v)});
      }
      std::sort(heavy.begin(), heavy.end(),
                [](const auto &x, const auto &y) { return x.first > y.first || (x.first == y.first && x.second < y.second); });
      for (auto &h : heavy) {
        I.heavy_degs.push_back(h.first);
        I.heavy_nodes.push_back(h.second);
      }
      upload(I.d_heavy, I.heavy_nodes.data(), I.heavy_nodes.size(), s);
      I.d_heavy_sel.alloc(I.heavy_nodes.size());
      I.set_heavy_degree(Instance::kDefaultHeavy);
    }
    I.stats.alloc(1);
    I.stat_slots.alloc(size_t(kStatSlots) * kStatSlotWords);
    I.work.alloc(kWorkCounters * kWorkStride);
    I.counter_memory = cfg->counter_memory;
    if (window_union_supported(int(r), int(cfg->log2m)) &&
        I.counter_memory == HB_COUNTERS_DEVICE && n > 0) {
      I.indeg.alloc(n);
      HB_HIP(launch_fill_u32(I.indeg.p, n, 0u, s));
      HB_HIP(launch_indegree(I.off.p, I.succ.p, n, I.indeg.p, s));
    }
    if (I.counter_memory == HB_COUNTERS_STREAM) {
      const uint64_t node_bytes = uint64_t(I.cstride) * 4;
      uint64_t t = cfg->stream_tile_nodes ? cfg->stream_tile_nodes
                                          : (uint64_t(1) << 30) / node_bytes;
      t = std::min<uint64_t>(std::max<uint64_t>(t, 128), (n + 127) / 128 * 128);
      t = (t + 127) / 128 * 128;
      I.tile_nodes = t;
      I.block_nodes = t / 4;  // a multiple of 32, as the summary kernel needs
      I.tiles = (n + t - 1) / t;
      I.blocks = (n + I.block_nodes - 1) / I.block_nodes;
      for (int k = 0; k < 2; k++) {
        I.s_own[k].alloc(t * I.cstride);
        I.s_part[k].alloc(t * I.cstride);
        I.s_src[k].alloc(I.block_nodes * I.cstride);
      }
      I.s_counts.alloc(I.tiles + I.blocks);
      void *hc = nullptr;
      HB_HIP(hipHostMalloc(&hc, std::max<uint64_t>(1, I.tiles + I.blocks) * 4,
                           hipHostMallocDefault));
      I.h_counts = static_cast<uint32_t *>(hc);
      HB_HIP(hipStreamCreateWithFlags(&I.s_up, hipStreamNonBlocking));
      HB_HIP(hipStreamCreateWithFlags(&I.s_down, hipStreamNonBlocking));
      for (auto &e : I.sev) HB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HB_HIP(hipStreamSynchronize(s));
    return inst;
}

// Staging memory of hb_get_counters_noalloc: a file-scope device array, so
// that entry point allocates nothing.  Each device holds its own copy from the
// first use until the process ends.  Host code uses only the address that
// hipGetSymbolAddress returns for the current device (scratch_address), never
// the array's name.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / 8];

constexpr int kScratchDevices = 64;

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
uint64_t *scratch_address(int device) {
  static std::mutex mu;
  static uint64_t *addr[kScratchDevices] = {};
  if (device < 0 || device >= kScratchDevices)
    fail(HB_E_NO_DEVICE, "device " + std::to_string(device) + " is beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    HB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(HB_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    addr[device] = static_cast<uint64_t *>(p);
  }
  return addr[device];
}

}  // namespace hb

using namespace hb;

extern "C" {

uint32_t hb_register_size(uint64_t n) { return hb::register_size(n); }
double hb_rel_std(uint32_t log2m) { return hb::rel_std(log2m); }
uint32_t hb_log2_num_registers(double rsd) { return hb::log2_num_registers(rsd); }

void hb_config_init(hb_config *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->log2m = 6;
  cfg->hash = HB_HASH_SIPHASH13;
  cfg->arc_granularity = 16 * 1024;
  cfg->acc_precision = HB_ACC_F64;
}

hb_status hb_create(const hb_csr *graph, const hb_csr *transpose, const hb_config *cfg,
                    hb_instance **out) {
  return guard([&] {
    if (!out) fail(HB_E_INVALID_ARG, "null output pointer");
    *out = nullptr;
    *out = reinterpret_cast<hb_instance *>(make_instance(graph, transpose, cfg, nullptr).release());
  });
}

hb_status hb_destroy(hb_instance *hb) {
  return guard([&] {
    Instance *I = reinterpret_cast<Instance *>(hb);
    if (I && I->in_batch) fail(HB_E_INVALID_ARG, "this instance belongs to a batch; destroy the batch");
    delete I;
  });
}

hb_status hb_add_discount_function(hb_instance *hb, hb_discount_fn f, void *ctx) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (!f) fail(HB_E_INVALID_ARG, "null discount function");
    if (I.discounts.size() >= size_t(kMaxDiscount))
      fail(HB_E_INVALID_ARG, "at most " + std::to_string(kMaxDiscount) + " discount functions");
    I.use_device();
    auto buf = std::make_unique<DevBuf<uint8_t>>();
    buf->alloc(I.n * I.acc_bytes());
    I.fill_zero(buf->p, I.n * I.acc_bytes());
    HB_HIP(hipStreamSynchronize(I.stream));
    I.dc.push_back(std::move(buf));
    I.discounts.push_back({f, ctx});
  });
}

hb_status hb_clear_discount_functions(hb_instance *hb) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.use_device();
    HB_HIP(hipStreamSynchronize(I.stream));
    I.discounts.clear();
    I.dc.clear();
  });
}

hb_status hb_run(hb_instance *hb, uint64_t upper_bound, double threshold) {
  return guard([&] { checked(hb)->run(upper_bound, threshold); });
}

hb_status hb_run_until_stable(hb_instance *hb, uint64_t upper_bound) {
  return guard([&] { checked(hb)->run(upper_bound, NAN); });
}

hb_status hb_run_until_done(hb_instance *hb) {
  return guard([&] { checked(hb)->run(UINT64_MAX, NAN); });
}

hb_status hb_init(hb_instance *hb) {
  return guard([&] { checked(hb)->init(); });
}

hb_status hb_iterate(hb_instance *hb) {
  return guard([&] { checked(hb)->iterate(); });
}

hb_status hb_neighborhood_function(hb_instance *hb, double *out, size_t cap, size_t *len) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    if (len) *len = I.nf.size();
    if (out) copy_out(I.nf, out, cap);
  });
}

hb_status hb_sum_of_distances(hb_instance *hb, double *out, size_t n) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    if (!I.do_sd)
      fail(HB_E_NOT_REQUESTED,
           "Sum of distances were not requested. Use builder.with_sum_of_distances(true) while "
           "building HyperBall to compute them");
    copy_out(I.read_acc(I.sd), out, n);
  });
}

hb_status hb_harmonic_centralities(hb_instance *hb, double *out, size_t n) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    if (!I.do_hc)
      fail(HB_E_NOT_REQUESTED,
           "Sum of inverse distances were not requested. Use "
           "builder.with_sum_of_inverse_distances(true) while building HyperBall to compute them");
    copy_out(I.read_acc(I.hc), out, n);
  });
}

hb_status hb_discounted_centrality(hb_instance *hb, size_t index, double *out, size_t n) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    if (index >= I.dc.size())
      fail(HB_E_NOT_REQUESTED, "Discount centrality of index " + std::to_string(index) +
                                   " does not exist");
    copy_out(I.read_acc(*I.dc[index]), out, n);
  });
}

static void derived(hb_instance *hb, int which, double *out, size_t cap) {
  Instance &I = *checked(hb);
  I.ensure_iteration();
  if (!I.do_sd) {
    static const char *what[3] = {"closeness centrality", "lin centrality", "lin centrality"};
    fail(HB_E_NOT_REQUESTED,
         std::string("Sum of distances were not requested. Use builder.with_sum_of_distances(true) "
                     "while building HyperBall to compute ") + what[which]);
  }
  std::vector<double> d = I.read_acc(I.sd);
  std::vector<double> c;
  if (which != 0) c = I.read_counts();
  for (uint64_t v = 0; v < I.n; v++) {
    if (which == 0) d[v] = d[v] == 0.0 ? 0.0 : 1.0 / d[v];
    else if (which == 1) d[v] = d[v] == 0.0 ? 1.0 : c[v] * c[v] / d[v];
    else d[v] = c[v] * c[v] - d[v];
  }
  copy_out(d, out, cap);
}

hb_status hb_closeness_centrality(hb_instance *hb, double *out, size_t n) {
  return guard([&] { derived(hb, 0, out, n); });
}
hb_status hb_lin_centrality(hb_instance *hb, double *out, size_t n) {
  return guard([&] { derived(hb, 1, out, n); });
}
hb_status hb_nieminen_centrality(hb_instance *hb, double *out, size_t n) {
  return guard([&] { derived(hb, 2, out, n); });
}

hb_status hb_reachable_nodes(hb_instance *hb, double *out, size_t n) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    copy_out(I.read_counts(), out, n);
  });
}

hb_status hb_reachable_nodes_from(hb_instance *hb, uint64_t node, double *out) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.ensure_iteration();
    if (!out) fail(HB_E_INVALID_ARG, "null output");
    if (node >= I.n) fail(HB_E_INVALID_ARG, "node " + std::to_string(node) + " out of range");
    I.use_device();
    HB_HIP(hipMemcpyAsync(out, I.cache.p + node, 8, hipMemcpyDeviceToHost, I.stream));
    HB_HIP(hipStreamSynchronize(I.stream));
  });
}

hb_status hb_get_params(hb_instance *hb, hb_params *out) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (!out) fail(HB_E_INVALID_ARG, "null output");
    out->log2m = I.log2m;
    out->num_registers = I.m;
    out->register_bits = I.r;
    out->words_per_counter = I.W;
    out->num_nodes = I.n;
    out->num_arcs = I.a;
    out->alpha_m_m = I.amm;
    out->counter_memory = I.counter_memory;
  });
}

// Counters in the reference layout, converted on the device in chunks of at
// most `stage_words` words through `stage` and copied to `out`.
static void copy_counters(Instance &I, int which, uint64_t *out, size_t words, uint64_t *stage,
                          uint64_t stage_words) {
  if (!out) fail(HB_E_INVALID_ARG, "null output");
  if (!I.inited) fail(HB_E_NOT_RUN, "counters are not initialised; call hb_init or hb_run");
  if (words < I.n * I.W) fail(HB_E_INVALID_ARG, "output buffer too small");
  if (which && I.iteration == 0) {  // the other buffer before the first iteration
    std::memset(out, 0, I.n * I.W * sizeof(uint64_t));
    return;
  }
  const uint64_t chunk = stage_words / I.W;
  if (chunk == 0)
    fail(HB_E_INVALID_ARG, "a counter of " + std::to_string(I.W * 8) +
                               " bytes does not fit the staging buffer");
  const uint32_t *src = I.cnt[which ? 1 - I.cur : I.cur].p;
  for (uint64_t b = 0; b < I.n; b += chunk) {
    const uint64_t c = std::min(chunk, I.n - b);
    HB_HIP(launch_to_reference_layout(src, b, c, I.r, I.m, stage, I.stream));
    HB_HIP(hipMemcpyAsync(out + b * I.W, stage, c * I.W * 8, hipMemcpyDeviceToHost, I.stream));
    HB_HIP(hipStreamSynchronize(I.stream));
  }
}

hb_status hb_get_counters(hb_instance *hb, int which, uint64_t *out, size_t words) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.use_device();
    const uint64_t stage_words = std::min<uint64_t>(
        std::max<uint64_t>(I.n, 1) * I.W, std::max<uint64_t>((uint64_t(64) << 20) / 8, I.W));
    DevBuf<uint64_t> tmp;
    if (out && I.inited && words >= I.n * I.W && !(which && I.iteration == 0))
      tmp.alloc(stage_words);
    copy_counters(I, which, out, words, tmp.p, stage_words);
  });
}

hb_status hb_get_counters_noalloc(hb_instance *hb, int which, uint64_t *out, size_t words) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.use_device();
    copy_counters(I, which, out, words, scratch_address(I.device), kNoAllocScratchBytes / 8);
  });
}

hb_status hb_get_modified(hb_instance *hb, uint64_t *out, size_t words) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (!out) fail(HB_E_INVALID_ARG, "null output");
    if (!I.inited) fail(HB_E_NOT_RUN, "not initialised; call hb_init or hb_run");
    const uint64_t need = (I.n + 63) / 64;
    if (words < need) fail(HB_E_INVALID_ARG, "output buffer too small");
    I.use_device();
    if (need) {
      HB_HIP(hipMemcpyAsync(out, I.mod[I.mcur].p, need * 8, hipMemcpyDeviceToHost, I.stream));
      HB_HIP(hipStreamSynchronize(I.stream));
    }
  });
}

hb_status hb_num_iterations(hb_instance *hb, uint64_t *out) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (!out) fail(HB_E_INVALID_ARG, "null output");
    *out = I.history.size();
  });
}

hb_status hb_get_iteration_stats(hb_instance *hb, uint64_t iteration, hb_iteration_stats *out) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (!out) fail(HB_E_INVALID_ARG, "null output");
    if (iteration >= I.history.size()) fail(HB_E_INVALID_ARG, "no such iteration");
    *out = I.history[iteration];
  });
}

hb_status hb_set_progress(hb_instance *hb, hb_progress_fn f, void *ctx) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.progress = f;
    I.progress_ctx = ctx;
  });
}

hb_status hb_set_stream(hb_instance *hb, void *stream) {
  return guard([&] {
    Instance &I = *checked(hb);
    if (I.in_batch) fail(HB_E_INVALID_ARG, "the instances of a batch share the batch's stream");
    I.use_device();
    HB_HIP(hipStreamSynchronize(I.stream));
    if (I.own_stream) HB_HIP(hipStreamDestroy(I.stream));
    I.stream = nullptr;
    if (stream) {
      I.stream = static_cast<hipStream_t>(stream);
      I.own_stream = false;
    } else {
      HB_HIP(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
      I.own_stream = true;
    }
  });
}

hb_status hb_set_timing(hb_instance *hb, int enabled) {
  return guard([&] { checked(hb)->timing = enabled != 0; });
}

hb_status hb_set_tuning(hb_instance *hb, uint32_t heavy_degree) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.set_heavy_degree(heavy_degree == 0 ? Instance::kDefaultHeavy
                       : heavy_degree == UINT32_MAX ? UINT64_MAX : heavy_degree);
  });
}

hb_status hb_set_delta_gathering(hb_instance *hb, int mode) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail(HB_E_INVALID_ARG, "delta gathering mode must be 0, 1 or 2");
    checked(hb)->delta_mode = mode;
  });
}

hb_status hb_set_window(hb_instance *hb, int mode, uint32_t block_nodes, uint32_t half_width,
                        uint32_t heavy_degree) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail(HB_E_INVALID_ARG, "window mode must be 0, 1 or 2");
    Instance &I = *checked(hb);
    I.window_mode = mode;
    I.win_block = block_nodes;
    I.win_half = half_width;
    I.win_heavy = heavy_degree == 0 ? 0
                  : heavy_degree == UINT32_MAX ? UINT64_MAX
                                               : std::max<uint64_t>(heavy_degree, Instance::kHeavyFloor);
  });
}

hb_status hb_set_streaming(hb_instance *hb, int mode) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail(HB_E_INVALID_ARG, "streaming mode must be 0, 1 or 2");
    checked(hb)->stream_mode = mode;
  });
}

hb_status hb_autotune(hb_instance *hb, uint32_t reps, uint32_t *thresholds, double *ms,
                      size_t cap, size_t *len, uint32_t *chosen) {
  return guard([&] {
    Instance &I = *checked(hb);
    I.use_device();
    const auto table = I.autotune(reps);
    if (len) *len = table.size();
    for (size_t k = 0; k < table.size() && k < cap; k++) {
      if (thresholds)
        thresholds[k] = table[k].first == UINT64_MAX ? UINT32_MAX : uint32_t(table[k].first);
      if (ms) ms[k] = table[k].second;
    }
    if (chosen) *chosen = I.heavy_degree == UINT64_MAX ? UINT32_MAX : uint32_t(I.heavy_degree);
  });
}

hb_status hb_graph_transpose_device(const hb_csr *g, int device, hb_graph *out) {
  return guard([&] {
    if (!out) fail(HB_E_INVALID_ARG, "null output graph");
    std::memset(out, 0, sizeof(*out));
    check_csr(g, "graph");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      fail(HB_E_NO_DEVICE, "no HIP device is visible");
    }
    if (device < 0 || device >= ndev)
      fail(HB_E_NO_DEVICE, "device " + std::to_string(device) + " does not exist");
    HB_HIP(hipSetDevice(device));
    const uint64_t n = g->num_nodes, a = g->num_arcs;
    hipStream_t s = nullptr;
    HB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DevBuf<uint64_t> off, toff;
    DevBuf<uint32_t> succ, tsucc;
    upload(off, g->offsets, n + 1, s);
    upload(succ, g->successors, a, s);
    toff.alloc(n + 1);
    tsucc.alloc(a);
    HB_HIP(launch_transpose(off.p, succ.p, n, a, toff.p, tsucc.p, s));
    auto *ho = static_cast<uint64_t *>(std::malloc((n + 1) * sizeof(uint64_t)));
    auto *hs = static_cast<uint32_t *>(std::malloc(std::max<uint64_t>(1, a) * sizeof(uint32_t)));
    if (!ho || !hs) {
      std::free(ho);
      std::free(hs);
      fail(HB_E_OOM, "host allocation of the transpose failed");
    }
    out->num_nodes = n;
    out->num_arcs = a;
    out->offsets = ho;
    out->successors = hs;
    HB_HIP(hipMemcpyAsync(ho, toff.p, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    if (a) HB_HIP(hipMemcpyAsync(hs, tsucc.p, a * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HB_HIP(hipStreamSynchronize(s));
  });
}

// ---- batches ---------------------------------------------------------------

hb_status hb_batch_create(size_t k, const hb_csr *graphs, const hb_csr *transposes,
                          const hb_config *cfgs, hb_batch **out) {
  return guard([&] {
    if (!out) fail(HB_E_INVALID_ARG, "null output pointer");
    *out = nullptr;
    if (k == 0 || !graphs || !cfgs) fail(HB_E_INVALID_ARG, "a batch needs k >= 1 graphs and configurations");
    auto b = std::make_unique<Batch>();
    b->device = cfgs[0].device;
    for (size_t i = 0; i < k; i++) {
      if (cfgs[i].device != b->device) fail(HB_E_INVALID_ARG, "all instances of a batch must use one device");
      if (cfgs[i].counter_memory == HB_COUNTERS_STREAM)
        fail(HB_E_INVALID_ARG, "instance " + std::to_string(i) +
                                   ": streamed counters are not available in a batch");
      const uint32_t r = cfgs[i].reg_bits ? cfgs[i].reg_bits : 5u;
      if (batch_shape(int(r), int(cfgs[i].log2m)) < 0)
        fail(HB_E_INVALID_ARG, "instance " + std::to_string(i) +
                                   ": a batch holds log2m 6, 8 or 10 with 5- or 6-bit registers");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      (void)hipGetLastError();
      fail(HB_E_NO_DEVICE, "no HIP device is visible");
    }
    HB_HIP(hipSetDevice(b->device));
    HB_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto &e : b->ev) HB_HIP(hipEventCreate(&e));
    uint64_t max_words = 1, max_nodes = 1;
    for (size_t i = 0; i < k; i++) {
      const hb_csr *t = (transposes && transposes[i].offsets) ? &transposes[i] : nullptr;
      auto I = make_instance(&graphs[i], t, &cfgs[i], b->stream);
      I->in_batch = true;
      max_words = std::max<uint64_t>(max_words, I->bitmap_words);
      max_nodes = std::max<uint64_t>(max_nodes, I->n);
      b->inst.push_back(std::move(I));
    }
    void *p = nullptr;
    HB_HIP(hipHostMalloc(&p, k * sizeof(BatchDesc), hipHostMallocDefault));
    b->h_desc = static_cast<BatchDesc *>(p);
    std::memset(b->h_desc, 0, k * sizeof(BatchDesc));
    HB_HIP(hipHostMalloc(&p, k * sizeof(Stats), hipHostMallocDefault));
    b->h_stats = static_cast<Stats *>(p);
    b->d_desc.alloc(k);
    b->d_stats.alloc(k);
    b->slots.alloc(k * size_t(kStatSlots) * kStatSlotWords);
    b->work.alloc(kBatchWorkWords);
    b->total_chunks.alloc(kBatchShapes);
    b->total_window.alloc(2 * kBatchShapes);
    for (size_t i = 0; i < k; i++) {
      Instance &I = *b->inst[i];
      BatchDesc &d = b->h_desc[i];
      d.shape = uint32_t(batch_shape(int(I.r), int(I.log2m)));
      d.words = I.bitmap_words;
      d.n = I.n;
      d.bsum = I.bsum.p;
    }
    // group the instances by shape, in instance order within a shape
    std::vector<uint32_t> off(kBatchShapes + 1, 0);
    for (size_t i = 0; i < k; i++) off[b->h_desc[i].shape + 1]++;
    for (int sh = 0; sh < kBatchShapes; sh++) off[sh + 1] += off[sh];
    b->h_shape_inst.assign(k, 0);
    {
      std::vector<uint32_t> pos(off.begin(), off.end() - 1);
      for (size_t i = 0; i < k; i++) b->h_shape_inst[pos[b->h_desc[i].shape]++] = uint32_t(i);
    }
    upload(b->shape_inst, b->h_shape_inst.data(), k, b->stream);
    upload(b->shape_off, off.data(), off.size(), b->stream);
    HB_HIP(hipStreamSynchronize(b->stream));
    BatchLaunch &L = b->L;
    L.d = b->d_desc.p;
    L.k = int(k);
    L.stats = b->d_stats.p;
    L.work = b->work.p;
    L.total_chunks = b->total_chunks.p;
    L.total_window = b->total_window.p;
    L.shape_inst = b->shape_inst.p;
    L.shape_off = b->shape_off.p;
    for (int sh = 0; sh < kBatchShapes; sh++) {
      L.first[sh] = off[sh];
      L.count[sh] = off[sh + 1] - off[sh];
    }
    L.prep_blocks = uint32_t(std::min<uint64_t>(64, (max_words + kBlock - 1) / kBlock));
    L.compact_blocks = uint32_t(std::max<uint64_t>(1, ((max_nodes + 31) / 32 + kBlock - 1) / kBlock));
    L.union_grid = b->inst[0]->max_grid;
    *out = reinterpret_cast<hb_batch *>(b.release());
  });
}

hb_status hb_batch_destroy(hb_batch *b) {
  return guard([&] { delete reinterpret_cast<Batch *>(b); });
}

static Batch *checked_batch(hb_batch *b) {
  if (!b) fail(HB_E_INVALID_ARG, "null batch");
  return reinterpret_cast<Batch *>(b);
}

hb_status hb_batch_instance(hb_batch *b, size_t i, hb_instance **out) {
  return guard([&] {
    Batch &B = *checked_batch(b);
    if (!out) fail(HB_E_INVALID_ARG, "null output pointer");
    if (i >= B.inst.size()) fail(HB_E_INVALID_ARG, "no such instance");
    *out = reinterpret_cast<hb_instance *>(B.inst[i].get());
  });
}

hb_status hb_batch_run(hb_batch *b, uint64_t upper_bound, double threshold) {
  return guard([&] { checked_batch(b)->run(upper_bound, threshold); });
}

hb_status hb_batch_run_until_done(hb_batch *b) {
  return guard([&] { checked_batch(b)->run(UINT64_MAX, NAN); });
}

hb_status hb_batch_autotune(hb_batch *b, uint32_t reps, uint32_t *thresholds, double *ms,
                            size_t cap, size_t *len, uint32_t *chosen) {
  return guard([&] {
    Batch &B = *checked_batch(b);
    HB_HIP(hipSetDevice(B.device));
    const auto table = B.autotune(reps);
    if (len) *len = table.size();
    for (size_t k = 0; k < table.size() && k < cap; k++) {
      if (thresholds)
        thresholds[k] = table[k].first == UINT64_MAX ? UINT32_MAX : uint32_t(table[k].first);
      if (ms) ms[k] = table[k].second;
    }
    const uint64_t h = B.inst.front()->heavy_degree;
    if (chosen) *chosen = h == UINT64_MAX ? UINT32_MAX : uint32_t(h);
  });
}

hb_status hb_batch_set_graph(hb_batch *b, int enabled) {
  return guard([&] {
    Batch &B = *checked_batch(b);
    B.use_graph = enabled != 0;
  });
}

hb_status hb_batch_step_times(hb_batch *b, double *out, size_t cap, size_t *len) {
  return guard([&] {
    Batch &B = *checked_batch(b);
    if (len) *len = B.step_ms.size();
    if (out) {
      if (cap < B.step_ms.size()) fail(HB_E_INVALID_ARG, "output buffer too small");
      std::memcpy(out, B.step_ms.data(), B.step_ms.size() * sizeof(double));
    }
  });
}

}  // extern "C"

