// Kernel launch interface of the HyperBall library (device side in
// hb_kernels.hip).
//
// Device counter layout ("slab" layout).  A counter of m registers of r bits
// is m/32 slabs; slab s holds registers 32s..32s+31 as r bit planes of 32
// bits (dword s*r + b holds bit b of each of the 32 registers, register
// 32s+q at bit q).  A counter is therefore m*r/32 dwords = exactly the bytes
// of the reference's packed layout (hyper_log_log_logic.rs:72-115), with the
// bits permuted so that a per-register maximum is a handful of bitwise
// operations per dword instead of the reference's broadword borrow chains
// (hyper_log_log_logic.rs:469-576).  kernel_to_reference_layout converts back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hb {

constexpr int kBlock = 256;

// Distance in dwords between consecutive counters of cdw = m*r/32 dwords.
// Counters sit on whole 64-byte sectors of HBM: below 16 dwords the stride is
// the next power of two (a counter never straddles a sector), from 16 dwords
// on a multiple of 16.  Random counter reads then move whole sectors only;
// unpadded 40- and 48-byte counters straddle two sectors most of the time.
__host__ __device__ constexpr uint32_t counter_stride(uint32_t cdw) {
  if (cdw >= 16) return (cdw + 15) / 16 * 16;
  uint32_t s = 1;
  while (s < cdw) s <<= 1;
  return s;
}
constexpr int kMaxDiscount = 8;

enum UnionFlags : uint32_t {
  F_SYSTOLIC = 1u,    // accumulate the neighbourhood function as post - pre; mark predecessors
  F_CHECK_ALL = 2u,   // every processed node enumerates its successors
  F_CENTRALITY = 4u,  // update the per-node accumulators
  F_LIST = 8u,        // nodes come from a compacted list whose length lives on the device
  F_DELTA = 16u,      // gather only the changed slabs of successors (change masks)
  F_SEEDED = 32u,     // counters are as seeded without weights: each holds the one register
                      // of its node's hash, taken from seed_items (or hashed) instead of read
  F_UNTESTED = 64u,   // window kernel: successors outside the tile (and those of rows that
                      // take a workgroup) are merged without testing their modified bits
};

enum AccMode : uint32_t { ACC_F64 = 0, ACC_F32_KAHAN = 1, ACC_F32 = 2 };

// Results of one iteration, accumulated with integer atomics so that they do
// not depend on the schedule.  The neighbourhood-function contribution is a
// fixed-point sum with 2^-64 resolution: a term x adds w = floor(x) to nf_int,
// the first 32 fraction bits of x - w to nf_hi and the next 32 (rounded) to
// nf_lo.  Integer sums are exact and order-independent, so the value is
// deterministic run to run and within n * 2^-65 of the exact sum.
struct Stats {
  int64_t nf_int, nf_hi, nf_lo;
  uint64_t visited, scanned, tscanned, checked, copied, modified;
  uint64_t gathered;  // successor slabs read
  uint64_t changed;   // set change-mask bits over the modified nodes
  uint64_t list_len;
};

// Statistics slots.  A kernel's wavefronts add their partial statistics (the
// first 11 Stats fields) to slot (global wavefront index % kStatSlots), one
// 128-byte line each, instead of all to one Stats: thousands of wavefronts
// end at nearly the same time, and their atomics on a single line serialize.
// launch_reduce_stats then writes the sums into the Stats.
constexpr int kStatSlots = 256;
constexpr int kStatSlotWords = 16;  // u64 per slot
constexpr int kStatFields = 11;

// Dynamic work counters, each on its own 256-byte line: [0, 8) one per XCD,
// [8] heavy rows, [9] heavy rows selected for a worklist iteration.
constexpr int kWorkStride = 64;
constexpr int kWorkCounters = 10;
constexpr int kHeavySelCounter = 9;  // select_heavy_kernel's count of selected heavy rows

inline double stats_nf(const Stats &s) {
  const int64_t hi = s.nf_hi + (s.nf_lo >> 32);
  const int64_t lo = s.nf_lo & 0xffffffffll;
  const int64_t whole = s.nf_int + (hi >> 32);
  const int64_t frac = hi & 0xffffffffll;
  return double(whole) + (double(frac) + double(lo) * 0x1p-32) * 0x1p-32;
}

struct UnionArgs {
  const uint64_t *off;
  const uint32_t *succ;
  const uint64_t *toff;
  const uint32_t *tsucc;
  const uint32_t *curr;  // counters, slab layout
  uint32_t *next;
  const uint32_t *curr_mod;
  uint32_t *next_mod;
  const uint32_t *must;
  uint32_t *next_must;
  double *cache;  // count() of every current counter
  const uint32_t *list;
  const uint32_t *list_len;
  const uint32_t *order;  // processing order when not in list mode (nullable: identity)
  uint64_t count;  // number of nodes when not in list mode
  uint32_t flags;
  uint32_t acc_mode;
  void *sd;
  void *hc;
  void *dc[kMaxDiscount];
  double disc[kMaxDiscount];
  int ndisc;
  double dist;  // iteration + 1
  double alpha_m_m;
  uint32_t m;
  const double *lc_table;  // m * ln(m / z) for z = 0..m
  // F_SEEDED: each node's seeded register, index | value << 16, or null to
  // hash the node ids again (hash, seed, log2m and sentinel as in SeedArgs)
  const uint32_t *seed_items;
  uint32_t log2m, hash;
  uint64_t seed, sentinel;
  Stats *stats;            // zeroed before the launch
  uint64_t *stat_slots;    // kStatSlots * kStatSlotWords, zeroed before the launch; null:
                           // wavefronts add to *stats directly
  uint32_t *work;          // kWorkCounters * kWorkStride words, zeroed before the launch
  uint64_t heavy_degree;         // rows longer than this take a workgroup each (heavy_phase)
  const uint32_t *heavy_list;    // the nodes with such rows
  uint64_t heavy_count;
  // nullable: the number of rows in heavy_list lives here on the device
  // (worklist iterations: select_heavy_kernel keeps the rows with work)
  const uint32_t *heavy_len;
  // Change masks: bit (slab >> chg_shift) of chg_cur[s] is set when a
  // register of that slab of node s changed in the last iteration (or, in
  // iteration 0, is not zero).  Read for modified successors, written for
  // modified nodes.
  const uint32_t *chg_cur;
  uint32_t *chg_next;
  uint32_t chg_shift;
  // Window kernel (launch_window_union): nodes per block and the half-width
  // of the counter tile around a block, at most the kernel's own (0: those);
  // indeg[s]: arcs into s from other nodes (self loops excluded).
  uint32_t win_block, win_half;
  const uint32_t *indeg;
};

// Streamed counters (counter arrays in host memory, worked on tile by tile
// in device memory).  One pass merges, into the partial next counters of the
// tile's checked nodes, their modified successors inside one source range
// [lo, hi) whose counters sit in `src`; the first pass of a tile also
// initialises the partial counters from the tile's own counters.
struct StreamPass {
  const uint32_t *src;  // counters of nodes lo .. hi - 1, at the counter stride
  uint64_t lo, hi;
  const uint32_t *own;  // current counters of the tile's nodes
  uint32_t *part;       // partial next counters of the tile's nodes
  uint64_t tile_lo, tile_hi;
  uint32_t first;
};

struct SeedArgs {
  uint32_t *counters;  // zeroed, slab layout
  uint64_t n;
  uint32_t cdw;        // dwords between counters (counter_stride)
  uint32_t r, log2m;
  uint64_t sentinel;
  uint32_t hash;
  uint64_t seed;
  const uint64_t *weights;       // nullable
  const uint64_t *weight_prefix; // exclusive prefix sums of weights
  uint64_t weight_seed;
  uint32_t *chg;        // per node: the slabs holding a non-zero register
  uint32_t chg_shift;
  uint32_t *items;      // nullable (seed_count only): each node's register, index | value << 16
};

// ---- batched execution --------------------------------------------------
// A batch runs many independent instances through one launch per step.  Each
// instance is described by a BatchDesc; kernels find the instance of a work
// item from prefix sums in the table and branch on its shape (uniform per
// wavefront).  Shapes in a batch: (r, log2m) in {5, 6} x {6, 8, 10}.
constexpr int kBatchShapes = 6;
int batch_shape(int r, int log2m);  // index, or -1 when a batch cannot hold it

struct BatchDesc {
  UnionArgs u;
  uint32_t shape, active, list_mode, clear_must;
  uint64_t words;  // 32-bit words of each bitmap
  uint64_t n;
  uint32_t *bsum;  // compaction scratch
  uint64_t chunk_base, chunk_end;  // in the shape's chunk space, written by the plan step
  uint64_t heavy_base, heavy_end;  // in the shape's heavy-row space
  uint32_t rounds;                 // rounds per chunk of the shape, written by the plan step
  uint32_t mark_first;             // mark the predecessors of the modified nodes into must first
  // This iteration runs as the window kernel (64-register counters, every
  // node checked; u carries its settings): the instance takes no chunks and
  // no heavy rows of batch_union_kernel, but blocks and heavy rows of the
  // shape's window spaces, written by the plan step.
  uint32_t window;
  uint64_t wblock_base, wblock_end;
  uint64_t wheavy_base, wheavy_end;
};

// Queue counters of a batch: one region of kWorkCounters per shape for the
// union kernels, then one per shape for the window kernels.
constexpr int kBatchWorkWords = 2 * kBatchShapes * kWorkCounters * kWorkStride;

// One batched iteration.  Instances are grouped by shape (`shape_inst`,
// `shape_off` on the device; `first`/`count` on the host): each shape has its
// own chunk space, heavy-row space and queue counters, and its own heavy and
// union kernels, compiled for that shape alone.
struct BatchLaunch {
  BatchDesc *d;
  int k;
  Stats *stats;
  uint32_t *work;           // kBatchWorkWords
  uint64_t *total_chunks;   // kBatchShapes, written by the plan step
  uint64_t *total_window;   // 2 * kBatchShapes (blocks, heavy rows), written by the plan step
  const uint32_t *shape_inst;
  const uint32_t *shape_off;  // kBatchShapes + 1
  uint32_t first[kBatchShapes], count[kBatchShapes];
  uint32_t prep_blocks, compact_blocks, union_grid;
  uint64_t total_heavy[kBatchShapes];
};

// Launches one batched iteration on s: clear the next bitmaps and statistics,
// compact the worklists of list-mode instances, plan the chunk spaces, then
// per shape the union kernel and, for 64-register shapes, the window kernel.
hipError_t launch_batch_iteration(const BatchLaunch &L, hipStream_t s);

// Chooses the lane group of a (r, log2m) shape: lanes per node, slabs per
// lane and segments per counter (a group works on one segment of lanes *
// slabs_per_lane slabs at a time).  Returns false for shapes without a
// kernel.
bool union_shape(int r, int log2m, int *lanes, int *slabs_per_lane, int *segments);

// Launches union_kernel (grid workgroups, capped at the resident ones): the
// heavy rows of a.heavy_list first, a workgroup each, then the chunks.
hipError_t launch_union(int r, int log2m, const UnionArgs &a, uint32_t grid, hipStream_t s);

// The window kernel: the union step of an iteration that checks every node
// (F_CHECK_ALL, not F_LIST, F_SEEDED or F_DELTA) for 64-register counters.
// A workgroup takes blocks of consecutive nodes and first copies the current
// counters of the nodes around its block into LDS; successors there are
// merged from LDS, the others read from a.curr.  Rows longer than
// a.heavy_degree (a.heavy_list, a.heavy_count) take the whole workgroup each,
// first.  window_union_supported tells the shapes it has; window_tile_nodes
// the counters its tile holds (block + both halves).
bool window_union_supported(int r, int log2m);
uint32_t window_block_nodes();
uint32_t window_tile_nodes(int r);
hipError_t launch_window_union(int r, int log2m, const UnionArgs &a, hipStream_t s);
// indeg[s] <- the arcs (v, s) with v != s of the CSR graph (off, succ); indeg
// must be zeroed.
hipError_t launch_indegree(const uint64_t *off, const uint32_t *succ, uint64_t n, uint32_t *indeg,
                           hipStream_t s);
hipError_t launch_count_all(int r, int log2m, const uint32_t *counters, uint64_t n,
                            double alpha_m_m, uint32_t m, const double *lc_table, double *out,
                            hipStream_t s);
hipError_t launch_seed(const SeedArgs &a, hipStream_t s);
// Seeding without weights (a.weights null) into a counter array that need not
// be zeroed, fused with count() of every counter into out (lc[m - 1], the
// linear-counting value of a counter with one non-zero register) and the
// change masks.
hipError_t launch_seed_count(const SeedArgs &a, uint32_t m, const double *lc, double *out,
                             hipStream_t s);
// Start of an iteration, one kernel (not memsets, so the sequence can be
// captured in a graph): zeroes next_mod and next_must (nullable), `words`
// words each, *stats, the work counters and the statistics slots (nullable).
hipError_t launch_prepare_iteration(uint32_t *next_mod, uint32_t *next_must, uint64_t words,
                                    Stats *stats, uint32_t *work, uint64_t *slots, hipStream_t s);
// *out <- the sums of the statistics slots (the first kStatFields fields) and
// stats->list_len.  out may be pinned host memory.
hipError_t launch_reduce_stats(const Stats *stats, const uint64_t *slots, Stats *out,
                               hipStream_t s);
// list <- ids of the set bits of (a | b) below n, ascending; *len <- count.
// Worklist iterations: the heavy rows to check or to copy (must or modified
// bit set), appended to out in any order, their number at *out_len (zeroed
// before the launch).
hipError_t launch_select_heavy(const uint32_t *heavy_list, uint64_t heavy_count,
                               const uint32_t *must, const uint32_t *curr_mod, uint32_t *out,
                               uint32_t *out_len, hipStream_t s);
hipError_t launch_compact(const uint32_t *a, const uint32_t *b, uint64_t n, uint32_t *block_sums,
                          uint32_t *list, uint32_t *len, hipStream_t s);
uint64_t compact_scratch_words(uint64_t n);
// must |= the predecessors (rows of the transpose toff/tsucc) of the nodes
// list[0, *len); grid workgroups.
hipError_t launch_mark_predecessors(const uint32_t *list, const uint32_t *len, const uint64_t *toff,
                                   const uint32_t *tsucc, uint32_t *must, uint32_t grid,
                                   hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t s);
// Streamed counters.  launch_stream_merge: one StreamPass over the tile's
// checked nodes.  launch_stream_finish: the rest of the per-node body for
// every node of the tile (changed test, count(), bookkeeping) with the
// complete next counters of the tile left in p.part (unchecked nodes get a
// copy of their own counters).  launch_stream_summary: counts[t] = nodes of
// tile t that are checked or were modified, counts[tiles + b] = modified nodes
// of block b (tiles of tile_nodes nodes, blocks of block_nodes nodes); counts
// must be zeroed.
hipError_t launch_stream_merge(int r, int log2m, const UnionArgs &a, const StreamPass &p,
                               hipStream_t s);
hipError_t launch_stream_finish(int r, int log2m, const UnionArgs &a, const StreamPass &p,
                                hipStream_t s);
hipError_t launch_stream_summary(const UnionArgs &a, uint64_t n, uint64_t tile_nodes,
                                 uint64_t block_nodes, uint64_t tiles, uint32_t *counts,
                                 hipStream_t s);
// Transpose of the CSR graph (off[n+1], succ[a]) into caller-allocated
// toff[n+1], tsucc[a], rows sorted (hb_transpose.hip).  Allocates and frees
// its own temporaries (about 8 bytes per arc plus the sort's workspace) and
// returns after the stream has finished them.
hipError_t launch_transpose(const uint64_t *off, const uint32_t *succ, uint64_t n, uint64_t a,
                            uint64_t *toff, uint32_t *tsucc, hipStream_t s);
// Sets bits [0, n) and clears the tail of the last word.
hipError_t launch_fill_bits(uint32_t *p, uint64_t n, hipStream_t s);
hipError_t launch_to_reference_layout(const uint32_t *counters, uint64_t node_begin,
                                      uint64_t node_count, uint32_t r, uint32_t m,
                                      uint64_t *out, hipStream_t s);

}  // namespace hb
