// HyperBall kernels for gfx950 (MI355X, CDNA4, wave64).
//
// Hot path: union_kernel, one iteration of HyperBall over the nodes it is
// given (all nodes, or a compacted worklist).  It restates the per-node body
// of the reference's parallel_task (src/algo/hyperball/hyperball_impl.rs:
// 1025-1145) on a lane group per node:
//   - the group reads G successor ids at a time, tests their modified bits,
//     and a group ballot yields the modified successors (hyperball_impl.rs:
//     1037-1049);
//   - each lane owns SPL slabs of the node's counter and merges the same
//     slabs of every modified successor (per-register maximum, the
//     definition of merge_hyperloglog_bitwise, hyper_log_log_logic.rs:
//     469-576), up to four successor loads in flight;
//   - the changed test, count(), the neighbourhood-function contribution,
//     centrality updates, modified bit and predecessor marking follow
//     :1052-1130; the new counter is stored when it changed or when the other
//     buffer holds a stale copy (:1132-1144).
// count() results are cached per node (count of the current counter), so a
// counter is counted once per change, and the pre-change count (:1066) is a
// read of the cache.
#include "hb_kernels.hpp"

#include <algorithm>
#include <atomic>

#include "hb_hll.hpp"

// Lane group per (r, log2m): G lanes per node and SPL slabs per lane, with
// G * SPL = m / 32.  HB_G_<log2m> / HB_SPL_<log2m> override a row at build
// time (the tuning sweeps use this).
#ifndef HB_G_10
#define HB_G_10 16
#define HB_SPL_10 2
#endif
#ifndef HB_G_8
#define HB_G_8 8
#define HB_SPL_8 1
#endif
#ifndef HB_G_6
#define HB_G_6 2
#define HB_SPL_6 1
#endif

namespace hb {
namespace {

__device__ __forceinline__ bool test_bit(const uint32_t *bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

template <int G>
__device__ __forceinline__ uint64_t group_ballot(bool pred) {
  uint64_t b = __ballot(pred);
  if constexpr (G == 64) {
    return b;
  } else {
    const int base = __lane_id() & ~(G - 1);
    return (b >> base) & ((uint64_t(1) << G) - 1);
  }
}

template <int G>
__device__ __forceinline__ uint64_t group_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <int G>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
  return v;
}

template <int G>
__device__ __forceinline__ uint32_t group_or_u32(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, G);
  return v;
}

// Per-register maximum of one slab: x = max(x, y) over 32 registers held as
// R bit planes.  gt marks the registers where x > y and eq those still tied,
// decided from the top plane down; d = x ^ y serves the tie update and the
// final select y ^ (d & gt).
template <int R>
__device__ __forceinline__ void slab_max(uint32_t *x, const uint32_t *y) {
  uint32_t d[R];
#pragma unroll
  for (int b = 0; b < R; b++) d[b] = x[b] ^ y[b];
  uint32_t gt = x[R - 1] & d[R - 1];
  uint32_t eq = ~d[R - 1];
#pragma unroll
  for (int b = R - 2; b >= 0; b--) {
    gt |= eq & x[b] & d[b];
    eq &= ~d[b];
  }
#pragma unroll
  for (int b = 0; b < R; b++) x[b] = y[b] ^ (d[b] & gt);
}

// count() pieces of one slab.
//   r = 5: S = sum 2^(31 - v) as a signed integer, with the reference's
//          -1 for v = 31 (hyper_log_log_logic.rs:178 shifts an i32);
//   r = 6: S = sum 2^(63 - v) in 128 bits (exact 2^-v).
// Registers are grouped by v >> 2 (class h) and v & 3; class masks come from
// the upper planes, so the sum is a few popcounts per class.
struct SlabSum5 {
  int64_t s;
  uint32_t zeros;
};

__device__ __forceinline__ SlabSum5 slab_sum5(const uint32_t *p) {
  SlabSum5 out;
  out.zeros = __popc(~(p[0] | p[1] | p[2] | p[3] | p[4]));
  const uint32_t n2 = ~p[2], n3 = ~p[3], n4 = ~p[4];
  const uint32_t q[4] = {n4 & n3, n4 & p[3], p[4] & n3, p[4] & p[3]};
  uint64_t s = 0;
  uint32_t quirk = 0;
#pragma unroll
  for (int h = 0; h < 8; h++) {
    const uint32_t mh = q[h >> 1] & ((h & 1) ? p[2] : n2);
    const uint32_t a = mh & ~p[1], b = mh & p[1];
    const uint32_t pa = __popc(a), pa0 = __popc(a & p[0]), pb = __popc(b), pb0 = __popc(b & p[0]);
    const uint64_t t = uint64_t(8 * pa - 4 * pa0 + 2 * pb - pb0);
    s += t << (28 - 4 * h);
    if (h == 7) quirk = pb0;
  }
  out.s = int64_t(s) - 2 * int64_t(quirk);
  return out;
}

struct SlabSum6 {
  uint64_t hi;  // sum of 2^(15 - v) over registers with v < 16
  uint64_t lo;  // sum of 2^(63 - v) over registers with v >= 16
  uint32_t zeros;
};

// S = hi * 2^48 + lo.  For up to 2^16 registers hi < 2^32 and lo <= 2^63,
// so neither part overflows 64 bits.
__device__ __forceinline__ SlabSum6 slab_sum6(const uint32_t *p) {
  SlabSum6 out;
  out.zeros = __popc(~(p[0] | p[1] | p[2] | p[3] | p[4] | p[5]));
  const uint32_t n[6] = {~p[0], ~p[1], ~p[2], ~p[3], ~p[4], ~p[5]};
  const uint32_t q45[4] = {n[5] & n[4], n[5] & p[4], p[5] & n[4], p[5] & p[4]};
  uint64_t hi = 0, lo = 0;
#pragma unroll
  for (int h = 0; h < 16; h++) {
    const uint32_t qh = q45[h >> 2] & ((h & 2) ? p[3] : n[3]) & ((h & 1) ? p[2] : n[2]);
    const uint32_t a = qh & n[1], b = qh & p[1];
    const uint32_t t = 8 * __popc(a) - 4 * __popc(a & p[0]) + 2 * __popc(b) - __popc(b & p[0]);
    if (h < 4) hi += uint64_t(t) << (12 - 4 * h);
    else lo += uint64_t(t) << (60 - 4 * h);
  }
  out.hi = hi;
  out.lo = lo;
  return out;
}

// Partial count() sums of one lane's slabs (SPL slabs of R planes at x),
// added into (s, hi, lo, zeros): s for 5-bit registers, (hi, lo) for 6-bit
// ones, as slab_sum5 / slab_sum6 define them.
template <int R, int SPL>
__device__ __forceinline__ void lane_sums(const uint32_t *x, int64_t &s, uint64_t &hi,
                                          uint64_t &lo, uint32_t &zeros) {
#pragma unroll
  for (int k = 0; k < SPL; k++) {
    if constexpr (R == 5) {
      const SlabSum5 t = slab_sum5(x + k * R);
      s += t.s;
      zeros += t.zeros;
    } else {
      const SlabSum6 t = slab_sum6(x + k * R);
      hi += t.hi;
      lo += t.lo;
      zeros += t.zeros;
    }
  }
}

// Zero registers among one lane's SPL slabs.
template <int R, int SPL>
__device__ __forceinline__ uint32_t lane_zeros(const uint32_t *x) {
  uint32_t z = 0;
#pragma unroll
  for (int k = 0; k < SPL; k++) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < R; b++) o |= x[k * R + b];
    z += __popc(~o);
  }
  return z;
}

// True when count() takes its linear-counting branch whatever the non-zero
// registers hold, so that the harmonic sum need not be formed.  The sum is
// at least zeros - 1: a zero register adds 1, a non-zero one a positive 2^-v
// or, a 5-bit register of 31, -2^-31 (at most 2^16 of them).  The sum is
// formed exactly and division rounds monotonically, so the raw estimate
// alpha_m_m / sum is at most alpha_m_m / (zeros - 1); when that is below
// 2.5 m the result is lc_table[zeros], as the full computation gives.
__device__ __forceinline__ bool linear_counting_certain(uint32_t zeros, double alpha_m_m,
                                                        uint32_t m) {
  return zeros > 1 && alpha_m_m / double(zeros - 1) < 2.5 * double(m);
}

// count() from the group's total of zero registers and the lane sums;
// every lane returns the result.  hyper_log_log_logic.rs:169-186.
template <int R, int G>
__device__ __forceinline__ double estimate_from_sums(int64_t s, uint64_t shi, uint64_t slo,
                                                     uint32_t zeros, double alpha_m_m,
                                                     uint32_t m, const double *lc_table) {
  double hm;
  if constexpr (R == 5) {
    s = int64_t(group_sum_u64<G>(uint64_t(s)));
    hm = double(s) * 0x1p-31;
  } else {
    shi = group_sum_u64<G>(shi);
    slo = group_sum_u64<G>(slo);
    // S = shi * 2^48 + slo as a 128-bit integer (hi:lo words)
    const uint64_t lo = slo + (shi << 48);
    const uint64_t hi = (shi >> 16) + (lo < slo ? 1 : 0);
    hm = (double(hi) * 18446744073709551616.0 + double(lo)) * 0x1p-63;
  }
  double est = alpha_m_m / hm;
  if (zeros != 0 && est < 2.5 * double(m)) est = lc_table[zeros];
  return est;
}

// count() from the lane sums of a whole counter spread over the G lanes of a
// group (zeros: this lane's zero registers); every lane returns the result.
template <int R, int G>
__device__ __forceinline__ double group_estimate(int64_t s, uint64_t shi, uint64_t slo,
                                                 uint32_t zeros, double alpha_m_m, uint32_t m,
                                                 const double *lc_table) {
  zeros = group_sum_u32<G>(zeros);
  if (linear_counting_certain(zeros, alpha_m_m, m)) return lc_table[zeros];
  return estimate_from_sums<R, G>(s, shi, slo, zeros, alpha_m_m, m, lc_table);
}

// count() of the counter whose slabs are spread over the G lanes of a group;
// every lane returns the result.  A counter with enough zero registers
// (linear_counting_certain) needs only their number.
template <int R, int G, int SPL>
__device__ __forceinline__ double group_count(const uint32_t *x, double alpha_m_m, uint32_t m,
                                              const double *lc_table) {
  const uint32_t zeros = group_sum_u32<G>(lane_zeros<R, SPL>(x));
  if (linear_counting_certain(zeros, alpha_m_m, m)) return lc_table[zeros];
  int64_t s = 0;
  uint64_t hi = 0, lo = 0;
  uint32_t lane_z = 0;
  lane_sums<R, SPL>(x, s, hi, lo, lane_z);
  return estimate_from_sums<R, G>(s, hi, lo, zeros, alpha_m_m, m, lc_table);
}

template <int D>
__device__ __forceinline__ void load_part(uint32_t *dst, const uint32_t *src) {
#pragma unroll
  for (int k = 0; k < D; k++) dst[k] = src[k];
}

template <int D>
__device__ __forceinline__ void store_part(uint32_t *dst, const uint32_t *src) {
#pragma unroll
  for (int k = 0; k < D; k++) dst[k] = src[k];
}

struct Acc {
  int64_t nf_int = 0, nf_hi = 0, nf_lo = 0;
  uint64_t visited = 0, scanned = 0, tscanned = 0, gathered = 0, changed = 0;
  uint32_t checked = 0, copied = 0, modified = 0;

  // Every step is exact except the final rint (|error| <= 2^-65).
  __device__ __forceinline__ void add_nf(double x) {
    const double w = floor(x);
    const double f = (x - w) * 4294967296.0;
    const double h = floor(f);
    nf_int += int64_t(w);
    nf_hi += int64_t(h);
    nf_lo += int64_t(rint((f - h) * 4294967296.0));
  }
};

__device__ __forceinline__ void kahan_add(float2 *p, float x) {
  float2 v = *p;
  const float y = x - v.y;
  const float t = v.x + y;
  v.y = (t - v.x) - y;
  v.x = t;
  *p = v;
}

__device__ __forceinline__ void acc_add(void *base, uint32_t mode, uint32_t v, double x) {
  if (mode == ACC_F64) {
    double *p = static_cast<double *>(base) + v;
    *p = *p + x;
  } else if (mode == ACC_F32_KAHAN) {
    kahan_add(static_cast<float2 *>(base) + v, float(x));
  } else {
    float *p = static_cast<float *>(base) + v;
    *p = *p + float(x);
  }
}

// acc_add onto an accumulator known to hold zero (iteration 0: init zeroes
// them): the same arithmetic on a zero operand, with no read.
__device__ __forceinline__ void acc_add_to_zero(void *base, uint32_t mode, uint32_t v, double x) {
  if (mode == ACC_F64) {
    static_cast<double *>(base)[v] = 0.0 + x;
  } else if (mode == ACC_F32_KAHAN) {
    float2 z;
    z.x = 0.0f;
    z.y = 0.0f;
    const float y = float(x) - z.y;
    const float t = z.x + y;
    z.y = (t - z.x) - y;
    z.x = t;
    static_cast<float2 *>(base)[v] = z;
  } else {
    static_cast<float *>(base)[v] = 0.0f + float(x);
  }
}

// Wavefront sum of the per-lane accumulators, added by lane 0 to the
// wavefront's statistics slot, or to *st when `slots` is null.  Every lane of
// the wavefront must call it.
__device__ __forceinline__ void flush_stats(const Acc &a, Stats *st, uint64_t *slots) {
  uint64_t v[kStatFields] = {uint64_t(a.nf_int), uint64_t(a.nf_hi), uint64_t(a.nf_lo), a.visited,
                             a.scanned, a.tscanned, a.checked, a.copied, a.modified, a.gathered,
                             a.changed};
#pragma unroll
  for (int k = 0; k < kStatFields; k++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
  if (__lane_id() == 0) {
    unsigned long long *p = reinterpret_cast<unsigned long long *>(st);  // Stats field order
    if (slots) {
      const uint32_t w = (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) % kStatSlots;
      p = reinterpret_cast<unsigned long long *>(slots + uint64_t(w) * kStatSlotWords);
    }
#pragma unroll
    for (int k = 0; k < kStatFields; k++)
      if (v[k]) atomicAdd(p + k, (unsigned long long)v[k]);
  }
}

// Dynamic work: chunks of consecutive positions are handed out through eight
// counters, one per XCD (workgroups are dealt round-robin to the 8 XCDs, so
// blockIdx.x % 8 names the XCD a workgroup runs on).  Each XCD walks its own
// contiguous eighth of the range, which keeps window-local successor reads in
// that XCD's L2, and steals from the other ranges when its own is exhausted.
// The first S chunks of each range go without atomics, one to each of the S
// wavefronts of the XCD's full rows of workgroups (S = gridDim.x / 8 * the
// wavefronts per workgroup; first_chunk); the counter hands out the rest,
// from S on (grab_chunk).  Both return UINT64_MAX when no work is left and
// are called by one lane.
__device__ __forceinline__ uint64_t static_waves_per_xcd() {
  return uint64_t(gridDim.x / 8) * (blockDim.x / 64);
}

// The same dealing for any work unit: `items` items, S of them static per
// XCD, the caller's static index idx.
__device__ __forceinline__ uint64_t grab_item(uint32_t *work, uint64_t items, uint64_t S) {
  const uint64_t per = (items + 7) / 8;
  const uint32_t home = blockIdx.x % 8;
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t q = (home + k) % 8;
    const uint64_t base = q * per;
    if (base >= items) continue;
    const uint64_t lim = per < items - base ? per : items - base;
    if (S >= lim) continue;
    uint32_t *ctr = work + q * kWorkStride;
    if (S + __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= lim) continue;
    const uint64_t i = S + atomicAdd(ctr, 1u);
    if (i < lim) return base + i;
  }
  return UINT64_MAX;
}

__device__ __forceinline__ uint64_t first_item(uint32_t *work, uint64_t items, uint64_t S,
                                               uint64_t idx) {
  const uint64_t per = (items + 7) / 8;
  const uint64_t base = (blockIdx.x % 8) * per;
  if (idx < S && base < items && idx < items - base && idx < per) return base + idx;
  return grab_item(work, items, S);
}

__device__ __forceinline__ uint64_t grab_chunk(uint32_t *work, uint64_t chunks) {
  return grab_item(work, chunks, static_waves_per_xcd());
}

__device__ __forceinline__ uint64_t first_chunk(uint32_t *work, uint64_t chunks) {
  const uint64_t idx = uint64_t(blockIdx.x / 8) * (blockDim.x / 64) + threadIdx.x / 64;
  return first_item(work, chunks, static_waves_per_xcd(), idx);
}

// Orders the LDS traffic of the W lanes of a lane group (W <= 64, one
// wavefront): writes before the call are visible to reads after it.
template <int W>
__device__ __forceinline__ void lane_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Successor ids are read in slices of slice_width<G>() consecutive positions:
// each lane of a G-lane group reads ids_per_lane<G>() of them, so narrow
// groups (small counters) still take 16 ids per ballot round.
#ifndef HB_ID_SLICE
#define HB_ID_SLICE 16
#endif
template <int G>
__host__ __device__ constexpr int ids_per_lane() {
  return G >= HB_ID_SLICE ? 1 : HB_ID_SLICE / G;
}
template <int G>
__host__ __device__ constexpr int slice_width() {
  return G * ids_per_lane<G>();
}

// The register a node's counter gets from seeding item `item` (no weights:
// the node id): index j and value val (hyper_log_log_logic.rs:151-167).
__device__ __forceinline__ void seed_register(uint32_t hash, uint64_t seed, uint32_t log2m,
                                              uint64_t sentinel, uint64_t item, uint32_t &j,
                                              uint32_t &val) {
  const uint64_t x = hash == 1 ? jenkins(item, seed) : siphash13(item);
  j = uint32_t(x & ((uint64_t(1) << log2m) - 1));
  val = uint32_t(__builtin_ctzll((x >> log2m) | sentinel)) + 1;
}

// The seeded register of node `item`, packed as index | value << 16.
__device__ __forceinline__ uint32_t seed_item(const UnionArgs &a, uint32_t item) {
  uint32_t reg = 0, val = 0;
  seed_register(a.hash, a.seed, a.log2m, a.sentinel, item, reg, val);
  return reg | (val << 16);
}

// Raises register (w & 0xffff) of a lane's slabs (slab0 .. slab0 + SPL - 1,
// R planes each, at cur) to the value w >> 16 if that is larger: one packed
// seeded item.
template <int R, int SPL>
__device__ __forceinline__ void seed_apply(uint32_t *cur, uint32_t w, uint32_t slab0) {
  const uint32_t reg = w & 0xffffu, val = w >> 16, q = (reg >> 5) - slab0, bit = reg & 31;
  uint32_t plane[R];
#pragma unroll
  for (int b = 0; b < R; b++) {
    plane[b] = cur[b];
#pragma unroll
    for (int qq = 1; qq < SPL; qq++) plane[b] = q == uint32_t(qq) ? cur[qq * R + b] : plane[b];
  }
  uint32_t cv = 0;
#pragma unroll
  for (int b = 0; b < R; b++) cv |= ((plane[b] >> bit) & 1u) << b;
  if (val > cv) {
#pragma unroll
    for (int b = 0; b < R; b++) {
      const uint32_t np = (plane[b] & ~(1u << bit)) | (((val >> b) & 1u) << bit);
#pragma unroll
      for (int qq = 0; qq < SPL; qq++) cur[qq * R + b] = q == uint32_t(qq) ? np : cur[qq * R + b];
    }
  }
}

// Merges into `cur` the counters of the modified successors (other than v)
// found in the slices of v's row that start at first, first + stride, ...
// (hyperball_impl.rs:1037-1049).  `prev` / `cur` are loaded from the node's
// own counter on the first hit.  The ids of the modified successors of a
// slice and their change masks are packed, in ballot order, into the lane
// group's LDS slot (`ids`: slice_width<G>() ids, then as many masks); they
// are then merged U at a time with no per-slot predication: a short last
// batch repeats its first id, and merging a counter twice changes nothing.
//
// Delta gathering: a lane reads only the slabs of a successor whose bit is
// set in the successor's change mask, and merges zeros for the others.  This
// is exact.  For t >= 1 every counter already dominates its successors'
// counters of iteration t - 1: iteration 0 checks all nodes; a node checked
// in iteration t - 1 merged them; a node not checked had no successor
// modified in iteration t - 2, so it still dominates their unchanged
// counters.  So only registers that changed in iteration t - 1 can raise it.
// In iteration 0 the mask holds the slabs with a non-zero register.
//
// Seeded counters (MODE 2, iteration 0 without weights): every successor's
// counter holds the single register of its own hash, which seeding left in
// seed_items as (register, value) when the array is small enough to stay in
// the caches (else the lanes hash the ids again); each lane puts those of the
// ids it read into the slot instead of the id, and the lane owning that
// register's slab takes the maximum.  No successor counter or change mask is
// read.
//
// Untested (MODE 3, window iterations): every successor other than v is
// merged, modified or not.  This is exact when the iteration checks every
// node, since each counter already dominates its successors' counters of the
// last iteration (the invariant above), so an unmodified successor raises
// nothing; the caller counts visited arcs from in-degrees instead.
//
// MODE: 0 whole counters, 1 delta gathering, 2 seeded counters, 3 untested.
template <int R, int G, int SPL, int MODE, int NSEG = 1>
__device__ __forceinline__ void merge_slices(const UnionArgs &a, uint32_t v, int gl,
                                                 uint64_t first, uint64_t end, uint64_t stride,
                                                 uint32_t *prev, uint32_t *cur, bool &loaded,
                                                 uint32_t &nvisit, uint32_t &acc_gathered,
                                                 uint32_t *ids, int seg = 0) {
  constexpr bool DELTA = MODE == 1, SEEDED = MODE == 2, UNTESTED = MODE == 3;
  constexpr int D = R * SPL;         // dwords per lane
  constexpr int CD = G * D;          // dwords per counter segment
  constexpr int K = ids_per_lane<G>();
#ifdef HB_LOADS_IN_FLIGHT
  constexpr int U = HB_LOADS_IN_FLIGHT;
#else
  constexpr int U = D <= 6 ? 8 : D <= 12 ? 4 : 2;  // successor loads in flight per lane
#endif
  constexpr uint32_t CS = counter_stride(NSEG * CD);  // dwords between counters
  const uint32_t *own = a.curr + uint64_t(v) * CS + seg * CD + gl * D;
  const uint32_t *lane_base = a.curr + seg * CD + gl * D;
  uint32_t *masks = ids + slice_width<G>();
  const uint64_t below = (uint64_t(1) << gl) - 1;
  // the counter's slab index of this lane's first slab
  const uint32_t slab0 = uint32_t(seg) * G * SPL + uint32_t(gl) * SPL;
  // this lane's slabs in change-mask bits
  uint32_t lane_bit[SPL];
#pragma unroll
  for (int q = 0; q < SPL; q++) lane_bit[q] = 1u << ((slab0 + q) >> a.chg_shift);
  // with every node checked most successors are modified: read the masks
  // together with the modified bits instead of after them
  const bool speculate = a.flags & F_CHECK_ALL;
  // Seeded counters need no counter reads and leave the mask half of the slot
  // unused: a pass takes two of the group's slices (KS ids per lane, packed
  // pairs filling the whole slot), and the node's own pair is read up front,
  // so that a short row costs two dependent reads (ids, then pairs).
  constexpr int KS = SEEDED ? 2 * K : K;
  const uint64_t step = SEEDED ? 2 * stride : stride;
  const uint32_t own_item = !SEEDED ? 0u : a.seed_items ? a.seed_items[v] : seed_item(a, v);
  for (uint64_t base = first; base < end; base += step) {
    uint32_t s[KS], mk[KS];
    bool f[KS];
#pragma unroll
    for (int j = 0; j < KS; j++) {
      const uint64_t idx = base + uint64_t(j / K) * stride + uint64_t(j % K) * G + gl;
      s[j] = idx < end ? a.succ[idx] : v;
    }
    if (!DELTA) {
#pragma unroll
      for (int j = 0; j < KS; j++) {
        // seeded counters: init marks every node modified
        f[j] = (s[j] != v) && (SEEDED || UNTESTED || test_bit(a.curr_mod, s[j]));
        mk[j] = ~0u;
      }
      if (SEEDED) {
        if (a.seed_items) {
#pragma unroll
          for (int j = 0; j < KS; j++) mk[j] = f[j] ? a.seed_items[s[j]] : 0u;
        } else {
#pragma unroll
          for (int j = 0; j < KS; j++) mk[j] = f[j] ? seed_item(a, s[j]) : 0u;
        }
      }
    } else if (speculate) {
#pragma unroll
      for (int j = 0; j < K; j++) mk[j] = s[j] != v ? a.chg_cur[s[j]] : 0u;
#pragma unroll
      for (int j = 0; j < K; j++) f[j] = (s[j] != v) && test_bit(a.curr_mod, s[j]);
    } else {
#pragma unroll
      for (int j = 0; j < K; j++) f[j] = (s[j] != v) && test_bit(a.curr_mod, s[j]);
#pragma unroll
      for (int j = 0; j < K; j++) mk[j] = f[j] ? a.chg_cur[s[j]] : 0u;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < KS; j++) {
      const uint64_t mask = group_ballot<G>(f[j]);
      if (f[j]) {
        const uint32_t pos = cnt + __popcll(mask & below);
        ids[pos] = SEEDED ? mk[j] : s[j];
        if (DELTA) masks[pos] = mk[j];
      }
      cnt += __popcll(mask);
    }
    if (!cnt) continue;
    lane_sync<G>();
    if (!loaded) {
      if constexpr (SEEDED) {
        // the node's own counter holds the one register of its own hash
        const uint32_t oval = own_item >> 16, oslab = (own_item & 0xffffu) >> 5,
                       obit = own_item & 31;
#pragma unroll
        for (int q = 0; q < SPL; q++)
#pragma unroll
          for (int b = 0; b < R; b++)
            prev[q * R + b] = oslab == slab0 + q ? ((oval >> b) & 1u) << obit : 0u;
      } else {
        load_part<D>(prev, own);
      }
#pragma unroll
      for (int k = 0; k < D; k++) cur[k] = prev[k];
      loaded = true;
    }
    nvisit += cnt;
    if (SEEDED) {
      // eight slots per batch, read before use; cnt is at most the slot's
      // 2 * slice_width<G>() words, a multiple of 8, so a batch that starts
      // below cnt reads inside the slot
      static_assert(slice_width<G>() % 4 == 0, "seeded batches read within the slot");
      for (uint32_t k0 = 0; k0 < cnt; k0 += 8) {
        uint32_t w8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) w8[u] = ids[k0 + u];
        if constexpr (G >= 8) {
          // each lane marks the slots whose register lies in its slabs (one in
          // G), then applies only those: a wavefront runs the register update
          // as often as its busiest lane has hits, not once per slot
          uint32_t hits = 0;
#pragma unroll
          for (int u = 0; u < 8; u++)
            if (k0 + u < cnt && ((w8[u] & 0xffffu) >> 5) - slab0 < uint32_t(SPL)) hits |= 1u << u;
          while (hits) {
            const uint32_t w = ids[k0 + __builtin_ctz(hits)];
            hits &= hits - 1;
            seed_apply<R, SPL>(cur, w, slab0);
          }
        } else {
#pragma unroll
          for (int u = 0; u < 8; u++) {
            if (k0 + u >= cnt) break;
            if (((w8[u] & 0xffffu) >> 5) - slab0 < uint32_t(SPL)) seed_apply<R, SPL>(cur, w8[u], slab0);
          }
        }
      }
      lane_sync<G>();
      continue;
    }
    for (uint32_t k = 0; k < cnt; k += U) {
      uint32_t y[U][D];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool real = k + u < cnt;
        const uint32_t slot = real ? k + u : k;
        const uint32_t t = ids[slot];
        const uint32_t *src = lane_base + uint64_t(t) * CS;
        if (!DELTA) {
          load_part<D>(y[u], src);
          acc_gathered += real ? uint32_t(SPL) : 0u;
          continue;
        }
        const uint32_t mt = masks[slot];
#pragma unroll
        for (int q = 0; q < SPL; q++) {
          if (mt & lane_bit[q]) {
            load_part<R>(y[u] + q * R, src + q * R);
            acc_gathered += real ? 1u : 0u;
          } else {
#pragma unroll
            for (int b = 0; b < R; b++) y[u][q * R + b] = 0u;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int q = 0; q < SPL; q++) slab_max<R>(cur + q * R, y[u] + q * R);
    }
    lane_sync<G>();
  }
}

template <int R, int G, int SPL, int NSEG = 1>
__device__ __forceinline__ void merge_successors(const UnionArgs &a, uint32_t v, int gl,
                                                 uint64_t first, uint64_t end, uint64_t stride,
                                                 uint32_t *prev, uint32_t *cur, bool &loaded,
                                                 uint32_t &nvisit, uint32_t &acc_gathered,
                                                 uint32_t *ids, int seg = 0) {
  if (a.flags & F_SEEDED)
    merge_slices<R, G, SPL, 2, NSEG>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                     acc_gathered, ids, seg);
  else if (a.flags & F_DELTA)
    merge_slices<R, G, SPL, 1, NSEG>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                     acc_gathered, ids, seg);
  else
    merge_slices<R, G, SPL, 0, NSEG>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                     acc_gathered, ids, seg);
}

// The per-node bookkeeping after the successor loop (hyperball_impl.rs:
// 1051-1130), once the group knows whether the counter changed (cm), its
// changed slabs (chg_bits, in change-mask bits, the same on every lane) and
// its new count (post, when cm): count cache, neighbourhood function,
// centralities, modified bit, change mask and predecessor marking.
template <int G>
__device__ __forceinline__ void node_update(const UnionArgs &a, uint32_t v, int gl, bool check,
                                            bool cm, uint32_t chg_bits, double post,
                                            uint32_t nvisit, uint64_t deg, Acc &acc_out) {
  const bool systolic = a.flags & F_SYSTOLIC;
  // seeded counters (iteration 0 without weights): every cached count is the
  // seeding's lc[m - 1] and every accumulator is zero, so neither is read
  const bool seeded = a.flags & F_SEEDED;
  if (gl == 0) {
    double pre = 0.0;
    if (cm) {
      a.chg_next[v] = chg_bits;
      acc_out.changed += __popc(chg_bits);
      pre = seeded ? a.lc_table[a.m - 1] : a.cache[v];
      a.cache[v] = post;
    } else if (!systolic && check) {
      post = seeded ? a.lc_table[a.m - 1] : a.cache[v];
    }
    if (!systolic && check) acc_out.add_nf(post);
    if (cm) {
      if (systolic) acc_out.add_nf(post - pre);
      if (a.flags & F_CENTRALITY) {
        const double delta = post - pre;
        if (delta > 0.0 && seeded) {
          if (a.sd) acc_add_to_zero(a.sd, a.acc_mode, v, delta * a.dist);
          if (a.hc) acc_add_to_zero(a.hc, a.acc_mode, v, delta / a.dist);
          for (int k = 0; k < a.ndisc; k++)
            acc_add_to_zero(a.dc[k], a.acc_mode, v, delta * a.disc[k]);
        } else if (delta > 0.0) {
          if (a.sd) acc_add(a.sd, a.acc_mode, v, delta * a.dist);
          if (a.hc) acc_add(a.hc, a.acc_mode, v, delta / a.dist);
          for (int k = 0; k < a.ndisc; k++) acc_add(a.dc[k], a.acc_mode, v, delta * a.disc[k]);
        }
      }
      atomicOr(a.next_mod + (v >> 5), 1u << (v & 31));
      acc_out.modified++;
    }
    acc_out.visited += nvisit;
    acc_out.scanned += deg;
    if (check) acc_out.checked++;
    else acc_out.copied++;
  }

  if (cm && systolic) {
    const uint64_t tb = a.toff[v], te = a.toff[v + 1];
    for (uint64_t k = tb + gl; k < te; k += G) {
      const uint32_t p = a.tsucc[k];
      atomicOr(a.next_must + (p >> 5), 1u << (p & 31));
    }
    if (gl == 0) acc_out.tscanned += te - tb;
  }
}

// Everything after the successor loop (hyperball_impl.rs:1051-1144) for a
// counter held whole by the group: changed test, count(), node_update, and
// the store into the next state.
template <int R, int G, int SPL>
__device__ __forceinline__ void finish_node(const UnionArgs &a, uint32_t v, int gl, bool check,
                                            bool was_mod, bool loaded, const uint32_t *prev,
                                            uint32_t *cur, uint32_t nvisit, uint64_t deg,
                                            Acc &acc_out) {
  constexpr int D = R * SPL, CD = G * D;
  bool cm = false;
  uint32_t chg_bits = 0;  // this lane's changed slabs, in change-mask bits
  if (loaded) {
#pragma unroll
    for (int q = 0; q < SPL; q++) {
      uint32_t d = 0;
#pragma unroll
      for (int b = 0; b < R; b++) d |= cur[q * R + b] ^ prev[q * R + b];
      if (d) chg_bits |= 1u << ((uint32_t(gl) * SPL + q) >> a.chg_shift);
    }
    cm = group_ballot<G>(chg_bits != 0) != 0;
  }
  double post = 0.0;
  if (cm) {
    chg_bits = group_or_u32<G>(chg_bits);
    post = group_count<R, G, SPL>(cur, a.alpha_m_m, a.m, a.lc_table);
  }
  node_update<G>(a, v, gl, check, cm, chg_bits, post, nvisit, deg, acc_out);

  if (cm || was_mod) {
    if (!loaded) load_part<D>(cur, a.curr + uint64_t(v) * counter_stride(CD) + gl * D);
    store_part<D>(a.next + uint64_t(v) * counter_stride(CD) + gl * D, cur);
  }
}

// Each wavefront takes chunks of `rounds` * (64 / G) positions and gives
// every lane group one node per round.  A launch uses the fewest rounds, at
// most kRoundsPerChunk, that give every wavefront of the grid at most one
// chunk; beyond kRoundsPerChunk * (64 / G) positions per wavefront the chunks
// are dealt dynamically.  A small node count or worklist is then spread over
// the whole grid, every chunk static, instead of walked by a few wavefronts
// round after round (each round is a chain of dependent loads).
constexpr int kRoundsPerChunk = 8;

template <int G>
__host__ __device__ constexpr uint64_t chunk_positions() {  // at most
  return uint64_t(kRoundsPerChunk) * (64 / G);
}

__host__ __device__ inline uint32_t chunk_rounds(uint64_t count, uint32_t nodes_per_wave,
                                                 uint64_t waves) {
  const uint64_t per = uint64_t(nodes_per_wave) * waves;
  const uint64_t r = (count + per - 1) / per;
  return r < 1 ? 1u : r > uint64_t(kRoundsPerChunk) ? uint32_t(kRoundsPerChunk) : uint32_t(r);
}

// Per-position node data of a chunk, staged in LDS by stage_chunk.
struct NodeMeta {
  uint64_t beg, end;  // CSR row
  uint32_t v;
  uint32_t flags;     // kMetaRun | kMetaCheck | kMetaWasMod
};
constexpr uint32_t kMetaRun = 1, kMetaCheck = 2, kMetaWasMod = 4;

// Loads the node data of all positions of a chunk in one pass (every lane of
// the wavefront, independent loads): the order or worklist entry, the
// frontier bits and the row bounds.  Positions past `count`, nodes neither
// checked nor modified, and rows left to heavy_kernel get no kMetaRun.
template <int G>
__device__ __forceinline__ void stage_chunk(const UnionArgs &a, uint64_t chunk, uint64_t count,
                                            uint32_t rounds, NodeMeta *meta) {
  const uint64_t CH = uint64_t(rounds) * (64 / G);
  for (uint32_t p = __lane_id(); p < CH; p += 64) {
    const uint64_t i = chunk * CH + p;
    NodeMeta m{0, 0, 0, 0};
    if (i < count) {
      const uint32_t v = (a.flags & F_LIST) ? a.list[i] : (a.order ? a.order[i] : uint32_t(i));
      const bool check = (a.flags & F_CHECK_ALL) || test_bit(a.must, v);
      const bool was_mod = test_bit(a.curr_mod, v);
      m.v = v;
      if (check || was_mod) {
        m.beg = a.off[v];
        m.end = a.off[v + 1];
        if (m.end - m.beg <= a.heavy_degree)
          m.flags = kMetaRun | (check ? kMetaCheck : 0u) | (was_mod ? kMetaWasMod : 0u);
      }
    }
    meta[p] = m;
  }
  lane_sync<64>();
}

template <int R, int G, int SPL, int NSEG>
__device__ __forceinline__ void process_node(const UnionArgs &a, const NodeMeta &m, int gl,
                                             Acc &acc_out, uint32_t *ids) {
  constexpr int D = R * SPL;
  const uint32_t v = m.v;
  const bool check = m.flags & kMetaCheck;
  const bool was_mod = m.flags & kMetaWasMod;
  const uint64_t beg = m.beg, end = m.end;

  uint32_t prev[D], cur[D];
  uint32_t nvisit = 0, gathered = 0;
  if constexpr (NSEG == 1) {
    bool loaded = false;
    if (check)
      merge_successors<R, G, SPL>(a, v, gl, beg, end, slice_width<G>(), prev, cur, loaded,
                                  nvisit, gathered, ids);
    acc_out.gathered += gathered;
    finish_node<R, G, SPL>(a, v, gl, check, was_mod, loaded, prev, cur, nvisit,
                           check ? end - beg : 0, acc_out);
  } else {
    // Counters of NSEG segments of G * D dwords: the row is merged segment by
    // segment.  The modified successors found are the same in every segment,
    // so a segment is loaded exactly when all are; count() sums accumulate
    // over the segments.  A lane stores its part of a segment when the node
    // was modified last iteration (the other buffer is stale) or when the
    // part changed; an unchanged part of an unmodified node already holds the
    // same registers in the other buffer.
    constexpr int CD = G * D;
    constexpr uint32_t CS = counter_stride(NSEG * CD);
    int64_t s5 = 0;
    uint64_t hi = 0, lo = 0;
    uint32_t zeros = 0, chg_bits = 0;
    bool loaded = false;
#pragma unroll 1
    for (int seg = 0; seg < NSEG; seg++) {
      uint32_t nv = 0;
      loaded = false;
      if (check)
        merge_successors<R, G, SPL, NSEG>(a, v, gl, beg, end, slice_width<G>(), prev, cur,
                                          loaded, nv, gathered, ids, seg);
      if (seg == 0) nvisit = nv;
      bool part_changed = false;
      if (loaded) {
#pragma unroll
        for (int q = 0; q < SPL; q++) {
          uint32_t d = 0;
#pragma unroll
          for (int b = 0; b < R; b++) d |= cur[q * R + b] ^ prev[q * R + b];
          if (d) {
            chg_bits |= 1u << ((uint32_t(seg) * G * SPL + uint32_t(gl) * SPL + q) >> a.chg_shift);
            part_changed = true;
          }
        }
        lane_sums<R, SPL>(cur, s5, hi, lo, zeros);
      }
      if (was_mod || part_changed) {
        const uint64_t at = uint64_t(v) * CS + seg * CD + gl * D;
        if (!loaded) load_part<D>(cur, a.curr + at);
        store_part<D>(a.next + at, cur);
      }
    }
    acc_out.gathered += gathered;
    const bool cm = loaded && group_ballot<G>(chg_bits != 0) != 0;
    double post = 0.0;
    if (cm) {
      chg_bits = group_or_u32<G>(chg_bits);
      post = group_estimate<R, G>(s5, hi, lo, zeros, a.alpha_m_m, a.m, a.lc_table);
    }
    node_update<G>(a, v, gl, check, cm, chg_bits, post, nvisit, check ? end - beg : 0, acc_out);
  }
}


// Chunk `chunk` of `rounds` rounds (at most kRoundsPerChunk).
// `wave_ids` and `meta` are the calling wavefront's own LDS regions, at least
// 2 * 64 * ids_per_lane<G>() words and chunk_positions<G>() entries; its lane
// groups take consecutive id-and-mask slots.  The regions are per wavefront, not per
// shape: in the batched kernel the wavefronts of one workgroup may run
// different shapes at the same time.
template <int R, int G, int SPL, int NSEG = 1>
__device__ __forceinline__ void run_chunk(const UnionArgs &a, uint64_t chunk, uint64_t count,
                                          uint32_t rounds, uint32_t *wave_ids, NodeMeta *meta,
                                          Acc &acc) {
  constexpr int NPW = 64 / G;
  const int g = __lane_id() / G, gl = __lane_id() % G;
  uint32_t *ids = wave_ids + g * 2 * slice_width<G>();
  stage_chunk<G>(a, chunk, count, rounds, meta);
#pragma unroll 1
  for (uint32_t round = 0; round < rounds; round++) {
    const NodeMeta m = meta[round * NPW + g];
    if (m.flags & kMetaRun) process_node<R, G, SPL, NSEG>(a, m, gl, acc, ids);
  }
  lane_sync<64>();  // the next chunk's staging overwrites meta
}

// Rows longer than heavy_degree: the whole workgroup (NT threads) works on
// one node.  Its NT/G lane groups take interleaved slices of the row, keep
// private maxima, and fold them in LDS; lane group 0 finishes the node.
// Every thread of the block calls it with the same v.  LDS: s_acc holds
// NT * D words, s_ids 2 * NT * ids_per_lane<G>() words (a slot per lane
// group, as in the chunk phase), s_sum two words (any lane group loaded the
// counter; visited arcs).  Without COUNT the node's visited arcs and gathered
// slabs are not counted (the window kernel counts them from in-degrees).  WIN
// (window iterations: neither seeded counters nor delta gathering) compiles
// only the whole-counter merges, untested under F_UNTESTED (merge_slices
// MODE 3).
template <int R, int G, int SPL, int NT = kBlock, bool COUNT = true, bool WIN = false>
__device__ __forceinline__ void heavy_node(const UnionArgs &a, uint32_t v, uint32_t *s_acc,
                                           uint32_t *s_ids, uint32_t *s_sum,
                                           Acc &acc) {
  constexpr int D = R * SPL, NG = NT / G, CD = G * D;
  const int g = threadIdx.x / G, gl = threadIdx.x % G;
  const bool in_list = !(a.flags & F_LIST) || test_bit(a.must, v) || test_bit(a.curr_mod, v);
  const bool check = (a.flags & F_CHECK_ALL) || test_bit(a.must, v);
  const bool was_mod = test_bit(a.curr_mod, v);
  if (!in_list || (!check && !was_mod)) return;  // uniform over the block
  const uint64_t beg = a.off[v], end = a.off[v + 1];
  uint32_t prev[D], cur[D];
  bool loaded = false;
  uint32_t nvisit = 0, gathered = 0;
  if (check) {
    const uint64_t first = beg + uint64_t(g) * slice_width<G>();
    const uint64_t stride = uint64_t(NG) * slice_width<G>();
    uint32_t *slot = s_ids + g * 2 * slice_width<G>();
    if (WIN && (a.flags & F_UNTESTED))
      merge_slices<R, G, SPL, 3>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                 gathered, slot);
    else if (WIN)
      merge_slices<R, G, SPL, 0>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                 gathered, slot);
    else
      merge_successors<R, G, SPL>(a, v, gl, first, end, stride, prev, cur, loaded, nvisit,
                                  gathered, slot);
    if (!COUNT) nvisit = gathered = 0;  // the caller counts them from in-degrees
  }
  acc.gathered += gathered;
  if (threadIdx.x == 0) s_sum[0] = s_sum[1] = 0;
  __syncthreads();
  if (loaded && gl == 0) atomicOr(s_sum, 1u);
  if (gl == 0 && nvisit) atomicAdd(s_sum + 1, nvisit);
#pragma unroll
  for (int k = 0; k < D; k++) s_acc[g * CD + gl * D + k] = loaded ? cur[k] : 0u;
  __syncthreads();
  for (int half = NG / 2; half > 0; half >>= 1) {
    if (g < half) {
      uint32_t x[D], y[D];
#pragma unroll
      for (int k = 0; k < D; k++) {
        x[k] = s_acc[g * CD + gl * D + k];
        y[k] = s_acc[(g + half) * CD + gl * D + k];
      }
#pragma unroll
      for (int k = 0; k < SPL; k++) slab_max<R>(x + k * R, y + k * R);
#pragma unroll
      for (int k = 0; k < D; k++) s_acc[g * CD + gl * D + k] = x[k];
    }
    __syncthreads();
  }
  if (g == 0) {
    const bool any = s_sum[0] != 0;
    const uint32_t total_visit = s_sum[1];
    if (any) {
      load_part<D>(prev, a.curr + uint64_t(v) * counter_stride(CD) + gl * D);
#pragma unroll
      for (int k = 0; k < D; k++) cur[k] = s_acc[gl * D + k];
      // the folded maxima never lose the node's own registers
#pragma unroll
      for (int k = 0; k < SPL; k++) slab_max<R>(cur + k * R, prev + k * R);
    }
    finish_node<R, G, SPL>(a, v, gl, check, was_mod, any, prev, cur, total_visit,
                           check ? end - beg : 0, acc);
  }
  __syncthreads();
}

// LDS of a union kernel's workgroup.  The heavy-row phase and the chunk
// phase share the successor-id slots; the heavy phase's fold buffer overlays
// the chunk phase's node data (a workgroup leaves the heavy phase at a
// barrier).
template <int R, int G, int SPL>
struct UnionLds {
  static constexpr int kWaveIds = 2 * 64 * ids_per_lane<G>();
  static constexpr int kWaveMeta = int(chunk_positions<G>());
  uint32_t ids[(kBlock / 64) * kWaveIds];
  union {
    NodeMeta meta[(kBlock / 64) * kWaveMeta];
    struct {
      uint32_t acc[kBlock * R * SPL];
      uint32_t sum[2];
      uint64_t h;  // the heavy row being worked on
    } heavy;
  } u;
};

// Work counter [8] of a.work hands out the heavy rows.  Every workgroup takes
// rows until none is left (the heaviest first: the list is sorted by
// decreasing degree; in worklist iterations the list holds only the selected
// rows, in no particular order), then turns to the chunk queues.  Returns
// when no heavy row is left; uniform over the workgroup.
template <int R, int G, int SPL>
__device__ __forceinline__ void heavy_phase(const UnionArgs &a, UnionLds<R, G, SPL> &L, Acc &acc) {
  const uint64_t count = a.heavy_len ? uint64_t(*a.heavy_len) : a.heavy_count;
  for (;;) {
    if (threadIdx.x == 0) L.u.heavy.h = atomicAdd(a.work + 8 * kWorkStride, 1u);
    __syncthreads();
    const uint64_t h = L.u.heavy.h;
    __syncthreads();
    if (h >= count) break;
    heavy_node<R, G, SPL>(a, a.heavy_list[h], L.u.heavy.acc, L.ids, L.u.heavy.sum, acc);
  }
}

// The union step of one iteration: first the rows longer than heavy_degree,
// a workgroup per row (heavy_phase), then chunks of the other nodes, a lane
// group per node.  Counters of several segments (NSEG > 1, more than 4096
// registers) have no heavy-row phase: the host leaves heavy_count at 0.
// Shapes of 5 dwords per lane fit four wavefronts per SIMD (at most 128
// VGPRs); the hint keeps the register allocator there.
template <int R, int G, int SPL, int NSEG>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(R * SPL <= 5 ? 4 : 1)))
void union_kernel(UnionArgs a) {
  __shared__ UnionLds<R, G, SPL> L;
  Acc acc;
  if (NSEG == 1 && a.heavy_count) heavy_phase<R, G, SPL>(a, L, acc);
  const uint64_t count = (a.flags & F_LIST) ? uint64_t(*a.list_len) : a.count;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.stats->list_len = count;
  const uint32_t rounds = __builtin_amdgcn_readfirstlane(
      chunk_rounds(count, 64 / G, uint64_t(gridDim.x) * (kBlock / 64)));
  const uint64_t cp = uint64_t(rounds) * (64 / G);
  const uint64_t chunks = (count + cp - 1) / cp;
  uint32_t *wave_ids = L.ids + (threadIdx.x / 64) * UnionLds<R, G, SPL>::kWaveIds;
  NodeMeta *meta = L.u.meta + (threadIdx.x / 64) * UnionLds<R, G, SPL>::kWaveMeta;
  // the loop body stays single: a first-iteration flag inside the loop gets
  // it peeled and costs 18 VGPRs (an occupancy step)
  uint64_t c = 0;
  if (__lane_id() == 0) c = first_chunk(a.work, chunks);
  c = __shfl(c, 0);
  while (c != UINT64_MAX) {
    run_chunk<R, G, SPL, NSEG>(a, c, count, rounds, wave_ids, meta, acc);
    if (__lane_id() == 0) c = grab_chunk(a.work, chunks);
    c = __shfl(c, 0);
  }
  flush_stats(acc, a.stats, a.stat_slots);
}

// ---- Window kernel: dense iterations with 64-register counters ------------
//
// With log2m 6 a counter is 40 (r = 5) or 48 (r = 6) bytes, and a random
// counter read moves a whole 128-byte line: in a dense iteration of
// union_kernel nearly every successor read misses the L2, so the step moves
// about three times its algorithmic bytes.  Web graphs take most successors
// from nearby ids.  Here a workgroup takes a block of consecutive nodes and
// first copies the current counters of [b0 - h, b0 + block + h) (the tile),
// their modified bits and the block's row offsets into LDS with coalesced
// reads; successors inside the tile are merged from LDS, the others read from
// a.curr, and a node's own counter comes from the tile.
//
// Work inside a block: rows longer than kWinMedium ids take a wavefront each
// (its 32 lane groups take interleaved runs of ids and fold their maxima
// across lanes), the others a lane group each, in rounds of 32 nodes sorted
// by out-degree (a counting sort per block), so the lane groups of a round
// walk rows of the same length; wavefronts take the wavefront rows first,
// then the rounds (longest rows first), from an LDS counter.  Rows longer than
// a.heavy_degree take the whole workgroup each before the blocks
// (heavy_node).  Blocks are dealt per XCD in node order (first_item /
// grab_item), so the workgroups of an XCD work on neighbouring blocks and
// successors just outside a tile tend to hit that XCD's L2.
//
// Results are those of union_kernel: the merge is a per-register maximum,
// and every statistic is an integer sum.
#ifndef HB_WIN_NT
#define HB_WIN_NT 1024
#endif
#ifndef HB_WIN_B
#define HB_WIN_B 1024
#endif
#ifndef HB_WIN_LDS
#define HB_WIN_LDS 163840
#endif
constexpr int kWinNT = HB_WIN_NT;   // threads per workgroup
constexpr int kWinB = HB_WIN_B;     // nodes per block, at most
constexpr int kWinMedium = 62;      // longer rows take a wavefront each
constexpr int kWinClasses = kWinMedium + 2;  // wavefront rows, then light rows by
                                              // decreasing out-degree (kWinMedium .. 0):
                                              // one class per lane of a wavefront
static_assert(kWinNT % 64 == 0 && kWinB % 32 == 0 && kWinB <= kWinNT, "window kernel shape");

template <int R>
struct WinLds {
  static constexpr int kCdw = 2 * R;  // dwords of a 64-register counter
  static constexpr int kFixed = 8 * (kWinB + 1) + 2 * kWinB + 8 * kWinClasses + 64;
  // tile counters: what the LDS budget leaves (with a modified bit each and
  // two spare bitmap words), a multiple of 32
  static constexpr int kTile = ((HB_WIN_LDS - kFixed - 16) * 8 / (32 * kCdw + 1)) / 32 * 32;
  union U {
    uint32_t tile[kTile * kCdw];
    struct {
      uint32_t acc[kWinNT * R];
      uint32_t ids[2 * kWinNT * ids_per_lane<2>()];
      uint32_t sum[2];
    } heavy;
  } u;
  uint32_t tmod[kTile / 32 + 2];   // curr_mod words from the tile's first node's word on
  uint64_t off[kWinB + 1];         // row offsets of the block
  uint16_t list[kWinB];            // block positions, by class
  uint32_t cls[kWinClasses];       // nodes per class
  uint32_t base[kWinClasses + 1];  // first list position of each class
  uint32_t q;                      // next work item
  uint64_t item;                   // block or heavy row being worked on
};

// Merges into cur the counters of v's successors (other than v) among the
// ids [base, base + WU) for base = first, first + stride, ... below end
// (hyperball_impl.rs:1037-1049): from the tile (LDS) when inside
// [tlo, tlo + tlen), else from a.curr.  Both lanes of the group read the same
// ids and merge the same successors, each into its own slab; WU loads are in
// flight per lane, and the next run's ids are read before this run's
// counters are merged.  Successors are merged untested, modified or not
// (exact when every node is checked: merge_slices MODE 3), and so is v
// itself (a self loop, or a slot past the row end), which changes nothing;
// with TEST an outside successor is read only when its modified bit is set,
// which saves its counter read in iterations where few nodes changed.  The
// tile holds v (its block lies inside the tile).  With COUNT, nvisit counts
// the modified successors other than v (the visited arcs of
// hyperball_impl.rs:1038-1039), from the tile's modified bits (tmod, from the
// word of tlo on) inside the tile.
#ifndef HB_WIN_U
#define HB_WIN_U 4
#endif
template <int R, bool TEST, bool COUNT = false>
__device__ __forceinline__ void win_merge(const UnionArgs &a, uint32_t v, int gl, uint64_t first,
                                          uint64_t end, uint64_t stride, const uint32_t *tile,
                                          const uint32_t *tmod, uint64_t tlo, uint32_t tlen,
                                          uint32_t *cur, uint32_t &nvisit) {
  static_assert(TEST || !COUNT, "counting visited arcs needs the modified bits");
  constexpr int D = R, CD = 2 * R, WU = HB_WIN_U;  // loads in flight per lane
  constexpr uint32_t CS = counter_stride(CD);
  const uint32_t *gbase = a.curr + gl * D;
  const uint32_t *lbase = tile + gl * D;
  const uint32_t self = v - uint32_t(tlo);  // v's place in the tile
  // ids past the row read v itself: its own counter (in the tile) changes
  // nothing, and neither does a self loop
  uint32_t nx[WU];
#pragma unroll
  for (int u = 0; u < WU; u++) nx[u] = first + u < end ? a.succ[first + u] : v;
#pragma unroll 1
  for (uint64_t base = first; base < end; base += stride) {
    uint32_t ids[WU];
#pragma unroll
    for (int u = 0; u < WU; u++) ids[u] = nx[u];
    const uint64_t nb = base + stride;
#pragma unroll
    for (int u = 0; u < WU; u++) nx[u] = nb + u < end ? a.succ[nb + u] : v;
    // The HBM reads of the run first, all in flight together, into registers
    // of their own: a tile read into the same registers would have to wait
    // for them.  Then the tile reads, one successor at a time, merged with
    // whichever of the two holds its counter; lanes whose successor lies
    // outside read v's own counter from the tile, which changes nothing.
    uint32_t yg[WU][D];
#pragma unroll
    for (int u = 0; u < WU; u++) {
      const uint32_t t = ids[u];
      // tlo + tlen <= n <= 2^32, so t - tlo wraps to at least tlen when t < tlo
      const uint32_t ti = t - uint32_t(tlo);
      const bool out = ti >= tlen;
      bool mod = true;
      if (TEST) {
        const uint32_t b = ti + uint32_t(tlo & 31u);  // bit of t in tmod (tlo's word first)
        mod = out ? test_bit(a.curr_mod, t) : ((tmod[b >> 5] >> (b & 31)) & 1u);
      }
      if (COUNT) nvisit += (mod && t != v) ? 1u : 0u;
#pragma unroll
      for (int b = 0; b < D; b++) yg[u][b] = 0u;
      if (out && mod) load_part<D>(yg[u], gbase + uint64_t(t) * CS);
    }
#pragma unroll
    for (int u = 0; u < WU; u++) {
      const uint32_t ti = ids[u] - uint32_t(tlo);
      const bool in = ti < tlen;
      const uint32_t *p = lbase + (in ? ti : self) * CD;
      uint32_t y[D];
#pragma unroll
      for (int b = 0; b < D; b++) y[b] = in ? p[b] : yg[u][b];
      slab_max<R>(cur, y);
    }
  }
}

// Node data that the end of a node's body needs, read by lane 0 of its group
// before the merge so that the read overlaps it: the cached count of the
// current counter.
struct WinNode {
  double pre;
};

__device__ __forceinline__ WinNode win_prefetch(const UnionArgs &a, uint32_t v) {
  WinNode w;
  w.pre = a.cache[v];
  return w;
}

// The end of the node body (hyperball_impl.rs:1051-1144) for a checked node
// of a window iteration: node_update and the store of finish_node, with the
// node data of win_prefetch (valid on lane 0 of the group).  Without SYS
// (every node checked, not systolic) visited arcs, scanned arcs and checked
// nodes are counted when the block is staged; with SYS (a worklist
// iteration, systolic or local) the node's nvisit modified successors count
// here, the neighbourhood function takes post - pre in systolic iterations,
// and a modified node marks its predecessors.
template <int R, bool SYS>
__device__ __forceinline__ void win_finish(const UnionArgs &a, uint32_t v, int gl, bool was_mod,
                                           const uint32_t *prev, const uint32_t *cur,
                                           uint32_t nvisit, const WinNode &w, Acc &acc) {
  constexpr int D = R;
  constexpr uint32_t CS = counter_stride(2 * R);
  const bool systolic = SYS && (a.flags & F_SYSTOLIC);
  uint32_t d = 0;
#pragma unroll
  for (int b = 0; b < D; b++) d |= cur[b] ^ prev[b];
  const bool cm = group_ballot<2>(d != 0) != 0;
  double post = 0.0;
  uint32_t chg_bits = 0;
  if (cm) {
    chg_bits = group_or_u32<2>(d ? 1u << (uint32_t(gl) >> a.chg_shift) : 0u);
    post = group_count<R, 2, 1>(cur, a.alpha_m_m, a.m, a.lc_table);
  }
  if (gl == 0) {
    if (cm) {
      a.chg_next[v] = chg_bits;
      acc.changed += __popc(chg_bits);
      a.cache[v] = post;
      acc.add_nf(systolic ? post - w.pre : post);
      if (a.flags & F_CENTRALITY) {
        const double delta = post - w.pre;
        if (delta > 0.0) {
          if (a.sd) acc_add(a.sd, a.acc_mode, v, delta * a.dist);
          if (a.hc) acc_add(a.hc, a.acc_mode, v, delta / a.dist);
          for (int k = 0; k < a.ndisc; k++) acc_add(a.dc[k], a.acc_mode, v, delta * a.disc[k]);
        }
      }
      atomicOr(a.next_mod + (v >> 5), 1u << (v & 31));
      acc.modified++;
    } else if (!systolic) {
      acc.add_nf(w.pre);
    }
    if (SYS) {
      acc.visited += nvisit;
      acc.gathered += 2 * uint64_t(nvisit);
    }
  }
  if (systolic && cm) {
    const uint64_t tb = a.toff[v], te = a.toff[v + 1];
    for (uint64_t k = tb + gl; k < te; k += 2) {
      const uint32_t p = a.tsucc[k];
      atomicOr(a.next_must + (p >> 5), 1u << (p & 31));
    }
    if (gl == 0) acc.tscanned += te - tb;
  }
  if (cm || was_mod) store_part<D>(a.next + uint64_t(v) * CS + gl * D, cur);
}

// A value the whole wavefront holds (read from LDS after a barrier), moved to
// scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  return uint64_t(hi) << 32 | lo;
}

// Window geometry of an instance: nodes per block (at most kWinB) and the
// tile's half-width around a block (at most what the tile leaves).
struct WinGeom {
  uint32_t B, half;
  uint64_t blocks;
};

template <int R>
__device__ __forceinline__ WinGeom win_geom(const UnionArgs &a) {
  constexpr int T = WinLds<R>::kTile;
  WinGeom w;
  w.B = a.win_block && a.win_block < uint32_t(kWinB) ? a.win_block : uint32_t(kWinB);
  const uint32_t hmax = uint32_t(T - w.B) / 2;
  w.half = a.win_half && a.win_half < hmax ? a.win_half : hmax;
  w.blocks = (a.count + w.B - 1) / w.B;
  return w;
}

// Rows longer than a.heavy_degree: heavy row h of a.heavy_list, the whole
// workgroup (visited arcs counted at staging, as for every node, except in
// worklist iterations).
template <int R>
__device__ __forceinline__ void win_heavy_row(const UnionArgs &a, uint64_t h, WinLds<R> &L,
                                              Acc &acc) {
  if (a.flags & (F_SYSTOLIC | F_LIST))  // worklist iterations count visited arcs per row
    heavy_node<R, 2, 1, kWinNT, true, true>(a, a.heavy_list[h], L.u.heavy.acc,
                                            L.u.heavy.ids, L.u.heavy.sum, acc);
  else
    heavy_node<R, 2, 1, kWinNT, false, true>(a, a.heavy_list[h], L.u.heavy.acc,
                                             L.u.heavy.ids, L.u.heavy.sum, acc);
}

// Block blk of the instance a: stage its tile, rank its nodes, work on them.
// Every thread of the workgroup calls it; L.cls and L.q are zero on entry and
// on return, and the workgroup has passed a barrier after the work.
template <int R, bool TEST, bool SYS>
__device__ __forceinline__ void win_block(const UnionArgs &a, uint64_t blk, const WinGeom &geo,
                                          WinLds<R> &L, Acc &acc) {
  static_assert(TEST || !SYS, "worklist iterations test modified bits");
  constexpr int D = R, CD = 2 * R, T = WinLds<R>::kTile;
  constexpr uint32_t CS = counter_stride(CD);
  const int tid = threadIdx.x, g = __lane_id() / 2, gl = __lane_id() & 1;
  const uint64_t n = a.count;
  const uint64_t b0 = blk * geo.B;
  const uint32_t bn = uint32_t(n - b0 < geo.B ? n - b0 : geo.B);
  const uint64_t tlo = b0 > geo.half ? b0 - geo.half : 0;
  const uint64_t thi = b0 + bn + geo.half < n ? b0 + bn + geo.half : n;
  const uint32_t tlen = uint32_t(thi - tlo);
  const uint64_t wlo = tlo & ~uint64_t(31);

  // stage the tile (64-bit loads, eight in flight per thread), its modified
  // bits and the block's offsets; rank the block's nodes by class
  {
    constexpr int H = CD / 2;  // 64-bit words per counter
    constexpr int PER = (T * H + kWinNT - 1) / kWinNT;
    const uint2 *src = reinterpret_cast<const uint2 *>(a.curr);
    uint2 *dst = reinterpret_cast<uint2 *>(L.u.tile);
    const uint32_t words = tlen * uint32_t(H);
    constexpr int STEP = 8;
#pragma unroll 1
    for (int j0 = 0; j0 < PER; j0 += STEP) {
      uint2 w[STEP];
#pragma unroll
      for (int j = 0; j < STEP; j++) {
        const uint32_t k = uint32_t((j0 + j) * kWinNT + tid);
        w[j] = k < words ? src[(tlo + k / H) * (CS / 2) + k % H] : make_uint2(0u, 0u);
      }
#pragma unroll
      for (int j = 0; j < STEP; j++) {
        const uint32_t k = uint32_t((j0 + j) * kWinNT + tid);
        if (k < words) dst[k] = w[j];
      }
    }
  }
  {
    const uint64_t w1 = (thi - 1) >> 5;
    for (uint64_t k = tid; k <= w1 - (wlo >> 5); k += kWinNT) L.tmod[k] = a.curr_mod[(wlo >> 5) + k];
  }
  for (uint32_t k = tid; k <= bn; k += kWinNT) L.off[k] = a.off[b0 + k];
  uint32_t my_cls = kWinClasses, my_rank = 0;
  if (uint32_t(tid) < bn) {
    const uint64_t v = b0 + tid;
    const uint64_t deg = a.off[v + 1] - a.off[v];
    const bool was_mod = test_bit(a.curr_mod, uint32_t(v));
    // worklist iterations (F_LIST): the listed nodes (must be checked, or
    // modified), all checked in local iterations (F_CHECK_ALL), and the
    // unchecked modified ones copied; otherwise every node, checked
    const bool must = SYS && test_bit(a.must, uint32_t(v));
    const bool listed = !SYS || !(a.flags & F_LIST) || must || was_mod;
    const bool check = !SYS || (a.flags & F_CHECK_ALL) || must;
    if (deg <= a.heavy_degree) {
      if (listed && check) {
        my_cls = deg > uint64_t(kWinMedium) ? 0u : uint32_t(kWinMedium + 1 - deg);
        my_rank = atomicAdd(&L.cls[my_cls], 1u);
        acc.scanned += deg;
        acc.checked++;
      } else if (SYS && was_mod) {
        constexpr int CDW = 2 * R;
        const uint32_t *src = a.curr + v * CS;
        uint32_t *dst = a.next + v * CS;
#pragma unroll
        for (int b = 0; b < CDW; b++) dst[b] = src[b];
        acc.copied++;
      }
    }
    if (!SYS && was_mod) {
      const uint32_t k = a.indeg[v];
      acc.visited += k;
      acc.gathered += 2 * uint64_t(k);
    }
  }
  __syncthreads();
  if (tid < 64) {  // exclusive prefix sums of the class sizes, one class per lane
    const uint32_t c = L.cls[tid];
    uint32_t x = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (tid >= o) x += y;
    }
    L.base[tid] = x - c;
    if (tid == 63) L.base[kWinClasses] = x;
  }
  __syncthreads();
  if (my_cls < uint32_t(kWinClasses)) L.list[L.base[my_cls] + my_rank] = uint16_t(tid);
  __syncthreads();

  // work items: the wavefront rows, then rounds of 32 light rows
  const uint32_t nw = __builtin_amdgcn_readfirstlane(L.cls[0]);
  const uint32_t nl = __builtin_amdgcn_readfirstlane(L.base[kWinClasses]) - nw;
  const uint32_t items = nw + (nl + 31) / 32;
  for (;;) {
    uint32_t it = 0;
    if (__lane_id() == 0) it = atomicAdd(&L.q, 1u);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= items) break;
    const bool wave_row = it < nw;
    const uint32_t pos = wave_row ? it : nw + (it - nw) * 32 + uint32_t(g);
    if (pos < nw + nl) {
      const uint32_t p = L.list[pos];
      const uint32_t v = uint32_t(b0 + p);
      WinNode wn{};
      if (gl == 0 && (!wave_row || g == 0)) wn = win_prefetch(a, v);
      const uint64_t beg = L.off[p], end = L.off[p + 1];
      const uint64_t own = uint64_t(v) - tlo;
      uint32_t prev[D], cur[D];
#pragma unroll
      for (int b = 0; b < D; b++) prev[b] = cur[b] = L.u.tile[own * CD + gl * D + b];
      uint32_t nvisit = 0;
      if (wave_row) {
        win_merge<R, TEST, SYS>(a, v, gl, beg + HB_WIN_U * uint64_t(g), end, HB_WIN_U * 32,
                                L.u.tile, L.tmod, tlo, tlen, cur, nvisit);
#pragma unroll
        for (int o = 2; o < 64; o <<= 1) {
          uint32_t y[D];
#pragma unroll
          for (int b = 0; b < D; b++) y[b] = __shfl_xor(cur[b], o);
          slab_max<R>(cur, y);
        }
        if (SYS) {
          nvisit = gl == 0 ? nvisit : 0u;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) nvisit += __shfl_xor(nvisit, o);
        }
      } else {
        win_merge<R, TEST, SYS>(a, v, gl, beg, end, HB_WIN_U, L.u.tile, L.tmod, tlo, tlen, cur,
                                nvisit);
      }
      const uint64_t ob = uint64_t(v) - wlo;
      const bool was_mod = (L.tmod[ob >> 5] >> (ob & 31)) & 1u;
      if (!wave_row || g == 0)
        win_finish<R, SYS>(a, v, gl, was_mod, prev, cur, nvisit, wn, acc);
    }
  }
  __syncthreads();
  if (tid < kWinClasses) L.cls[tid] = 0;
  if (tid == 0) L.q = 0;
}

template <int R, bool TEST, bool SYS>
__global__ __launch_bounds__(kWinNT) void window_kernel(UnionArgs a) {
  using LT = WinLds<R>;
  static_assert(LT::kTile >= kWinB + 64, "the tile holds the block");
  static_assert(sizeof(typename LT::U) >= sizeof(uint32_t) * (kWinNT * (R + 2 * ids_per_lane<2>()) + 2),
                "the heavy rows' buffers fit the tile");
  static_assert(kWinClasses == 64 && kWinNT >= 64, "one class per lane of the first wavefront");
  __shared__ LT L;
  Acc acc;
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid == 0) a.stats->list_len = a.count;

  // rows longer than heavy_degree: the whole workgroup each
  if (a.heavy_count) {
    for (;;) {
      if (tid == 0) L.item = atomicAdd(a.work + 8 * kWorkStride, 1u);
      __syncthreads();
      const uint64_t h = L.item;
      __syncthreads();
      if (h >= a.heavy_count) break;
      win_heavy_row<R>(a, h, L, acc);
    }
  }

  const WinGeom geo = win_geom<R>(a);
  const uint64_t S = gridDim.x / 8;
  if (tid < kWinClasses) L.cls[tid] = 0;
  if (tid == 0) {
    L.item = first_item(a.work, geo.blocks, S, blockIdx.x / 8);
    L.q = 0;
  }
  __syncthreads();
  for (;;) {
    const uint64_t blk = uniform_u64(L.item);
    if (blk == UINT64_MAX) break;
    win_block<R, TEST, SYS>(a, blk, geo, L, acc);
    if (tid == 0) L.item = grab_item(a.work, geo.blocks, S);
    __syncthreads();
  }
  flush_stats(acc, a.stats, a.stat_slots);
}

__global__ __launch_bounds__(kBlock) void indegree_kernel(const uint64_t *off, const uint32_t *succ,
                                                          uint64_t n, uint32_t *indeg) {
  const uint64_t waves = uint64_t(gridDim.x) * (kBlock / 64);
  for (uint64_t v = uint64_t(blockIdx.x) * (kBlock / 64) + threadIdx.x / 64; v < n; v += waves) {
    const uint64_t e = off[v + 1];
    for (uint64_t k = off[v] + __lane_id(); k < e; k += 64) {
      const uint32_t s = succ[k];
      if (s != v) atomicAdd(indeg + s, 1u);
    }
  }
}

// Seeding without weights, fused with the count: G lanes per node write the
// whole counter (zero but for the register of the node's hash, as
// seed_kernel leaves it in a zeroed array) and the change mask, and store
// count() of it.  With one non-zero register a counter has m - 1 zeros and a
// harmonic sum above m - 1, so the raw estimate alpha_m_m / sum stays below
// 0.7214 (m + 2) < 2.5 m: count() is the linear-counting value
// lc[m - 1] for every node, as group_count returns it.  Padding dwords beyond
// the counter are not written; nothing reads them.  Counters of NSEG
// segments are written segment by segment.
template <int R, int G, int SPL, int NSEG>
__global__ __launch_bounds__(kBlock) void seed_count_kernel(SeedArgs a, uint32_t m,
                                                           const double *lc, double *out) {
  constexpr int D = R * SPL, CD = G * D, NPB = kBlock / G;
  const int g = threadIdx.x / G, gl = threadIdx.x % G;
  for (uint64_t batch = blockIdx.x; batch * NPB < a.n; batch += gridDim.x) {
    const uint64_t i = batch * NPB + g;
    if (i < a.n) {
      uint32_t reg = 0, val = 0;
      seed_register(a.hash, a.seed, a.log2m, a.sentinel, i, reg, val);
#pragma unroll 1
      for (int seg = 0; seg < NSEG; seg++) {
        uint32_t x[D];
#pragma unroll
        for (int q = 0; q < SPL; q++) {
          const bool mine = (reg >> 5) == uint32_t(seg) * G * SPL + uint32_t(gl) * SPL + q;
#pragma unroll
          for (int b = 0; b < R; b++)
            x[q * R + b] = mine && ((val >> b) & 1u) ? 1u << (reg & 31) : 0u;
        }
        store_part<D>(a.counters + i * a.cdw + seg * CD + gl * D, x);
      }
      if (gl == 0) {
        out[i] = lc[m - 1];
        a.chg[i] = 1u << ((reg >> 5) >> a.chg_shift);
        if (a.items) a.items[i] = reg | (val << 16);
      }
    }
  }
}

template <int R, int G, int SPL, int NSEG>
__global__ __launch_bounds__(kBlock) void count_all_kernel(const uint32_t *cnt, uint64_t n,
                                                          double alpha_m_m, uint32_t m,
                                                          const double *lc, double *out) {
  constexpr int D = R * SPL, CD = G * D, NPB = kBlock / G;
  const int g = threadIdx.x / G, gl = threadIdx.x % G;
  for (uint64_t batch = blockIdx.x; batch * NPB < n; batch += gridDim.x) {
    const uint64_t i = batch * NPB + g;
    if (i < n) {
      int64_t s = 0;
      uint64_t hi = 0, lo = 0;
      uint32_t zeros = 0;
#pragma unroll 1
      for (int seg = 0; seg < NSEG; seg++) {
        uint32_t x[D];
        load_part<D>(x, cnt + i * counter_stride(NSEG * CD) + seg * CD + gl * D);
        lane_sums<R, SPL>(x, s, hi, lo, zeros);
      }
      const double c = group_estimate<R, G>(s, hi, lo, zeros, alpha_m_m, m, lc);
      if (gl == 0) out[i] = c;
    }
  }
}

// HyperBall::init seeding (hyperball_impl.rs:1167-1178) with
// HyperLogLog::add (hyper_log_log_logic.rs:151-167), one thread per node.
__global__ __launch_bounds__(kBlock) void seed_kernel(SeedArgs a) {
  for (uint64_t v = uint64_t(blockIdx.x) * kBlock + threadIdx.x; v < a.n;
       v += uint64_t(gridDim.x) * kBlock) {
    uint32_t *c = a.counters + v * a.cdw;
    const uint64_t items = a.weights ? a.weights[v] : 1;
    const uint64_t first = a.weights ? a.weight_prefix[v] : 0;
    uint32_t chg = 0;
    for (uint64_t k = 0; k < items; k++) {
      const uint64_t item = a.weights ? weight_item(a.weight_seed, first + k) : v;
      uint32_t j = 0, val = 0;
      seed_register(a.hash, a.seed, a.log2m, a.sentinel, item, j, val);
      chg |= 1u << ((j >> 5) >> a.chg_shift);
      uint32_t *slab = c + (j >> 5) * a.r;
      const uint32_t q = j & 31;
      uint32_t cur = 0;
      for (uint32_t b = 0; b < a.r; b++) cur |= ((slab[b] >> q) & 1u) << b;
      if (val > cur)
        for (uint32_t b = 0; b < a.r; b++)
          slab[b] = (slab[b] & ~(1u << q)) | (((val >> b) & 1u) << q);
    }
    a.chg[v] = chg;
  }
}

// Start of an iteration in one launch: clears next_mod and (when non-null)
// next_must, `words` words each; workgroup 0 also zeroes the statistics, the
// work counters and the statistics slots (nullable).
__global__ __launch_bounds__(kBlock) void prepare_iteration_kernel(uint32_t *next_mod,
                                                                   uint32_t *next_must,
                                                                   uint64_t words, Stats *st,
                                                                   uint32_t *work,
                                                                   uint64_t *slots) {
  for (uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x; w < words;
       w += uint64_t(gridDim.x) * kBlock) {
    next_mod[w] = 0;
    if (next_must) next_must[w] = 0;
  }
  if (blockIdx.x != 0) return;
  uint32_t *p = reinterpret_cast<uint32_t *>(st);
  for (uint32_t i = threadIdx.x; i < sizeof(Stats) / 4; i += blockDim.x) p[i] = 0;
  for (uint32_t i = threadIdx.x; i < kWorkCounters * kWorkStride; i += blockDim.x) work[i] = 0;
  if (slots)
    for (uint32_t i = threadIdx.x; i < kStatSlots * kStatSlotWords; i += blockDim.x) slots[i] = 0;
}

// One workgroup of kStatSlots threads: thread t reads slot t; wavefront sums,
// then the kStatSlots / 64 partial sums of each field are added in LDS.
__device__ __forceinline__ void reduce_slots(Stats *st, const uint64_t *slots) {
  __shared__ uint64_t part[kStatSlots / 64][kStatFields];
  uint64_t v[kStatFields];
#pragma unroll
  for (int k = 0; k < kStatFields; k++) v[k] = slots[uint64_t(threadIdx.x) * kStatSlotWords + k];
#pragma unroll
  for (int k = 0; k < kStatFields; k++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
  if (__lane_id() == 0)
#pragma unroll
    for (int k = 0; k < kStatFields; k++) part[threadIdx.x / 64][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kStatFields) {
    uint64_t sum = 0;
    for (int w = 0; w < kStatSlots / 64; w++) sum += part[w][threadIdx.x];
    reinterpret_cast<uint64_t *>(st)[threadIdx.x] = sum;  // Stats field order
  }
}

// out <- the slot sums, and st's list_len (written by the union kernel).  out
// may be pinned host memory: the caller then reads it after synchronising,
// with no copy.
__global__ __launch_bounds__(kStatSlots) void reduce_stats_kernel(const Stats *st,
                                                                  const uint64_t *slots,
                                                                  Stats *out) {
  reduce_slots(out, slots);
  if (threadIdx.x == kStatFields) out->list_len = st->list_len;
}

// ---- bitmap compaction ---------------------------------------------------

__device__ __forceinline__ uint32_t union_word(const uint32_t *a, const uint32_t *b, uint64_t w,
                                               uint64_t n) {
  const uint64_t nw = (n + 31) / 32;
  if (w >= nw) return 0;
  uint32_t x = a[w] | (b ? b[w] : 0u);
  if (w == nw - 1 && (n & 31)) x &= (1u << (n & 31)) - 1;
  return x;
}

// Exclusive block scan of one value per thread; returns the block total too.
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t s_w[kBlock / 64];
  const int lane = __lane_id(), w = threadIdx.x / 64;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int k = 0; k < kBlock / 64; k++) {
    if (k < w) before += s_w[k];
    tot += s_w[k];
  }
  __syncthreads();
  *total = tot;
  return before + inc - v;
}

// Worklist iterations skip most heavy rows: this keeps those whose node must
// be checked or was modified, so that union_kernel's heavy phase does not
// line its workgroups up on the row counter while they skip the rest.  One
// atomic per wavefront appends its selected rows.
__global__ __launch_bounds__(kBlock) void select_heavy_kernel(const uint32_t *heavy_list,
                                                              uint64_t heavy_count,
                                                              const uint32_t *must,
                                                              const uint32_t *curr_mod,
                                                              uint32_t *out, uint32_t *out_len) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t base = uint64_t(blockIdx.x) * kBlock + (threadIdx.x & ~63u); base < heavy_count;
       base += stride) {
    const uint64_t h = base + __lane_id();
    uint32_t v = 0;
    bool sel = false;
    if (h < heavy_count) {
      v = heavy_list[h];
      sel = test_bit(must, v) || test_bit(curr_mod, v);
    }
    uint32_t x = sel ? 1u : 0u;  // inclusive prefix count over the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (__lane_id() >= uint32_t(o)) x += y;
    }
    const uint32_t total = __shfl(x, 63);
    uint32_t at = 0;
    if (__lane_id() == 0 && total) at = atomicAdd(out_len, total);
    at = __shfl(at, 0);
    if (sel) out[at + x - 1] = v;
  }
}

__global__ __launch_bounds__(kBlock) void compact_count_kernel(const uint32_t *a, const uint32_t *b,
                                                               uint64_t n, uint32_t *bsum) {
  const uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t tot;
  block_exclusive_scan(__popc(union_word(a, b, w, n)), &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of bsum[0, nb) in place by one 1024-thread block; *len gets
// the total.
__device__ void scan_block_sums(uint32_t *bsum, uint32_t nb, uint32_t *len) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const int lane = __lane_id(), w = threadIdx.x / 64;
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o);
      if (lane >= o) inc += t;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = s_carry, tot = 0;
    for (int k = 0; k < 16; k++) {
      if (k < w) before += s_w[k];
      tot += s_w[k];
    }
    if (i < nb) bsum[i] = before + inc - v;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *len = s_carry;
}

__global__ __launch_bounds__(1024) void compact_scan_kernel(uint32_t *bsum, uint32_t nb,
                                                            uint32_t *len) {
  scan_block_sums(bsum, nb, len);
}

__global__ __launch_bounds__(kBlock) void compact_emit_kernel(const uint32_t *a, const uint32_t *b,
                                                              uint64_t n, const uint32_t *boff,
                                                              uint32_t *list) {
  const uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t x = union_word(a, b, w, n);
  uint32_t tot;
  uint32_t pos = boff[blockIdx.x] + block_exclusive_scan(__popc(x), &tot);
  while (x) {
    const int bit = __builtin_ctz(x);
    x &= x - 1;
    list[pos++] = uint32_t(w * 32 + bit);
  }
}

// must |= the predecessors (transpose rows) of the nodes list[0, *len): one
// wavefront per listed node, its lanes striding over the row.
__global__ __launch_bounds__(kBlock) void mark_predecessors_kernel(const uint32_t *list,
                                                                   const uint32_t *len,
                                                                   const uint64_t *toff,
                                                                   const uint32_t *tsucc,
                                                                   uint32_t *must) {
  const uint64_t count = *len;
  const uint64_t waves = uint64_t(gridDim.x) * (kBlock / 64);
  for (uint64_t i = uint64_t(blockIdx.x) * (kBlock / 64) + threadIdx.x / 64; i < count; i += waves) {
    const uint32_t v = list[i];
    const uint64_t e = toff[v + 1];
    for (uint64_t k = toff[v] + __lane_id(); k < e; k += 64) {
      const uint32_t p = tsucc[k];
      atomicOr(must + (p >> 5), 1u << (p & 31));
    }
  }
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t words, uint32_t value) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = value;
}

__global__ void fill_bits_kernel(uint32_t *p, uint64_t n) {
  const uint64_t nw = (n + 31) / 32;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nw;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = (i == nw - 1 && (n & 31)) ? (1u << (n & 31)) - 1 : ~0u;
}

// Slab layout -> reference layout, one thread per output word.
__global__ void to_reference_kernel(const uint32_t *cnt, uint64_t node_begin, uint64_t node_count,
                                    uint32_t r, uint32_t m, uint64_t *out) {
  const uint32_t W = m * r / 64, cdw = m * r / 32;
  const uint64_t total = node_count * W;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t v = node_begin + i / W;
    const uint32_t k = uint32_t(i % W);
    const uint32_t *c = cnt + v * counter_stride(cdw);
    uint64_t word = 0;
    for (uint32_t p = 0; p < 64; p++) {
      const uint32_t g = 64 * k + p, j = g / r, b = g % r;
      word |= uint64_t((c[(j >> 5) * r + b] >> (j & 31)) & 1u) << p;
    }
    out[i] = word;
  }
}

// ---- batched execution ----------------------------------------------------

#define HB_BATCH_SHAPES(X)                                                              \
  X(0, 5, HB_G_6, HB_SPL_6) X(1, 5, HB_G_8, HB_SPL_8) X(2, 5, HB_G_10, HB_SPL_10)             \
  X(3, 6, HB_G_6, HB_SPL_6) X(4, 6, HB_G_8, HB_SPL_8) X(5, 6, HB_G_10, HB_SPL_10)

__device__ __forceinline__ uint32_t shape_nodes_per_wave(uint32_t shape) {
  switch (shape) {
#define X(ID, RR, GG, SS) \
  case ID: return 64 / GG;
    HB_BATCH_SHAPES(X)
#undef X
  }
  return 1;
}

// Instance whose [base, end) range of work items contains x, among the
// instances of one shape (`inst`, in range order).
__device__ __forceinline__ uint32_t find_in_shape(const BatchDesc *d, const uint32_t *inst, int cnt,
                                                  uint64_t x, bool heavy) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    const uint64_t e = heavy ? d[inst[mid]].heavy_end : d[inst[mid]].chunk_end;
    if (e > x) hi = mid;
    else lo = mid + 1;
  }
  return inst[lo];
}

// grid (x, instance): clears the next modified bitmap (and next must-be-checked
// bitmap in systolic iterations) of active instances, and the statistics.
__global__ __launch_bounds__(kBlock) void batch_prep_kernel(const BatchDesc *d, Stats *stats,
                                                            uint32_t *work) {
  const BatchDesc &x = d[blockIdx.y];
  if (blockIdx.x == 0) {
    uint32_t *p = reinterpret_cast<uint32_t *>(stats + blockIdx.y);
    for (uint32_t i = threadIdx.x; i < sizeof(Stats) / 4; i += blockDim.x) p[i] = 0;
    if (blockIdx.y == 0)
      for (uint32_t i = threadIdx.x; i < kBatchWorkWords; i += blockDim.x) work[i] = 0;
    if (x.active && x.u.stat_slots)
      for (uint32_t i = threadIdx.x; i < kStatSlots * kStatSlotWords; i += blockDim.x)
        x.u.stat_slots[i] = 0;
  }
  if (!x.active) return;
  for (uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; w < x.words;
       w += uint64_t(gridDim.x) * blockDim.x) {
    x.u.next_mod[w] = 0;
    if (x.clear_must) x.u.next_must[w] = 0;
    if (x.mark_first) const_cast<uint32_t *>(x.u.must)[w] = 0;
  }
}

// The batched worklist compaction runs in two phases.  Marking phase: the
// modified nodes of the instances that mark predecessors first.  Worklist
// phase: must | modified of the instances in list mode.
__device__ __forceinline__ bool compact_this(const BatchDesc &x, int marking) {
  return x.active && (marking ? x.mark_first != 0 : x.list_mode != 0);
}
__device__ __forceinline__ uint32_t compact_word(const BatchDesc &x, int marking, uint64_t w) {
  return marking ? union_word(x.u.curr_mod, nullptr, w, x.n)
                 : union_word(x.u.must, x.u.curr_mod, w, x.n);
}

__global__ __launch_bounds__(kBlock) void batch_compact_count_kernel(const BatchDesc *d,
                                                                     int marking) {
  const BatchDesc &x = d[blockIdx.y];
  if (!compact_this(x, marking)) return;
  const uint64_t nb = ((x.n + 31) / 32 + kBlock - 1) / kBlock;
  if (blockIdx.x >= nb) return;
  const uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t tot;
  block_exclusive_scan(__popc(compact_word(x, marking, w)), &tot);
  if (threadIdx.x == 0) x.bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void batch_compact_scan_kernel(const BatchDesc *d, int marking) {
  const BatchDesc &x = d[blockIdx.x];
  if (!compact_this(x, marking)) return;
  const uint32_t nb = uint32_t(((x.n + 31) / 32 + kBlock - 1) / kBlock);
  scan_block_sums(x.bsum, nb, const_cast<uint32_t *>(x.u.list_len));
}

__global__ __launch_bounds__(kBlock) void batch_compact_emit_kernel(const BatchDesc *d,
                                                                    int marking) {
  const BatchDesc &x = d[blockIdx.y];
  if (!compact_this(x, marking)) return;
  const uint64_t nb = ((x.n + 31) / 32 + kBlock - 1) / kBlock;
  if (blockIdx.x >= nb) return;
  const uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t bits = compact_word(x, marking, w);
  uint32_t tot;
  uint32_t pos = x.bsum[blockIdx.x] + block_exclusive_scan(__popc(bits), &tot);
  uint32_t *list = const_cast<uint32_t *>(x.u.list);
  while (bits) {
    const int b = __builtin_ctz(bits);
    bits &= bits - 1;
    list[pos++] = uint32_t(w * 32 + b);
  }
}

// grid (x, instance): must |= the predecessors of the listed nodes of the
// instances that mark them first (mark_predecessors_kernel, per instance).
__global__ __launch_bounds__(kBlock) void batch_mark_kernel(const BatchDesc *d) {
  const BatchDesc &x = d[blockIdx.y];
  if (!x.active || !x.mark_first) return;
  const uint64_t count = *x.u.list_len;
  const uint64_t waves = uint64_t(gridDim.x) * (kBlock / 64);
  uint32_t *must = const_cast<uint32_t *>(x.u.must);
  for (uint64_t i = uint64_t(blockIdx.x) * (kBlock / 64) + threadIdx.x / 64; i < count; i += waves) {
    const uint32_t v = x.u.list[i];
    const uint64_t e = x.u.toff[v + 1];
    for (uint64_t k = x.u.toff[v] + __lane_id(); k < e; k += 64) {
      const uint32_t p = x.u.tsucc[k];
      atomicOr(must + (p >> 5), 1u << (p & 31));
    }
  }
}

// One workgroup per instance: the statistics slots of an active instance
// summed into its Stats.
__global__ __launch_bounds__(kStatSlots) void batch_reduce_stats_kernel(const BatchDesc *d) {
  const BatchDesc &x = d[blockIdx.x];
  if (!x.active || !x.u.stat_slots) return;
  reduce_slots(x.u.stats, x.u.stat_slots);
}

// Wavefronts of each shape's union grid.
struct ShapeWaves {
  uint64_t w[kBatchShapes];
};

// One thread: the positions of every active instance and, shape by shape,
// the rounds per chunk (from the shape's total positions over the wavefronts
// of its union grid) and the prefix sums of the instances' chunks (each shape
// has its own chunk space).
__global__ void batch_plan_kernel(BatchDesc *d, const uint32_t *shape_inst,
                                  const uint32_t *shape_off, Stats *stats, uint64_t *total_chunks,
                                  uint64_t *total_window, ShapeWaves waves) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int sh = 0; sh < kBatchShapes; sh++) {
    uint64_t total = 0;
    for (uint32_t j = shape_off[sh]; j < shape_off[sh + 1]; j++) {
      const BatchDesc &x = d[shape_inst[j]];
      if (x.active && !x.window) total += x.list_mode ? uint64_t(*x.u.list_len) : x.n;
    }
    const uint32_t npw = shape_nodes_per_wave(uint32_t(sh));
    const uint32_t rounds = chunk_rounds(total, npw, waves.w[sh]);
    const uint64_t ch = uint64_t(rounds) * npw;
    uint64_t acc = 0, wacc = 0, hacc = 0;
    for (uint32_t j = shape_off[sh]; j < shape_off[sh + 1]; j++) {
      BatchDesc &x = d[shape_inst[j]];
      uint64_t count = 0;
      if (x.active) count = x.list_mode ? uint64_t(*x.u.list_len) : x.n;
      stats[shape_inst[j]].list_len = count;
      const bool win = x.active && x.window;
      x.rounds = rounds;
      x.chunk_base = acc;
      acc += win ? 0 : (count + ch - 1) / ch;
      x.chunk_end = acc;
      const uint64_t B = x.u.win_block && x.u.win_block < uint32_t(kWinB) ? x.u.win_block : kWinB;
      x.wblock_base = wacc;
      wacc += win ? (count + B - 1) / B : 0;
      x.wblock_end = wacc;
      x.wheavy_base = hacc;
      hacc += win ? x.u.heavy_count : 0;
      x.wheavy_end = hacc;
    }
    total_chunks[sh] = acc;
    total_window[2 * sh] = wacc;
    total_window[2 * sh + 1] = hacc;
  }
}

// The union step of one batched shape: first the shape's heavy rows, a
// workgroup per row, from the shape's heavy-row space (prefix sums over its
// instances); then per-XCD queues over the shape's chunk space, a wavefront
// finding the instance of its chunk.  Statistics go to the instance they
// belong to.  The hint keeps every shape at three wavefronts per SIMD (at
// most 168 VGPRs, no spills); left alone, log2m 10 with 6-bit registers takes
// 175.
template <int R, int G, int SPL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
void batch_union_kernel(const BatchDesc *d, const uint32_t *inst, int cnt,
                        const uint64_t *total_chunks, uint64_t total_heavy, uint32_t *work) {
  __shared__ UnionLds<R, G, SPL> L;
  Acc acc;
  int64_t current = -1;
  for (;;) {
    if (!total_heavy) break;
    if (threadIdx.x == 0) L.u.heavy.h = atomicAdd(work + 8 * kWorkStride, 1u);
    __syncthreads();
    const uint64_t h = L.u.heavy.h;
    __syncthreads();
    if (h >= total_heavy) break;
    const uint32_t i = find_in_shape(d, inst, cnt, h, true);
    const BatchDesc &x = d[i];
    if (!x.active || x.window) continue;
    if (int64_t(i) != current) {
      if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
      acc = Acc();
      current = i;
    }
    heavy_node<R, G, SPL>(x.u, x.u.heavy_list[h - x.heavy_base], L.u.heavy.acc, L.ids,
                          L.u.heavy.sum, acc);
  }
  uint32_t *wave_ids = L.ids + (threadIdx.x / 64) * UnionLds<R, G, SPL>::kWaveIds;
  NodeMeta *meta = L.u.meta + (threadIdx.x / 64) * UnionLds<R, G, SPL>::kWaveMeta;
  const uint64_t chunks = *total_chunks;
  uint64_t c = 0;
  if (__lane_id() == 0) c = first_chunk(work, chunks);
  c = __shfl(c, 0);
  for (; c != UINT64_MAX; c = __shfl(c, 0)) {
    const uint32_t i = find_in_shape(d, inst, cnt, c, false);
    if (int64_t(i) != current) {
      if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
      acc = Acc();
      current = i;
    }
    const BatchDesc &x = d[i];
    const uint64_t count = x.list_mode ? uint64_t(*x.u.list_len) : x.n;
    run_chunk<R, G, SPL>(x.u, c - x.chunk_base, count, x.rounds, wave_ids, meta, acc);
    if (__lane_id() == 0) c = grab_chunk(work, chunks);
  }
  if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
}

// Instance whose window block range (or window heavy-row range) contains x,
// among the instances of one shape.
__device__ __forceinline__ uint32_t find_window(const BatchDesc *d, const uint32_t *inst, int cnt,
                                                uint64_t x, bool heavy) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    const uint64_t e = heavy ? d[inst[mid]].wheavy_end : d[inst[mid]].wblock_end;
    if (e > x) hi = mid;
    else lo = mid + 1;
  }
  return inst[lo];
}

// The window kernel for the instances of one 64-register shape whose
// iteration runs as a window iteration (BatchDesc::window): first their rows
// longer than the window threshold, a workgroup each, then their blocks, from
// the shape's window spaces (prefix sums over its instances, written by the
// plan step), dealt per XCD.  The other instances have empty ranges.
template <int R>
__global__ __launch_bounds__(kWinNT) void batch_window_kernel(const BatchDesc *d,
                                                              const uint32_t *inst, int cnt,
                                                              const uint64_t *total_window,
                                                              uint32_t *work) {
  __shared__ WinLds<R> L;
  Acc acc;
  int64_t current = -1;
  const int tid = threadIdx.x;
  const uint64_t heavy = total_window[1];
  while (heavy) {
    if (tid == 0) L.item = atomicAdd(work + 8 * kWorkStride, 1u);
    __syncthreads();
    const uint64_t h = L.item;
    __syncthreads();
    if (h >= heavy) break;
    const uint32_t i = find_window(d, inst, cnt, h, true);
    if (int64_t(i) != current) {
      if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
      acc = Acc();
      current = i;
    }
    win_heavy_row<R>(d[i].u, h - d[i].wheavy_base, L, acc);
  }
  const uint64_t blocks = total_window[0];
  const uint64_t S = gridDim.x / 8;
  if (tid < kWinClasses) L.cls[tid] = 0;
  if (tid == 0) {
    L.item = first_item(work, blocks, S, blockIdx.x / 8);
    L.q = 0;
  }
  __syncthreads();
  for (;;) {
    const uint64_t blk = uniform_u64(L.item);
    if (blk == UINT64_MAX) break;
    const uint32_t i = __builtin_amdgcn_readfirstlane(find_window(d, inst, cnt, blk, false));
    if (int64_t(i) != current) {
      if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
      acc = Acc();
      current = i;
    }
    const BatchDesc &x = d[i];
    const WinGeom geo = win_geom<R>(x.u);
    if (x.u.flags & (F_SYSTOLIC | F_LIST))
      win_block<R, true, true>(x.u, blk - x.wblock_base, geo, L, acc);
    else if (x.u.flags & F_UNTESTED)
      win_block<R, false, false>(x.u, blk - x.wblock_base, geo, L, acc);
    else
      win_block<R, true, false>(x.u, blk - x.wblock_base, geo, L, acc);
    if (tid == 0) L.item = grab_item(work, blocks, S);
    __syncthreads();
  }
  if (current >= 0) flush_stats(acc, d[current].u.stats, d[current].u.stat_slots);
}

// ---- streamed counters ------------------------------------------------------
// The counter arrays stay in host memory; an iteration walks the nodes tile by
// tile.  For each tile the host uploads the tile's own counters, then every
// source block that holds modified nodes; one merge pass per source range
// (the tile itself first) merges the modified successors inside that range
// into the partial next counters of the tile's checked nodes; a finish pass
// does the rest of the per-node body and leaves the tile's next counters for
// the download.  Registers do not depend on the order in which a node's
// successors are merged (a per-register maximum), so the passes give the
// union kernel's registers exactly.

// Whether node v is checked this iteration (the union kernel's rule: in list
// mode the list holds must | modified, and F_CHECK_ALL checks all of it).
__device__ __forceinline__ bool stream_checked(const UnionArgs &a, uint32_t v, bool was_mod) {
  const bool must = test_bit(a.must, v);
  if (a.flags & F_LIST) return (a.flags & F_CHECK_ALL) ? (must || was_mod) : must;
  return (a.flags & F_CHECK_ALL) || must;
}

// First position in succ[b, e) (sorted) holding a value >= x.
__device__ __forceinline__ uint64_t row_lower_bound(const uint32_t *succ, uint64_t b, uint64_t e,
                                                    uint64_t x) {
  while (b < e) {
    const uint64_t mid = b + (e - b) / 2;
    if (succ[mid] < x) b = mid + 1;
    else e = mid;
  }
  return b;
}

template <int R, int G, int SPL, int NSEG>
__global__ __launch_bounds__(kBlock) void stream_merge_kernel(UnionArgs a, StreamPass p) {
  constexpr int D = R * SPL, CD = G * D, K = ids_per_lane<G>(), SW = slice_width<G>();
  constexpr int NG = kBlock / G, U = D <= 6 ? 8 : D <= 12 ? 4 : 2;
  constexpr uint32_t CS = counter_stride(NSEG * CD);
  __shared__ uint32_t s_ids[NG * SW];
  const int g = threadIdx.x / G, gl = threadIdx.x % G;
  uint32_t *ids = s_ids + g * SW;
  const uint64_t below = (uint64_t(1) << gl) - 1;
  Acc acc;
  const uint64_t nt = p.tile_hi - p.tile_lo;
  for (uint64_t i = uint64_t(blockIdx.x) * NG + g; i < nt; i += uint64_t(gridDim.x) * NG) {
    const uint32_t v = uint32_t(p.tile_lo + i);
    if (!stream_checked(a, v, test_bit(a.curr_mod, v))) continue;
    const uint64_t rb = a.off[v], re = a.off[v + 1];
    const uint64_t b = row_lower_bound(a.succ, rb, re, p.lo);
    const uint64_t e = row_lower_bound(a.succ, b, re, p.hi);
    uint32_t nvisit = 0;
#pragma unroll 1
    for (int seg = 0; seg < NSEG; seg++) {
      uint32_t cur[D];
      const uint64_t at = i * CS + seg * CD + gl * D;
      load_part<D>(cur, (p.first ? p.own : p.part) + at);
      bool any = false;
      const uint32_t *lane_base = p.src + seg * CD + gl * D;
      for (uint64_t base = b; base < e; base += SW) {
        uint32_t sj[K];
        bool f[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
          const uint64_t idx = base + uint64_t(j) * G + gl;
          sj[j] = idx < e ? a.succ[idx] : v;
          f[j] = sj[j] != v && test_bit(a.curr_mod, sj[j]);
        }
        uint32_t cnt = 0;
#pragma unroll
        for (int j = 0; j < K; j++) {
          const uint64_t mask = group_ballot<G>(f[j]);
          if (f[j]) ids[cnt + __popcll(mask & below)] = sj[j];
          cnt += __popcll(mask);
        }
        if (!cnt) continue;
        lane_sync<G>();
        any = true;
        if (seg == 0) nvisit += cnt;
        acc.gathered += uint64_t(cnt) * SPL;
        for (uint32_t k = 0; k < cnt; k += U) {
          uint32_t y[U][D];
#pragma unroll
          for (int u = 0; u < U; u++) {
            const uint32_t t = ids[k + u < cnt ? k + u : k];
            load_part<D>(y[u], lane_base + uint64_t(t - p.lo) * CS);
          }
#pragma unroll
          for (int u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < SPL; q++) slab_max<R>(cur + q * R, y[u] + q * R);
        }
        lane_sync<G>();
      }
      if (p.first || any) store_part<D>(p.part + at, cur);
    }
    if (gl == 0) acc.visited += nvisit;
  }
  flush_stats(acc, a.stats, a.stat_slots);
}

template <int R, int G, int SPL, int NSEG>
__global__ __launch_bounds__(kBlock) void stream_finish_kernel(UnionArgs a, StreamPass p) {
  constexpr int D = R * SPL, CD = G * D, NG = kBlock / G;
  constexpr uint32_t CS = counter_stride(NSEG * CD);
  const int g = threadIdx.x / G, gl = threadIdx.x % G;
  Acc acc;
  const uint64_t nt = p.tile_hi - p.tile_lo;
  for (uint64_t i = uint64_t(blockIdx.x) * NG + g; i < nt; i += uint64_t(gridDim.x) * NG) {
    const uint32_t v = uint32_t(p.tile_lo + i);
    const bool was_mod = test_bit(a.curr_mod, v);
    const bool check = stream_checked(a, v, was_mod);
    int64_t s5 = 0;
    uint64_t hi = 0, lo = 0;
    uint32_t zeros = 0, chg_bits = 0;
#pragma unroll 1
    for (int seg = 0; seg < NSEG; seg++) {
      const uint64_t at = i * CS + seg * CD + gl * D;
      uint32_t prev[D];
      load_part<D>(prev, p.own + at);
      if (check) {
        uint32_t cur[D];
        load_part<D>(cur, p.part + at);
#pragma unroll
        for (int q = 0; q < SPL; q++) {
          uint32_t d = 0;
#pragma unroll
          for (int b = 0; b < R; b++) d |= cur[q * R + b] ^ prev[q * R + b];
          if (d) chg_bits |= 1u << ((uint32_t(seg) * G * SPL + uint32_t(gl) * SPL + q) >> a.chg_shift);
        }
        lane_sums<R, SPL>(cur, s5, hi, lo, zeros);
      } else {
        store_part<D>(p.part + at, prev);  // the tile's next counters are downloaded whole
      }
    }
    if (!check && !was_mod) continue;
    const bool cm = check && group_ballot<G>(chg_bits != 0) != 0;
    double post = 0.0;
    if (cm) {
      chg_bits = group_or_u32<G>(chg_bits);
      post = group_estimate<R, G>(s5, hi, lo, zeros, a.alpha_m_m, a.m, a.lc_table);
    }
    node_update<G>(a, v, gl, check, cm, chg_bits, post, 0u,
                   check ? a.off[v + 1] - a.off[v] : 0, acc);
  }
  flush_stats(acc, a.stats, a.stat_slots);
}

__global__ __launch_bounds__(kBlock) void stream_summary_kernel(UnionArgs a, uint64_t n,
                                                                uint64_t tile_nodes,
                                                                uint64_t block_nodes,
                                                                uint64_t tiles, uint32_t *counts) {
  const uint64_t words = (n + 31) / 32;
  const bool all = (a.flags & F_CHECK_ALL) && !(a.flags & F_LIST);
  for (uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x; w < words;
       w += uint64_t(gridDim.x) * kBlock) {
    const uint32_t valid = (w == words - 1 && (n & 31)) ? (1u << (n & 31)) - 1 : ~0u;
    const uint32_t mod = a.curr_mod[w] & valid;
    const uint32_t work = all ? valid : (a.must[w] | mod) & valid;
    if (work) atomicAdd(counts + (w * 32) / tile_nodes, uint32_t(__popc(work)));
    if (mod) atomicAdd(counts + tiles + (w * 32) / block_nodes, uint32_t(__popc(mod)));
  }
}

uint32_t grid_for(uint64_t items, uint32_t per_block, uint32_t cap) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return uint32_t(g);
}

// Workgroups of `kernel` resident on the current device at once (its
// workgroups per CU, from registers and LDS, times the CUs), a multiple of 8;
// 0 when unknown.  The kernels with static first chunks launch at most this
// many: a workgroup beyond it would start only when others have exited, and
// its static chunk would then run last.  Cached per kernel (`cache`) and
// device.
constexpr int kMaxDevices = 64;
template <class K>
uint32_t resident_grid(K kernel, std::atomic<uint32_t> *cache) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return 0;
  }
  uint32_t g = cache[dev].load(std::memory_order_relaxed);
  if (g) return g;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  g = uint32_t(std::max(1, per_cu) * std::max(1, cus));
  g = std::max<uint32_t>(8, g / 8 * 8);
  cache[dev].store(g, std::memory_order_relaxed);
  return g;
}

template <int R, int G, int SPL, int NSEG>
hipError_t launch_union_t(const UnionArgs &a, uint32_t grid, hipStream_t s) {
  static std::atomic<uint32_t> cache[kMaxDevices];
  const uint32_t cap = resident_grid(union_kernel<R, G, SPL, NSEG>, cache);
  if (cap && grid > cap) grid = cap;
  union_kernel<R, G, SPL, NSEG><<<grid, kBlock, 0, s>>>(a);
  return hipGetLastError();
}

template <int R, int G, int SPL, int NSEG>
hipError_t launch_seed_count_t(const SeedArgs &a, uint32_t m, const double *lc, double *out,
                               hipStream_t s) {
  seed_count_kernel<R, G, SPL, NSEG><<<grid_for(a.n, kBlock / G, 65536), kBlock, 0, s>>>(a, m, lc,
                                                                                        out);
  return hipGetLastError();
}

template <int R, int G, int SPL, int NSEG>
hipError_t launch_count_t(const uint32_t *c, uint64_t n, double amm, uint32_t m, const double *lc,
                          double *out, hipStream_t s) {
  count_all_kernel<R, G, SPL, NSEG><<<grid_for(n, kBlock / G, 65536), kBlock, 0, s>>>(c, n, amm, m,
                                                                                     lc, out);
  return hipGetLastError();
}

}  // namespace

// (r, log2m, lanes per node, slabs per lane, segments per counter).  Counters
// above 4096 registers are worked on one 64-lane segment of 128 slabs at a
// time.
#define HB_SHAPES(X)                                                                             \
  X(5, 6, HB_G_6, HB_SPL_6, 1) X(5, 7, 4, 1, 1) X(5, 8, HB_G_8, HB_SPL_8, 1) X(5, 9, 8, 2, 1)     \
  X(5, 10, HB_G_10, HB_SPL_10, 1) X(5, 11, 32, 2, 1) X(5, 12, 64, 2, 1) X(5, 13, 64, 2, 2)       \
  X(5, 14, 64, 2, 4) X(5, 15, 64, 2, 8) X(5, 16, 64, 2, 16) X(6, 5, 1, 1, 1)                    \
  X(6, 6, HB_G_6, HB_SPL_6, 1) X(6, 7, 4, 1, 1) X(6, 8, HB_G_8, HB_SPL_8, 1) X(6, 9, 8, 2, 1)     \
  X(6, 10, HB_G_10, HB_SPL_10, 1) X(6, 11, 32, 2, 1) X(6, 12, 64, 2, 1) X(6, 13, 64, 2, 2)       \
  X(6, 14, 64, 2, 4) X(6, 15, 64, 2, 8) X(6, 16, 64, 2, 16)

bool union_shape(int r, int log2m, int *lanes, int *spl, int *nseg) {
#define X(RR, LM, GG, SS, NS) \
  if (r == RR && log2m == LM) {   \
    *lanes = GG;                  \
    *spl = SS;                    \
    *nseg = NS;                   \
    return true;                  \
  }
  HB_SHAPES(X)
#undef X
  return false;
}

// The seeded iteration 0 reads no counters, so at log2m 8 and from log2m 10
// to 12 it takes half the lanes per node, with twice the slabs each: half as
// many lanes wait on each short row (log2m 8, r = 6: 3.84 -> 3.32 ms; log2m
// 10, r = 6: 7.17 -> 6.54 ms; log2m 12, r = 5: 22.3 -> 16.4 ms; a quarter of
// the lanes was slower at log2m 10, one lane instead of two at log2m 6 too).
#define HB_SEED_SHAPES(X)                                                                        \
  X(5, 8, 4, 2, 1) X(6, 8, 4, 2, 1) X(5, 10, 8, 4, 1) X(6, 10, 8, 4, 1) X(5, 11, 16, 4, 1)      \
  X(6, 11, 16, 4, 1) X(5, 12, 32, 4, 1) X(6, 12, 32, 4, 1)

hipError_t launch_union(int r, int log2m, const UnionArgs &a, uint32_t grid, hipStream_t s) {
#define X(RR, LM, GG, SS, NS) \
  if ((a.flags & F_SEEDED) && r == RR && log2m == LM) return launch_union_t<RR, GG, SS, NS>(a, grid, s);
  HB_SEED_SHAPES(X)
#undef X
#define X(RR, LM, GG, SS, NS) \
  if (r == RR && log2m == LM) return launch_union_t<RR, GG, SS, NS>(a, grid, s);
  HB_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

bool window_union_supported(int r, int log2m) { return (r == 5 || r == 6) && log2m == 6; }

uint32_t window_block_nodes() { return uint32_t(kWinB); }

uint32_t window_tile_nodes(int r) {
  return r == 5 ? uint32_t(WinLds<5>::kTile) : r == 6 ? uint32_t(WinLds<6>::kTile) : 0u;
}

template <int R, bool TEST, bool SYS>
static hipError_t launch_window_t(const UnionArgs &a, hipStream_t s) {
  static std::atomic<uint32_t> cache[kMaxDevices];
  uint32_t grid = 0;
  {
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDevices) {
      grid = cache[dev].load(std::memory_order_relaxed);
      if (!grid) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, window_kernel<R, TEST, SYS>,
                                                         kWinNT, 0) ==
                hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                hipSuccess) {
          grid = uint32_t(std::max(1, per_cu) * std::max(1, cus));
          cache[dev].store(grid, std::memory_order_relaxed);
        }
      }
    }
    (void)hipGetLastError();
  }
  const uint64_t B = a.win_block && a.win_block < uint32_t(kWinB) ? a.win_block : kWinB;
  const uint64_t blocks = (a.count + B - 1) / B;
  uint64_t g = grid ? grid : 256;
  g = std::min<uint64_t>(g, std::max<uint64_t>(blocks, a.heavy_count));
  g = std::max<uint64_t>(8, (g + 7) / 8 * 8);
  window_kernel<R, TEST, SYS><<<uint32_t(g), kWinNT, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_window_union(int r, int log2m, const UnionArgs &a, hipStream_t s) {
  if (!window_union_supported(r, log2m) || !a.indeg) return hipErrorInvalidValue;
  const bool sys = a.flags & (F_SYSTOLIC | F_LIST);
  const bool test = a.flags & F_UNTESTED ? false : true;
  if (r == 5)
    return sys ? launch_window_t<5, true, true>(a, s)
               : test ? launch_window_t<5, true, false>(a, s) : launch_window_t<5, false, false>(a, s);
  return sys ? launch_window_t<6, true, true>(a, s)
             : test ? launch_window_t<6, true, false>(a, s) : launch_window_t<6, false, false>(a, s);
}

hipError_t launch_indegree(const uint64_t *off, const uint32_t *succ, uint64_t n, uint32_t *indeg,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  indegree_kernel<<<grid_for(n, kBlock / 64, 65536), kBlock, 0, s>>>(off, succ, n, indeg);
  return hipGetLastError();
}

int batch_shape(int r, int log2m) {
  if ((r != 5 && r != 6) || (log2m != 6 && log2m != 8 && log2m != 10)) return -1;
  return (r - 5) * 3 + (log2m - 6) / 2;
}

// Grid of the batched union kernel of shape `sh`: L.union_grid, capped at
// the kernel's resident workgroups.
static uint32_t batch_union_grid(int sh, uint32_t grid) {
  uint32_t cap = 0;
  switch (sh) {
#define X(ID, RR, GG, SS)                                                       \
  case ID: {                                                                    \
    static std::atomic<uint32_t> cache[kMaxDevices];                            \
    cap = resident_grid(batch_union_kernel<RR, GG, SS>, cache);                 \
    break;                                                                      \
  }
    HB_BATCH_SHAPES(X)
#undef X
  }
  return cap && grid > cap ? cap : grid;
}

// Workgroups of a batched window kernel: the resident ones, a multiple of 8.
static uint32_t batch_window_grid(int r) {
  static std::atomic<uint32_t> cache[2][kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return 256;
  }
  std::atomic<uint32_t> &c = cache[r - 5][dev];
  uint32_t g = c.load(std::memory_order_relaxed);
  if (g) return g;
  int per_cu = 0, cus = 0;
  const hipError_t e = r == 5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                    &per_cu, batch_window_kernel<5>, kWinNT, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                    &per_cu, batch_window_kernel<6>, kWinNT, 0);
  if (e != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 256;
  }
  g = std::max<uint32_t>(8, uint32_t(std::max(1, per_cu) * std::max(1, cus)) / 8 * 8);
  c.store(g, std::memory_order_relaxed);
  return g;
}

hipError_t launch_batch_iteration(const BatchLaunch &L, hipStream_t s) {
  ShapeWaves waves{};
  uint32_t grid[kBatchShapes];
  for (int sh = 0; sh < kBatchShapes; sh++) {
    grid[sh] = batch_union_grid(sh, L.union_grid);
    waves.w[sh] = uint64_t(grid[sh]) * (kBlock / 64);
  }
  batch_prep_kernel<<<dim3(L.prep_blocks, L.k), kBlock, 0, s>>>(L.d, L.stats, L.work);
  for (int marking = 1; marking >= 0; marking--) {
    batch_compact_count_kernel<<<dim3(L.compact_blocks, L.k), kBlock, 0, s>>>(L.d, marking);
    batch_compact_scan_kernel<<<L.k, 1024, 0, s>>>(L.d, marking);
    batch_compact_emit_kernel<<<dim3(L.compact_blocks, L.k), kBlock, 0, s>>>(L.d, marking);
    if (marking) batch_mark_kernel<<<dim3(64, L.k), kBlock, 0, s>>>(L.d);
  }
  batch_plan_kernel<<<1, 64, 0, s>>>(L.d, L.shape_inst, L.shape_off, L.stats, L.total_chunks,
                                     L.total_window, waves);
  for (int sh = 0; sh < kBatchShapes; sh++) {
    if (!L.count[sh]) continue;
    const uint32_t *inst = L.shape_inst + L.first[sh];
    const int cnt = int(L.count[sh]);
    uint32_t *work = L.work + size_t(sh) * kWorkCounters * kWorkStride;
    switch (sh) {
#define X(ID, RR, GG, SS)                                                                      \
  case ID:                                                                                     \
    batch_union_kernel<RR, GG, SS><<<grid[sh], kBlock, 0, s>>>(L.d, inst, cnt,                \
                                                                L.total_chunks + sh,           \
                                                                L.total_heavy[sh], work);      \
    break;
      HB_BATCH_SHAPES(X)
#undef X
    }
    // 64-register shapes: the instances in a window iteration
    uint32_t *wwork = L.work + (kBatchShapes + size_t(sh)) * kWorkCounters * kWorkStride;
    const uint64_t *tw = L.total_window + 2 * sh;
    if (sh == batch_shape(5, 6))
      batch_window_kernel<5><<<batch_window_grid(5), kWinNT, 0, s>>>(L.d, inst, cnt, tw, wwork);
    else if (sh == batch_shape(6, 6))
      batch_window_kernel<6><<<batch_window_grid(6), kWinNT, 0, s>>>(L.d, inst, cnt, tw, wwork);
  }
  batch_reduce_stats_kernel<<<L.k, kStatSlots, 0, s>>>(L.d);
  return hipGetLastError();
}

hipError_t launch_count_all(int r, int log2m, const uint32_t *c, uint64_t n, double amm, uint32_t m,
                            const double *lc, double *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
#define X(RR, LM, GG, SS, NS) \
  if (r == RR && log2m == LM) return launch_count_t<RR, GG, SS, NS>(c, n, amm, m, lc, out, s);
  HB_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_seed_count(const SeedArgs &a, uint32_t m, const double *lc, double *out,
                             hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.weights) return hipErrorInvalidValue;
#define X(RR, LM, GG, SS, NS)                                      \
  if (int(a.r) == RR && int(a.log2m) == LM)                          \
    return launch_seed_count_t<RR, GG, SS, NS>(a, m, lc, out, s);
  HB_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_seed(const SeedArgs &a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  seed_kernel<<<grid_for(a.n, kBlock, 65536), kBlock, 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_prepare_iteration(uint32_t *next_mod, uint32_t *next_must, uint64_t words,
                                    Stats *stats, uint32_t *work, uint64_t *slots, hipStream_t s) {
  prepare_iteration_kernel<<<grid_for(words, kBlock, 16384), kBlock, 0, s>>>(next_mod, next_must,
                                                                            words, stats, work,
                                                                            slots);
  return hipGetLastError();
}

hipError_t launch_reduce_stats(const Stats *stats, const uint64_t *slots, Stats *out,
                               hipStream_t s) {
  reduce_stats_kernel<<<1, kStatSlots, 0, s>>>(stats, slots, out);
  return hipGetLastError();
}

uint64_t compact_scratch_words(uint64_t n) {
  const uint64_t nw = (n + 31) / 32;
  return (nw + kBlock - 1) / kBlock + 1;
}

hipError_t launch_select_heavy(const uint32_t *heavy_list, uint64_t heavy_count,
                               const uint32_t *must, const uint32_t *curr_mod, uint32_t *out,
                               uint32_t *out_len, hipStream_t s) {
  if (!heavy_count) return hipSuccess;
  select_heavy_kernel<<<grid_for(heavy_count, kBlock, 1024), kBlock, 0, s>>>(
      heavy_list, heavy_count, must, curr_mod, out, out_len);
  return hipGetLastError();
}

hipError_t launch_compact(const uint32_t *a, const uint32_t *b, uint64_t n, uint32_t *bsum,
                          uint32_t *list, uint32_t *len, hipStream_t s) {
  const uint64_t nw = (n + 31) / 32;
  const uint32_t nb = uint32_t((nw + kBlock - 1) / kBlock);
  if (nb == 0) {
    fill_u32_kernel<<<1, 64, 0, s>>>(len, 1, 0);
    return hipGetLastError();
  }
  compact_count_kernel<<<nb, kBlock, 0, s>>>(a, b, n, bsum);
  compact_scan_kernel<<<1, 1024, 0, s>>>(bsum, nb, len);
  compact_emit_kernel<<<nb, kBlock, 0, s>>>(a, b, n, bsum, list);
  return hipGetLastError();
}

hipError_t launch_mark_predecessors(const uint32_t *list, const uint32_t *len, const uint64_t *toff,
                                   const uint32_t *tsucc, uint32_t *must, uint32_t grid,
                                   hipStream_t s) {
  mark_predecessors_kernel<<<grid, kBlock, 0, s>>>(list, len, toff, tsucc, must);
  return hipGetLastError();
}

hipError_t launch_stream_merge(int r, int log2m, const UnionArgs &a, const StreamPass &p,
                               hipStream_t s) {
  const uint64_t nt = p.tile_hi - p.tile_lo;
  if (nt == 0) return hipSuccess;
#define X(RR, LM, GG, SS, NS)                                                                \
  if (r == RR && log2m == LM) {                                                              \
    stream_merge_kernel<RR, GG, SS, NS><<<grid_for(nt, kBlock / GG, 16384), kBlock, 0, s>>>(a, \
                                                                                       p);   \
    return hipGetLastError();                                                                \
  }
  HB_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_stream_finish(int r, int log2m, const UnionArgs &a, const StreamPass &p,
                                hipStream_t s) {
  const uint64_t nt = p.tile_hi - p.tile_lo;
  if (nt == 0) return hipSuccess;
#define X(RR, LM, GG, SS, NS)                                                                 \
  if (r == RR && log2m == LM) {                                                               \
    stream_finish_kernel<RR, GG, SS, NS><<<grid_for(nt, kBlock / GG, 16384), kBlock, 0, s>>>(a, \
                                                                                        p);   \
    return hipGetLastError();                                                                 \
  }
  HB_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_stream_summary(const UnionArgs &a, uint64_t n, uint64_t tile_nodes,
                                 uint64_t block_nodes, uint64_t tiles, uint32_t *counts,
                                 hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (tile_nodes % 32 || block_nodes % 32) return hipErrorInvalidValue;
  stream_summary_kernel<<<grid_for((n + 31) / 32, kBlock, 4096), kBlock, 0, s>>>(
      a, n, tile_nodes, block_nodes, tiles, counts);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t s) {
  if (words == 0) return hipSuccess;
  fill_u32_kernel<<<grid_for(words, kBlock, 16384), kBlock, 0, s>>>(p, words, value);
  return hipGetLastError();
}

hipError_t launch_fill_bits(uint32_t *p, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  fill_bits_kernel<<<grid_for((n + 31) / 32, kBlock, 16384), kBlock, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_to_reference_layout(const uint32_t *c, uint64_t node_begin, uint64_t node_count,
                                      uint32_t r, uint32_t m, uint64_t *out, hipStream_t s) {
  const uint64_t total = node_count * (uint64_t(m) * r / 64);
  if (total == 0) return hipSuccess;
  to_reference_kernel<<<grid_for(total, kBlock, 65536), kBlock, 0, s>>>(c, node_begin, node_count,
                                                                        r, m, out);
  return hipGetLastError();
}

}  // namespace hb
