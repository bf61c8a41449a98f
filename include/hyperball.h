/*
 * hyperball.h -- C ABI of the MI355X (gfx950) HyperBall engine.
 *
 * Drop-in boundary for the HyperBall path of the reference crate
 * vigna/webgraph-algo-rs (webgraph-algo 0.1.0).  Every entry point names the
 * reference item it replaces (paths relative to the reference's root).  Plain
 * pointers and sizes only; no C++ exception crosses this boundary.  Every
 * function returns an hb_status; hb_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Ownership: graphs and weights passed to hb_create are borrowed only for the
 * duration of the call (they are copied to device memory).  Counters,
 * frontier bitmaps and accumulators live in device memory owned by the
 * instance and are released by hb_destroy.  Output buffers belong to the
 * caller.
 *
 * Threading: an instance is not thread safe (the reference takes &mut self);
 * distinct instances may run concurrently, each on its own HIP stream.
 */
#ifndef HYPERBALL_H
#define HYPERBALL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (anyhow::Result / panics in the reference) ------------ */
typedef int hb_status;
#define HB_OK 0
#define HB_E_NOT_RUN 1       /* ensure_iteration, hyperball_impl.rs:624-630 */
#define HB_E_NOT_REQUESTED 2 /* bail! for outputs not requested, :645 :655 :669 :684 :710 :730 */
#define HB_E_INVALID_ARG 3   /* builder ensure! failures and former panics */
#define HB_E_OOM 4           /* device or host allocation failed */
#define HB_E_HIP 5           /* HIP runtime error */
#define HB_E_NO_DEVICE 6     /* no usable gfx950 device */

/* Message of the last failing call on this thread ("" if none). */
const char *hb_last_error(void);

/* ---- graphs -------------------------------------------------------------- */

/* A graph in CSR form: successors of node v are succ[offsets[v] ..
 * offsets[v+1]), sorted increasingly.  Replaces the reference's
 * RandomAccessGraph::successors / num_nodes / num_arcs (external crate
 * webgraph 0.2.0, used at hyperball_impl.rs:1037, :1119, :1123) and the
 * cumulative outdegree function `D: Succ` (hyperball_impl.rs:1012), which is
 * `offsets` itself. */
typedef struct {
  uint64_t num_nodes;
  uint64_t num_arcs;
  const uint64_t *offsets;   /* num_nodes + 1 entries */
  const uint32_t *successors; /* num_arcs entries */
} hb_csr;

/* A graph whose arrays were allocated by this library (free with
 * hb_graph_free). */
typedef struct {
  uint64_t num_nodes;
  uint64_t num_arcs;
  uint64_t *offsets;
  uint32_t *successors;
} hb_graph;

/* Decodes a BVGraph (version 0, big-endian bit streams) from
 * <basename>.properties and <basename>.graph into CSR.  Replaces
 * BvGraph::with_basename(..).load() + successor decoding (webgraph 0.2.0;
 * tests/test_hyperball.rs:226-227).  Codes: gamma, delta, unary, zeta_k. */
hb_status hb_bvgraph_load(const char *basename, hb_graph *out);
/* The same with an explicit host thread count (0: all cores).  With more than
 * one thread and a node index, node ranges are decoded in parallel from their
 * indexed bit offsets (the random access of BvGraph::with_basename(..).load()
 * in webgraph 0.2.0); references that leave a range are decoded at their own
 * offsets.  The index is <basename>.ef (webgraph's Elias-Fano file) when it
 * exists, else <basename>.offsets; every range must end where the index
 * says.  One thread: a single sequential pass, no index needed. */
hb_status hb_bvgraph_load_threads(const char *basename, unsigned threads, hb_graph *out);

/* Writes a graph in CSR form (rows strictly increasing) as a BVGraph,
 * version 0, big-endian: <basename>.graph, .offsets and .properties, in the
 * published format that hb_bvgraph_load and the reference's
 * BvGraph::with_basename(..).load() (webgraph 0.2.0) read.  No counterpart in
 * the reference, which only loads graphs: it stores synthetic graphs and is
 * the writer of the round-trip tests.  For each node it picks the reference
 * within the window (at most max_ref_count references deep) whose record
 * takes the fewest bits. */
#define HB_CODE_GAMMA 0
#define HB_CODE_DELTA 1
#define HB_CODE_UNARY 2
#define HB_CODE_ZETA 3
typedef struct {
  uint32_t window_size;         /* 0: no references */
  uint32_t max_ref_count;
  uint32_t min_interval_length; /* 0: no intervals */
  uint32_t zeta_k;              /* parameter of the zeta codes, 1..16 */
  uint32_t outdegrees, references, blocks, block_count, residuals, offsets; /* HB_CODE_* */
} hb_bvgraph_params;
/* The defaults of the Java BVGraph compressor: window 7, max_ref_count 3,
 * intervals of at least 4, zeta_3 residuals, gamma elsewhere, unary
 * references. */
void hb_bvgraph_params_init(hb_bvgraph_params *p);
hb_status hb_bvgraph_store(const hb_csr *g, const char *basename, const hb_bvgraph_params *p);

/* Decodes <basename>.offsets (gamma-coded bit-offset gaps) into num_nodes+1
 * bit offsets; used to check a decode against the offsets the writer
 * recorded. */
hb_status hb_bvgraph_load_offsets(const char *basename, uint64_t num_nodes, uint64_t *out);

/* Bit offset at which node v's record starts, as observed by the decoder of
 * hb_bvgraph_load; num_nodes+1 entries. */
hb_status hb_bvgraph_decoder_offsets(const char *basename, uint64_t *out, uint64_t cap);

/* Decodes an Elias-Fano sequence stored by the sux crate through epserde:
 * <basename>.ef (the num_nodes+1 bit offsets of the node records, which
 * BvGraph::with_basename(..).load() reads in webgraph 0.2.0) or
 * <basename>.dcf (the cumulative out-degrees the reference loads with
 * DCF::load_mmap, hyperball_impl.rs:1270, main.rs:57).  *count receives the
 * number of elements; out (cap entries) receives them unless it is NULL.
 * Errors: HB_E_INVALID_ARG for a file that is not such a sequence, or out
 * too small. */
hb_status hb_elias_fano_load(const char *path, uint64_t *out, uint64_t cap, uint64_t *count);

/* Transpose by counting sort; rows of the result are sorted.  Replaces the
 * `-t` transposed BvGraph the reference loads for systolic iterations
 * (hyperball_impl.rs:204-256). */
hb_status hb_graph_transpose(const hb_csr *g, hb_graph *out);
/* The same transpose built on HIP device `device` (in-degree atomics, a scan
 * and a stable radix sort of (successor, source) pairs) and copied back: the
 * arrays equal hb_graph_transpose's.  hb_config.device_transpose builds it in
 * place inside hb_create instead, with no host copy. */
hb_status hb_graph_transpose_device(const hb_csr *g, int device, hb_graph *out);

/* Synthetic graphs (SplitMix64-seeded xoshiro256**, per-node streams, so the
 * result does not depend on the host thread count).
 *   uniform: every node gets exactly k distinct successors drawn uniformly
 *            from [0,n) minus itself (Floyd sampling).
 *   web:     out-degree Zipf(alpha) on [dmin, dmax]; each drawn successor is
 *            uniform in the window [v-window, v+window] minus v (wrapping
 *            mod n) with probability p_window, otherwise the target of a
 *            uniformly random earlier arc (copy model, preferential
 *            attachment on in-degree); rows are sorted, deduplicated, self
 *            loops dropped. */
hb_status hb_graph_generate_uniform(uint64_t n, uint32_t k, uint64_t seed, hb_graph *out);
hb_status hb_graph_generate_web(uint64_t n, double alpha, uint32_t dmin, uint32_t dmax,
                                uint64_t window, double p_window, uint64_t seed, hb_graph *out);
void hb_graph_free(hb_graph *g);

/* ---- HyperLogLog parameters (hyper_log_log_logic.rs:275-303) ------------- */
uint32_t hb_register_size(uint64_t n);       /* HyperLogLog::register_size */
double hb_rel_std(uint32_t log2m);            /* HyperLogLog::rel_std */
uint32_t hb_log2_num_registers(double rsd);  /* HyperLogLog::log_2_num_of_registers */

/* ---- configuration (HyperBallBuilder, hyperball_impl.rs:26-304) ---------- */
#define HB_HASH_SIPHASH13 0 /* std DefaultHasher (SipHash-1-3, zero keys): the default of
                               with_hyper_log_log, hyperball_impl.rs:69-71 */
#define HB_HASH_JENKINS 1   /* the Jenkins hasher of tests/test_hyperball.rs:26-172 */

#define HB_ACC_F64 0        /* per-node f64 accumulators, as hyperball_impl.rs:1072-1093 */
#define HB_ACC_F32_KAHAN 1  /* f32 (sum, compensation) pairs */
#define HB_ACC_F32 2        /* plain f32 sums (precision study only) */

/* Where the two counter arrays live (the role of the reference's
 * `mmap_options: TempMmapOptions` argument, hyperball_impl.rs:89-117,
 * mmap_helper.rs:282-293, which backs them by RAM or by temporary files). */
#define HB_COUNTERS_DEVICE 0 /* HBM (TempMmapOptions::Default) */
#define HB_COUNTERS_HOST 1   /* pinned host memory that the kernels read and write over the
                                host link: instances whose counters exceed HBM; every other
                                buffer stays in HBM.  No file-backed mode: the memory must be
                                allocated pinned. */
#define HB_COUNTERS_STREAM 2 /* pinned host memory, streamed through device memory: each
                                iteration walks the nodes in tiles of stream_tile_nodes nodes;
                                a tile's own counters and then every block (a quarter of a
                                tile) of the current array that holds modified nodes are
                                copied to the device, merged there, and the tile's next
                                counters copied back.  Device memory holds about 4.5 tiles of
                                counters; every other buffer stays in HBM. */

typedef struct {
  uint32_t log2m;        /* log2 of registers per counter (with_hyper_log_log `log2m`) */
  uint32_t reg_bits;     /* 0: HyperLogLog::register_size(num_elements); else 5 or 6 */
  uint32_t hash;         /* HB_HASH_* */
  uint64_t seed;         /* Jenkins seed */
  uint64_t num_elements; /* 0: n, or the sum of the weights when weights are given */
  const uint64_t *weights;   /* nullable: HyperBallBuilder::weights (:281-287) */
  uint64_t weight_seed;  /* seeds the stand-in for rand::random in weighted init */
  uint64_t arc_granularity; /* HyperBallBuilder::arc_granularity (:271-274); accepted, unused */
  int sum_of_distances;  /* HyperBallBuilder::sum_of_distances (:259-262) */
  int sum_of_inverse_distances; /* HyperBallBuilder::sum_of_inverse_distances (:265-268) */
  uint32_t acc_precision; /* HB_ACC_* */
  int device;            /* HIP device ordinal */
  int device_transpose;  /* with no transpose given: build it on the device (systolic mode) */
  uint32_t counter_memory; /* HB_COUNTERS_* */
  uint64_t stream_tile_nodes; /* HB_COUNTERS_STREAM: nodes per tile, rounded up to a multiple of
                                 128 (0: the tile whose counters take 1 GiB) */
} hb_config;

/* Fills the defaults of HyperBallBuilder::with_hyper_log_log (:91-133):
 * log2m 4 is NOT a default there, so log2m is set to 6; SipHash-1-3; no
 * outputs requested; f64 accumulators; device 0. */
void hb_config_init(hb_config *cfg);

typedef struct hb_instance hb_instance;

/* HyperBallBuilder::with_hyper_log_log(graph, transposed, cumul_outdeg, log2m,
 * weights, mmap) + .sum_of_distances(..) + .sum_of_inverse_distances(..) +
 * .build(pl)  (hyperball_impl.rs:91-133, :321-403).  `transpose` may be NULL
 * (no systolic iterations).  Errors: HB_E_INVALID_ARG for log2m < 4, counters
 * that are not whole 64-bit words, wrong weight/transposed sizes, or shapes
 * this library does not hold (log2m above 16, register width other than 5 or
 * 6). */
hb_status hb_create(const hb_csr *graph, const hb_csr *transpose, const hb_config *cfg,
                    hb_instance **out);
hb_status hb_destroy(hb_instance *hb);

/* HyperBallBuilder::discount_function (:291-297).  The reference calls the
 * closure only as f(iteration + 1) (:1089); here it is evaluated once per
 * iteration on the host. */
typedef double (*hb_discount_fn)(uint64_t distance, void *ctx);
hb_status hb_add_discount_function(hb_instance *hb, hb_discount_fn f, void *ctx);
/* HyperBallBuilder::no_discount_function (:300-303) */
hb_status hb_clear_discount_functions(hb_instance *hb);

/* ---- running (hyperball_impl.rs:534-621) --------------------------------- */
/* HyperBall::run(upper_bound, threshold): threshold NaN means None. */
hb_status hb_run(hb_instance *hb, uint64_t upper_bound, double threshold);
hb_status hb_run_until_stable(hb_instance *hb, uint64_t upper_bound);
hb_status hb_run_until_done(hb_instance *hb);

/* Step-wise driving for parity checks: HyperBall::init (:1155-1213) and one
 * HyperBall::iterate (:775-949). */
hb_status hb_init(hb_instance *hb);
hb_status hb_iterate(hb_instance *hb);

/* ---- results (hyperball_impl.rs:633-757) --------------------------------- */
hb_status hb_neighborhood_function(hb_instance *hb, double *out, size_t cap, size_t *len);
hb_status hb_sum_of_distances(hb_instance *hb, double *out, size_t n);
hb_status hb_harmonic_centralities(hb_instance *hb, double *out, size_t n);
hb_status hb_discounted_centrality(hb_instance *hb, size_t index, double *out, size_t n);
hb_status hb_closeness_centrality(hb_instance *hb, double *out, size_t n);
hb_status hb_lin_centrality(hb_instance *hb, double *out, size_t n);
hb_status hb_nieminen_centrality(hb_instance *hb, double *out, size_t n);
hb_status hb_reachable_nodes(hb_instance *hb, double *out, size_t n);
hb_status hb_reachable_nodes_from(hb_instance *hb, uint64_t node, double *out);

/* ---- inspection ---------------------------------------------------------- */
typedef struct {
  uint32_t log2m, num_registers, register_bits, words_per_counter;
  uint64_t num_nodes, num_arcs;
  double alpha_m_m;
  uint32_t counter_memory;  /* HB_COUNTERS_* */
} hb_params;
hb_status hb_get_params(hb_instance *hb, hb_params *out);

/* Counter arrays in the reference layout (slice_counter_array.rs:166-169,
 * hyper_log_log_logic.rs:72-115): counter v is words [v*W, (v+1)*W), register
 * j at bit j*r, LSB first.  which = 0: current state, 1: the other buffer
 * (all zero before the first iteration). */
hb_status hb_get_counters(hb_instance *hb, int which, uint64_t *out, size_t words);
/* hb_get_counters that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * conversion to the reference layout goes through a 16 MiB staging array in
 * device memory that the library holds per GPU (from the first call on that
 * GPU until the process ends) instead of a temporary hipMalloc; counters are
 * converted and copied in chunks of at most 16 MiB, so every instance size is
 * accepted (a single counter larger than 16 MiB would be rejected with
 * HB_E_INVALID_ARG).  All calls on one GPU share that array: calls that
 * overlap on one GPU (from several threads) race and must be serialised by
 * the caller.  Same arguments, checks and results as hb_get_counters. */
hb_status hb_get_counters_noalloc(hb_instance *hb, int which, uint64_t *out, size_t words);
/* Modified-counter bits of the last iteration (AtomicBitVec layout: bit v of
 * word v/64). */
hb_status hb_get_modified(hb_instance *hb, uint64_t *out, size_t words);

typedef struct {
  uint64_t iteration;     /* 0-based */
  uint32_t systolic, local, pre_local;
  uint64_t checked;       /* nodes whose successors were enumerated (K_t) */
  uint64_t visited_arcs;  /* merges performed (V_t), hyperball_impl.rs:1038-1039 */
  uint64_t scanned_arcs;  /* successor ids read (A_t) */
  uint64_t transpose_arcs;/* transpose ids read (T_t) */
  uint64_t copied;        /* counters copied without check (Y_t) */
  uint64_t modified;      /* modified counters (M_t) */
  double nf;              /* neighbourhood-function value pushed */
  double union_ms;        /* device time of the union kernel(s), when timing is on */
  double iteration_ms;    /* device time of the whole iteration, when timing is on */
  uint64_t gathered_slabs;/* successor slabs (32 registers each) actually read */
  uint32_t delta_gathering; /* 1 when only the slabs that changed in the previous iteration
                               were gathered (chosen per iteration) */
  uint32_t streamed;      /* HB_COUNTERS_STREAM: 1 when the iteration streamed the counters
                             through device memory, 0 when its kernels read them in place */
  uint32_t window;        /* 1 when the union step ran as the window kernel (hb_set_window) */
  uint32_t reserved;
} hb_iteration_stats;
hb_status hb_num_iterations(hb_instance *hb, uint64_t *out);
hb_status hb_get_iteration_stats(hb_instance *hb, uint64_t iteration, hb_iteration_stats *out);
/* Turns per-kernel HIP-event timing on or off (off by default). */
hb_status hb_set_timing(hb_instance *hb, int enabled);
/* Per-iteration progress (the reference's `pl: ConcurrentProgressLog`
 * argument of run, hyperball_impl.rs:534-583): f(stats, ctx) is called on the
 * calling thread after every iteration; NULL turns it off. */
typedef void (*hb_progress_fn)(const hb_iteration_stats *stats, void *ctx);
hb_status hb_set_progress(hb_instance *hb, hb_progress_fn f, void *ctx);
/* Runs the instance's work on the caller's HIP stream (a hipStream_t of the
 * instance's device), e.g. to order it with the caller's own kernels; NULL
 * gives the instance a private stream again.  Waits for the current stream
 * first.  Not for the instances of a batch. */
hb_status hb_set_stream(hb_instance *hb, void *stream);
/* Overrides the lane-group split of the union kernel (0 restores the
 * default); for the autotuner.  heavy_degree: rows with more successors than
 * this go to the workgroup-per-node kernel. */
hb_status hb_set_tuning(hb_instance *hb, uint32_t heavy_degree);
/* Delta gathering: read only the slabs (32 registers) of a modified successor
 * that changed in the previous iteration; exact, since every counter already
 * dominates its successors' previous counters.  mode 0 (default): per
 * iteration, when counters have at least 1024 registers and the expected
 * share of changed slabs is below 0.7; 1: always; 2: never. */
hb_status hb_set_delta_gathering(hb_instance *hb, int mode);
/* Window kernel, for 64-register counters (log2m 6) in device memory: in an
 * iteration that reads the counters (not the seeded first one, no delta
 * gathering), each workgroup takes blocks of consecutive nodes and first
 * copies the current counters of the nodes around its block (the tile:
 * half_width nodes on each side) into LDS, where successors inside it are
 * read; the others are read from HBM.  Exact, like every kernel choice.
 * mode 0 (default): the iterations that check every node, and worklist
 * (systolic or local) iterations expected to check at least an eighth of the
 * nodes; 1: every iteration it can run; 2: never.  block_nodes and
 * half_width cap the kernel's own block and tile (0: the kernel's largest);
 * rows longer than heavy_degree take a workgroup each (0: the larger of
 * hb_set_tuning's threshold and 8192; UINT32_MAX: none), shorter ones a
 * wavefront (above 64 successors) or a lane pair inside their block.  For
 * tests and tuning.  No counterpart in the reference (its union step is
 * hyperball_impl.rs:1025-1145 on CPU threads). */
hb_status hb_set_window(hb_instance *hb, int mode, uint32_t block_nodes, uint32_t half_width,
                        uint32_t heavy_degree);
/* Streamed counters (HB_COUNTERS_STREAM): per iteration, either stream the
 * counter arrays through device memory tile by tile, or let the union kernel
 * read and write them in place over the host link, as HB_COUNTERS_HOST does.
 * mode 0 (default): in place for the first iteration without weights (it
 * reads no counter: the seeded registers are computed from hashes) and for
 * iterations expected to visit fewer arcs than max(1, tiles - 2) * n
 * (estimated as the last iteration's modified nodes times the mean
 * out-degree), streamed otherwise; 1: always streamed; 2: always in place.
 * Results do not depend on the mode. */
hb_status hb_set_streaming(hb_instance *hb, int mode);
/* Autotuner of that split (in the role of the reference's arc_granularity
 * tuning knob, hyperball_impl.rs:271-274, which sizes CPU work units): times
 * the first iteration that gathers whole counters (iteration 1, every node
 * checked, delta gathering off) under each threshold 64 * 2^k that changes the
 * set of heavy rows, and "none" (UINT32_MAX), best of `reps`; keeps the
 * fastest.  Writes up to `cap` (threshold, ms) pairs, their count to *len and
 * the chosen threshold to *chosen (each pointer may be NULL).  Resets the run
 * state as hb_init does; results never depend on the threshold. */
hb_status hb_autotune(hb_instance *hb, uint32_t reps, uint32_t *thresholds, double *ms,
                      size_t cap, size_t *len, uint32_t *chosen);

/* ---- batches ------------------------------------------------------------
 * No counterpart in the reference, which runs one HyperBall at a time: many
 * independent instances (each its own graph, configuration and stop rule)
 * advance together, one fused launch sequence per batched iteration, replayed
 * from a HIP graph.  Each instance keeps the single-instance semantics and
 * accessors.  A batch holds log2m 6, 8 or 10 with 5- or 6-bit registers. */
typedef struct hb_batch hb_batch;
/* transposes: NULL, or k entries where offsets == NULL means "no transpose". */
hb_status hb_batch_create(size_t k, const hb_csr *graphs, const hb_csr *transposes,
                          const hb_config *cfgs, hb_batch **out);
hb_status hb_batch_destroy(hb_batch *b);
/* Borrowed handle of instance i for the accessors and hb_add_discount_function;
 * hb_destroy refuses it. */
hb_status hb_batch_instance(hb_batch *b, size_t i, hb_instance **out);
/* HyperBall::run semantics per instance; the batch ends when every instance
 * has stopped. */
hb_status hb_batch_run(hb_batch *b, uint64_t upper_bound, double threshold);
hb_status hb_batch_run_until_done(hb_batch *b);
/* hb_autotune for a batch: one threshold for every instance, timed on the
 * second batched iteration (same outputs and reset as hb_autotune). */
hb_status hb_batch_autotune(hb_batch *b, uint32_t reps, uint32_t *thresholds, double *ms,
                            size_t cap, size_t *len, uint32_t *chosen);
/* HIP-graph replay of the batched iteration on (default) or off. */
hb_status hb_batch_set_graph(hb_batch *b, int enabled);
/* Device time of every batched iteration of the last run, in milliseconds. */
hb_status hb_batch_step_times(hb_batch *b, double *out, size_t cap, size_t *len);

#ifdef __cplusplus
}
#endif
#endif
