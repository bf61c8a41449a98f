/*
 * CPU oracle for the HyperBall hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it to compute
 * the expected answer.  The product library never links, loads or calls it.
 *
 * Every function restates the behaviour of the reference crate
 * (vigna/webgraph-algo-rs, crate webgraph-algo 0.1.0); each definition cites
 * the file:line it follows.  Two families are provided:
 *   - the sequential Jacobi HyperBall of the reference's own white-box test
 *     (src/algo/hyperball/hyperball_impl.rs:1230-1261), which fixes the
 *     register contents after every iteration;
 *   - a restatement of the reference's parallel engine (iterate /
 *     parallel_task / init, hyperball_impl.rs:775-1213) with its systolic,
 *     local and pre-local modes, per-thread Kahan neighbourhood-function sums
 *     and per-node f64 accumulators, run on OpenMP threads.  It is the CPU
 *     baseline and the source of the expected float outputs.
 *
 * What pins each function is listed in the header of hb_oracle.c.
 */
#ifndef HB_ORACLE_H
#define HB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_HASH_SIPHASH13 = 0, OR_HASH_JENKINS = 1 };

/* Largest counter the oracle holds, in 64-bit words: 2^16 registers of 6
 * bits. */
#define OR_MAX_WORDS 6144

typedef struct {
  uint32_t log2m;     /* base-2 logarithm of the number of registers */
  uint32_t m;         /* registers per counter */
  uint32_t r;         /* register width in bits */
  uint32_t W;         /* 64-bit words per counter */
  uint64_t sentinel;  /* 1 << (2^r - 2) */
  double alpha_m_m;   /* alpha * m^2 */
  int hash;           /* OR_HASH_* */
  uint64_t seed;      /* Jenkins seed */
  uint64_t msb[OR_MAX_WORDS];  /* H_r: top bit of every register (W words) */
  uint64_t lsb[OR_MAX_WORDS];  /* L_r: bottom bit of every register (W words) */
} or_hll;

/* HyperLogLogBuilder statics, hyper_log_log_logic.rs:275-303 */
uint32_t or_register_size(uint64_t n);
double or_rel_std(uint32_t log2m);
uint32_t or_log2_num_registers(double rsd);

/* HyperLogLogBuilder::build, hyper_log_log_logic.rs:374-428.  reg_bits = 0
 * uses the reference's formula.  Returns 0, or -1 for log2m < 4, -2 when the
 * counter is not a whole number of 64-bit words, -3 for sizes this oracle does
 * not hold (W > OR_MAX_WORDS or r outside [4, 8]). */
int or_hll_init(or_hll *h, uint32_t log2m, uint32_t reg_bits, uint64_t num_elements,
                int hash, uint64_t seed);

/* Hashes. */
uint64_t or_jenkins(uint64_t x, uint64_t seed);
uint64_t or_siphash(const uint8_t *msg, size_t len, uint64_t k0, uint64_t k1, int c_rounds,
                    int d_rounds);
uint64_t or_siphash13_u64(uint64_t x);
uint64_t or_hash_item(const or_hll *h, uint64_t item);

/* Counter logic on one counter of W words (hyper_log_log_logic.rs:72-229). */
uint64_t or_get_register(const or_hll *h, const uint64_t *c, uint32_t i);
void or_set_register(const or_hll *h, uint64_t *c, uint32_t i, uint64_t v);
void or_add(const or_hll *h, uint64_t *c, uint64_t item);
double or_count(const or_hll *h, const uint64_t *c);
void or_merge_broadword(const or_hll *h, uint64_t *x, const uint64_t *y);
void or_merge_max(const or_hll *h, uint64_t *x, const uint64_t *y);

/* Seeded stand-in for rand::random::<u64>() in weighted initialisation. */
uint64_t or_weight_item(uint64_t seed, uint64_t global_index);

/* Sequential Jacobi HyperBall (hyperball_impl.rs:1230-1261). */
void or_jacobi_init(const or_hll *h, uint64_t n, uint64_t *curr, const uint64_t *weights,
                    uint64_t weight_seed);
uint64_t or_jacobi_iterate(const or_hll *h, uint64_t n, const uint64_t *offsets,
                           const uint32_t *succ, const uint64_t *curr, uint64_t *next,
                           uint64_t *modified_bits);

/* Reference-engine restatement. */
typedef struct or_hb or_hb;
typedef double (*or_discount_fn)(uint64_t distance, void *ctx);

or_hb *or_hb_create(uint64_t n, uint64_t a, const uint64_t *offsets, const uint32_t *succ,
                    const uint64_t *t_offsets, const uint32_t *t_succ, uint32_t log2m,
                    uint32_t reg_bits, int hash, uint64_t seed, int sum_of_distances,
                    int sum_of_inverse_distances, const uint64_t *weights, uint64_t weight_seed,
                    int threads);
int or_hb_add_discount(or_hb *hb, or_discount_fn f, void *ctx);
void or_hb_destroy(or_hb *hb);
void or_hb_init(or_hb *hb);
void or_hb_iterate(or_hb *hb);
int or_hb_run(or_hb *hb, uint64_t upper_bound, double threshold /* NaN = None */);
uint64_t or_hb_iteration(const or_hb *hb);
uint64_t or_hb_modified(const or_hb *hb);
uint64_t or_hb_visited_arcs(const or_hb *hb);
double or_hb_relative_increment(const or_hb *hb);
int or_hb_mode(const or_hb *hb); /* bit0 systolic, bit1 local, bit2 pre_local */
size_t or_hb_nf(const or_hb *hb, double *out, size_t cap);
const uint64_t *or_hb_counters(const or_hb *hb, int which); /* 0 current, 1 next */
const uint64_t *or_hb_modified_bits(const or_hb *hb);
const double *or_hb_sum_of_distances(const or_hb *hb);
const double *or_hb_harmonic(const or_hb *hb);
const double *or_hb_discounted(const or_hb *hb, size_t k);
/* hyperball_impl.rs:674-757: 0 closeness, 1 lin, 2 nieminen, 3 reachable */
int or_hb_derived(const or_hb *hb, int which, double *out);

#ifdef __cplusplus
}
#endif
#endif
