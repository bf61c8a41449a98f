/*
 * CPU oracle for the HyperBall hot path.  TEST INFRASTRUCTURE ONLY: the
 * product library never links or calls this file; the tests, the smoke check
 * and the benchmark's CPU-baseline leg load it as the checker.
 *
 * Citations are to the reference crate (vigna/webgraph-algo-rs), paths
 * relative to its root.  What pins each part (tests/test_oracle.py holds the
 * checks):
 *   or_register_size / or_rel_std / or_log2_num_registers / or_hll_init
 *       closed forms at the cited lines (r = 5 for every n < 2^64, alpha
 *       table, sentinel, divisibility error); the cnr-2000 golden vectors.
 *   or_jenkins      end to end by the five cnr-2000 golden vectors (Jenkins,
 *                   seed 42, log2m 8); a wrong step moves them far past 1e-6.
 *   or_siphash      the SipHash-2-4 test vector of the SipHash paper
 *                   (Aumasson & Bernstein 2012, appendix A), and CPython
 *                   3.10's own SipHash-2-4 of bytes with a zero secret
 *                   (PYTHONHASHSEED=0) on 8-byte little-endian messages: the
 *                   zero-key, one-word path Rust's DefaultHasher takes for a
 *                   usize, through the same round-parameterised code.  The
 *                   1-3 round counts themselves are PARITY UNPINNED: no
 *                   reference test and no library here holds a 1-3 value.
 *   or_get/set_register, or_add
 *                   round trips and brute force on tiny counters.
 *   or_merge_broadword
 *                   equals a per-register maximum (or_merge_max) on random
 *                   counters for r in {4,5,6}.
 *   or_count        closed forms (empty counter, singleton = m ln(m/(m-1)),
 *                   the golden minimum of the nieminen array).
 *   or_jacobi_*     is itself the reference's test oracle (SeqHyperBall).
 *   or_hb_*         registers and modified bits equal the Jacobi oracle after
 *                   every iteration (the reference's own white-box test);
 *                   outputs equal the cnr-2000 golden vectors at 1e-6.  Its
 *                   neighbourhood function is checked against the exact sum
 *                   of counts within 1e-12; the reference's own value is not
 *                   reproducible bit for bit (thread-order Kahan sums), and
 *                   kahan 0.1.4's exact variant is PARITY UNPINNED.
 *   or_weight_item  stands in for rand::random (hyperball_impl.rs:1172),
 *                   which is unseeded in the reference: PARITY UNPINNED by
 *                   construction.
 */
#include "hb_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * HyperLogLogBuilder statics and build
 * ------------------------------------------------------------------------- */

/* Rust's `f64 as usize`: NaN and negatives give 0, large values saturate. */
static uint64_t rust_f64_to_usize(double x) {
  if (isnan(x) || x <= 0.0) return 0;
  if (x >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)x;
}

/* hyper_log_log_logic.rs:301-303:
 * max(5, ceil(ln((ln n / ln 2) / ln 2))) -- the parenthesisation makes the
 * result 5 for every n < 2^64. */
uint32_t or_register_size(uint64_t n) {
  const double ln2 = 0.6931471805599453; /* std::f64::consts::LN_2 */
  double x = (log((double)n) / ln2) / ln2;
  uint64_t v = rust_f64_to_usize(ceil(log(x)));
  return v > 5 ? (uint32_t)v : 5u;
}

/* hyper_log_log_logic.rs:285-294 */
double or_rel_std(uint32_t log2m) {
  double t;
  switch (log2m) {
    case 4: t = 1.106; break;
    case 5: t = 1.070; break;
    case 6: t = 1.054; break;
    case 7: t = 1.046; break;
    default: t = 1.04; break;
  }
  return t / sqrt((double)(1ull << log2m));
}

/* hyper_log_log_logic.rs:275-277 */
uint32_t or_log2_num_registers(double rsd) {
  return (uint32_t)rust_f64_to_usize(ceil(log2(pow(1.106 / rsd, 2.0))));
}

/* hyper_log_log_logic.rs:374-428 */
int or_hll_init(or_hll *h, uint32_t log2m, uint32_t reg_bits, uint64_t num_elements, int hash,
                uint64_t seed) {
  memset(h, 0, sizeof(*h));
  if (log2m < 4) return -1;
  if (log2m > 20) return -3;
  uint32_t r = reg_bits ? reg_bits : or_register_size(num_elements);
  if (r < 4 || r > 6) return -3;
  uint64_t m = 1ull << log2m;
  uint64_t bits = m * r;
  if (bits % 64 != 0) return -2;
  if (bits / 64 > OR_MAX_WORDS) return -3;
  h->log2m = log2m;
  h->m = (uint32_t)m;
  h->r = r;
  h->W = (uint32_t)(bits / 64);
  h->sentinel = 1ull << ((1u << r) - 2);
  double alpha;
  switch (log2m) {
    case 4: alpha = 0.673; break;
    case 5: alpha = 0.697; break;
    case 6: alpha = 0.709; break;
    default: alpha = 0.7213 / (1.0 + 1.079 / (double)m); break;
  }
  h->alpha_m_m = alpha * ((double)m * (double)m);
  h->hash = hash;
  h->seed = seed;
  for (uint64_t i = 0; i < m; i++) {
    uint64_t lo = i * r, hi = i * r + r - 1;
    h->lsb[lo / 64] |= 1ull << (lo % 64);
    h->msb[hi / 64] |= 1ull << (hi % 64);
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Hashes
 * ------------------------------------------------------------------------- */

/* The 64-bit Jenkins mix of the reference's test hasher
 * (tests/test_hyperball.rs:33-41 for the state, :52-139 for the steps,
 * :143-145 for finish = c).  A usize is hashed as one 8-byte word, so one
 * digest step runs. */
uint64_t or_jenkins(uint64_t x, uint64_t seed) {
  uint64_t a = seed + x, b = seed, c = 0x9e3779b97f4a7c13ull;
  a -= b; a -= c; a ^= c >> 43;
  b -= c; b -= a; b ^= a << 9;
  c -= a; c -= b; c ^= b >> 8;
  a -= b; a -= c; a ^= c >> 38;
  b -= c; b -= a; b ^= a << 23;
  c -= a; c -= b; c ^= b >> 5;
  a -= b; a -= c; a ^= c >> 35;
  b -= c; b -= a; b ^= a << 49;
  c -= a; c -= b; c ^= b >> 11;
  a -= b; a -= c; a ^= c >> 12;
  b -= c; b -= a; b ^= a << 18;
  c -= a; c -= b; c ^= b >> 22;
  return c;
}

static inline uint64_t rotl64(uint64_t x, int b) { return (x << b) | (x >> (64 - b)); }

#define SIPROUND                                                   \
  do {                                                             \
    v0 += v1; v1 = rotl64(v1, 13); v1 ^= v0; v0 = rotl64(v0, 32); \
    v2 += v3; v3 = rotl64(v3, 16); v3 ^= v2;                       \
    v0 += v3; v3 = rotl64(v3, 21); v3 ^= v0;                       \
    v2 += v1; v1 = rotl64(v1, 17); v1 ^= v2; v2 = rotl64(v2, 32); \
  } while (0)

/* SipHash-c-d as published (Aumasson & Bernstein, "SipHash: a fast
 * short-input PRF", 2012).  Rust's std DefaultHasher is SipHash-1-3 with both
 * keys zero (hyperball_impl.rs:16, :69-71; hyper_log_log_logic.rs:241-251). */
uint64_t or_siphash(const uint8_t *msg, size_t len, uint64_t k0, uint64_t k1, int c_rounds,
                    int d_rounds) {
  uint64_t v0 = k0 ^ 0x736f6d6570736575ull, v1 = k1 ^ 0x646f72616e646f6dull;
  uint64_t v2 = k0 ^ 0x6c7967656e657261ull, v3 = k1 ^ 0x7465646279746573ull;
  size_t full = len / 8;
  for (size_t i = 0; i < full; i++) {
    uint64_t m = 0;
    for (int k = 0; k < 8; k++) m |= (uint64_t)msg[8 * i + k] << (8 * k);
    v3 ^= m;
    for (int r = 0; r < c_rounds; r++) SIPROUND;
    v0 ^= m;
  }
  uint64_t b = (uint64_t)(len & 0xff) << 56;
  for (size_t k = 0; k < len % 8; k++) b |= (uint64_t)msg[8 * full + k] << (8 * k);
  v3 ^= b;
  for (int r = 0; r < c_rounds; r++) SIPROUND;
  v0 ^= b;
  v2 ^= 0xff;
  for (int r = 0; r < d_rounds; r++) SIPROUND;
  return v0 ^ v1 ^ v2 ^ v3;
}

/* `usize::hash` writes the 8 native-endian (little-endian) bytes. */
uint64_t or_siphash13_u64(uint64_t x) {
  uint8_t bytes[8];
  for (int k = 0; k < 8; k++) bytes[k] = (uint8_t)(x >> (8 * k));
  return or_siphash(bytes, 8, 0, 0, 1, 3);
}

uint64_t or_hash_item(const or_hll *h, uint64_t item) {
  return h->hash == OR_HASH_JENKINS ? or_jenkins(item, h->seed) : or_siphash13_u64(item);
}

/* SplitMix64 (Steele, Lea & Flood 2014) over a global item index. */
uint64_t or_weight_item(uint64_t seed, uint64_t global_index) {
  uint64_t z = seed + (global_index + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* ---------------------------------------------------------------------------
 * Counter logic
 * ------------------------------------------------------------------------- */

/* hyper_log_log_logic.rs:72-87: register i sits at bit i*r, LSB first, and
 * may straddle two words. */
uint64_t or_get_register(const or_hll *h, const uint64_t *c, uint32_t i) {
  uint64_t mask = (1ull << h->r) - 1;
  uint64_t pos = (uint64_t)i * h->r;
  uint64_t w = pos / 64, b = pos % 64;
  if (b + h->r <= 64) return (c[w] >> b) & mask;
  return ((c[w] >> b) | (c[w + 1] << (64 - b))) & mask;
}

/* hyper_log_log_logic.rs:91-115 */
void or_set_register(const or_hll *h, uint64_t *c, uint32_t i, uint64_t v) {
  uint64_t mask = (1ull << h->r) - 1;
  uint64_t pos = (uint64_t)i * h->r;
  uint64_t w = pos / 64, b = pos % 64;
  if (b + h->r <= 64) {
    c[w] = (c[w] & ~(mask << b)) | (v << b);
  } else {
    c[w] = (c[w] & ((1ull << b) - 1)) | (v << b);
    c[w + 1] = (c[w + 1] & ~(mask >> (64 - b))) | (v >> (64 - b));
  }
}

/* hyper_log_log_logic.rs:151-167 */
void or_add(const or_hll *h, uint64_t *c, uint64_t item) {
  uint64_t x = or_hash_item(h, item);
  uint32_t j = (uint32_t)(x & (h->m - 1));
  uint64_t rho = (uint64_t)__builtin_ctzll((x >> h->log2m) | h->sentinel);
  uint64_t cur = or_get_register(h, c, j);
  if (rho + 1 > cur) or_set_register(h, c, j, rho + 1);
}

/* hyper_log_log_logic.rs:169-186.
 * For r <= 5 this keeps the reference's arithmetic: `1.0 / (1 << value) as
 * f64` (:178) shifts an i32 literal, so a register of value 31 contributes
 * -2^-31.  Every partial sum is a multiple of 2^-31 below 2^13, so the f64
 * sum is exact and independent of order.
 * For r = 6 (beyond the reference, whose register width is always 5: its
 * register_size() is 5 for every usize, weight sums included) the
 * terms are the exact 2^-v, summed as the integer S = sum 2^(63-v) in 128
 * bits and converted as (double)hi * 2^64 + (double)lo, then scaled by
 * 2^-63; the GPU kernels use the same definition. */
double or_count(const or_hll *h, const uint64_t *c) {
  double hm = 0.0;
  uint32_t zeros = 0;
  if (h->r <= 5) {
    for (uint32_t i = 0; i < h->m; i++) {
      uint64_t v = or_get_register(h, c, i);
      if (v == 0) zeros++;
      hm += 1.0 / (double)(int32_t)(1u << v);
    }
  } else {
    unsigned __int128 s = 0;
    for (uint32_t i = 0; i < h->m; i++) {
      uint64_t v = or_get_register(h, c, i);
      if (v == 0) zeros++;
      s += (unsigned __int128)1 << (63 - v);
    }
    double hi = (double)(uint64_t)(s >> 64), lo = (double)(uint64_t)s;
    hm = (hi * 18446744073709551616.0 + lo) * 0x1p-63;
  }
  double est = h->alpha_m_m / hm;
  if (zeros != 0 && est < 2.5 * (double)h->m)
    est = (double)h->m * log((double)h->m / (double)zeros);
  return est;
}

/* hyper_log_log_logic.rs:451-467: multiword x -= y, least significant word
 * first. */
static void subtract_words(uint64_t *x, const uint64_t *y, uint32_t n) {
  int borrow = 0;
  for (uint32_t i = 0; i < n; i++) {
    uint64_t a = x[i], b = y[i];
    uint64_t d = a - b - (uint64_t)borrow;
    borrow = borrow ? (a <= b) : (a < b);
    x[i] = d;
  }
}

/* hyper_log_log_logic.rs:469-576: broadword per-register maximum.
 *   z = ((((y | H) - (x & ~H)) | (y ^ x)) ^ (y | ~x)) & H   (x < y per register)
 *   mask = (((z >> (r-1) | H) - L) | H) ^ z                  (all-ones where x < y)
 *   x ^= (x ^ y) & mask */
void or_merge_broadword(const or_hll *h, uint64_t *x, const uint64_t *y) {
  uint32_t W = h->W, r1 = h->r - 1;
  uint64_t acc[OR_MAX_WORDS], mask[OR_MAX_WORDS];
  for (uint32_t i = 0; i < W; i++) {
    acc[i] = y[i] | h->msb[i];
    mask[i] = x[i] & ~h->msb[i];
  }
  subtract_words(acc, mask, W);
  for (uint32_t i = 0; i < W; i++)
    acc[i] = ((acc[i] | (y[i] ^ x[i])) ^ (y[i] | ~x[i])) & h->msb[i];
  for (uint32_t i = 0; i + 1 < W; i++)
    mask[i] = (acc[i] >> r1) | (acc[i + 1] << (64 - r1)) | h->msb[i];
  mask[W - 1] = (acc[W - 1] >> r1) | h->msb[W - 1];
  subtract_words(mask, h->lsb, W);
  for (uint32_t i = 0; i < W; i++) mask[i] = (mask[i] | h->msb[i]) ^ acc[i];
  for (uint32_t i = 0; i < W; i++) x[i] = x[i] ^ ((x[i] ^ y[i]) & mask[i]);
}

/* The definition the broadword merge implements: register-wise maximum. */
void or_merge_max(const or_hll *h, uint64_t *x, const uint64_t *y) {
  for (uint32_t i = 0; i < h->m; i++) {
    uint64_t a = or_get_register(h, x, i), b = or_get_register(h, y, i);
    if (b > a) or_set_register(h, x, i, b);
  }
}

/* ---------------------------------------------------------------------------
 * Sequential Jacobi HyperBall: hyperball_impl.rs:1230-1261 (SeqHyperBall).
 * next(v) = curr(v) merged with curr(s) for every successor s, self loops
 * included (they change nothing).  The merge is the reference's broadword
 * merge, which the tests pin to or_merge_max.
 * ------------------------------------------------------------------------- */

static void seed_counters(const or_hll *h, uint64_t n, uint64_t *curr, const uint64_t *weights,
                          uint64_t weight_seed) {
  memset(curr, 0, n * h->W * sizeof(uint64_t));
  if (weights) {
    /* hyperball_impl.rs:1167-1174 */
    uint64_t g = 0;
    for (uint64_t i = 0; i < n; i++)
      for (uint64_t k = 0; k < weights[i]; k++, g++)
        or_add(h, curr + i * h->W, or_weight_item(weight_seed, g));
  } else {
    /* hyperball_impl.rs:1176-1178 */
    for (uint64_t i = 0; i < n; i++) or_add(h, curr + i * h->W, i);
  }
}

void or_jacobi_init(const or_hll *h, uint64_t n, uint64_t *curr, const uint64_t *weights,
                    uint64_t weight_seed) {
  seed_counters(h, n, curr, weights, weight_seed);
}

uint64_t or_jacobi_iterate(const or_hll *h, uint64_t n, const uint64_t *offsets,
                           const uint32_t *succ, const uint64_t *curr, uint64_t *next,
                           uint64_t *modified_bits) {
  uint64_t W = h->W, modified = 0;
  memset(modified_bits, 0, ((n + 63) / 64) * sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : modified)
  for (uint64_t v = 0; v < n; v++) {
    uint64_t *dst = next + v * W;
    memcpy(dst, curr + v * W, W * sizeof(uint64_t));
    for (uint64_t k = offsets[v]; k < offsets[v + 1]; k++)
      or_merge_broadword(h, dst, curr + (uint64_t)succ[k] * W);
    if (memcmp(dst, curr + v * W, W * sizeof(uint64_t)) != 0) {
#pragma omp atomic
      modified_bits[v / 64] |= 1ull << (v % 64);
      modified++;
    }
  }
  return modified;
}

/* ---------------------------------------------------------------------------
 * Reference-engine restatement: hyperball_impl.rs:534-1213.
 * ------------------------------------------------------------------------- */

#define OR_MAX_DISCOUNT 16
#define DEFAULT_GRANULARITY (16 * 1024) /* hyperball_impl.rs:193 */

struct or_hb {
  or_hll h;
  uint64_t n, a;
  const uint64_t *offsets, *t_offsets;
  const uint32_t *succ, *t_succ;
  int threads;
  uint64_t *curr, *next;
  uint64_t *curr_mod, *next_mod, *must, *next_must;
  uint64_t bitmap_words;
  double *sd, *hc, *dc[OR_MAX_DISCOUNT];
  or_discount_fn dfn[OR_MAX_DISCOUNT];
  void *dctx[OR_MAX_DISCOUNT];
  int ndisc;
  const uint64_t *weights;
  uint64_t weight_seed;
  uint64_t iteration;
  int systolic, local, pre_local;
  double last, rel_inc, current_nf;
  double *nf;
  size_t nf_len, nf_cap;
  uint64_t modified_counters, visited_arcs;
  uint64_t granularity, base_granularity;
  uint64_t *checklist, *local_next;
  size_t cl_len, cl_cap, ln_len, ln_cap;
};

static int bit_get(const uint64_t *b, uint64_t i) { return (int)((b[i / 64] >> (i % 64)) & 1); }

static void bit_set_atomic(uint64_t *b, uint64_t i) {
  __atomic_fetch_or(&b[i / 64], 1ull << (i % 64), __ATOMIC_RELAXED);
}

static void push_local(or_hb *hb, const uint64_t *items, size_t k) {
  if (hb->ln_len + k > hb->ln_cap) {
    size_t cap = hb->ln_cap ? hb->ln_cap : 1024;
    while (cap < hb->ln_len + k) cap *= 2;
    hb->local_next = (uint64_t *)realloc(hb->local_next, cap * sizeof(uint64_t));
    hb->ln_cap = cap;
  }
  memcpy(hb->local_next + hb->ln_len, items, k * sizeof(uint64_t));
  hb->ln_len += k;
}

or_hb *or_hb_create(uint64_t n, uint64_t a, const uint64_t *offsets, const uint32_t *succ,
                    const uint64_t *t_offsets, const uint32_t *t_succ, uint32_t log2m,
                    uint32_t reg_bits, int hash, uint64_t seed, int sum_of_distances,
                    int sum_of_inverse_distances, const uint64_t *weights, uint64_t weight_seed,
                    int threads) {
  or_hb *hb = (or_hb *)calloc(1, sizeof(or_hb));
  uint64_t num_elements = n;
  if (weights) {
    num_elements = 0;
    for (uint64_t i = 0; i < n; i++) num_elements += weights[i];
  }
  if (or_hll_init(&hb->h, log2m, reg_bits, num_elements, hash, seed) != 0) {
    free(hb);
    return NULL;
  }
  hb->n = n;
  hb->a = a;
  hb->offsets = offsets;
  hb->succ = succ;
  hb->t_offsets = t_offsets;
  hb->t_succ = t_succ;
  hb->threads = threads > 0 ? threads : omp_get_max_threads();
  size_t cw = (size_t)n * hb->h.W;
  hb->curr = (uint64_t *)calloc(cw ? cw : 1, sizeof(uint64_t));
  hb->next = (uint64_t *)calloc(cw ? cw : 1, sizeof(uint64_t));
  hb->bitmap_words = (n + 63) / 64;
  size_t bw = hb->bitmap_words ? hb->bitmap_words : 1;
  hb->curr_mod = (uint64_t *)calloc(bw, 8);
  hb->next_mod = (uint64_t *)calloc(bw, 8);
  hb->must = (uint64_t *)calloc(bw, 8);
  hb->next_must = (uint64_t *)calloc(bw, 8);
  if (sum_of_distances) hb->sd = (double *)calloc(n ? n : 1, sizeof(double));
  if (sum_of_inverse_distances) hb->hc = (double *)calloc(n ? n : 1, sizeof(double));
  hb->weights = weights;
  hb->weight_seed = weight_seed;
  hb->base_granularity = DEFAULT_GRANULARITY;
  return hb;
}

int or_hb_add_discount(or_hb *hb, or_discount_fn f, void *ctx) {
  if (hb->ndisc >= OR_MAX_DISCOUNT) return -1;
  hb->dfn[hb->ndisc] = f;
  hb->dctx[hb->ndisc] = ctx;
  hb->dc[hb->ndisc] = (double *)calloc(hb->n ? hb->n : 1, sizeof(double));
  hb->ndisc++;
  return 0;
}

void or_hb_destroy(or_hb *hb) {
  if (!hb) return;
  free(hb->curr); free(hb->next);
  free(hb->curr_mod); free(hb->next_mod); free(hb->must); free(hb->next_must);
  free(hb->sd); free(hb->hc);
  for (int k = 0; k < hb->ndisc; k++) free(hb->dc[k]);
  free(hb->nf); free(hb->checklist); free(hb->local_next);
  free(hb);
}

static void nf_push(or_hb *hb, double v) {
  if (hb->nf_len == hb->nf_cap) {
    hb->nf_cap = hb->nf_cap ? 2 * hb->nf_cap : 64;
    hb->nf = (double *)realloc(hb->nf, hb->nf_cap * sizeof(double));
  }
  hb->nf[hb->nf_len++] = v;
}

/* hyperball_impl.rs:1155-1213 */
void or_hb_init(or_hb *hb) {
  seed_counters(&hb->h, hb->n, hb->curr, hb->weights, hb->weight_seed);
  memset(hb->next, 0, (size_t)hb->n * hb->h.W * sizeof(uint64_t));
  hb->iteration = 0;
  hb->systolic = hb->local = hb->pre_local = 0;
  hb->granularity = hb->base_granularity;
  hb->modified_counters = hb->visited_arcs = 0;
  if (hb->sd) memset(hb->sd, 0, hb->n * sizeof(double));
  if (hb->hc) memset(hb->hc, 0, hb->n * sizeof(double));
  for (int k = 0; k < hb->ndisc; k++) memset(hb->dc[k], 0, hb->n * sizeof(double));
  hb->last = (double)hb->n;
  hb->nf_len = 0;
  nf_push(hb, hb->last);
  hb->ln_len = 0;
  hb->cl_len = 0;
  for (uint64_t w = 0; w < hb->bitmap_words; w++) hb->curr_mod[w] = ~0ull;
  if (hb->n % 64) hb->curr_mod[hb->bitmap_words - 1] = (1ull << (hb->n % 64)) - 1;
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* First node whose cumulative outdegree reaches `target` (DCF::succ). */
static uint64_t cumul_succ(const uint64_t *offsets, uint64_t n, uint64_t target) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (offsets[mid] >= target) hi = mid; else lo = mid + 1;
  }
  return lo;
}

/* hyperball_impl.rs:960-1152, run by every thread of the parallel region. */
static void parallel_task(or_hb *hb, uint64_t *node_cursor, uint64_t *arc_cursor_node,
                          uint64_t *arc_cursor_arc, double *disc_now) {
  const or_hll *h = &hb->h;
  const uint64_t W = h->W;
  const uint64_t limit = hb->local ? hb->cl_len : hb->n;
  const uint64_t arc_gran = hb->n
      ? (uint64_t)ceil((double)hb->a * (double)hb->granularity / (double)hb->n) : 1;
  const int do_centrality = hb->sd || hb->hc || hb->ndisc > 0;
  uint64_t visited = 0, modified = 0;
  double ks = 0.0, kc = 0.0; /* Kahan sum of the neighbourhood-function delta */
  uint64_t next_counter[OR_MAX_WORDS];
  uint64_t *pushes = NULL;
  size_t np = 0, npc = 0;
#define KAHAN_ADD(x)                    \
  do {                                  \
    double y_ = (x) - kc;               \
    double t_ = ks + y_;                \
    kc = (t_ - ks) - y_;                \
    ks = t_;                            \
  } while (0)
#define PUSH(x)                                                       \
  do {                                                                \
    if (np == npc) {                                                  \
      npc = npc ? 2 * npc : 256;                                      \
      pushes = (uint64_t *)realloc(pushes, npc * sizeof(uint64_t));   \
    }                                                                 \
    pushes[np++] = (x);                                               \
  } while (0)

  for (;;) {
    uint64_t start, end;
    if (hb->local) {
      start = __atomic_fetch_add(node_cursor, 1, __ATOMIC_RELAXED);
      if (start > limit) start = limit;
      end = start + 1 < limit ? start + 1 : limit;
    } else {
#pragma omp critical(or_arc_cursor)
      {
        uint64_t next_node = *arc_cursor_node, next_arc = *arc_cursor_arc;
        if (next_node >= limit) {
          start = end = limit;
        } else {
          start = next_node;
          uint64_t target = next_arc + arc_gran;
          if (target >= hb->a) {
            next_node = limit;
          } else {
            next_node = cumul_succ(hb->offsets, hb->n + 1, target);
            next_arc = hb->offsets[next_node];
          }
          end = next_node;
          *arc_cursor_node = next_node;
          *arc_cursor_arc = next_arc;
        }
      }
    }
    if (start == limit) break;

    for (uint64_t i = start; i < end; i++) {
      uint64_t node = hb->local ? hb->checklist[i] : i;
      const uint64_t *prev = hb->curr + node * W;
      if (!hb->systolic || hb->local || bit_get(hb->must, node)) {
        memcpy(next_counter, prev, W * sizeof(uint64_t));
        int mod = 0;
        for (uint64_t k = hb->offsets[node]; k < hb->offsets[node + 1]; k++) {
          uint64_t s = hb->succ[k];
          if (s != node && bit_get(hb->curr_mod, s)) {
            visited++;
            mod = 1;
            or_merge_broadword(h, next_counter, hb->curr + s * W);
          }
        }
        double post = NAN;
        int counter_modified = mod && memcmp(next_counter, prev, W * sizeof(uint64_t)) != 0;
        if (!hb->systolic || counter_modified) post = or_count(h, next_counter);
        if (!hb->systolic) KAHAN_ADD(post);
        if (counter_modified && (hb->systolic || do_centrality)) {
          double pre = or_count(h, prev);
          if (hb->systolic) {
            KAHAN_ADD(-pre);
            KAHAN_ADD(post);
          }
          if (do_centrality) {
            double delta = post - pre;
            if (delta > 0.0) {
              double d = (double)(hb->iteration + 1);
              if (hb->sd) hb->sd[node] += delta * d;
              if (hb->hc) hb->hc[node] += delta / d;
              for (int k = 0; k < hb->ndisc; k++) hb->dc[k][node] += delta * disc_now[k];
            }
          }
        }
        if (counter_modified) {
          if (hb->pre_local) PUSH(node);
          bit_set_atomic(hb->next_mod, node);
          if (hb->systolic) {
            for (uint64_t k = hb->t_offsets[node]; k < hb->t_offsets[node + 1]; k++) {
              if (hb->pre_local) PUSH(hb->t_succ[k]);
              else bit_set_atomic(hb->next_must, hb->t_succ[k]);
            }
          }
          modified++;
        }
        memcpy(hb->next + node * W, next_counter, W * sizeof(uint64_t));
      } else if (bit_get(hb->curr_mod, node)) {
        memcpy(hb->next + node * W, prev, W * sizeof(uint64_t));
      }
    }
  }
#pragma omp critical(or_merge_results)
  {
    hb->current_nf += ks;
    hb->visited_arcs += visited;
    hb->modified_counters += modified;
    if (np) push_local(hb, pushes, np);
  }
  free(pushes);
#undef KAHAN_ADD
#undef PUSH
}

/* hyperball_impl.rs:775-949 */
void or_hb_iterate(or_hb *hb) {
  const uint64_t n = hb->n, a = hb->a;
  const uint64_t modified = hb->modified_counters;
  const int prev_systolic = hb->systolic, prev_local = hb->local;
  hb->systolic = hb->t_offsets != NULL && hb->iteration > 0 && modified < n / 4;
  hb->current_nf = hb->systolic ? hb->last : 0.0;
  hb->local = hb->pre_local;
  hb->pre_local = hb->systolic && a > 0 && modified < (n * n) / (a * 10);

  if (prev_local) {
    for (size_t i = 0; i < hb->cl_len; i++)
      hb->next_mod[hb->checklist[i] / 64] &= ~(1ull << (hb->checklist[i] % 64));
  } else {
    memset(hb->next_mod, 0, hb->bitmap_words * 8);
  }
  if (hb->local) {
    /* sort + dedup the pushed list into the checklist (:830-846) */
    qsort(hb->local_next, hb->ln_len, sizeof(uint64_t), cmp_u64);
    size_t k = 0;
    for (size_t i = 0; i < hb->ln_len; i++)
      if (k == 0 || hb->local_next[i] != hb->local_next[k - 1])
        hb->local_next[k++] = hb->local_next[i];
    uint64_t *tmp = hb->checklist;
    size_t tcap = hb->cl_cap;
    hb->checklist = hb->local_next;
    hb->cl_len = k;
    hb->cl_cap = hb->ln_cap;
    hb->local_next = tmp;
    hb->ln_cap = tcap;
    hb->ln_len = 0;
  } else if (hb->systolic) {
    memset(hb->next_must, 0, hb->bitmap_words * 8);
    if (!prev_systolic) memset(hb->must, 0xff, hb->bitmap_words * 8);
  }

  /* adaptive granularity (:863-878) */
  uint64_t g = hb->granularity;
  if (hb->threads > 1 && !hb->local && hb->iteration > 0) {
    double cap = (double)(n / (uint64_t)hb->threads > 1 ? n / (uint64_t)hb->threads : 1);
    double grown = (double)g * ((double)n / (double)(modified > 1 ? modified : 1));
    g = (uint64_t)(cap < grown ? cap : grown);
  }
  hb->granularity = g ? g : 1;
  hb->modified_counters = 0;
  hb->visited_arcs = 0;

  double disc_now[OR_MAX_DISCOUNT];
  for (int k = 0; k < hb->ndisc; k++) disc_now[k] = hb->dfn[k](hb->iteration + 1, hb->dctx[k]);

  uint64_t node_cursor = 0, arc_node = 0, arc_arc = 0;
#pragma omp parallel num_threads(hb->threads)
  parallel_task(hb, &node_cursor, &arc_node, &arc_arc, disc_now);

  uint64_t *t;
  t = hb->curr; hb->curr = hb->next; hb->next = t;
  t = hb->curr_mod; hb->curr_mod = hb->next_mod; hb->next_mod = t;
  if (hb->systolic) {
    t = hb->must; hb->must = hb->next_must; hb->next_must = t;
  }
  hb->last = hb->current_nf;
  double last_output = hb->nf[hb->nf_len - 1];
  double cur = hb->current_nf < last_output ? last_output : hb->current_nf;
  hb->rel_inc = cur / last_output;
  nf_push(hb, cur);
  hb->iteration++;
}

/* hyperball_impl.rs:534-583; returns the number of iterations run */
int or_hb_run(or_hb *hb, uint64_t upper_bound, double threshold) {
  if (upper_bound > hb->n) upper_bound = hb->n;
  or_hb_init(hb);
  uint64_t i;
  for (i = 0; i < upper_bound; i++) {
    or_hb_iterate(hb);
    if (hb->modified_counters == 0) { i++; break; }
    if (!isnan(threshold) && i > 3 && hb->rel_inc < 1.0 + threshold) { i++; break; }
  }
  return (int)i;
}

uint64_t or_hb_iteration(const or_hb *hb) { return hb->iteration; }
uint64_t or_hb_modified(const or_hb *hb) { return hb->modified_counters; }
uint64_t or_hb_visited_arcs(const or_hb *hb) { return hb->visited_arcs; }
double or_hb_relative_increment(const or_hb *hb) { return hb->rel_inc; }
int or_hb_mode(const or_hb *hb) { return hb->systolic | (hb->local << 1) | (hb->pre_local << 2); }

size_t or_hb_nf(const or_hb *hb, double *out, size_t cap) {
  size_t k = hb->nf_len < cap ? hb->nf_len : cap;
  if (out) memcpy(out, hb->nf, k * sizeof(double));
  return hb->nf_len;
}

const uint64_t *or_hb_counters(const or_hb *hb, int which) { return which ? hb->next : hb->curr; }
const uint64_t *or_hb_modified_bits(const or_hb *hb) { return hb->curr_mod; }
const double *or_hb_sum_of_distances(const or_hb *hb) { return hb->sd; }
const double *or_hb_harmonic(const or_hb *hb) { return hb->hc; }
const double *or_hb_discounted(const or_hb *hb, size_t k) {
  return k < (size_t)hb->ndisc ? hb->dc[k] : NULL;
}

/* hyperball_impl.rs:674-757 */
int or_hb_derived(const or_hb *hb, int which, double *out) {
  if (which != 3 && !hb->sd) return -1;
  for (uint64_t v = 0; v < hb->n; v++) {
    double c = (which == 0) ? 0.0 : or_count(&hb->h, hb->curr + v * hb->h.W);
    double d = hb->sd ? hb->sd[v] : 0.0;
    switch (which) {
      case 0: out[v] = d == 0.0 ? 0.0 : 1.0 / d; break;
      case 1: out[v] = d == 0.0 ? 1.0 : c * c / d; break;
      case 2: out[v] = c * c - d; break;
      default: out[v] = c; break;
    }
  }
  return 0;
}
