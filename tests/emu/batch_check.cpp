// Emulated run of a batch against single-instance runs of the same
// configurations (test infrastructure, built by build_emu.py): every
// instance's registers, modified bits, neighbourhood function and
// accumulators must be identical.  Usage:
//   batch_check K GRAPH(0/1) SEED NMIN NMAX DMAX [HEAVY] [AUTOTUNE(0/1)]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "hyperball.h"

static void check(hb_status s, const char *what) {
  if (s != HB_OK) {
    std::fprintf(stderr, "%s failed: %s\n", what, hb_last_error());
    std::exit(2);
  }
}

struct Out {
  std::vector<uint64_t> c0, mod;
  std::vector<double> nf, sd;
};

static Out outputs(hb_instance *h, size_t words, size_t n) {
  Out o;
  o.c0.resize(words);
  check(hb_get_counters(h, 0, o.c0.data(), words), "hb_get_counters");
  o.mod.resize((n + 63) / 64);
  check(hb_get_modified(h, o.mod.data(), o.mod.size()), "hb_get_modified");
  size_t len = 0;
  check(hb_neighborhood_function(h, nullptr, 0, &len), "nf length");
  o.nf.resize(len);
  check(hb_neighborhood_function(h, o.nf.data(), len, &len), "hb_neighborhood_function");
  o.sd.resize(n);
  check(hb_sum_of_distances(h, o.sd.data(), n), "hb_sum_of_distances");
  return o;
}

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s K GRAPH SEED NMIN NMAX DMAX [HEAVY]\n", argv[0]);
    return 2;
  }
  const int k = std::atoi(argv[1]);
  const int graph = std::atoi(argv[2]);
  const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
  const double nmin = std::atof(argv[4]), nmax = std::atof(argv[5]);
  const uint32_t dmax = uint32_t(std::atoi(argv[6]));
  const uint32_t heavy = argc > 7 ? uint32_t(std::atoi(argv[7])) : 0;
  const int tune = argc > 8 ? std::atoi(argv[8]) : 0;

  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(std::log(nmin), std::log(nmax));
  std::vector<hb_graph> gs(k), ts(k);
  std::vector<hb_csr> g(k), t(k);
  std::vector<hb_config> cfg(k);
  for (int i = 0; i < k; i++) {
    const uint64_t n = uint64_t(std::exp(u(rng)));
    check(hb_graph_generate_web(n, 2.1, 1, i % 3 == 0 ? dmax : 40, 100, 0.7, seed + i, &gs[i]),
          "generate");
    g[i] = hb_csr{gs[i].num_nodes, gs[i].num_arcs, gs[i].offsets, gs[i].successors};
    t[i] = hb_csr{0, 0, nullptr, nullptr};
    if (i % 4 != 3) {
      check(hb_graph_transpose(&g[i], &ts[i]), "transpose");
      t[i] = hb_csr{ts[i].num_nodes, ts[i].num_arcs, ts[i].offsets, ts[i].successors};
    }
    hb_config_init(&cfg[i]);
    cfg[i].log2m = (i % 3 == 0) ? 6 : (i % 3 == 1) ? 8 : 10;
    cfg[i].reg_bits = (i % 2 == 0) ? 5 : 6;
    cfg[i].hash = HB_HASH_JENKINS;
    cfg[i].seed = uint64_t(i);
    cfg[i].sum_of_distances = 1;
    cfg[i].sum_of_inverse_distances = i % 2 == 0;
    std::printf("instance %d: n=%llu arcs=%llu log2m=%u r=%u transpose=%d\n", i,
                (unsigned long long)g[i].num_nodes, (unsigned long long)g[i].num_arcs,
                cfg[i].log2m, cfg[i].reg_bits, t[i].offsets != nullptr);
  }
  std::fflush(stdout);

  hb_batch *b = nullptr;
  check(hb_batch_create(size_t(k), g.data(), t.data(), cfg.data(), &b), "hb_batch_create");
  check(hb_batch_set_graph(b, graph), "hb_batch_set_graph");
  if (heavy)
    for (int i = 0; i < k; i++) {
      hb_instance *h = nullptr;
      check(hb_batch_instance(b, size_t(i), &h), "hb_batch_instance");
      check(hb_set_tuning(h, heavy), "hb_set_tuning");
    }
  if (tune) {
    uint32_t chosen = 0;
    size_t len = 0;
    check(hb_batch_autotune(b, 1, nullptr, nullptr, 0, &len, &chosen), "hb_batch_autotune");
    std::printf("autotune: %zu candidates, chose %u\n", len, chosen);
  }
  check(hb_batch_run_until_done(b), "hb_batch_run_until_done");
  std::printf("batch run done\n");
  std::fflush(stdout);

  int bad = 0;
  for (int i = 0; i < k; i++) {
    hb_instance *hb = nullptr;
    check(hb_batch_instance(b, size_t(i), &hb), "hb_batch_instance");
    hb_params p;
    check(hb_get_params(hb, &p), "hb_get_params");
    const size_t words = size_t(p.num_nodes) * p.words_per_counter;
    const Out ob = outputs(hb, words, p.num_nodes);
    hb_instance *s = nullptr;
    check(hb_create(&g[i], t[i].offsets ? &t[i] : nullptr, &cfg[i], &s), "hb_create");
    if (heavy) check(hb_set_tuning(s, heavy), "hb_set_tuning");
    check(hb_run_until_done(s), "hb_run_until_done");
    const Out os = outputs(s, words, p.num_nodes);
    check(hb_destroy(s), "hb_destroy");
    const bool same = ob.c0 == os.c0 && ob.mod == os.mod && ob.nf == os.nf && ob.sd == os.sd;
    uint64_t iters = 0, windows = 0;
    check(hb_num_iterations(hb, &iters), "hb_num_iterations");
    for (uint64_t it = 0; it < iters; it++) {
      hb_iteration_stats st;
      check(hb_get_iteration_stats(hb, it, &st), "hb_get_iteration_stats");
      windows += st.window;
    }
    std::printf("instance %d: %s (%zu iterations, %llu window)\n", i,
                same ? "identical" : "DIFFERENT", ob.nf.size() - 1, (unsigned long long)windows);
    bad += !same;
  }
  check(hb_batch_destroy(b), "hb_batch_destroy");
  for (int i = 0; i < k; i++) {
    hb_graph_free(&gs[i]);
    if (t[i].offsets) hb_graph_free(&ts[i]);
  }
  std::printf("%s\n", bad ? "FAIL" : "OK");
  return bad ? 1 : 0;
}
