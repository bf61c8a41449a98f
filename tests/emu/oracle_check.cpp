// The emulated library against the oracle's restatement of the reference
// engine, in lockstep (test infrastructure, built by build_emu.py): after
// every iteration the current registers (reference layout), the modified
// bits and the iteration mode must be identical; at the end, the sums of
// distances.  Visited arcs and modified counts of every iteration must equal
// the oracle's too.  Usage:
//   oracle_check LOG2M REG_BITS N DMAX SEED HEAVY [DELTA_MODE [COUNTER_MEMORY TILE_NODES
//                [STREAMING_MODE]]]
// HB_CHECK_WINDOW="MODE BLOCK HALF HEAVY" in the environment sets hb_set_window.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hyperball.h"

extern "C" {
#include "hb_oracle.h"
}

static void check(hb_status s, const char *what) {
  if (s != HB_OK) {
    std::fprintf(stderr, "%s failed: %s\n", what, hb_last_error());
    std::exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s LOG2M REG_BITS N DMAX SEED HEAVY\n", argv[0]);
    return 2;
  }
  const uint32_t log2m = uint32_t(std::atoi(argv[1])), r = uint32_t(std::atoi(argv[2]));
  const uint64_t n = std::strtoull(argv[3], nullptr, 10);
  const uint32_t dmax = uint32_t(std::atoi(argv[4]));
  const uint64_t seed = std::strtoull(argv[5], nullptr, 10);
  const uint32_t heavy = uint32_t(std::atoi(argv[6]));
  const int delta_mode = argc > 7 ? std::atoi(argv[7]) : 0;
  const uint32_t counter_memory = argc > 8 ? uint32_t(std::atoi(argv[8])) : HB_COUNTERS_DEVICE;
  const uint64_t tile_nodes = argc > 9 ? std::strtoull(argv[9], nullptr, 10) : 0;
  const int streaming = argc > 10 ? std::atoi(argv[10]) : 0;

  hb_graph gg, gt;
  check(hb_graph_generate_web(n, 2.1, 1, dmax, 100, 0.7, seed, &gg), "generate");
  hb_csr g{gg.num_nodes, gg.num_arcs, gg.offsets, gg.successors};
  check(hb_graph_transpose(&g, &gt), "transpose");
  hb_csr t{gt.num_nodes, gt.num_arcs, gt.offsets, gt.successors};
  hb_config cfg;
  hb_config_init(&cfg);
  cfg.log2m = log2m;
  cfg.reg_bits = r;
  cfg.hash = HB_HASH_JENKINS;
  cfg.seed = seed;
  cfg.sum_of_distances = 1;
  cfg.counter_memory = counter_memory;
  cfg.stream_tile_nodes = tile_nodes;
  hb_instance *hb = nullptr;
  check(hb_create(&g, &t, &cfg, &hb), "hb_create");
  if (heavy) check(hb_set_tuning(hb, heavy), "hb_set_tuning");
  check(hb_set_delta_gathering(hb, delta_mode), "hb_set_delta_gathering");
  if (counter_memory == HB_COUNTERS_STREAM) check(hb_set_streaming(hb, streaming), "hb_set_streaming");
  if (const char *w = std::getenv("HB_CHECK_WINDOW")) {
    unsigned wm = 0, wb = 0, wh = 0, wy = 0;
    if (std::sscanf(w, "%u %u %u %u", &wm, &wb, &wh, &wy) != 4) {
      std::fprintf(stderr, "HB_CHECK_WINDOW: MODE BLOCK HALF HEAVY\n");
      return 2;
    }
    check(hb_set_window(hb, int(wm), wb, wh, wy), "hb_set_window");
    std::printf("window mode %u block %u half %u heavy %u\n", wm, wb, wh, wy);
  }
  or_hb *ref = or_hb_create(n, gg.num_arcs, gg.offsets, gg.successors, gt.offsets, gt.successors,
                            log2m, r, OR_HASH_JENKINS, seed, 1, 0, nullptr, 0, 1);
  if (!ref) {
    std::fprintf(stderr, "or_hb_create failed\n");
    return 2;
  }
  hb_params p;
  check(hb_get_params(hb, &p), "hb_get_params");
  const size_t words = size_t(n) * p.words_per_counter;
  std::vector<uint64_t> c(words), c2(words), mod((n + 63) / 64);
  std::printf("n=%llu arcs=%llu log2m=%u r=%u heavy=%u counter memory %u tile %llu\n",
              (unsigned long long)n, (unsigned long long)gg.num_arcs, log2m, r, heavy,
              counter_memory, (unsigned long long)tile_nodes);
  check(hb_init(hb), "hb_init");
  or_hb_init(ref);
  int bad = 0;
  for (int it = 0; it < 200; it++) {
    check(hb_iterate(hb), "hb_iterate");
    or_hb_iterate(ref);
    check(hb_get_counters(hb, 0, c.data(), words), "hb_get_counters");
    check(hb_get_counters_noalloc(hb, it & 1, c2.data(), words), "hb_get_counters_noalloc");
    std::vector<uint64_t> c1(words);
    check(hb_get_counters(hb, it & 1, c1.data(), words), "hb_get_counters");
    if (std::memcmp(c1.data(), c2.data(), words * 8) != 0) {
      std::printf("hb_get_counters_noalloc DIFFERENT at iteration %d\n", it);
      bad++;
    }
    check(hb_get_modified(hb, mod.data(), mod.size()), "hb_get_modified");
    hb_iteration_stats st;
    check(hb_get_iteration_stats(hb, uint64_t(it), &st), "hb_get_iteration_stats");
    const int mode = int(st.systolic) | int(st.local) << 1 | int(st.pre_local) << 2;
    const bool same_c = std::memcmp(c.data(), or_hb_counters(ref, 0), words * 8) == 0;
    const bool same_m = std::memcmp(mod.data(), or_hb_modified_bits(ref), mod.size() * 8) == 0;
    const bool same_mode = mode == or_hb_mode(ref);
    const bool same_stats =
        st.visited_arcs == or_hb_visited_arcs(ref) && st.modified == or_hb_modified(ref);
    std::printf("iteration %d: mode %d modified %llu visited %llu (oracle %llu) delta %u "
                "streamed %u window %u gathered %llu of %llu: %s\n",
                it, mode, (unsigned long long)st.modified, (unsigned long long)st.visited_arcs,
                (unsigned long long)or_hb_visited_arcs(ref), st.delta_gathering, st.streamed,
                st.window, (unsigned long long)st.gathered_slabs,
                (unsigned long long)(st.visited_arcs * (p.num_registers / 32)),
                same_c && same_m && same_mode && same_stats ? "identical" : "DIFFERENT");
    bad += !(same_c && same_m && same_mode && same_stats);
    if (or_hb_modified(ref) == 0 || bad) break;
  }
  std::vector<double> sd(n);
  check(hb_sum_of_distances(hb, sd.data(), n), "hb_sum_of_distances");
  if (std::memcmp(sd.data(), or_hb_sum_of_distances(ref), n * 8) != 0) {
    std::printf("sum of distances DIFFERENT\n");
    bad++;
  }
  or_hb_destroy(ref);
  check(hb_destroy(hb), "hb_destroy");
  hb_graph_free(&gg);
  hb_graph_free(&gt);
  std::printf("%s\n", bad ? "FAIL" : "OK");
  return bad ? 1 : 0;
}
