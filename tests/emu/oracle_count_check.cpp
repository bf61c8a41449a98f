// The library's count() kernel (count_all_kernel through launch_count_all,
// emulated) against the oracle's or_count on crafted counters (test
// infrastructure, built by build_emu.py).  Random graphs almost never seed a
// register above 27, so this driver writes the high register classes
// directly: for 5-bit registers values 28..31 (a 31 register adds -2^-31 in
// the reference, hyper_log_log_logic.rs:178), for 6-bit registers 28..63,
// mixed with zeros (the linear-counting branch) and uniform values, for every
// shape the library holds.  Counters are built in the reference layout, then
// permuted into the device slab layout; the f64 results must be identical.
// Usage: oracle_count_check [COUNTERS_PER_SHAPE]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "hb_kernels.hpp"

extern "C" {
#include "hb_oracle.h"
}

namespace {

// Reference layout (register j at bits j*r.., LSB first across u64 words) to
// the slab layout at the library's counter stride: dword s*r + b of a counter
// holds bit b of registers 32s..32s+31.
void to_slabs(const or_hll &h, const uint64_t *ref, uint32_t *dev) {
  const uint32_t cdw = h.m * h.r / 32;
  std::memset(dev, 0, hb::counter_stride(cdw) * 4);
  for (uint32_t j = 0; j < h.m; j++) {
    const uint64_t v = or_get_register(&h, ref, j);
    for (uint32_t b = 0; b < h.r; b++)
      if ((v >> b) & 1) dev[(j >> 5) * h.r + b] |= 1u << (j & 31);
  }
}

// Register value for counter `kind`: each kind stresses a different part of
// the sum.
uint32_t draw(uint32_t kind, uint32_t r, std::mt19937_64 &rng) {
  const uint32_t top = (1u << r) - 1;
  switch (kind % 6) {
    case 0:  // only the top values
      return top - uint32_t(rng() % 4);
    case 1:  // the high classes, every value from 28 up
      return 28 + uint32_t(rng() % (top - 27));
    case 2:  // high values and zeros: the linear-counting branch
      return rng() % 3 == 0 ? 0u : 28 + uint32_t(rng() % (top - 27));
    case 3:  // uniform over every value
      return uint32_t(rng() % (top + 1));
    case 4:  // mostly small, a few at the top
      return rng() % 16 == 0 ? top : uint32_t(rng() % 6);
    default:  // all equal to one value of the class
      return 24 + kind % (top - 23);
  }
}

}  // namespace

int main(int argc, char **argv) {
  const uint32_t per_shape = argc > 1 ? uint32_t(std::atoi(argv[1])) : 36;
  int bad = 0, shapes = 0;
  std::mt19937_64 rng(12345);
  for (uint32_t r = 5; r <= 6; r++) {
    for (uint32_t log2m = 5; log2m <= 16; log2m++) {
      int lanes = 0, spl = 0;
      int nseg = 0;
      if (!hb::union_shape(int(r), int(log2m), &lanes, &spl, &nseg)) continue;
      or_hll h;
      if (or_hll_init(&h, log2m, r, 0, OR_HASH_JENKINS, 0) != 0) continue;
      shapes++;
      const uint32_t cdw = h.m * h.r / 32, cs = hb::counter_stride(cdw);
      std::vector<uint64_t> ref(size_t(per_shape) * h.W, 0);
      std::vector<uint32_t> dev(size_t(per_shape) * cs, 0);
      uint32_t max_value = 0;
      for (uint32_t c = 0; c < per_shape; c++) {
        uint64_t *x = ref.data() + size_t(c) * h.W;
        for (uint32_t j = 0; j < h.m; j++) {
          const uint32_t v = draw(c, r, rng);
          max_value = v > max_value ? v : max_value;
          or_set_register(&h, x, j, v);
        }
        to_slabs(h, x, dev.data() + size_t(c) * cs);
      }
      std::vector<double> lc(h.m + 1, 0.0);
      for (uint32_t z = 1; z <= h.m; z++) lc[z] = double(h.m) * std::log(double(h.m) / double(z));
      uint32_t *d_cnt = nullptr;
      double *d_lc = nullptr, *d_out = nullptr;
      if (hipMalloc(reinterpret_cast<void **>(&d_cnt), dev.size() * 4) != hipSuccess ||
          hipMalloc(reinterpret_cast<void **>(&d_lc), lc.size() * 8) != hipSuccess ||
          hipMalloc(reinterpret_cast<void **>(&d_out), per_shape * 8) != hipSuccess) {
        std::fprintf(stderr, "allocation failed\n");
        return 2;
      }
      hipMemcpyAsync(d_cnt, dev.data(), dev.size() * 4, hipMemcpyHostToDevice, nullptr);
      hipMemcpyAsync(d_lc, lc.data(), lc.size() * 8, hipMemcpyHostToDevice, nullptr);
      if (hb::launch_count_all(int(r), int(log2m), d_cnt, per_shape, h.alpha_m_m, h.m, d_lc, d_out,
                               nullptr) != hipSuccess) {
        std::fprintf(stderr, "launch_count_all failed for r=%u log2m=%u\n", r, log2m);
        return 2;
      }
      std::vector<double> got(per_shape);
      hipMemcpyAsync(got.data(), d_out, per_shape * 8, hipMemcpyDeviceToHost, nullptr);
      hipStreamSynchronize(nullptr);
      int wrong = 0;
      for (uint32_t c = 0; c < per_shape; c++) {
        const double exp = or_count(&h, ref.data() + size_t(c) * h.W);
        if (std::memcmp(&exp, &got[c], 8) != 0) {
          if (wrong < 3)
            std::printf("  counter %u (kind %u): device %.17g, oracle %.17g\n", c, c % 6, got[c],
                        exp);
          wrong++;
        }
      }
      std::printf("r=%u log2m=%u (G=%d, %d slabs per lane): %u counters, registers up to %u: %s\n",
                  r, log2m, lanes, spl, per_shape, max_value, wrong ? "DIFFERENT" : "identical");
      bad += wrong != 0;
      hipFree(d_cnt);
      hipFree(d_lc);
      hipFree(d_out);
    }
  }
  std::printf("%d shapes\n%s\n", shapes, bad ? "FAIL" : "OK");
  return bad ? 1 : 0;
}
