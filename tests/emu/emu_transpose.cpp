// Host stand-in for hb::launch_transpose in the emulation build, whose
// device version is built on rocPRIM (device-only code): a counting sort
// that yields the same arrays (rows sorted by source).  Test infrastructure.
#include <vector>

#include "hb_kernels.hpp"

namespace hb {
hipError_t launch_transpose(const uint64_t *off, const uint32_t *succ, uint64_t n, uint64_t a,
                            uint64_t *toff, uint32_t *tsucc, hipStream_t) {
  std::vector<uint64_t> cnt(n + 1, 0);
  for (uint64_t k = 0; k < a; k++) cnt[succ[k]]++;
  toff[0] = 0;
  for (uint64_t v = 0; v < n; v++) toff[v + 1] = toff[v] + cnt[v];
  std::vector<uint64_t> cur(toff, toff + n);
  for (uint64_t u = 0; u < n; u++)
    for (uint64_t k = off[u]; k < off[u + 1]; k++) tsucc[cur[succ[k]]++] = uint32_t(u);
  return hipSuccess;
}
}  // namespace hb
