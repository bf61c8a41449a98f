// HyperLogLog arithmetic shared by host and device code.
//
// Counter layout (the reference's SliceCounterArray of usize words,
// src/utils/slice_counter_array.rs:166-169, with the packed registers of
// src/utils/hyper_log_log/hyper_log_log_logic.rs:72-115): counter v occupies
// 64-bit words [v*W, (v+1)*W); register j sits at bit j*r, least significant
// bit first, and may straddle two words.
//
// The device kernels cut a counter into chunks of lcm(64, r) bits, so that no
// register crosses a chunk: r = 5 -> 5 words / 64 registers, r = 6 -> 3 words
// / 32 registers.  Inside a chunk the register pattern is fixed, so the
// broadword masks are compile-time constants.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HB_HD __host__ __device__ __forceinline__
#else
#define HB_HD inline
#endif

namespace hb {

// ---- parameters (hyper_log_log_logic.rs:275-303, :374-428) -------------

// HyperLogLog::register_size: max(5, ceil(ln((ln n / ln 2) / ln 2))).
inline uint32_t register_size(uint64_t n) {
  const double ln2 = 0.6931471805599453;
  double y = std::ceil(std::log((std::log(double(n)) / ln2) / ln2));
  uint64_t v = (std::isnan(y) || y <= 0.0) ? 0 : (y >= 1.8446744073709552e19 ? UINT64_MAX : uint64_t(y));
  return v > 5 ? uint32_t(v) : 5u;
}

inline double rel_std(uint32_t log2m) {
  double t = log2m == 4 ? 1.106 : log2m == 5 ? 1.070 : log2m == 6 ? 1.054 : log2m == 7 ? 1.046 : 1.04;
  return t / std::sqrt(double(uint64_t(1) << log2m));
}

inline uint32_t log2_num_registers(double rsd) {
  double y = std::ceil(std::log2(std::pow(1.106 / rsd, 2.0)));
  return (std::isnan(y) || y <= 0.0) ? 0u : uint32_t(y);
}

inline double alpha_m_m(uint32_t log2m) {
  double m = double(uint64_t(1) << log2m);
  double alpha = log2m == 4 ? 0.673 : log2m == 5 ? 0.697 : log2m == 6 ? 0.709 : 0.7213 / (1.0 + 1.079 / m);
  return alpha * (m * m);
}

// ---- hashes ---------------------------------------------------------------

// Jenkins 64-bit mix of one word (tests/test_hyperball.rs:52-139, finish = c
// at :143-145; a = b = seed, c = 0x9e3779b97f4a7c13 at :33-41).
HB_HD uint64_t jenkins(uint64_t x, uint64_t seed) {
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

HB_HD uint64_t rotl64(uint64_t x, int b) { return (x << b) | (x >> (64 - b)); }

// SipHash-1-3 with zero keys over the 8 little-endian bytes of x: Rust's std
// DefaultHasher on a usize (hyperball_impl.rs:16, :69-71).  The algorithm is
// the published SipHash (Aumasson & Bernstein 2012) with c = 1, d = 3.
HB_HD uint64_t siphash13(uint64_t x) {
  uint64_t v0 = 0x736f6d6570736575ull, v1 = 0x646f72616e646f6dull;
  uint64_t v2 = 0x6c7967656e657261ull, v3 = 0x7465646279746573ull;
  auto round = [&]() {
    v0 += v1; v1 = rotl64(v1, 13); v1 ^= v0; v0 = rotl64(v0, 32);
    v2 += v3; v3 = rotl64(v3, 16); v3 ^= v2;
    v0 += v3; v3 = rotl64(v3, 21); v3 ^= v0;
    v2 += v1; v1 = rotl64(v1, 17); v1 ^= v2; v2 = rotl64(v2, 32);
  };
  v3 ^= x; round(); v0 ^= x;
  const uint64_t b = uint64_t(8) << 56;
  v3 ^= b; round(); v0 ^= b;
  v2 ^= 0xff;
  round(); round(); round();
  return v0 ^ v1 ^ v2 ^ v3;
}

// Seeded stand-in for rand::random::<u64>() (hyperball_impl.rs:1172): the
// SplitMix64 output (Steele, Lea & Flood 2014) for a global item index.
HB_HD uint64_t weight_item(uint64_t seed, uint64_t index) {
  uint64_t z = seed + (index + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// ---- chunk geometry -------------------------------------------------------

template <int R> struct ChunkGeom;
template <> struct ChunkGeom<5> { static constexpr int CW = 5, REGS = 64; };
template <> struct ChunkGeom<6> { static constexpr int CW = 3, REGS = 32; };

// H_r / L_r restricted to word k of a chunk.
template <int R>
HB_HD constexpr uint64_t msb_word(int k) {
  uint64_t m = 0;
  for (int b = 0; b < 64; b++)
    if ((64 * k + b) % R == R - 1) m |= uint64_t(1) << b;
  return m;
}
template <int R>
HB_HD constexpr uint64_t lsb_word(int k) {
  uint64_t m = 0;
  for (int b = 0; b < 64; b++)
    if ((64 * k + b) % R == 0) m |= uint64_t(1) << b;
  return m;
}

}  // namespace hb
