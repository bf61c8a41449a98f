// Host-side graph input of the HyperBall library: BVGraph decoding, CSR
// transposition and the synthetic generators.  No device code here.
//
// BVGraph: the reference obtains successors through the external crate
// webgraph 0.2.0 (BvGraph::with_basename(..).load(), tests/test_hyperball.rs:
// 226-227; successor iteration at src/algo/hyperball/hyperball_impl.rs:1037).
// This decoder follows the published BVGraph format (Boldi & Vigna, "The
// WebGraph Framework I", WWW 2004; format version 0): per node, the outdegree,
// a reference into a window of previous nodes with copy blocks, intervals of
// consecutive successors and zeta_k-coded residual gaps, all in MSB-first bit
// streams.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

#include "hb_common.hpp"

namespace hb {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

// ------------------------------------------------------------------------
// Bit input, MSB first (dsiutils InputBitStream conventions).
// ------------------------------------------------------------------------
// Reads a BitFile's bytes (shared, read-only; one reader per thread).
class BitReader {
 public:
  BitReader(const uint8_t *data, uint64_t nbits) : buf_(data), nbits_(nbits) {}
  uint64_t position() const { return pos_; }
  void seek(uint64_t bit) { pos_ = bit; }

  // 64-bit window starting at the current position.
  uint64_t window() const {
    uint64_t byte = pos_ >> 3;
    uint64_t w = 0;
    for (int k = 0; k < 8; k++) w = (w << 8) | buf_[byte + k];
    unsigned sh = pos_ & 7;
    if (sh) w = (w << sh) | (uint64_t(buf_[byte + 8]) >> (8 - sh));
    return w;
  }
  uint64_t read_int(unsigned len) {  // len <= 64
    if (len == 0) return 0;
    check(len);
    uint64_t v = window() >> (64 - len);
    pos_ += len;
    return v;
  }
  uint64_t read_bit() { return read_int(1); }
  uint64_t read_unary() {
    uint64_t zeros = 0;
    for (;;) {
      check(1);
      uint64_t w = window();
      if (w) {
        unsigned lz = __builtin_clzll(w);
        zeros += lz;
        pos_ += lz + 1;
        return zeros;
      }
      zeros += 64;
      pos_ += 64;
    }
  }
  uint64_t read_gamma() {
    uint64_t msb = read_unary();
    if (msb >= 64) fail(HB_E_INVALID_ARG, "BVGraph: gamma code out of range");
    return ((uint64_t(1) << msb) | read_int(unsigned(msb))) - 1;
  }
  uint64_t read_delta() {
    uint64_t msb = read_gamma();
    if (msb >= 64) fail(HB_E_INVALID_ARG, "BVGraph: delta code out of range");
    return ((uint64_t(1) << msb) | read_int(unsigned(msb))) - 1;
  }
  // zeta_k: h in unary, then a minimal binary code of x + 1 - 2^(hk) over
  // [0, 2^((h+1)k) - 2^(hk)).
  uint64_t read_zeta(unsigned k) {
    if (k == 1) return read_gamma();
    uint64_t h = read_unary();
    if ((h + 1) * k > 63) fail(HB_E_INVALID_ARG, "BVGraph: zeta code out of range");
    uint64_t left = uint64_t(1) << (h * k);
    uint64_t m = read_int(unsigned(h * k + k - 1));
    if (m < left) return m + left - 1;
    return (m << 1) + read_bit() - 1;
  }

 private:
  void check(unsigned len) const {
    if (pos_ + len > nbits_) fail(HB_E_INVALID_ARG, "BVGraph: read past the end of the bit stream");
  }
  const uint8_t *buf_;
  uint64_t nbits_ = 0, pos_ = 0;
};

// A bit-stream file in memory, padded for the reader's 8-byte windows.
struct BitFile {
  std::vector<uint8_t> bytes;
  uint64_t nbits = 0;
  explicit BitFile(std::vector<uint8_t> b) : bytes(std::move(b)) {
    nbits = uint64_t(bytes.size()) * 8;
    bytes.resize(bytes.size() + 16, 0);
  }
  BitReader reader() const { return BitReader(bytes.data(), nbits); }
};

enum Code { GAMMA, DELTA, UNARY, ZETA };

struct BvProps {
  uint64_t nodes = 0, arcs = 0;
  unsigned window = 7, min_interval = 4, zeta_k = 3;
  int version = 0;
  Code outdegrees = GAMMA, references = UNARY, blocks = GAMMA, block_count = GAMMA,
       residuals = ZETA, offsets = GAMMA;
};

std::vector<uint8_t> read_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(HB_E_INVALID_ARG, "cannot open " + path);
  f.seekg(0, std::ios::end);
  std::streamoff sz = f.tellg();
  f.seekg(0);
  std::vector<uint8_t> v(size_t(sz > 0 ? sz : 0));
  if (sz > 0) f.read(reinterpret_cast<char *>(v.data()), sz);
  return v;
}

Code parse_code(const std::string &name, const std::string &flag) {
  if (name == "GAMMA") return GAMMA;
  if (name == "DELTA") return DELTA;
  if (name == "UNARY") return UNARY;
  if (name == "ZETA") return ZETA;
  fail(HB_E_INVALID_ARG, "BVGraph: unsupported compression flag " + flag);
}

BvProps read_properties(const std::string &basename) {
  std::ifstream f(basename + ".properties");
  if (!f) fail(HB_E_INVALID_ARG, "cannot open " + basename + ".properties");
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    auto eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string val = line.substr(eq + 1);
    while (!val.empty() && (val.back() == '\r' || val.back() == ' ')) val.pop_back();
    kv[line.substr(0, eq)] = val;
  }
  BvProps p;
  auto num = [&](const char *k, uint64_t def) -> uint64_t {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::stoull(it->second);
  };
  if (!kv.count("nodes") || !kv.count("arcs"))
    fail(HB_E_INVALID_ARG, "BVGraph: properties lack nodes/arcs");
  p.nodes = num("nodes", 0);
  p.arcs = num("arcs", 0);
  p.window = unsigned(num("windowsize", 7));
  p.min_interval = unsigned(num("minintervallength", 4));
  p.zeta_k = unsigned(num("zetak", 3));
  p.version = int(num("version", 0));
  if (p.version != 0) fail(HB_E_INVALID_ARG, "BVGraph: only format version 0 is supported");
  if (kv.count("endianness") && kv["endianness"] != "big")
    fail(HB_E_INVALID_ARG, "BVGraph: only big-endian bit streams are supported");
  if (p.nodes > 0xffffffffull) fail(HB_E_INVALID_ARG, "BVGraph: more than 2^32 nodes");
  std::string flags = kv.count("compressionflags") ? kv["compressionflags"] : "";
  std::stringstream ss(flags);
  std::string tok;
  while (std::getline(ss, tok, '|')) {
    while (!tok.empty() && tok.front() == ' ') tok.erase(tok.begin());
    while (!tok.empty() && tok.back() == ' ') tok.pop_back();
    if (tok.empty()) continue;
    auto us = tok.rfind('_');
    if (us == std::string::npos) fail(HB_E_INVALID_ARG, "BVGraph: bad compression flag " + tok);
    std::string comp = tok.substr(0, us), code = tok.substr(us + 1);
    Code c = parse_code(code, tok);
    if (comp == "OUTDEGREES") p.outdegrees = c;
    else if (comp == "REFERENCES") p.references = c;
    else if (comp == "BLOCKS") p.blocks = c;
    else if (comp == "BLOCK_COUNT") p.block_count = c;
    else if (comp == "RESIDUALS") p.residuals = c;
    else if (comp == "OFFSETS") p.offsets = c;
    else fail(HB_E_INVALID_ARG, "BVGraph: bad compression flag " + tok);
  }
  return p;
}

uint64_t read_code(BitReader &in, Code c, unsigned zk) {
  switch (c) {
    case GAMMA: return in.read_gamma();
    case DELTA: return in.read_delta();
    case UNARY: return in.read_unary();
    case ZETA: return in.read_zeta(zk);
  }
  return 0;
}

// int2nat inverse: 0,1,2,3,4 -> 0,-1,1,-2,2.
int64_t nat2int(uint64_t x) { return int64_t(x >> 1) ^ -int64_t(x & 1); }

// Decodes the successor list of node x, the reader being positioned at x's
// code, into `out` (sorted).  ref(r) gives the decoded list of an earlier
// node r as (pointer, length).  Format: outdegree; if the window is not 0, a
// reference (x - r); for a reference, a block count and copy/skip block
// lengths over r's list; then, if min_interval > 0, intervals of consecutive
// successors; then gap-coded residuals (the first relative to x, signed).
template <class RefFn>
void decode_list(BitReader &in, const BvProps &p, uint64_t x, RefFn &&ref,
                 std::vector<uint32_t> &out, std::vector<uint64_t> &blocks) {
  const uint64_t n = p.nodes;
  out.clear();
  const uint64_t d = read_code(in, p.outdegrees, p.zeta_k);
  if (d == 0) return;
  const uint64_t refd = p.window > 0 ? read_code(in, p.references, p.zeta_k) : 0;
  if (refd > x || refd > p.window) fail(HB_E_INVALID_ARG, "BVGraph: reference out of the window");
  if (refd > 0) {
    const auto rl = ref(x - refd);
    const uint32_t *rs = rl.first;
    const uint64_t rdeg = rl.second;
    const uint64_t bc = read_code(in, p.block_count, p.zeta_k);
    blocks.assign(bc, 0);
    for (uint64_t i = 0; i < bc; i++)
      blocks[i] = read_code(in, p.blocks, p.zeta_k) + (i == 0 ? 0 : 1);
    // Blocks alternate copy / skip, starting with copy; after the last
    // block the rest of the reference list is copied iff the count is even.
    uint64_t pos = 0;
    for (uint64_t i = 0; i < bc; i++) {
      const uint64_t len = blocks[i];
      if (pos + len > rdeg) fail(HB_E_INVALID_ARG, "BVGraph: copy block past reference list");
      if ((i & 1) == 0) out.insert(out.end(), rs + pos, rs + pos + len);
      pos += len;
    }
    if ((bc & 1) == 0) out.insert(out.end(), rs + pos, rs + rdeg);
  }
  if (out.size() > d) fail(HB_E_INVALID_ARG, "BVGraph: more copied successors than outdegree");
  uint64_t extra_count = d - out.size();
  if (extra_count > 0 && p.min_interval > 0) {
    const uint64_t ic = in.read_gamma();
    int64_t prev = 0;
    for (uint64_t i = 0; i < ic; i++) {
      int64_t left;
      if (i == 0) left = nat2int(in.read_gamma()) + int64_t(x);
      else left = int64_t(in.read_gamma()) + prev + 1;
      const uint64_t len = in.read_gamma() + p.min_interval;
      if (left < 0 || uint64_t(left) + len > n || len > extra_count)
        fail(HB_E_INVALID_ARG, "BVGraph: interval out of range");
      for (uint64_t k = 0; k < len; k++) out.push_back(uint32_t(left + int64_t(k)));
      prev = left + int64_t(len);
      extra_count -= len;
    }
  }
  int64_t prev = 0;
  for (uint64_t i = 0; i < extra_count; i++) {
    int64_t s;
    if (i == 0) s = int64_t(x) + nat2int(read_code(in, p.residuals, p.zeta_k));
    else s = prev + int64_t(read_code(in, p.residuals, p.zeta_k)) + 1;
    if (s < 0 || uint64_t(s) >= n) fail(HB_E_INVALID_ARG, "BVGraph: residual out of range");
    out.push_back(uint32_t(s));
    prev = s;
  }
  std::sort(out.begin(), out.end());
}

// Sequential decode of every successor list.
void decode_bvgraph(const std::string &basename, std::vector<uint64_t> &offsets,
                    std::vector<uint32_t> &succ, std::vector<uint64_t> *bit_offsets) {
  BvProps p = read_properties(basename);
  const BitFile gf(read_file(basename + ".graph"));
  BitReader in = gf.reader();
  const uint64_t n = p.nodes;
  offsets.assign(n + 1, 0);
  succ.clear();
  succ.reserve(p.arcs);
  if (bit_offsets) bit_offsets->assign(n + 1, 0);
  std::vector<uint64_t> blocks;
  std::vector<uint32_t> merged;
  auto ref = [&](uint64_t r) {
    return std::make_pair(succ.data() + offsets[r], offsets[r + 1] - offsets[r]);
  };
  for (uint64_t x = 0; x < n; x++) {
    if (bit_offsets) (*bit_offsets)[x] = in.position();
    offsets[x] = succ.size();
    decode_list(in, p, x, ref, merged, blocks);
    succ.insert(succ.end(), merged.begin(), merged.end());
  }
  offsets[n] = succ.size();
  if (bit_offsets) (*bit_offsets)[n] = in.position();
  if (succ.size() != p.arcs)
    fail(HB_E_INVALID_ARG, "BVGraph: decoded " + std::to_string(succ.size()) +
                               " arcs, properties say " + std::to_string(p.arcs));
}

// Bit offset of every node's code in <basename>.graph, and the stream's end,
// from <basename>.offsets (gap-coded with the OFFSETS code).
std::vector<uint64_t> read_bit_offsets(const std::string &basename, const BvProps &p) {
  const BitFile of(read_file(basename + ".offsets"));
  BitReader in = of.reader();
  std::vector<uint64_t> out(p.nodes + 1);
  uint64_t acc = 0;
  for (uint64_t i = 0; i <= p.nodes; i++) {
    acc += read_code(in, p.offsets, p.zeta_k);
    out[i] = acc;
  }
  return out;
}

// ------------------------------------------------------------------------
// Elias-Fano sequences in the epserde files of the sux crate (<basename>.ef,
// the bit offsets that webgraph's BvGraph::with_basename(..).load() reads for
// random access; <basename>.dcf, the cumulative degrees that the reference
// loads with DCF::load_mmap at hyperball_impl.rs:1270 and main.rs:57).
//
// File: the 8-byte magic "epserde ", major and minor version (u16 each), the
// size of usize (u8), two 64-bit type hashes, the type name (u64 length and
// bytes), then the fields in declaration order, little-endian.  Scalars
// follow each other unaligned; a slice is its u64 length, zero padding to
// an 8-byte file offset, and its words.  EliasFano { n, u, l, low_bits:
// BitFieldVec { words, bit_width, mask, len }, high_bits }, where the
// high-bits structure (SelectAdaptConst, or SelectZeroAdaptConst over it)
// serialises its BitVec { words, len } first; the selection inventories
// after it are not needed for a full decode.  Element i is
// ((position of the i-th one in the high bits) - i) << l | low_bits[i].
// The type hashes are not checked; the structure is (lengths, widths, the
// high-bits length n + (u >> l) + 1, monotone values up to u).
struct EfReader {
  const std::vector<uint8_t> &b;
  const std::string &path;
  uint64_t o = 0;
  [[noreturn]] void bad(const std::string &why) const {
    fail(HB_E_INVALID_ARG, path + ": not an Elias-Fano epserde file (" + why + ")");
  }
  void need(uint64_t k) const {
    if (o > b.size() || b.size() - o < k) bad("truncated");
  }
  uint64_t u64() {
    need(8);
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | b[o + i];
    o += 8;
    return v;
  }
  // a slice of words: its offset in the file and its length
  std::pair<uint64_t, uint64_t> words() {
    const uint64_t len = u64();
    o = (o + 7) & ~uint64_t(7);
    if (len > (b.size() - std::min<uint64_t>(o, b.size())) / 8) bad("slice past the end");
    const uint64_t at = o;
    o += len * 8;
    return {at, len};
  }
  uint64_t word(uint64_t at, uint64_t i) const {
    uint64_t v = 0;
    for (int k = 7; k >= 0; k--) v = (v << 8) | b[at + i * 8 + k];
    return v;
  }
};

std::vector<uint64_t> read_elias_fano(const std::string &path) {
  const std::vector<uint8_t> b = read_file(path);
  EfReader r{b, path};
  r.need(13);
  if (std::memcmp(b.data(), "epserde ", 8) != 0) r.bad("no epserde magic");
  const unsigned major = b[8] | (b[9] << 8);
  if (major != 1) r.bad("epserde major version " + std::to_string(major));
  if (b[12] != 8) r.bad("usize of " + std::to_string(b[12]) + " bytes");
  r.o = 13 + 16;  // the two type hashes
  const uint64_t name_len = r.u64();
  r.need(name_len);
  const std::string name(reinterpret_cast<const char *>(b.data() + r.o), size_t(name_len));
  r.o += name_len;
  if (name.rfind("sux::dict::elias_fano::EliasFano<", 0) != 0) r.bad("type " + name);
  const uint64_t n = r.u64(), u = r.u64(), l = r.u64();
  if (l > 63) r.bad("low width " + std::to_string(l));
  const auto low = r.words();
  const uint64_t bit_width = r.u64(), mask = r.u64(), low_len = r.u64();
  if (bit_width != l || low_len != n || mask != (l ? (~uint64_t(0) >> (64 - l)) : 0))
    r.bad("low bits do not match n and l");
  if (low.second < (n * l + 63) / 64) r.bad("low bits too short");
  const auto high = r.words();
  const uint64_t high_len = r.u64();
  if (high_len != n + (u >> l) + 1 || high.second < (high_len + 63) / 64)
    r.bad("high bits do not match n, u and l");
  std::vector<uint64_t> out(n);
  uint64_t i = 0;
  for (uint64_t w = 0; w < high.second && i < n; w++) {
    if (w * 64 >= high_len) break;
    uint64_t x = r.word(high.first, w);
    if (high_len - w * 64 < 64) x &= (uint64_t(1) << (high_len - w * 64)) - 1;
    while (x && i < n) {
      const uint64_t pos = w * 64 + uint64_t(__builtin_ctzll(x));
      x &= x - 1;
      uint64_t lo = 0;
      if (l) {
        const uint64_t bit = i * l, wi = bit >> 6, sh = bit & 63;
        lo = r.word(low.first, wi) >> sh;
        if (sh + l > 64) lo |= r.word(low.first, wi + 1) << (64 - sh);
        lo &= mask;
      }
      out[i] = ((pos - i) << l) | lo;
      if (out[i] > u || (i && out[i] < out[i - 1])) r.bad("values not monotone up to u");
      i++;
    }
  }
  if (i != n) r.bad("fewer ones in the high bits than elements");
  return out;
}

// The bit offsets of <basename>.graph's node records, n + 1 of them: from
// <basename>.ef when it exists (webgraph's own index), else from .offsets.
std::vector<uint64_t> read_node_index(const std::string &basename, const BvProps &p) {
  if (std::ifstream(basename + ".ef").good()) {
    std::vector<uint64_t> bits = read_elias_fano(basename + ".ef");
    if (bits.size() != p.nodes + 1)
      fail(HB_E_INVALID_ARG, basename + ".ef holds " + std::to_string(bits.size()) +
                                 " offsets, the graph has " + std::to_string(p.nodes) + " nodes");
    return bits;
  }
  return read_bit_offsets(basename, p);
}

// Parallel decode through the node index (.ef or .offsets): node ranges balanced by bit
// length are decoded by separate threads.  A reference to a node before a
// thread's range is resolved by decoding that node at its indexed bit offset
// (recursively through its own references), with a per-thread cache.  Every
// range must end exactly at the next range's indexed offset.
void decode_bvgraph_parallel(const std::string &basename, unsigned threads,
                             std::vector<uint64_t> &offsets, std::vector<uint32_t> &succ) {
  const BvProps p = read_properties(basename);
  const uint64_t n = p.nodes;
  const std::vector<uint64_t> bits = read_node_index(basename, p);
  const BitFile gf(read_file(basename + ".graph"));
  if (bits[n] > gf.nbits)
    fail(HB_E_INVALID_ARG, "BVGraph: the node index points past the .graph stream");
  for (uint64_t x = 0; x < n; x++)
    if (bits[x + 1] < bits[x]) fail(HB_E_INVALID_ARG, "BVGraph: the node index is not monotone");
  const uint64_t want = std::max<uint64_t>(1, std::min<uint64_t>(n, uint64_t(threads) * 8));
  std::vector<uint64_t> bounds{0};
  for (uint64_t c = 1; c < want; c++) {
    const uint64_t target = bits[n] / want * c;
    const uint64_t x = uint64_t(std::lower_bound(bits.begin(), bits.begin() + n, target) - bits.begin());
    if (x > bounds.back() && x < n) bounds.push_back(x);
  }
  bounds.push_back(n);
  const size_t chunks = bounds.size() - 1;
  std::vector<std::vector<uint64_t>> loff(chunks);
  std::vector<std::vector<uint32_t>> lsucc(chunks);
  std::atomic<size_t> next{0};
  std::mutex err_mu;
  std::exception_ptr err;
  auto worker = [&] {
    try {
      for (size_t c; (c = next.fetch_add(1)) < chunks;) {
        const uint64_t b = bounds[c], e = bounds[c + 1];
        std::vector<uint64_t> &lo = loff[c];
        std::vector<uint32_t> &ls = lsucc[c];
        lo.reserve(e - b + 1);
        std::map<uint64_t, std::vector<uint32_t>> outside;
        std::vector<uint64_t> blocks;
        std::function<std::pair<const uint32_t *, uint64_t>(uint64_t)> ref =
            [&](uint64_t r) -> std::pair<const uint32_t *, uint64_t> {
          if (r >= b) return {ls.data() + lo[r - b], lo[r - b + 1] - lo[r - b]};
          auto it = outside.find(r);
          if (it == outside.end()) {
            BitReader rin = gf.reader();
            rin.seek(bits[r]);
            std::vector<uint32_t> lst;
            std::vector<uint64_t> blk;
            decode_list(rin, p, r, ref, lst, blk);
            it = outside.emplace(r, std::move(lst)).first;
          }
          return {it->second.data(), it->second.size()};
        };
        BitReader in = gf.reader();
        in.seek(bits[b]);
        std::vector<uint32_t> merged;
        for (uint64_t x = b; x < e; x++) {
          lo.push_back(ls.size());  // also the end of x - 1's list, for ref()
          if (in.position() != bits[x])
            fail(HB_E_INVALID_ARG, "BVGraph: the node index disagrees with the .graph stream at node " +
                                       std::to_string(x));
          decode_list(in, p, x, ref, merged, blocks);
          ls.insert(ls.end(), merged.begin(), merged.end());
        }
        lo.push_back(ls.size());
        if (in.position() != bits[e])
          fail(HB_E_INVALID_ARG, "BVGraph: the node index disagrees with the .graph stream at node " +
                                     std::to_string(e));
      }
    } catch (...) {
      std::lock_guard<std::mutex> l(err_mu);
      if (!err) err = std::current_exception();
    }
  };
  std::vector<std::thread> pool;
  const unsigned t = std::max(1u, std::min<unsigned>(threads, unsigned(chunks)));
  for (unsigned i = 0; i < t; i++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  if (err) std::rethrow_exception(err);
  offsets.assign(n + 1, 0);
  uint64_t total = 0;
  for (size_t c = 0; c < chunks; c++) total += lsucc[c].size();
  if (total != p.arcs)
    fail(HB_E_INVALID_ARG, "BVGraph: decoded " + std::to_string(total) + " arcs, properties say " +
                               std::to_string(p.arcs));
  succ.assign(total, 0);
  uint64_t base = 0;
  for (size_t c = 0; c < chunks; c++) {
    const uint64_t b = bounds[c], e = bounds[c + 1];
    for (uint64_t x = b; x < e; x++) offsets[x] = base + loff[c][x - b];
    std::copy(lsucc[c].begin(), lsucc[c].end(), succ.begin() + base);
    base += lsucc[c].size();
  }
  offsets[n] = base;
}

// ------------------------------------------------------------------------
// BVGraph writer: the same format, MSB-first bit streams, with the
// compressor's choices of the published format: for each node the reference
// within the window (at most maxrefcount references deep) whose encoding
// takes the fewest bits, ties to the nearer candidate, no reference counted as
// distance 0; copy blocks over the reference's list; maximal runs of
// consecutive extra successors of at least min_interval as intervals; the
// rest as residual gaps.
// ------------------------------------------------------------------------
class BitWriter {
 public:
  void write_int(uint64_t v, unsigned len) {  // the low `len` bits of v, MSB first
    for (unsigned i = len; i-- > 0;) put((v >> i) & 1);
  }
  void write_unary(uint64_t x) {
    for (uint64_t i = 0; i < x; i++) put(0);
    put(1);
  }
  void write_gamma(uint64_t x) {
    const uint64_t y = x + 1;
    const unsigned msb = 63 - __builtin_clzll(y);
    write_unary(msb);
    write_int(y, msb);
  }
  void write_delta(uint64_t x) {
    const uint64_t y = x + 1;
    const unsigned msb = 63 - __builtin_clzll(y);
    write_gamma(msb);
    write_int(y, msb);
  }
  // The inverse of BitReader::read_zeta: h = floor(log2(x + 1) / k) in
  // unary, then x + 1 - 2^(hk) in a minimal binary code over
  // [0, 2^((h+1)k) - 2^(hk)).
  void write_zeta(uint64_t x, unsigned k) {
    if (k == 1) return write_gamma(x);
    const uint64_t y = x + 1;
    const unsigned msb = 63 - __builtin_clzll(y);
    const unsigned h = msb / k;
    if ((h + 1) * k > 63) fail(HB_E_INVALID_ARG, "BVGraph writer: zeta code out of range");
    write_unary(h);
    const uint64_t left = uint64_t(1) << (h * k);
    const unsigned s = (h + 1) * k;
    if (y < 2 * left) write_int(y - left, s - 1);
    else write_int(y, s);
  }
  uint64_t bits() const { return nbits_; }
  // The bytes, zero-padded to a whole number of 32-bit words as the Java
  // writer leaves its files.
  std::vector<uint8_t> take() {
    if (nbits_ & 7) bytes_.push_back(uint8_t(cur_ << (8 - (nbits_ & 7))));
    while (bytes_.size() % 4) bytes_.push_back(0);
    cur_ = 0;
    return std::move(bytes_);
  }

 private:
  void put(uint64_t b) {
    cur_ = uint8_t((cur_ << 1) | b);
    if ((++nbits_ & 7) == 0) {
      bytes_.push_back(cur_);
      cur_ = 0;
    }
  }
  std::vector<uint8_t> bytes_;
  uint8_t cur_ = 0;
  uint64_t nbits_ = 0;
};

// Counts the bits a code would take, for the reference choice.
class BitCounter {
 public:
  void write_int(uint64_t, unsigned len) { n_ += len; }
  void write_unary(uint64_t x) { n_ += x + 1; }
  void write_gamma(uint64_t x) { n_ += 2 * (63 - __builtin_clzll(x + 1)) + 1; }
  void write_delta(uint64_t x) {
    const unsigned msb = 63 - __builtin_clzll(x + 1);
    write_gamma(msb);
    n_ += msb;
  }
  void write_zeta(uint64_t x, unsigned k) {
    if (k == 1) return write_gamma(x);
    const uint64_t y = x + 1;
    const unsigned h = (63 - __builtin_clzll(y)) / k;
    const uint64_t left = uint64_t(1) << (h * k);
    n_ += h + 1 + ((h + 1) * k - (y < 2 * left ? 1 : 0));
  }
  uint64_t bits() const { return n_; }

 private:
  uint64_t n_ = 0;
};

template <class W>
void write_code(W &w, Code c, unsigned zk, uint64_t x) {
  switch (c) {
    case GAMMA: w.write_gamma(x); break;
    case DELTA: w.write_delta(x); break;
    case UNARY: w.write_unary(x); break;
    case ZETA: w.write_zeta(x, zk); break;
  }
}

uint64_t int2nat(int64_t v) { return v >= 0 ? uint64_t(v) << 1 : (uint64_t(-(v + 1)) << 1) | 1; }

// Everything of node x's record after the outdegree, given its list and
// the reference r (0: none) with the referenced node's list.
template <class W>
void write_tail(W &w, const BvProps &p, uint64_t x, const uint32_t *list, uint64_t d, uint64_t r,
                const uint32_t *rl, uint64_t rd, std::vector<uint32_t> &extra) {
  if (p.window > 0) write_code(w, p.references, p.zeta_k, r);
  extra.clear();
  if (r > 0) {
    // runs over the referenced list of "copied" / "not copied", from a
    // copy run (possibly empty); the last run is implied by the parity
    std::vector<uint64_t> runs;
    uint64_t i = 0, j = 0, run = 0;
    bool copying = true;
    for (; i < rd; i++) {
      while (j < d && list[j] < rl[i]) extra.push_back(list[j++]);
      const bool in = j < d && list[j] == rl[i];
      if (in) j++;
      if (in != copying) {
        runs.push_back(run);
        run = 0;
        copying = in;
      }
      run++;
    }
    while (j < d) extra.push_back(list[j++]);
    write_code(w, p.block_count, p.zeta_k, runs.size());
    for (size_t b = 0; b < runs.size(); b++) write_code(w, p.blocks, p.zeta_k, runs[b] - (b ? 1 : 0));
  } else {
    extra.assign(list, list + d);
  }
  if (extra.empty()) return;
  std::vector<uint32_t> resid;
  if (p.min_interval > 0) {
    std::vector<std::pair<uint64_t, uint64_t>> iv;
    for (size_t a = 0; a < extra.size();) {
      size_t b = a + 1;
      while (b < extra.size() && extra[b] == extra[b - 1] + 1) b++;
      if (b - a >= p.min_interval) iv.push_back({extra[a], b - a});
      else resid.insert(resid.end(), extra.begin() + a, extra.begin() + b);
      a = b;
    }
    w.write_gamma(iv.size());
    uint64_t prev = 0;
    for (size_t k = 0; k < iv.size(); k++) {
      if (k == 0) w.write_gamma(int2nat(int64_t(iv[k].first) - int64_t(x)));
      else w.write_gamma(iv[k].first - prev - 1);
      w.write_gamma(iv[k].second - p.min_interval);
      prev = iv[k].first + iv[k].second;
    }
  } else {
    resid = extra;
  }
  uint64_t prev = 0;
  for (size_t k = 0; k < resid.size(); k++) {
    if (k == 0) write_code(w, p.residuals, p.zeta_k, int2nat(int64_t(resid[k]) - int64_t(x)));
    else write_code(w, p.residuals, p.zeta_k, resid[k] - prev - 1);
    prev = resid[k];
  }
}

const char *code_name(Code c) {
  switch (c) {
    case GAMMA: return "GAMMA";
    case DELTA: return "DELTA";
    case UNARY: return "UNARY";
    case ZETA: return "ZETA";
  }
  return "";
}

void write_file(const std::string &path, const std::vector<uint8_t> &bytes) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) fail(HB_E_INVALID_ARG, "cannot create " + path);
  f.write(reinterpret_cast<const char *>(bytes.data()), std::streamsize(bytes.size()));
  if (!f) fail(HB_E_INVALID_ARG, "cannot write " + path);
}

void store_bvgraph(const hb_csr &g, const std::string &basename, const BvProps &p,
                   uint64_t maxref) {
  const uint64_t n = g.num_nodes;
  BitWriter gw, ow;
  std::vector<uint64_t> refcount(n, 0);
  std::vector<uint32_t> extra;
  uint64_t last = 0;
  for (uint64_t x = 0; x < n; x++) {
    const uint64_t pos = gw.bits();
    write_code(ow, p.offsets, p.zeta_k, pos - last);  // gaps between record starts
    last = pos;
    const uint32_t *list = g.successors + g.offsets[x];
    const uint64_t d = g.offsets[x + 1] - g.offsets[x];
    for (uint64_t k = 1; k < d; k++)
      if (list[k] <= list[k - 1]) fail(HB_E_INVALID_ARG, "BVGraph writer: rows must be strictly increasing");
    write_code(gw, p.outdegrees, p.zeta_k, d);
    if (d == 0) continue;
    uint64_t best_r = 0, best_bits = UINT64_MAX;
    for (uint64_t r = 0; r <= std::min<uint64_t>(p.window, x); r++) {
      const uint64_t y = x - r;
      const uint64_t rd = r ? g.offsets[y + 1] - g.offsets[y] : 0;
      if (r > 0 && (rd == 0 || refcount[y] >= maxref)) continue;
      BitCounter c;
      write_tail(c, p, x, list, d, r, r ? g.successors + g.offsets[y] : nullptr, rd, extra);
      if (c.bits() < best_bits) {
        best_bits = c.bits();
        best_r = r;
      }
    }
    const uint64_t y = x - best_r;
    write_tail(gw, p, x, list, d, best_r, best_r ? g.successors + g.offsets[y] : nullptr,
               best_r ? g.offsets[y + 1] - g.offsets[y] : 0, extra);
    refcount[x] = best_r ? refcount[y] + 1 : 0;
  }
  write_code(ow, p.offsets, p.zeta_k, gw.bits() - last);
  std::ostringstream props;
  props << "#BVGraph properties\n"
        << "graphclass=it.unimi.dsi.webgraph.BVGraph\n"
        << "version=0\n"
        << "endianness=big\n"
        << "nodes=" << n << "\n"
        << "arcs=" << g.num_arcs << "\n"
        << "windowsize=" << p.window << "\n"
        << "maxrefcount=" << maxref << "\n"
        << "minintervallength=" << p.min_interval << "\n"
        << "zetak=" << p.zeta_k << "\n";
  std::string flags;
  const std::pair<const char *, Code> comps[] = {
      {"OUTDEGREES", p.outdegrees}, {"REFERENCES", p.references}, {"BLOCKS", p.blocks},
      {"BLOCK_COUNT", p.block_count}, {"RESIDUALS", p.residuals}, {"OFFSETS", p.offsets}};
  const Code defaults[] = {GAMMA, UNARY, GAMMA, GAMMA, ZETA, GAMMA};
  for (int i = 0; i < 6; i++) {
    if (comps[i].second == defaults[i]) continue;
    if (!flags.empty()) flags += "|";
    flags += std::string(comps[i].first) + "_" + code_name(comps[i].second);
  }
  props << "compressionflags=" << flags << "\n";
  const std::string pstr = props.str();
  write_file(basename + ".graph", gw.take());
  write_file(basename + ".offsets", ow.take());
  write_file(basename + ".properties", std::vector<uint8_t>(pstr.begin(), pstr.end()));
}

template <class T>
T *copy_out(const std::vector<T> &v) {
  T *p = static_cast<T *>(std::malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!p) throw std::bad_alloc();
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

// ------------------------------------------------------------------------
// Random numbers: xoshiro256** (Blackman & Vigna 2018) seeded by SplitMix64.
// ------------------------------------------------------------------------
inline uint64_t splitmix_next(uint64_t &z) {
  uint64_t r = (z += 0x9e3779b97f4a7c15ull);
  r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
  r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
  return r ^ (r >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Rng {
  uint64_t s[4];
  Rng(uint64_t seed, uint64_t stream) {
    uint64_t z = seed ^ (stream * 0xd1b54a32d192ed03ull);
    for (auto &w : s) w = splitmix_next(z);
  }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  uint64_t below(uint64_t n) { return uint64_t((unsigned __int128)next() * n >> 64); }
  double uniform() { return double(next() >> 11) * 0x1p-53; }
};

void fill_graph(hb_graph *out, uint64_t n, std::vector<uint64_t> &&off, std::vector<uint32_t> &&succ) {
  out->num_nodes = n;
  out->num_arcs = succ.size();
  out->offsets = copy_out(off);
  out->successors = copy_out(succ);
}

}  // namespace
}  // namespace hb

using namespace hb;

extern "C" {

const char *hb_last_error(void) { return hb::g_last_error.c_str(); }

hb_status hb_bvgraph_load_threads(const char *basename, unsigned threads, hb_graph *out) {
  return guard([&] {
    if (!basename || !out) fail(HB_E_INVALID_ARG, "null argument");
    std::vector<uint64_t> off;
    std::vector<uint32_t> succ;
    if (threads == 0) threads = host_threads();
    const std::string base(basename);
    if (threads > 1 &&
        (std::ifstream(base + ".ef").good() || std::ifstream(base + ".offsets").good()))
      decode_bvgraph_parallel(base, threads, off, succ);
    else
      decode_bvgraph(base, off, succ, nullptr);
    fill_graph(out, off.size() - 1, std::move(off), std::move(succ));
  });
}

void hb_bvgraph_params_init(hb_bvgraph_params *p) {
  if (!p) return;
  p->window_size = 7;
  p->max_ref_count = 3;
  p->min_interval_length = 4;
  p->zeta_k = 3;
  p->outdegrees = HB_CODE_GAMMA;
  p->references = HB_CODE_UNARY;
  p->blocks = HB_CODE_GAMMA;
  p->block_count = HB_CODE_GAMMA;
  p->residuals = HB_CODE_ZETA;
  p->offsets = HB_CODE_GAMMA;
}

hb_status hb_bvgraph_store(const hb_csr *g, const char *basename, const hb_bvgraph_params *params) {
  return guard([&] {
    if (!g || !basename || !params) fail(HB_E_INVALID_ARG, "null argument");
    if (!g->offsets || (g->num_arcs && !g->successors)) fail(HB_E_INVALID_ARG, "null graph arrays");
    const uint64_t n = g->num_nodes;
    if (n > 0xffffffffull) fail(HB_E_INVALID_ARG, "more than 2^32 - 1 nodes");
    if (g->offsets[0] != 0 || g->offsets[n] != g->num_arcs)
      fail(HB_E_INVALID_ARG, "offsets must run from 0 to num_arcs");
    for (uint64_t v = 0; v < n; v++)
      if (g->offsets[v + 1] < g->offsets[v]) fail(HB_E_INVALID_ARG, "offsets are not monotone");
    for (uint64_t k = 0; k < g->num_arcs; k++)
      if (g->successors[k] >= n) fail(HB_E_INVALID_ARG, "successor out of range");
    const uint32_t codes[6] = {params->outdegrees, params->references, params->blocks,
                               params->block_count, params->residuals, params->offsets};
    for (uint32_t c : codes)
      if (c > HB_CODE_ZETA) fail(HB_E_INVALID_ARG, "unknown code");
    if (params->zeta_k < 1 || params->zeta_k > 16) fail(HB_E_INVALID_ARG, "zeta_k must be 1..16");
    if (params->window_size > 1u << 20) fail(HB_E_INVALID_ARG, "window size too large");
    BvProps p;
    p.nodes = n;
    p.arcs = g->num_arcs;
    p.window = params->window_size;
    p.min_interval = params->min_interval_length;
    p.zeta_k = params->zeta_k;
    p.outdegrees = Code(params->outdegrees);
    p.references = Code(params->references);
    p.blocks = Code(params->blocks);
    p.block_count = Code(params->block_count);
    p.residuals = Code(params->residuals);
    p.offsets = Code(params->offsets);
    store_bvgraph(*g, basename, p, params->max_ref_count);
  });
}

hb_status hb_bvgraph_load(const char *basename, hb_graph *out) {
  return hb_bvgraph_load_threads(basename, 0, out);
}

hb_status hb_bvgraph_decoder_offsets(const char *basename, uint64_t *out, uint64_t cap) {
  return guard([&] {
    if (!basename || !out) fail(HB_E_INVALID_ARG, "null argument");
    std::vector<uint64_t> off, bits;
    std::vector<uint32_t> succ;
    decode_bvgraph(basename, off, succ, &bits);
    if (cap < bits.size()) fail(HB_E_INVALID_ARG, "output buffer too small");
    std::memcpy(out, bits.data(), bits.size() * sizeof(uint64_t));
  });
}

hb_status hb_bvgraph_load_offsets(const char *basename, uint64_t num_nodes, uint64_t *out) {
  return guard([&] {
    if (!basename || !out) fail(HB_E_INVALID_ARG, "null argument");
    BvProps p = read_properties(basename);
    if (num_nodes != p.nodes)
      fail(HB_E_INVALID_ARG, "BVGraph: the graph has " + std::to_string(p.nodes) + " nodes");
    const std::vector<uint64_t> bits = read_bit_offsets(basename, p);
    std::memcpy(out, bits.data(), bits.size() * sizeof(uint64_t));
  });
}

hb_status hb_elias_fano_load(const char *path, uint64_t *out, uint64_t cap, uint64_t *count) {
  return guard([&] {
    if (!path || !count) fail(HB_E_INVALID_ARG, "null argument");
    const std::vector<uint64_t> v = read_elias_fano(path);
    *count = v.size();
    if (!out) return;
    if (cap < v.size()) fail(HB_E_INVALID_ARG, "output buffer too small");
    std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
  });
}

hb_status hb_graph_transpose(const hb_csr *g, hb_graph *out) {
  return guard([&] {
    if (!g || !out || !g->offsets || (g->num_arcs && !g->successors))
      fail(HB_E_INVALID_ARG, "null argument");
    const uint64_t n = g->num_nodes, a = g->num_arcs;
    std::vector<uint64_t> toff(n + 1, 0);
    std::vector<uint32_t> tsucc(a);
    std::unique_ptr<std::atomic<uint64_t>[]> cnt(new std::atomic<uint64_t>[n + 1]);
    for (uint64_t i = 0; i <= n; i++) cnt[i].store(0, std::memory_order_relaxed);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t u = b; u < e; u++)
        for (uint64_t k = g->offsets[u]; k < g->offsets[u + 1]; k++) {
          uint32_t v = g->successors[k];
          if (v >= n) fail(HB_E_INVALID_ARG, "successor out of range");
          cnt[v].fetch_add(1, std::memory_order_relaxed);
        }
    });
    for (uint64_t v = 0; v < n; v++) toff[v + 1] = toff[v] + cnt[v].load();
    for (uint64_t v = 0; v < n; v++) cnt[v].store(toff[v], std::memory_order_relaxed);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t u = b; u < e; u++)
        for (uint64_t k = g->offsets[u]; k < g->offsets[u + 1]; k++)
          tsucc[cnt[g->successors[k]].fetch_add(1, std::memory_order_relaxed)] = uint32_t(u);
    });
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; v++) std::sort(tsucc.begin() + toff[v], tsucc.begin() + toff[v + 1]);
    });
    fill_graph(out, n, std::move(toff), std::move(tsucc));
  });
}

hb_status hb_graph_generate_uniform(uint64_t n, uint32_t k, uint64_t seed, hb_graph *out) {
  return guard([&] {
    if (!out) fail(HB_E_INVALID_ARG, "null argument");
    if (n > 0xffffffffull) fail(HB_E_INVALID_ARG, "more than 2^32 nodes");
    if (n > 0 && uint64_t(k) > n - 1) fail(HB_E_INVALID_ARG, "k exceeds n - 1");
    std::vector<uint64_t> off(n + 1);
    for (uint64_t v = 0; v <= n; v++) off[v] = v * k;
    std::vector<uint32_t> succ(n * k);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      std::vector<uint64_t> s;
      for (uint64_t v = b; v < e; v++) {
        Rng rng(seed, v);
        s.clear();
        // Floyd's sampling of k distinct values from [0, n-1), then skip v.
        const uint64_t N = n - 1;
        for (uint64_t j = N - k; j < N; j++) {
          uint64_t t = rng.below(j + 1);
          if (std::find(s.begin(), s.end(), t) != s.end()) s.push_back(j);
          else s.push_back(t);
        }
        for (auto &x : s) x += (x >= v);
        std::sort(s.begin(), s.end());
        for (uint32_t i = 0; i < k; i++) succ[v * k + i] = uint32_t(s[i]);
      }
    });
    fill_graph(out, n, std::move(off), std::move(succ));
  });
}

hb_status hb_graph_generate_web(uint64_t n, double alpha, uint32_t dmin, uint32_t dmax,
                                uint64_t window, double p_window, uint64_t seed, hb_graph *out) {
  return guard([&] {
    if (!out) fail(HB_E_INVALID_ARG, "null argument");
    if (n < 2 || n > 0xffffffffull) fail(HB_E_INVALID_ARG, "n must lie in [2, 2^32]");
    if (dmin < 1 || dmax < dmin) fail(HB_E_INVALID_ARG, "bad degree range");
    if (window < 1) fail(HB_E_INVALID_ARG, "window must be positive");
    // Zipf(alpha) on [dmin, dmax] by inverse CDF.
    std::vector<double> cdf(dmax - dmin + 1);
    double acc = 0;
    for (uint32_t d = dmin; d <= dmax; d++) cdf[d - dmin] = (acc += std::pow(double(d), -alpha));
    for (auto &c : cdf) c /= acc;
    std::vector<uint64_t> deg(n);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; v++) {
        Rng rng(seed, v);
        double u = rng.uniform();
        uint64_t idx = uint64_t(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        deg[v] = dmin + std::min<uint64_t>(idx, dmax - dmin);
      }
    });
    std::vector<uint64_t> slot(n + 1, 0);
    for (uint64_t v = 0; v < n; v++) slot[v + 1] = slot[v] + deg[v];
    const uint64_t total = slot[n];
    // Draw every arc slot.  High bit set: copy the target of an earlier slot.
    const uint64_t COPY = uint64_t(1) << 63;
    std::vector<uint64_t> draw(total);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; v++) {
        Rng rng(seed, v);
        rng.uniform();  // the degree draw
        for (uint64_t k = slot[v]; k < slot[v + 1]; k++) {
          if (rng.uniform() < p_window || k == 0) {
            uint64_t off = rng.below(2 * window) ;  // [0, 2w) -> [-w, -1] U [1, w]
            int64_t delta = off < window ? -int64_t(window - off) : int64_t(off - window + 1);
            int64_t t = (int64_t(v) + delta) % int64_t(n);
            if (t < 0) t += int64_t(n);
            draw[k] = uint64_t(t);
          } else {
            draw[k] = COPY | rng.below(k);
          }
        }
      }
    });
    std::vector<uint32_t> target(total);
    parallel_for(total, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t k = b; k < e; k++) {
        uint64_t j = k;
        while (draw[j] & COPY) j = draw[j] & ~COPY;
        target[k] = uint32_t(draw[j]);
      }
    });
    std::vector<uint64_t>().swap(draw);
    // Sort, deduplicate and drop self loops per row.
    std::vector<uint64_t> keep(n, 0);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; v++) {
        auto first = target.begin() + slot[v], last = target.begin() + slot[v + 1];
        std::sort(first, last);
        auto end = std::unique(first, last);
        end = std::remove(first, end, uint32_t(v));
        keep[v] = uint64_t(end - first);
      }
    });
    std::vector<uint64_t> off(n + 1, 0);
    for (uint64_t v = 0; v < n; v++) off[v + 1] = off[v] + keep[v];
    std::vector<uint32_t> succ(off[n]);
    parallel_for(n, [&](uint64_t b, uint64_t e, unsigned) {
      for (uint64_t v = b; v < e; v++)
        std::copy(target.begin() + slot[v], target.begin() + slot[v] + keep[v], succ.begin() + off[v]);
    });
    fill_graph(out, n, std::move(off), std::move(succ));
  });
}

void hb_graph_free(hb_graph *g) {
  if (!g) return;
  std::free(g->offsets);
  std::free(g->successors);
  g->offsets = nullptr;
  g->successors = nullptr;
  g->num_nodes = g->num_arcs = 0;
}

}  // extern "C"
