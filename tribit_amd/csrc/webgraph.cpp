// webgraph.cpp -- sequential reader of BVGraph (WebGraph) compressed graphs.
//
// The reference's multi-GPU host loads its inputs with the webgraph crate
// (webgraph = "0.6", src/multi_gpu/Cargo.toml; BvGraphSeq in
// src/multi_gpu/src/main.rs:88) and turns every arc (u, v), u != v, into the
// key (min << 32) | max (main.rs:150-161). This reads the same files,
// <basename>.properties and <basename>.graph, and returns those keys.
//
// BVGraph format (Boldi & Vigna, "The WebGraph Framework I: Compression
// Techniques", WWW 2004; instantaneous codes from Boldi & Vigna, "Codes for
// the World-Wide Web", Internet Math. 2005):
//   - bits are read most significant first, bytes in file order;
//   - unary(n): n zeros, then a one; gamma(n): unary(b), then the b low bits
//     of n + 1, b = floor(log2(n + 1)); delta(n): gamma(b), then the same b
//     bits; zeta_k(n): unary(h) with n + 1 in [2^(hk), 2^((h+1)k)), then
//     n + 1 - 2^(hk) in minimal binary over that interval (hk + k - 1 bits,
//     one more for the upper part of the range);
//   - per node x in order: its outdegree d; if d > 0 and windowsize > 0, a
//     reference r <= windowsize (r > 0: the successors of node x - r are the
//     reference list); with a reference, the number of copy blocks and their
//     lengths (the first as is, the others minus one), blocks alternating
//     copy / skip over the reference list starting with copy, the implicit
//     last block copying when the count is even (so a count of 0 copies the
//     whole list); then, if successors remain and minintervallength > 0, the
//     number of intervals, each a left extreme (the first as the signed
//     offset from x, the others as the gap from the previous interval's end
//     minus one) and a length minus minintervallength; then the remaining
//     successors (residuals), the first as the signed offset from x, the
//     others as gaps minus one. Signed offsets v are coded as 2v (v >= 0) or
//     -2v - 1 (v < 0). The successors of x are the merge of the copied,
//     interval and residual lists.
//   - codes per field come from the properties' compressionflags (defaults:
//     gamma outdegrees, unary references, gamma block counts, blocks and
//     intervals, zeta_k residuals with k = zetak).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tribit.h"

namespace tb {

namespace {

enum Code { kGamma, kDelta, kUnary, kZeta };

struct Props {
  uint64_t nodes = 0, arcs = 0;
  int window = 7, maxref = 3, min_interval = 4, zetak = 3;
  Code outdeg = kGamma, refs = kUnary, block_count = kGamma, blocks = kGamma, intervals = kGamma, residuals = kZeta;
};

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && isspace((unsigned char)s[a])) ++a;
  while (b > a && isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool parse_int(const std::string& v, uint64_t* out) {
  if (v.empty()) return false;
  uint64_t x = 0;
  for (char c : v) {
    if (c < '0' || c > '9') return false;
    if (x > (UINT64_MAX - 9) / 10) return false;
    x = x * 10 + (uint64_t)(c - '0');
  }
  *out = x;
  return true;
}

// Java-style properties: "key=value" or "key: value", '#' and '!' comments.
int read_props(const std::string& path, Props* p) {
  std::ifstream f(path);
  if (!f) return TRIBIT_ERR_IO;
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#' || line[0] == '!') continue;
    const size_t eq = line.find_first_of("=:");
    if (eq == std::string::npos) continue;
    kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  uint64_t v = 0;
  if (!kv.count("nodes") || !parse_int(kv["nodes"], &p->nodes)) return TRIBIT_ERR_FORMAT;
  if (kv.count("arcs") && !parse_int(kv["arcs"], &p->arcs)) return TRIBIT_ERR_FORMAT;
  if (kv.count("version") && kv["version"] != "0") return TRIBIT_ERR_FORMAT;
  if (kv.count("endianness") && kv["endianness"] != "big" && kv["endianness"] != "BE") return TRIBIT_ERR_FORMAT;
  struct {
    const char* key;
    int* dst;
  } ints[] = {{"windowsize", &p->window}, {"maxrefcount", &p->maxref}, {"minintervallength", &p->min_interval},
              {"zetak", &p->zetak}};
  for (auto& e : ints)
    if (kv.count(e.key)) {
      if (!parse_int(kv[e.key], &v) || v > 1u << 20) return TRIBIT_ERR_FORMAT;
      *e.dst = (int)v;
    }
  if (p->zetak < 1 || p->zetak > 16) return TRIBIT_ERR_FORMAT;
  if (kv.count("compressionflags")) {
    std::stringstream ss(kv["compressionflags"]);
    std::string tok;
    while (std::getline(ss, tok, '|')) {
      tok = trim(tok);
      if (tok.empty()) continue;
      const size_t us = tok.rfind('_');
      if (us == std::string::npos) return TRIBIT_ERR_FORMAT;
      const std::string field = tok.substr(0, us), code = tok.substr(us + 1);
      Code c;
      if (code == "GAMMA")
        c = kGamma;
      else if (code == "DELTA")
        c = kDelta;
      else if (code == "UNARY")
        c = kUnary;
      else if (code == "ZETA")
        c = kZeta;
      else
        return TRIBIT_ERR_FORMAT;  // e.g. nibble codes: not supported
      if (field == "OUTDEGREES")
        p->outdeg = c;
      else if (field == "REFERENCES")
        p->refs = c;
      else if (field == "BLOCK_COUNT")
        p->block_count = c;
      else if (field == "BLOCKS")
        p->blocks = c;
      else if (field == "INTERVALS")
        p->intervals = c;
      else if (field == "RESIDUALS")
        p->residuals = c;
      else if (field != "OFFSETS")  // offsets are for random access; a sequential read ignores them
        return TRIBIT_ERR_FORMAT;
    }
  }
  return TRIBIT_OK;
}

// Most-significant-bit-first reader over a byte buffer; reads past the end
// see zeros and set `over`.
struct Bits {
  const uint8_t* data = nullptr;
  uint64_t nbytes = 0, pos = 0;
  bool over = false;

  uint64_t peek64() const {
    const uint64_t byte = pos >> 3;
    const unsigned sh = (unsigned)(pos & 7);
    uint64_t w = 0;
    if (byte + 9 <= nbytes) {
      memcpy(&w, data + byte, 8);
      w = __builtin_bswap64(w);
      if (sh) w = (w << sh) | (uint64_t)(data[byte + 8] >> (8 - sh));
      return w;
    }
    for (int i = 0; i < 9; ++i) {
      const uint64_t b = byte + (uint64_t)i < nbytes ? data[byte + i] : 0u;
      if (i < 8)
        w = (w << 8) | b;
      else if (sh)
        w = (w << sh) | (b >> (8 - sh));
    }
    return w;
  }
  void advance(uint64_t n) {
    pos += n;
    if (pos > nbytes * 8) over = true;
  }
  uint64_t read(unsigned n) {  // n <= 63
    if (!n) return 0;
    const uint64_t v = peek64() >> (64 - n);
    advance(n);
    return v;
  }
  uint64_t unary() {
    uint64_t n = 0;
    for (;;) {
      if (pos >= nbytes * 8) {
        over = true;
        return n;
      }
      const uint64_t w = peek64();
      if (w) {
        const unsigned z = (unsigned)__builtin_clzll(w);
        n += z;
        advance(z + 1);
        return n;
      }
      n += 64;
      advance(64);
    }
  }
  uint64_t gamma() {
    const uint64_t b = unary();
    if (b > 62) {
      over = true;
      return 0;
    }
    return ((1ull << b) | read((unsigned)b)) - 1;
  }
  uint64_t delta() {
    const uint64_t b = gamma();
    if (b > 62) {
      over = true;
      return 0;
    }
    return ((1ull << b) | read((unsigned)b)) - 1;
  }
  uint64_t zeta(int k) {
    const uint64_t h = unary();
    if (h * (uint64_t)k + (uint64_t)k > 62) {
      over = true;
      return 0;
    }
    const uint64_t left = 1ull << (h * k);
    const uint64_t m = read((unsigned)(h * k + k - 1));
    if (m < left) return m + left - 1;
    return (m << 1) + read(1) - 1;
  }
  uint64_t code(Code c, int k) {
    switch (c) {
      case kGamma: return gamma();
      case kDelta: return delta();
      case kUnary: return unary();
      default: return zeta(k);
    }
  }
};

inline int64_t nat2int(uint64_t v) { return (v & 1) ? -(int64_t)((v + 1) >> 1) : (int64_t)(v >> 1); }

}  // namespace

int read_webgraph(const char* basename, uint64_t** keys_out, uint64_t* nkeys_out, uint64_t* n_out) {
  Props pr;
  const std::string base(basename);
  int st = read_props(base + ".properties", &pr);
  if (st != TRIBIT_OK) return st;
  if (pr.nodes > 0xFFFFFFFFull) return TRIBIT_ERR_TOO_LARGE;
  const int fd = open((base + ".graph").c_str(), O_RDONLY);
  if (fd < 0) return TRIBIT_ERR_IO;
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return TRIBIT_ERR_IO;
  }
  const uint64_t size = (uint64_t)sb.st_size;
  void* map = size ? mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
  close(fd);
  if (size && map == MAP_FAILED) return TRIBIT_ERR_IO;
  Bits in;
  in.data = static_cast<const uint8_t*>(map);
  in.nbytes = size;

  const uint64_t n = pr.nodes;
  const int W = pr.window;
  std::vector<std::vector<uint32_t>> win((size_t)W + 1);  // successors of the last W + 1 nodes, by x mod (W + 1)
  std::vector<uint64_t> keys;
  keys.reserve(pr.arcs ? (size_t)std::min<uint64_t>(pr.arcs, 1ull << 31) : 1024);
  std::vector<uint32_t> copied, extra, merged;
  for (uint64_t x = 0; x < n && st == TRIBIT_OK; ++x) {
    const uint64_t d = in.code(pr.outdeg, pr.zetak);
    copied.clear();
    extra.clear();
    if (d > n) {
      st = TRIBIT_ERR_FORMAT;
      break;
    }
    if (d > 0) {
      uint64_t r = W > 0 ? in.code(pr.refs, pr.zetak) : 0;
      if (r > (uint64_t)W || r > x) {
        st = TRIBIT_ERR_FORMAT;
        break;
      }
      if (r > 0) {
        const std::vector<uint32_t>& ref = win[(size_t)((x - r) % (uint64_t)(W + 1))];
        const uint64_t nb = in.code(pr.block_count, pr.zetak);
        uint64_t at = 0;
        bool copy = true;
        for (uint64_t i = 0; i < nb && st == TRIBIT_OK; ++i) {
          const uint64_t len = in.code(pr.blocks, pr.zetak) + (i ? 1 : 0);
          if (in.over || len > ref.size() || at + len > ref.size()) {
            st = TRIBIT_ERR_FORMAT;
            break;
          }
          if (copy) copied.insert(copied.end(), ref.begin() + (ptrdiff_t)at, ref.begin() + (ptrdiff_t)(at + len));
          at += len;
          copy = !copy;
        }
        if (st != TRIBIT_OK) break;
        if (nb % 2 == 0) copied.insert(copied.end(), ref.begin() + (ptrdiff_t)at, ref.end());
        if (copied.size() > d) {
          st = TRIBIT_ERR_FORMAT;
          break;
        }
      }
      uint64_t left_over = d - copied.size();
      if (left_over > 0 && pr.min_interval > 0) {
        const uint64_t ni = in.code(pr.intervals, pr.zetak);
        int64_t prev = 0;  // one past the previous interval's end
        for (uint64_t i = 0; i < ni; ++i) {
          // every decoded value is checked against n before it enters a sum
          const uint64_t g = in.code(pr.intervals, pr.zetak);
          if (g > 2 * n) {
            st = TRIBIT_ERR_FORMAT;
            break;
          }
          int64_t left;
          if (i == 0)
            left = (int64_t)x + nat2int(g);
          else
            left = prev + 1 + (int64_t)g;  // intervals are at least one id apart
          const uint64_t len = in.code(pr.intervals, pr.zetak) + (uint64_t)pr.min_interval;
          if (in.over || left < 0 || len > n || (uint64_t)left + len > n || len > left_over) {
            st = TRIBIT_ERR_FORMAT;
            break;
          }
          for (uint64_t j = 0; j < len; ++j) extra.push_back((uint32_t)((uint64_t)left + j));
          prev = left + (int64_t)len;
          left_over -= len;
        }
      }
      if (st != TRIBIT_OK) break;
      int64_t prev = 0;  // the previous residual
      for (uint64_t i = 0; i < left_over; ++i) {
        const uint64_t g = in.code(pr.residuals, pr.zetak);
        int64_t v = -1;
        if (g <= 2 * n) v = i == 0 ? (int64_t)x + nat2int(g) : prev + 1 + (int64_t)g;
        if (in.over || v < 0 || (uint64_t)v >= n) {
          st = TRIBIT_ERR_FORMAT;
          break;
        }
        extra.push_back((uint32_t)v);
        prev = v;
      }
      if (st != TRIBIT_OK) break;
    }
    if (in.over) {
      st = TRIBIT_ERR_FORMAT;  // truncated bit stream
      break;
    }
    // successors = merge of the copied list and the (sorted) intervals and residuals
    std::sort(extra.begin(), extra.end());
    merged.resize(copied.size() + extra.size());
    std::merge(copied.begin(), copied.end(), extra.begin(), extra.end(), merged.begin());
    for (uint32_t y : merged)
      if (y != x) {
        const uint64_t a = std::min<uint64_t>(x, y), b = std::max<uint64_t>(x, y);
        keys.push_back((a << 32) | b);
      }
    if (W > 0) win[(size_t)(x % (uint64_t)(W + 1))].swap(merged);
  }
  if (map) munmap(map, size);
  if (st != TRIBIT_OK) return st;
  uint64_t* out = static_cast<uint64_t*>(malloc((keys.size() ? keys.size() : 1) * sizeof(uint64_t)));
  if (!out) return TRIBIT_ERR_NOMEM;
  if (!keys.empty()) memcpy(out, keys.data(), keys.size() * sizeof(uint64_t));
  *keys_out = out;
  *nkeys_out = keys.size();
  *n_out = n;
  return TRIBIT_OK;
}

}  // namespace tb
