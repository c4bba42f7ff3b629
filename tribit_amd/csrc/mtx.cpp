// mtx.cpp -- Matrix Market coordinate reader for the tribit C ABI.
//
// Semantics of read_mtx, src/single_gpu/tribit.cu:228-269:
//   - the first line (banner) is skipped; the qualifier (general/symmetric,
//     pattern/real/...) is ignored: both give the same undirected graph;
//   - '%' lines are skipped; the first other line is "rows cols nnz";
//   - rows != cols is an error (tribit.cu:241-245);
//   - each entry line is "r c [values...]", 1-based; value columns ignored;
//   - self-loops are dropped; every other entry becomes (min<<32)|max;
//   - blank lines between entries are skipped and reading ends at the first
//     other line that does not start with two integers (the reference's
//     `file >> r >> c` loop skips whitespace and stops there too).
// Divergences, on malformed input only: blank lines before the size line
// are skipped; an id outside [1, n] is TRIBIT_ERR_RANGE (reference: UB); an
// entry is one line (the reference would take r and c from two lines).
//
// The file is memory-mapped and the entry region is parsed by several host
// threads, each on a line-aligned slice (the reference parses with one
// iostream thread, its wall-clock bottleneck).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/tribit.h"

namespace tb {

namespace {

inline bool is_sp(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Parses one unsigned decimal at p (after blanks); returns false if none.
inline bool parse_u64(const char*& p, const char* e, uint64_t& out) {
  while (p < e && is_sp(*p)) ++p;
  if (p >= e || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    const uint64_t d = (uint64_t)(*p - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
    ++p;
  }
  out = v;
  return true;
}

inline const char* line_end(const char* p, const char* e) {
  const void* q = memchr(p, '\n', (size_t)(e - p));
  return q ? static_cast<const char*>(q) : e;
}

struct Slice {
  std::vector<uint64_t> keys;
  uint64_t entries = 0;
  const char* stop = nullptr;     // first non-entry line in this slice
  const char* bad_id = nullptr;   // first out-of-range entry in this slice
};

void parse_slice(const char* p, const char* e, uint64_t n, Slice* out) {
  out->keys.reserve((size_t)((e - p) / 12 + 16));
  while (p < e) {
    const char* le = line_end(p, e);
    const char* q = p;
    while (q < le && is_sp(*q)) ++q;
    if (q == le) {  // blank line: `file >> r >> c` skips it
      p = le + 1;
      continue;
    }
    uint64_t r, c;
    if (!parse_u64(q, le, r) || !parse_u64(q, le, c)) {
      out->stop = p;
      return;
    }
    if (r < 1 || c < 1 || r > n || c > n) {
      out->bad_id = p;
      return;
    }
    ++out->entries;
    if (r != c) {
      const uint64_t lo = (r < c ? r : c) - 1, hi = (r < c ? c : r) - 1;
      out->keys.push_back((lo << 32) | hi);
    }
    p = le + 1;
  }
}

}  // namespace

int read_mtx(const char* path, uint64_t** keys_out, uint64_t* nkeys_out, uint64_t* n_out, uint64_t* entries_out) {
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return TRIBIT_ERR_IO;
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return TRIBIT_ERR_IO;
  }
  const size_t size = (size_t)sb.st_size;
  const char* base = nullptr;
  if (size) {
    void* mp = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (mp == MAP_FAILED) {
      close(fd);
      return TRIBIT_ERR_IO;
    }
    base = static_cast<const char*>(mp);
  }
  close(fd);
  const char* e = base + size;
  const char* p = base;
  int st = TRIBIT_OK;
  uint64_t nr = 0, nc = 0, nnz = 0;
  bool have_size = false;
  if (!size) st = TRIBIT_ERR_FORMAT;
  if (st == TRIBIT_OK) {
    p = line_end(p, e);  // banner
    p = (p < e) ? p + 1 : e;
    while (p < e) {
      const char* le = line_end(p, e);
      if (*p == '%') {
        p = le + 1;
        continue;
      }
      const char* q = p;
      while (q < le && (is_sp(*q))) ++q;
      if (q == le) {  // blank line
        p = le + 1;
        continue;
      }
      if (!parse_u64(q, le, nr) || !parse_u64(q, le, nc) || !parse_u64(q, le, nnz)) {
        st = TRIBIT_ERR_FORMAT;
        break;
      }
      have_size = true;
      p = le + 1;
      break;
    }
    if (st == TRIBIT_OK && !have_size) st = TRIBIT_ERR_FORMAT;
    if (st == TRIBIT_OK && nr != nc) st = TRIBIT_ERR_SHAPE;
    if (st == TRIBIT_OK && nr > 0xFFFFFFFEull) st = TRIBIT_ERR_RANGE;
  }
  std::vector<Slice> slices;
  if (st == TRIBIT_OK && p < e) {
    const size_t bytes = (size_t)(e - p);
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 32) nt = 32;
    if (bytes < (size_t)(1u << 20)) nt = 1;
    std::vector<const char*> cut(nt + 1);
    cut[0] = p;
    cut[nt] = e;
    for (unsigned t = 1; t < nt; ++t) {
      const char* c = p + bytes * t / nt;
      if (c < cut[t - 1]) c = cut[t - 1];
      c = line_end(c, e);
      cut[t] = (c < e) ? c + 1 : e;
    }
    slices.resize(nt);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back(parse_slice, cut[t], cut[t + 1], nr, &slices[t]);
    for (auto& x : th) x.join();
    // The read ends at the first slice that stopped; an id error counts only
    // if it precedes that stop.
    for (auto& s : slices) {
      if (s.bad_id) {
        st = TRIBIT_ERR_RANGE;
        break;
      }
      if (s.stop) break;
    }
  }
  uint64_t total = 0, entries = 0;
  if (st == TRIBIT_OK) {
    for (auto& s : slices) {
      total += s.keys.size();
      entries += s.entries;
      if (s.stop) break;
    }
    uint64_t* keys = static_cast<uint64_t*>(malloc((total ? total : 1) * sizeof(uint64_t)));
    if (!keys) {
      st = TRIBIT_ERR_NOMEM;
    } else {
      uint64_t o = 0;
      for (auto& s : slices) {
        if (!s.keys.empty()) memcpy(keys + o, s.keys.data(), s.keys.size() * sizeof(uint64_t));
        o += s.keys.size();
        if (s.stop) break;
      }
      *keys_out = keys;
      *nkeys_out = total;
      *n_out = nr;
      if (entries_out) *entries_out = entries;
    }
  }
  if (base) munmap(const_cast<char*>(base), size);
  (void)nnz;
  return st;
}

}  // namespace tb
