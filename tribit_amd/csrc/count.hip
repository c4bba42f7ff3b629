// count.hip -- scheduler + counting kernels (gfx950, wave64).
//
// Replaces compute_binary_popcount_mma_9 and its launch,
// src/single_gpu/tribit.cu:519-586 and :765-802. The reference forms, per
// (U word, 32-neighbour window), one b1 tensor-core AND+POPC; here the
// count is the degree-oriented edge iterator
//     triangles = sum over u, v in N+(u) of |N+(u) & N+(v)|
// where, for each oriented edge u->v, one of the two lists is staged as an
// LDS set (the task "owner" x) and the other is streamed from HBM with
// coalesced int32 loads, one branch-free LDS probe per element. Either
// direction counts the same triangles; the schedule picks it per edge.
//
// The owner's outdeg d picks the kernel variant (class):
//   class 0: d <= 32      wave task,  cuckoo set 2 x 64 slots
//   class 1: d <= 256     wave task,  cuckoo set 2 x 512 slots
//   class 2: d <= 2048    block task, cuckoo set 2 x 4096 slots
//   class 3: d >  2048    block task per 2048-id chunk of N+(x); each
//                         streamed range is cut to the chunk's id interval.
// and the owner's streamed work picks how many tasks its lists split into.
#include <cstring>
#include <algorithm>
#include <chrono>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kBlock = 256;
// Every kernel takes its tasks from the END of its list. Lists hold owners in
// id order, and with degree-rank ids the highest ids are the hubs, whose
// tasks are the longest: taken first, they overlap with the many short
// tasks instead of forming the launch's tail.
constexpr uint32_t kWaveBatch = 16;      // most wave tasks taken per queue atomic
constexpr uint32_t kClaimsPerWave = 8;   // queue claims per wave of a wave launch (when tasks allow)
constexpr uint32_t kChunk = 2048;        // probe-set chunk of class-3 rows
constexpr uint64_t kWaveCap = 1u << 15;  // target streamed elements per wave task
constexpr uint64_t kBlockCap = 1u << 18; // target streamed elements per block task
constexpr uint32_t kWaveListCap = 1024;   // most lists per wave task
constexpr uint32_t kBucketShift = 21;   // task.w = chunk | bucket << 21
constexpr uint32_t kChunkMask = (1u << kBucketShift) - 1;
// Chunk field of a cursor task: one block task takes every chunk of its
// owner's row in turn, each of its lists keeping a cursor (at most kChunk
// lists per task, staged once), instead of one task per chunk whose lists are
// each cut to the chunk by two binary searches.
constexpr uint32_t kAllChunks = kChunkMask;
constexpr uint64_t kCursorCap = 1u << 20;  // target streamed elements per cursor task (2^20: 62.8 ms for the s25 rank
                                            // off-diagonal task, 2^22: 66.4, 2^24: 64.0)

// ------------------------------------------------------------ LDS sets
// N+(u) (or one id-range chunk of it) is held in LDS as a two-table cuckoo
// set: tables [0,S) and [S,2S), key x lives at h_a(x) or S + h_b(x), so a
// membership probe is two LDS reads and two compares, with no data-dependent
// loop (every lane of a wave does the same work). Tables are filled to at
// most 1/4 of their 2S slots. A build that fails (an insertion chain longer
// than kCuckooIters) is retried with the next multiplier; after
// kCuckooSeeds failures the task falls back to a linear-probing table in the
// same 2S slots.

constexpr uint32_t kNoKey = 0xFFFFFFFEu;  // never a vertex id (n <= 2^32-2) and never kEmpty
// Range entries (begin, end | flags): for wedge ranges of graphs with
// m < 2^30, bit 31 of the end marks a narrow range, whose ids share their high
// 16 bits or (bit 30 too) cross one 64K-id boundary, at a position kept in
// sfx_split; block tasks stream narrow ranges from the 16-bit low halves
// (col16). range_end_mask(m) strips the flags.
constexpr uint32_t kNarrowFlag = 0x80000000u;
constexpr uint32_t kTwoWinFlag = 0x40000000u;
constexpr int kNoBitmapBit = 1 << 9;       // kernel mode bit: block tasks never use bitmap sets
constexpr int kWideBit = 1 << 10;          // launch-time flag: narrow ranges streamed as 32-bit ids too
constexpr int kCuckooIters = 64;
constexpr int kCuckooSeeds = 4;
__device__ __constant__ uint32_t kSeedA[kCuckooSeeds] = {0x9E3779B1u, 0x85EBCA77u, 0xC2B2AE3Du, 0x27D4EB2Fu};

// Both slots of x come from one product h = x * a (one 32-bit multiply per
// probe): table A takes the top LOGS bits of h, table B the next LOGS bits
// XOR the top ones. The next bits alone depend only on the low bits of x, and
// R-MAT hubs (ids with few set bits) often share those, for every multiplier;
// the XOR makes both slots depend on every bit of x.
template <int LOGS>
__device__ __forceinline__ uint32_t slot_a(uint32_t h) {
  return h >> (32 - LOGS);
}
template <int LOGS>
__device__ __forceinline__ uint32_t slot_b(uint32_t h) {
  return (1u << LOGS) + (((h >> (32 - 2 * LOGS)) ^ (h >> (32 - LOGS))) & ((1u << LOGS) - 1));
}

template <int LOGS>
__device__ __forceinline__ uint32_t hmul(uint32_t x, uint32_t a) {
  return (x * a) >> (32 - LOGS);
}

template <int LOGS>
__device__ __forceinline__ bool cuckoo_insert(uint32_t* T, uint32_t x, uint32_t a) {
  uint32_t pos = slot_a<LOGS>(x * a);
  for (int it = 0; it < kCuckooIters; ++it) {
    const uint32_t y = atomicExch(&T[pos], x);
    if (y == kEmpty) return true;
    // y sat in one of its two slots; it moves to the other one
    x = y;
    const uint32_t h = x * a;
    const uint32_t pa = slot_a<LOGS>(h);
    pos = (pos == pa) ? slot_b<LOGS>(h) : pa;
  }
  return false;
}

template <int LOGS>
struct CuckooProbe {
  static constexpr bool kZeroSafe = false;  // id 0 may be in the set
  static constexpr bool kRelative = false;
  const uint32_t* T;
  uint32_t a;
  __device__ __forceinline__ uint32_t operator()(uint32_t x, uint32_t = 0) const {
    const uint32_t h = x * a;
    return (uint32_t)((T[slot_a<LOGS>(h)] == x) | (T[slot_b<LOGS>(h)] == x));
  }
};

// Several cuckoo sets of 2S slots side by side (multi-task waves): the
// selector picks the set of the range being streamed.
template <int LOGS>
struct MultiCuckooProbe {
  static constexpr bool kZeroSafe = false;
  static constexpr bool kRelative = false;
  const uint32_t* T;
  uint32_t a;
  __device__ __forceinline__ uint32_t operator()(uint32_t x, uint32_t sel) const {
    const uint32_t* t = T + (sel << (LOGS + 1));
    const uint32_t h = x * a;
    return (uint32_t)((t[slot_a<LOGS>(h)] == x) | (t[slot_b<LOGS>(h)] == x));
  }
};

// Linear probing over all 2S slots (fallback only).
template <int LOGS2>
__device__ __forceinline__ void linear_insert(uint32_t* T, uint32_t w) {
  uint32_t h = hmul<LOGS2>(w, 0x9E3779B1u);
  while (atomicCAS(&T[h], kEmpty, w) != kEmpty) h = (h + 1) & ((1u << LOGS2) - 1);
}

template <int LOGS2>
struct LinearProbe {
  static constexpr bool kZeroSafe = false;
  static constexpr bool kRelative = false;
  const uint32_t* T;
  __device__ __forceinline__ uint32_t operator()(uint32_t w, uint32_t sel = 0) const {
    const uint32_t* t = T + (sel << LOGS2);  // sets of 2^LOGS2 slots side by side (selector 0: one set)
    uint32_t h = hmul<LOGS2>(w, 0x9E3779B1u);
    for (;;) {
      const uint32_t k = t[h];
      if (k == w) return 1u;
      if (k == kEmpty) return 0u;
      h = (h + 1) & ((1u << LOGS2) - 1);
    }
  }
};

// Bitmap over the id window [base, base + 32 * words): one LDS read per
// probe. Block tasks use it instead of the cuckoo set when the owner's ids
// fit the set's 2S words (2^18 ids for S = 4096), as they do for almost all
// owners once ids follow the degree rank. kNoKey and ids below base wrap to
// indexes past the window; every such index is clamped to the zero word that
// follows the window, so a probe is one unconditional LDS read (a branch
// around the read would serialize the probes of a step). Block tasks use the
// bitmap only for sets without id 0, so the 0 that a masked load returns
// misses as well (kZeroSafe).
template <int LOGW>
struct BitmapProbe {
  static constexpr bool kZeroSafe = true;
  static constexpr bool kRelative = true;  // also probes by window index i = x - base (rel)
  const uint32_t* T;  // 2^LOGW words, then a zero word
  uint32_t base;
  __device__ __forceinline__ uint32_t rel(uint32_t i) const {
    const uint32_t w = min(i >> 5, 1u << LOGW);
    return __builtin_amdgcn_ubfe(T[w], i, 1u);  // the bit field offset is i mod 32
  }
  __device__ __forceinline__ uint32_t operator()(uint32_t x, uint32_t = 0) const { return rel(x - base); }
};

template <int G>
__host__ __device__ constexpr uint64_t leader_mask() {
  uint64_t m = 0;
  for (int i = 0; i < 64; i += G) m |= 1ull << i;
  return m;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// --------------------------------------------------------- range streaming
// The lists to stream are staged in LDS as ranges R[k] = [begin, end) of col.
// 64/G groups of G lanes each own one range at a time; per step a group
// covers G*U consecutive elements, each lane issuing U independent loads
// before probing (U loads in flight per lane). With BUF the loads are raw
// buffer loads: one 32-bit byte offset per lane, the U loads differ only by
// an immediate offset, no branches (an offset past the buffer reads 0 and is
// masked anyway). A group whose range is exhausted takes the next index:
// wave-private counter, or, for block tasks (SHARED), an LDS counter shared
// by the block's waves.

// Raw buffer loads address col with 32-bit byte offsets: the buffer path is
// taken when the streamed array's bytes stay at or below kBufMax, and a
// masked slot uses offset kBufOob (plus at most 4 KiB of immediate offset),
// past every such buffer.
constexpr uint64_t kBufMax = 0xFFFFE000ull;
constexpr uint32_t kBufOob = 0xFFFFE000u;

struct ColSrc {
  const uint32_t* col;
  __amdgpu_buffer_rsrc_t rsrc;    // raw buffer over col
  __amdgpu_buffer_rsrc_t rsrc16;  // raw buffer over col16 (narrow ranges), when given
  uint32_t emask;                 // range_end_mask(m)
};

__device__ __forceinline__ ColSrc make_col_src(const uint32_t* col, uint64_t m, const uint16_t* col16 = nullptr) {
  ColSrc c;
  c.col = col;
  const uint64_t bytes = m * 4 < kBufMax ? m * 4 : kBufMax;
  c.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(col), (short)0, (int)(uint32_t)bytes, 0x00020000);
  const uint64_t bytes16 = col16 ? 4 * ((m + 1) / 2 + 1) : 0;  // col16 is padded to whole dwords plus one
  c.rsrc16 = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(col16 ? col16 : reinterpret_cast<const uint16_t*>(col)),
                                               (short)0, (int)(uint32_t)(bytes16 < kBufMax ? bytes16 : kBufMax),
                                               0x00020000);
  c.emask = range_end_mask(m);
  return c;
}

// ATTR (multi-task waves of a batch): the hits of a range go to s_acc[its
// set] when the group leaves the range (an LDS atomic per lane with hits),
// so that every task's count can be added to its own graph; the return value
// is then 0.
template <int G, int U, bool SHARED, bool BUF, bool ATTR = false, typename Probe>
__device__ __forceinline__ uint32_t stream_ranges(const ColSrc& src, const uint2* R, uint32_t nk, uint32_t* s_next,
                                                  uint32_t lane, const Probe& probe, const uint8_t* Rsel = nullptr,
                                                  unsigned long long* s_acc = nullptr) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint32_t base = 0, end = 0, cnt = 0, next = 0, sel = 0;
  for (;;) {
    const bool need = base >= end;  // uniform within a group
    const uint64_t want = __ballot(need && gl == 0);
    const uint32_t nw = (uint32_t)__popcll(want);
    uint32_t kb = 0;
    if (nw) {
      if (SHARED) {
        if (lane == 0) kb = atomicAdd(s_next, nw);
        kb = __shfl(kb, 0);
      } else {
        kb = next;
        next = min(next + nw, nk);
      }
      if (want == kLeads && kb >= nk) break;
    }
    if (need) {
      if (ATTR && cnt) {  // the range just left
        atomicAdd(&s_acc[sel], (unsigned long long)cnt);
        cnt = 0;
      }
      const uint32_t k = kb + (uint32_t)__popcll(want & below);
      if (k < nk) {
        const uint2 r = R[k];
        base = r.x;
        end = r.y & src.emask;
        if (Rsel) sel = Rsel[k];
      }
    }
    const uint32_t p0 = base + gl;
    const int rem = (int)(end - p0);  // elements left for this lane's first load (may be <= 0)
    uint32_t x[U];
    if (BUF) {
      // an invalid element gets an offset past the buffer (kBufOob plus the
      // slot's constant offset stays below 2^32): no memory access, reads 0.
      // The slot's offset j G 4 goes in the instruction's scalar offset, so a
      // slot costs one compare and one select.
      const uint32_t off = p0 * 4u;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = j * G < rem;
        x[j] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, (int)(ok ? off : kBufOob), j * G * 4, 0);
        if (!Probe::kZeroSafe) x[j] = ok ? x[j] : kNoKey;
      }
      __builtin_amdgcn_sched_barrier(0);  // all U loads issue before the first probe waits
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) x[j] = (j * G < rem) ? src.col[p0 + j * G] : kNoKey;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) cnt += probe(x[j], sel);
    base += (uint32_t)(G * U);
  }
  if (ATTR) {
    if (cnt) atomicAdd(&s_acc[sel], (unsigned long long)cnt);
    return 0;
  }
  return cnt;
}

// Cursor streaming (block tasks, SHARED hand-out) for an owner whose row is
// counted one chunk at a time: range k runs from its cursor R[k].x up to its
// first element >= bound (the next chunk's first id), and the cursor is left
// there for the next chunk. The lists are sorted, so the elements below the
// bound in a step are a prefix of it; elements at or past the bound are not in
// this chunk's set, so probing them adds nothing.
template <int G, int U, bool BUF, typename Probe>
__device__ __forceinline__ uint32_t stream_bounded(const ColSrc& src, uint2* R, uint32_t nk, uint32_t* s_next,
                                                   uint32_t lane, const Probe& probe, uint32_t bound) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint32_t base = 0, end = 0, cnt = 0, k = 0;
  for (;;) {
    const bool need = base >= end;  // uniform within a group
    const uint64_t want = __ballot(need && gl == 0);
    const uint32_t nw = (uint32_t)__popcll(want);
    uint32_t kb = 0;
    if (nw) {
      if (lane == 0) kb = atomicAdd(s_next, nw);
      kb = __shfl(kb, 0);
      if (want == kLeads && kb >= nk) break;
    }
    if (need) {
      k = kb + (uint32_t)__popcll(want & below);
      base = end = 0;
      if (k < nk) {
        const uint2 r = R[k];
        base = r.x;
        end = r.y;
      }
    }
    const uint32_t p0 = base + gl;
    const int rem = (int)(end - p0);
    uint32_t x[U];
    if (BUF) {
      const uint32_t off = p0 * 4u;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = j * G < rem;
        x[j] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, (int)(ok ? off : kBufOob), j * G * 4, 0);
        if (!Probe::kZeroSafe) x[j] = ok ? x[j] : kNoKey;
      }
      __builtin_amdgcn_sched_barrier(0);
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) x[j] = (j * G < rem) ? src.col[p0 + j * G] : kNoKey;
    }
    uint32_t under = 0;  // this lane's valid elements below the bound
#pragma unroll
    for (int j = 0; j < U; ++j) {
      cnt += probe(x[j], 0);
      under += (j * G < rem && x[j] < bound) ? 1u : 0u;
    }
#pragma unroll
    for (int o = 1; o < G; o <<= 1) under += (uint32_t)__shfl_xor((int)under, o);  // group sum
    if (base < end) {
      const uint32_t step = min((uint32_t)(G * U), end - base);
      if (under < step) {  // the bound lies inside this step
        if (gl == 0) R[k].x = base + under;
        base = end;
      } else {
        base += (uint32_t)(G * U);
        if (base >= end && gl == 0) R[k].x = end;
      }
    }
  }
  return cnt;
}

// Narrow ranges (ids sharing their high 16 bits, flagged by the range
// builder) streamed from col16, the low halves of all ids: one 32-bit load
// returns two consecutive ids, the range's high half (staged in Rh) is or-ed
// back. A slot of lane gl covers dword q = qb + gl + j G, positions 2q and
// 2q + 1; the range [b, e) may start at an odd position (the first half of
// its first dword lies outside it) and end at one. Block tasks only (a
// block-shared counter hands out the ranges).
template <int G, int U, typename Probe>
__device__ __forceinline__ uint32_t stream_narrow(const ColSrc& src, const uint2* R, const uint16_t* Rh, uint32_t nk,
                                                  uint32_t* s_next, uint32_t lane, const Probe& probe) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint32_t qb = 0, b = 0, e = 0, hi = 0, cnt = 0;
  for (;;) {
    const bool need = 2 * qb >= e;  // uniform within a group
    const uint64_t want = __ballot(need && gl == 0);
    const uint32_t nw = (uint32_t)__popcll(want);
    uint32_t kb = 0;
    if (nw) {
      if (lane == 0) kb = atomicAdd(s_next, nw);
      kb = __shfl(kb, 0);
      if (want == kLeads && kb >= nk) break;
    }
    if (need) {
      const uint32_t k = kb + (uint32_t)__popcll(want & below);
      if (k < nk) {
        const uint2 r = R[k];
        b = r.x;
        e = r.y;
        qb = b >> 1;
        hi = (uint32_t)Rh[k] << 16;
      }
    }
    const uint32_t q0 = qb + gl;
    const int rem = (int)(e - 2 * q0);  // positions left from this lane's first dword (may be <= 0)
    const bool skip0 = 2 * q0 < b;      // first dword of a range starting at an odd position
    const uint32_t off = 4 * q0;
    uint32_t w[U];
#pragma unroll
    for (int j = 0; j < U; ++j)
      w[j] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc16, (int)(2 * j * G < rem ? off : kBufOob), j * G * 4, 0);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (Probe::kRelative) {
      // window indexes hi + half - base: one add per id; a masked slot or
      // half gets an offset that puts it past the window
      const uint32_t dh = hi - probe.base;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const uint32_t d = 2 * j * G < rem ? dh : 0xFFFF0000u;
        uint32_t i0 = (w[j] & 0xFFFFu) + d;
        uint32_t i1 = (w[j] >> 16) + d;
        if (j == 0) i0 = skip0 ? 0xFFFFFFFFu : i0;
        i1 = 2 * j * G + 1 < rem ? i1 : 0xFFFFFFFFu;
        cnt += probe.rel(i0) + probe.rel(i1);
      }
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = 2 * j * G < rem;  // a masked slot read 0: both halves become 0 (or kNoKey)
        const uint32_t h = ok ? hi : 0u;
        uint32_t x0 = (w[j] & 0xFFFFu) | h;
        uint32_t x1 = (w[j] >> 16) | h;
        if (j == 0) x0 = skip0 ? kNoKey : x0;
        x1 = 2 * j * G + 1 < rem ? x1 : kNoKey;
        if (!Probe::kZeroSafe) x0 = ok ? x0 : kNoKey;
        cnt += probe(x0) + probe(x1);
      }
    }
    qb += (uint32_t)(G * U);
  }
  return cnt;
}

// Narrow ranges crossing one 64K-id boundary: staged as (begin, split << 16 |
// length), ids before begin + split have high half hi, the others hi + 1.
// Same slots as stream_narrow; each half picks its high half by comparing its
// position with the split (t = split position - the lane's first position).
template <int G, int U, typename Probe>
__device__ __forceinline__ uint32_t stream_narrow2(const ColSrc& src, const uint2* R, const uint16_t* Rh, uint32_t nk,
                                                   uint32_t* s_next, uint32_t lane, const Probe& probe) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint32_t qb = 0, b = 0, e = 0, p = 0, hi = 0, cnt = 0;
  for (;;) {
    const bool need = 2 * qb >= e;
    const uint64_t want = __ballot(need && gl == 0);
    const uint32_t nw = (uint32_t)__popcll(want);
    uint32_t kb = 0;
    if (nw) {
      if (lane == 0) kb = atomicAdd(s_next, nw);
      kb = __shfl(kb, 0);
      if (want == kLeads && kb >= nk) break;
    }
    if (need) {
      const uint32_t k = kb + (uint32_t)__popcll(want & below);
      if (k < nk) {
        const uint2 r = R[k];
        b = r.x;
        e = b + (r.y & 0xFFFFu);
        p = b + (r.y >> 16);
        qb = b >> 1;
        hi = (uint32_t)Rh[k] << 16;
      }
    }
    const uint32_t q0 = qb + gl;
    const int rem = (int)(e - 2 * q0);
    const int t = (int)(p - 2 * q0);  // slot j's first half is past the split iff 2 j G >= t
    const bool skip0 = 2 * q0 < b;
    const uint32_t off = 4 * q0;
    uint32_t w[U];
#pragma unroll
    for (int j = 0; j < U; ++j)
      w[j] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc16, (int)(2 * j * G < rem ? off : kBufOob), j * G * 4, 0);
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (Probe::kRelative) {
      const uint32_t dh = hi - probe.base, dh2 = dh + 0x10000u;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = 2 * j * G < rem;
        const uint32_t d0 = ok ? (2 * j * G < t ? dh : dh2) : 0xFFFF0000u;
        const uint32_t d1 = 2 * j * G + 1 < t ? dh : dh2;
        uint32_t i0 = (w[j] & 0xFFFFu) + d0;
        uint32_t i1 = (w[j] >> 16) + d1;
        if (j == 0) i0 = skip0 ? 0xFFFFFFFFu : i0;
        i1 = 2 * j * G + 1 < rem ? i1 : 0xFFFFFFFFu;
        cnt += probe.rel(i0) + probe.rel(i1);
      }
    } else {
      const uint32_t hi2 = hi + 0x10000u;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = 2 * j * G < rem;
        uint32_t x0 = (w[j] & 0xFFFFu) | (2 * j * G < t ? hi : hi2);
        uint32_t x1 = (w[j] >> 16) | (2 * j * G + 1 < t ? hi : hi2);
        x0 = ok && !(j == 0 && skip0) ? x0 : kNoKey;
        x1 = 2 * j * G + 1 < rem ? x1 : kNoKey;
        cnt += probe(x0) + probe(x1);
      }
    }
    qb += (uint32_t)(G * U);
  }
  return cnt;
}

// Tile streaming for block tasks whose set is a bitmap: each item is a tile
// (blk, mask) of a streamed list; its matches are popcount(mask & the
// bitmap's 64-bit word for columns [64 blk, 64 blk + 64)), one LDS read per
// 64 columns. Same range hand-out as stream_ranges (block-shared counter).
struct TileSrc {
  const uint32_t* blk;
  const unsigned long long* mask;   // 64-column tiles
  const uint2* trng;
  const uint32_t* mask32 = nullptr;  // 32-column tiles (then mask is null)
};

// TW: tile width, 32 (uint32 masks against one 32-bit bitmap word) or 64.
template <int G, int U, int TW>
__device__ __forceinline__ uint32_t stream_tiles(const TileSrc& ts, const uint2* R, uint32_t nk, uint32_t* s_next,
                                                 uint32_t lane, const uint32_t* T, uint32_t base, uint32_t span) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  const unsigned long long* W = reinterpret_cast<const unsigned long long*>(T);
  uint32_t base_r = 0, end = 0, cnt = 0;
  for (;;) {
    const bool need = base_r >= end;
    const uint64_t want = __ballot(need && gl == 0);
    const uint32_t nw = (uint32_t)__popcll(want);
    uint32_t kb = 0;
    if (nw) {
      if (lane == 0) kb = atomicAdd(s_next, nw);
      kb = __shfl(kb, 0);
      if (want == kLeads && kb >= nk) break;
    }
    if (need) {
      const uint32_t k = kb + (uint32_t)__popcll(want & below);
      if (k < nk) {
        const uint2 r = R[k];
        base_r = r.x;
        end = r.y;  // tile ranges carry no flag
      }
    }
    const uint32_t p0 = base_r + gl;
    const int rem = (int)(end - p0);
    uint32_t b[U];
    unsigned long long mk[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const bool ok = j * G < rem;
      b[j] = ok ? ts.blk[p0 + j * G] : 0xFFFFFFFFu;
      if constexpr (TW == 32)
        mk[j] = ok ? ts.mask32[p0 + j * G] : 0u;
      else
        mk[j] = ok ? ts.mask[p0 + j * G] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = b[j] * (uint32_t)TW - base;  // words past the window: the zero word after it
      if constexpr (TW == 32) {
        cnt += (uint32_t)__popc((uint32_t)mk[j] & T[min(i >> 5, span >> 5)]);
      } else {
        const unsigned long long w = W[min(i >> 6, span >> 6)];
        cnt += (uint32_t)__popcll(mk[j] & w);
      }
    }
    base_r += (uint32_t)(G * U);
  }
  return cnt;
}

// ------------------------------------------------------ wave-level tasks
// A task (x, rbase, rcount, chunk) pairs the set N+(x) with the lists whose
// ranges are rng[rbase, rbase + rcount). A wave owns one task: it builds the
// set in its LDS slice once, then stages the ranges KMAX at a time.

// Hands out the tasks of a wave launch to its waves in units of `unit`
// consecutive tasks. The units are split over kQueues sub-queues (unit k goes
// to sub-queue k % kQueues, in order: each sub-queue also starts with its
// heaviest) and claimed `claim` units at a time; a block starts on sub-queue
// blockIdx.x % kQueues and moves on to the next one when that one runs dry.
// Claims on one address serialise (about 13 ns each on MI355X): with one
// queue, a launch of ~100K tasks spent most of its time claiming. Every value
// is wave-uniform.
//
// Contiguous mode (batches, whose tasks are grouped by graph): sub-queue q
// holds the q-th eighth of the units instead. Workgroup b runs on XCD b % 8
// and starts on sub-queue b % kQueues, so each XCD streams the rows of its own
// graphs and its L2 is not filled with the rows the other seven stream.
struct TaskFeed {
  uint32_t* queue;
  uint32_t ntasks, unit, claim;
  bool contig;
  uint32_t q, left;      // current sub-queue, sub-queues not yet found empty
  uint32_t c, cend, nq;  // claimed units [c, cend) of sub-queue q, its unit count
  uint32_t q0;           // contiguous mode: the first unit of sub-queue q

  __device__ TaskFeed(uint32_t* qs, uint32_t nt, uint32_t unit_, uint32_t claim_, bool contig_ = false)
      : queue(qs), ntasks(nt), unit(unit_), claim(claim_), contig(contig_), q(blockIdx.x % kQueues),
        left(kQueues), c(0), cend(0), nq(units_of(blockIdx.x % kQueues)), q0(first_unit(blockIdx.x % kQueues)) {}

  __device__ uint32_t first_unit(uint32_t sq) const {
    const uint64_t nu = (ntasks + unit - 1) / unit;
    return contig ? (uint32_t)(nu * sq / kQueues) : 0u;
  }

  __device__ uint32_t units_of(uint32_t sq) const {
    const uint32_t nu = (ntasks + unit - 1) / unit;
    if (contig) return first_unit(sq + 1) - first_unit(sq);
    return nu > sq ? (nu - sq + kQueues - 1) / kQueues : 0u;
  }

  // The next unit [tb, tend) of this wave; false when the launch has none left.
  __device__ bool get(uint32_t lane, uint32_t& tb, uint32_t& tend) {
    for (;;) {
      if (c < cend) {
        tb = (contig ? q0 + c : c * kQueues + q) * unit;
        tend = min(tb + unit, ntasks);
        ++c;
        return true;
      }
      if (!left) return false;
      uint32_t v = 0;
      if (lane == 0) v = atomicAdd(queue + q * kQStride, claim);
      v = __shfl(v, 0);
      if (v < nq) {
        c = v;
        cend = min(v + claim, nq);
      } else {
        --left;
        q = (q + 1) % kQueues;
        nq = units_of(q);
        q0 = first_unit(q);
      }
    }
  }
};

// A wave's per-graph counts (batches), summed while consecutive tasks belong
// to one graph and added to that graph's counter when the graph changes: one
// atomic per run of tasks instead of one per task, so that tasks grouped by
// graph do not all add to the same address at once. Wave-uniform.
struct GraphSum {
  uint32_t g = 0xFFFFFFFFu;
  uint64_t sum = 0;
  __device__ void flush(unsigned long long* gcount, uint32_t lane) {
    if (lane == 0 && sum) atomicAdd(&gcount[g], (unsigned long long)sum);
    sum = 0;
  }
  __device__ void add(unsigned long long* gcount, uint32_t graph, uint64_t v, uint32_t lane) {
    if (graph != g) {
      flush(gcount, lane);
      g = graph;
    }
    sum += v;
  }
};

template <int LOGS, int KMAX, int G, int U, bool BUF>
__global__ __launch_bounds__(256) void k_count_wave(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, uint64_t m,
                                                    const uint2* __restrict__ rng,
                                                    const uint4* __restrict__ tasks, uint32_t ntasks,
                                                    uint32_t* __restrict__ queue,
                                                    unsigned long long* __restrict__ result, int attempts,
                                                    uint32_t* __restrict__ tdur,
                                                    const uint32_t* __restrict__ gid,
                                                    unsigned long long* __restrict__ gcount, uint32_t wb) {
  constexpr uint32_t S = 1u << LOGS;
  const ColSrc src = make_col_src(col, m);
  __shared__ uint32_t sT[4][2 * S];
  __shared__ uint2 sR[4][KMAX];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t* T = sT[wv];
  uint2* R = sR[wv];
  uint64_t acc = 0;
  uint32_t fallbacks = 0;
  GraphSum gsum;
  TaskFeed feed(queue, ntasks, 1, wb, gcount != nullptr);
  uint32_t tb, tend;
  while (feed.get(lane, tb, tend)) {
    for (uint32_t t = tb; t < tend; ++t) {
      const uint4 tk = tasks[ntasks - 1 - t];
      const uint32_t x = tk.x, rbase = tk.y, rcount = tk.z;
      const uint64_t t0 = tdur ? (uint64_t)wall_clock64() : 0;
      const uint32_t r0 = row_ptr[x], d = row_ptr[x + 1] - r0;
      uint32_t sa = 0;
      bool ok = false;
      for (int attempt = 0; attempt < attempts && !ok; ++attempt) {
        sa = kSeedA[attempt];
        for (uint32_t i = lane; i < 2 * S; i += 64) T[i] = kEmpty;
        wave_lds_sync();
        bool fail = false;
        for (uint32_t i = lane; i < d; i += 64) fail |= !cuckoo_insert<LOGS>(T, col[r0 + i], sa);
        wave_lds_sync();
        ok = !__any(fail);
      }
      if (!ok) {
        ++fallbacks;
        for (uint32_t i = lane; i < 2 * S; i += 64) T[i] = kEmpty;
        wave_lds_sync();
        for (uint32_t i = lane; i < d; i += 64) linear_insert<LOGS + 1>(T, col[r0 + i]);
      }
      uint32_t cnt = 0;
      for (uint32_t c0 = 0; c0 < rcount; c0 += KMAX) {
        const uint32_t nk = min((uint32_t)KMAX, rcount - c0);
        for (uint32_t k = lane; k < nk; k += 64) R[k] = rng[rbase + c0 + k];
        wave_lds_sync();
        cnt += ok ? stream_ranges<G, U, false, BUF>(src, R, nk, nullptr, lane, CuckooProbe<LOGS>{T, sa})
                  : stream_ranges<G, U, false, BUF>(src, R, nk, nullptr, lane, LinearProbe<LOGS + 1>{T});
        wave_lds_sync();
      }
      acc += cnt;
      if (gcount) gsum.add(gcount, gid[x], wave_sum_u64(cnt), lane);  // per-graph counts of a batch
      if (tdur && lane == 0) tdur[ntasks - 1 - t] = (uint32_t)(wall_clock64() - t0);
    }
  }
  if (gcount) gsum.flush(gcount, lane);
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
  if (lane == 0 && fallbacks) atomicAdd(result + 1, (unsigned long long)fallbacks);
}

// Multi-task waves for owners with tiny sets: a wave takes MT tasks at once,
// builds their MT sets (2S slots each) side by side in its LDS slice, stages
// the concatenation of their range lists with the set of each range, and
// streams them together, so a wave keeps up to 64 short lists in flight
// instead of the few of one small owner. With per-graph counts (batches)
// every range's hits are added to its task's LDS counter and each task's
// count to its owner's graph; timing logs split a group's time evenly over
// its tasks.
template <int LOGS, int KMAX, int MT, int G, int U, bool BUF>
__global__ __launch_bounds__(256) void k_count_wave_multi(const uint32_t* __restrict__ row_ptr,
                                                          const uint32_t* __restrict__ col, uint64_t m,
                                                          const uint2* __restrict__ rng,
                                                          const uint4* __restrict__ tasks, uint32_t ntasks,
                                                          uint32_t* __restrict__ queue,
                                                          unsigned long long* __restrict__ result, int attempts,
                                                          uint32_t* __restrict__ tdur, uint32_t wb,
                                                          const uint32_t* __restrict__ gid,
                                                          unsigned long long* __restrict__ gcount) {
  constexpr uint32_t S = 1u << LOGS;
  static_assert(MT <= 16 && (MT & (MT - 1)) == 0, "MT: power of two up to 16");
  const ColSrc src = make_col_src(col, m);
  __shared__ uint32_t sT[4][MT * 2 * S];
  __shared__ uint2 sR[4][KMAX];
  __shared__ uint8_t sSel[4][KMAX];
  __shared__ uint32_t sInfo[4][4][MT];  // per task: outdeg prefix, row start, range prefix, range base
  __shared__ unsigned long long sAcc[4][MT];  // per task: its count (batches: each task's graph gets its own)
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t* T = sT[wv];
  uint2* R = sR[wv];
  uint8_t* Sel = sSel[wv];
  uint32_t* DP = sInfo[wv][0];
  uint32_t* R0 = sInfo[wv][1];
  uint32_t* KP = sInfo[wv][2];
  uint32_t* RB = sInfo[wv][3];
  unsigned long long* A = sAcc[wv];
  uint64_t acc = 0;
  uint32_t fallbacks = 0;
  GraphSum gsum;
  TaskFeed feed(queue, ntasks, MT, wb / MT, gcount != nullptr);
  uint32_t tb, tend;
  while (feed.get(lane, tb, tend)) {
    for (uint32_t t0 = tb; t0 < tend; t0 += MT) {
      const uint32_t nt = min((uint32_t)MT, tend - t0);
      const uint64_t c0t = tdur ? (uint64_t)wall_clock64() : 0;
      // lane j < nt holds task j: owner row, outdeg, ranges
      uint4 tk = make_uint4(0u, 0u, 0u, 0u);
      uint32_t r0 = 0, d = 0;
      if (lane < nt) {
        tk = tasks[ntasks - 1 - (t0 + lane)];
        r0 = row_ptr[tk.x];
        d = row_ptr[tk.x + 1] - r0;
      }
      // exclusive prefixes of outdeg and range counts over the nt tasks
      uint32_t dp = d, kp = tk.z;
#pragma unroll
      for (int o = 1; o < MT; o <<= 1) {
        const uint32_t a = __shfl_up(dp, o), b = __shfl_up(kp, o);
        if (lane >= (uint32_t)o) {
          dp += a;
          kp += b;
        }
      }
      const uint32_t D = __shfl(dp, MT - 1), K = __shfl(kp, MT - 1);
      dp -= d;
      kp -= tk.z;
      if (lane < (uint32_t)MT) {  // tasks >= nt get empty slots past every index
        DP[lane] = lane < nt ? dp : 0xFFFFFFFFu;
        R0[lane] = r0;
        KP[lane] = lane < nt ? kp : 0xFFFFFFFFu;
        RB[lane] = tk.y;
        A[lane] = 0ull;
      }
      wave_lds_sync();
      uint32_t sa = 0;
      bool ok = false;
      for (int attempt = 0; attempt < attempts && !ok; ++attempt) {
        sa = kSeedA[attempt];
        for (uint32_t i = lane; i < nt * 2 * S; i += 64) T[i] = kEmpty;
        wave_lds_sync();
        bool fail = false;
        for (uint32_t f = lane; f < D; f += 64) {
          uint32_t j = 0;
#pragma unroll
          for (int q = 1; q < MT; ++q) j += f >= DP[q] ? 1u : 0u;  // DP ascending
          fail |= !cuckoo_insert<LOGS>(T + j * 2 * S, col[R0[j] + f - DP[j]], sa);
        }
        wave_lds_sync();
        ok = !__any(fail);
      }
      if (!ok) {
        ++fallbacks;
        for (uint32_t i = lane; i < nt * 2 * S; i += 64) T[i] = kEmpty;
        wave_lds_sync();
        for (uint32_t f = lane; f < D; f += 64) {
          uint32_t j = 0;
#pragma unroll
          for (int q = 1; q < MT; ++q) j += f >= DP[q] ? 1u : 0u;
          linear_insert<LOGS + 1>(T + j * 2 * S, col[R0[j] + f - DP[j]]);
        }
      }
      uint32_t cnt = 0;
      for (uint32_t c0 = 0; c0 < K; c0 += KMAX) {
        const uint32_t nk = min((uint32_t)KMAX, K - c0);
        for (uint32_t k = lane; k < nk; k += 64) {
          const uint32_t f = c0 + k;
          uint32_t j = 0;
#pragma unroll
          for (int q = 1; q < MT; ++q) j += f >= KP[q] ? 1u : 0u;
          R[k] = rng[RB[j] + f - KP[j]];
          Sel[k] = (uint8_t)j;
        }
        wave_lds_sync();
        if (gcount) {  // hits of each range go to its task's counter
          if (ok)
            stream_ranges<G, U, false, BUF, true>(src, R, nk, nullptr, lane, MultiCuckooProbe<LOGS>{T, sa}, Sel, A);
          else
            stream_ranges<G, U, false, BUF, true>(src, R, nk, nullptr, lane, LinearProbe<LOGS + 1>{T}, Sel, A);
        } else {
          cnt += ok ? stream_ranges<G, U, false, BUF>(src, R, nk, nullptr, lane, MultiCuckooProbe<LOGS>{T, sa}, Sel)
                    : stream_ranges<G, U, false, BUF>(src, R, nk, nullptr, lane, LinearProbe<LOGS + 1>{T}, Sel);
        }
        wave_lds_sync();
      }
      acc += cnt;
      if (gcount) {  // each task's count to its owner's graph, in task order
        const unsigned long long v = lane < nt ? A[lane] : 0ull;
        const uint32_t g = lane < nt ? gid[tk.x] : 0u;
        acc += v;
        for (uint32_t j = 0; j < nt; ++j)
          gsum.add(gcount, (uint32_t)__shfl((int)g, (int)j), (uint64_t)__shfl((long long)v, (int)j), lane);
      }
      if (tdur && lane < nt) tdur[ntasks - 1 - (t0 + lane)] = (uint32_t)((wall_clock64() - c0t) / nt);
      wave_lds_sync();  // the next batch rewrites the task table and sets
    }
  }
  if (gcount) gsum.flush(gcount, lane);
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
  if (lane == 0 && fallbacks) atomicAdd(result + 1, (unsigned long long)fallbacks);
}

// ----------------------------------------------------- block-level tasks
// A 512-thread block owns one task: the set is the chunk
// N+(x)[chunk*kChunk, ...) (at most kChunk ids) and the ranges are staged
// kChunk at a time; its 8 waves share each staged batch through an LDS
// counter. When x has more than one chunk, every staged range is cut to the
// chunk's id interval [lo, hi] (lists are sorted by id).

constexpr int kBlockThreads = 512;

template <int LOGS, int G, int U, bool BUF, int TW, bool N16>
__global__ __launch_bounds__(kBlockThreads) void k_count_block(const uint32_t* __restrict__ row_ptr,
                                                               const uint32_t* __restrict__ col, uint64_t m,
                                                               const uint16_t* __restrict__ col16,
                                                               const uint8_t* __restrict__ dh,
                                                               const uint16_t* __restrict__ split,
                                                               const uint2* __restrict__ rng, TileSrc ts,
                                                               const uint4* __restrict__ tasks, uint32_t ntasks,
                                                               uint32_t* __restrict__ queue,
                                                               unsigned long long* __restrict__ result,
                                                               int mode, uint32_t* __restrict__ tdur,
                                                               const uint32_t* __restrict__ gid,
                                                               unsigned long long* __restrict__ gcount) {
  constexpr uint32_t S = 1u << LOGS;
  constexpr uint32_t kSpan = 64u * S;  // ids covered by the 2S-word bitmap
  const int attempts = mode & 0xFF;
  const bool bitmaps = !(mode & kNoBitmapBit) && attempts > 0;
  __shared__ __attribute__((aligned(8))) uint32_t T[2 * S + 2];  // bitmap mode: 2S words, then a zero 64-bit word
  __shared__ uint2 R[kChunk];
  __shared__ uint16_t Rh[N16 ? kChunk : 1];  // high halves of narrow ranges
  __shared__ uint32_t s_task, s_next, s_fail, s_next16, s_next2, s_cat[3];
  __shared__ unsigned long long s_part[kBlockThreads / 64];
  const uint32_t lane = threadIdx.x & 63;
  const ColSrc src = make_col_src(col, m, N16 ? col16 : nullptr);
  uint64_t acc = 0;
  uint32_t fallbacks = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t t = s_task;
    if (t >= ntasks) break;
    const uint4 tk = tasks[ntasks - 1 - t];
    const uint32_t x = tk.x, rbase = tk.y, rcount = tk.z, c = tk.w & kChunkMask;
    const uint64_t t0 = tdur ? (uint64_t)wall_clock64() : 0;
    const uint32_t r0 = row_ptr[x], d = row_ptr[x + 1] - r0;
    // The set of N+(x)[q0, q1) in T: a bitmap over [base, base + kSpan) when
    // the ids fit, else a cuckoo set (a linear-probing table when every seed fails).
    auto build_set = [&](uint32_t q0, uint32_t q1, bool& bm, bool& ck, uint32_t& sa, uint32_t& base) {
      const uint32_t lo = col[r0 + q0], hi = col[r0 + q1 - 1];
      base = lo & ~63u;
      bm = bitmaps && (hi - base) < kSpan && hi < 0xFFF00000u && lo != 0;
      sa = 0;
      ck = bm;
      if (bm) {
        for (uint32_t i = threadIdx.x; i < 2 * S + 2; i += kBlockThreads) T[i] = 0u;
        __syncthreads();
        for (uint32_t i = q0 + threadIdx.x; i < q1; i += kBlockThreads) {
          const uint32_t y = col[r0 + i] - base;
          atomicOr(&T[y >> 5], 1u << (y & 31));
        }
        __syncthreads();
      }
      for (int attempt = 0; attempt < attempts && !ck; ++attempt) {
        sa = kSeedA[attempt];
        for (uint32_t i = threadIdx.x; i < 2 * S; i += kBlockThreads) T[i] = kEmpty;
        if (threadIdx.x == 0) s_fail = 0;
        __syncthreads();
        bool fail = false;
        for (uint32_t i = q0 + threadIdx.x; i < q1; i += kBlockThreads)
          fail |= !cuckoo_insert<LOGS>(T, col[r0 + i], sa);
        if (fail) s_fail = 1;
        __syncthreads();
        ck = (s_fail == 0);
        __syncthreads();
      }
      if (!ck) {
        ++fallbacks;
        for (uint32_t i = threadIdx.x; i < 2 * S; i += kBlockThreads) T[i] = kEmpty;
        __syncthreads();
        for (uint32_t i = q0 + threadIdx.x; i < q1; i += kBlockThreads) linear_insert<LOGS + 1>(T, col[r0 + i]);
      }
    };
    uint32_t cnt = 0;
    if (c == kAllChunks) {
      // cursor task: every chunk of N+(x) in turn against the same (at most
      // kChunk) lists, staged once with their cursors
      for (uint32_t k = threadIdx.x; k < rcount; k += kBlockThreads) {
        const uint2 r = rng[rbase + k];
        R[k] = make_uint2(r.x, r.y & src.emask);
      }
      const uint32_t nch = (d + kChunk - 1) / kChunk;
      for (uint32_t ch = 0; ch < nch; ++ch) {
        const uint32_t q0 = ch * kChunk, q1 = min(q0 + kChunk, d);
        const uint32_t bound = q1 < d ? col[r0 + q1] : 0xFFFFFFFFu;
        bool bm, ck;
        uint32_t sa, sbase;
        build_set(q0, q1, bm, ck, sa, sbase);
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        if (bm)
          cnt += stream_bounded<G, U, BUF>(src, R, rcount, &s_next, lane, BitmapProbe<LOGS + 1>{T, sbase}, bound);
        else if (ck)
          cnt += stream_bounded<G, U, BUF>(src, R, rcount, &s_next, lane, CuckooProbe<LOGS>{T, sa}, bound);
        else
          cnt += stream_bounded<G, U, BUF>(src, R, rcount, &s_next, lane, LinearProbe<LOGS + 1>{T}, bound);
        __syncthreads();
      }
    }
    const uint32_t p0 = c * kChunk, p1 = min(p0 + kChunk, d);
    const bool ranged = (p0 > 0) || (p1 < d);
    const uint32_t lo = c == kAllChunks ? 0u : col[r0 + p0], hi = c == kAllChunks ? 0u : col[r0 + p1 - 1];
    bool bitmap = false, ok = false;
    uint32_t sa = 0, base = 0;
    if (c != kAllChunks) build_set(p0, p1, bitmap, ok, sa, base);
    const bool tiles = TW != 0 && bitmap;  // block-uniform
    const uint2* rs = tiles ? ts.trng : rng;
    for (uint32_t c0 = 0; c0 < (c == kAllChunks ? 0u : rcount); c0 += kChunk) {
      const uint32_t nk = min((uint32_t)kChunk, rcount - c0);
      uint32_t n1 = 0, n2 = 0;  // with N16: R[0, n1) narrow, R[n1, n1 + n2) narrow across two windows, then wide
      if (N16 && !tiles) {
        constexpr int kPer = kChunk / kBlockThreads;
        if (threadIdx.x < 3) s_cat[threadIdx.x] = 0;
        __syncthreads();
        uint2 rv[kPer];
        uint32_t cv[kPer], iv[kPer], hv[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
          const uint32_t k = threadIdx.x + i * kBlockThreads;
          cv[i] = 3u;
          if (k < nk) {
            uint2 r = rng[rbase + c0 + k];
            const uint32_t fl = r.y & ~src.emask, b0 = r.x;
            uint32_t e = r.y & src.emask;
            if (ranged) {
              r.x = lower_bound_u32(col, r.x, e, lo);
              e = upper_bound_u32(col, r.x, e, hi);
            }
            uint32_t c = 0, h = 0;
            if ((fl & kNarrowFlag) && r.x < e) {
              const uint32_t d8 = dh[rbase + c0 + k];  // high half relative to the owner's: no read of col
              h = d8 < 255u ? (x >> 16) + d8 : col[b0] >> 16;
              c = 1;
              if (fl & kTwoWinFlag) {
                const uint32_t sp = b0 + split[rbase + c0 + k];  // first position with high half h + 1
                if (r.x >= sp)
                  ++h;  // the cut range lies past the boundary
                else if (e > sp)
                  c = 2;
              }
            }
            rv[i] = c == 2 ? make_uint2(r.x, ((b0 + split[rbase + c0 + k] - r.x) << 16) | (e - r.x))
                           : make_uint2(r.x, e);
            cv[i] = c;
            hv[i] = h;
            iv[i] = atomicAdd(&s_cat[c], 1u);
          }
        }
        __syncthreads();
        n1 = s_cat[1];
        n2 = s_cat[2];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
          if (cv[i] > 2u) continue;
          const uint32_t at = (cv[i] == 1u ? 0u : cv[i] == 2u ? n1 : n1 + n2) + iv[i];
          R[at] = rv[i];
          if (cv[i]) Rh[at] = (uint16_t)hv[i];
        }
      } else {
        for (uint32_t k = threadIdx.x; k < nk; k += kBlockThreads) {
          uint2 r = rs[rbase + c0 + k];
          if (!tiles) {
            uint32_t e = r.y & src.emask;
            if (ranged) {  // tiles outside the chunk's window miss in the bitmap
              r.x = lower_bound_u32(col, r.x, e, lo);
              e = upper_bound_u32(col, r.x, e, hi);
            }
            r.y = e;
          }
          R[k] = r;
        }
      }
      if (threadIdx.x == 0) s_next = s_next16 = s_next2 = 0;
      __syncthreads();
      const uint32_t nn = n1 + n2;
      if (tiles) {
        if constexpr (TW != 0) cnt += stream_tiles<G, U, TW>(ts, R, nk, &s_next, lane, T, base, kSpan);
      } else if (bitmap) {
        const BitmapProbe<LOGS + 1> pr{T, base};
        if (N16 && n1) cnt += stream_narrow<G, U>(src, R, Rh, n1, &s_next16, lane, pr);
        if (N16 && n2) cnt += stream_narrow2<G, U>(src, R + n1, Rh + n1, n2, &s_next2, lane, pr);
        cnt += stream_ranges<G, U, true, BUF>(src, R + nn, nk - nn, &s_next, lane, pr);
      } else if (ok) {
        const CuckooProbe<LOGS> pr{T, sa};
        if (N16 && n1) cnt += stream_narrow<G, U>(src, R, Rh, n1, &s_next16, lane, pr);
        if (N16 && n2) cnt += stream_narrow2<G, U>(src, R + n1, Rh + n1, n2, &s_next2, lane, pr);
        cnt += stream_ranges<G, U, true, BUF>(src, R + nn, nk - nn, &s_next, lane, pr);
      } else {
        const LinearProbe<LOGS + 1> pr{T};
        if (N16 && n1) cnt += stream_narrow<G, U>(src, R, Rh, n1, &s_next16, lane, pr);
        if (N16 && n2) cnt += stream_narrow2<G, U>(src, R + n1, Rh + n1, n2, &s_next2, lane, pr);
        cnt += stream_ranges<G, U, true, BUF>(src, R + nn, nk - nn, &s_next, lane, pr);
      }
      __syncthreads();
    }
    acc += cnt;
    if (gcount) {  // per-graph counts of a batch: the owner's graph
      const uint64_t tc = wave_sum_u64(cnt);
      if (lane == 0) s_part[threadIdx.x >> 6] = tc;
      __syncthreads();
      if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlockThreads / 64; ++w) t += s_part[w];
        if (t) atomicAdd(&gcount[gid[x]], t);
      }
    }
    if (tdur && threadIdx.x == 0) tdur[ntasks - 1 - t] = (uint32_t)(wall_clock64() - t0);
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
  if (threadIdx.x == 0 && fallbacks) atomicAdd(result + 1, (unsigned long long)fallbacks);
}

// Kernel variants: LDS set size (LOGS), staged ranges per batch (KMAX), lane
// group width G and loads in flight per lane U. In the DIRECTED schedule a
// streamed list is never longer than its owner's list, so the 64-slot sets
// stream lists of at most 32 elements and use narrow groups.
constexpr int kC0Logs = 6, kC0K = 64;    // d <= 32, 2x64 slots
constexpr int kC1Logs = 9, kC1K = 256;   // d <= 256, 2x512 slots
constexpr int kC2Logs = 12;              // block tasks: <= 2048 ids, 2x4096 slots
constexpr int kU = 8;
constexpr int kForceGlobal = 1 << 8;  // launch-time flag folded into the mode word (stripped before the kernel)

// ------------------------------------------------------------- scheduler
// A schedule is a set of owner lists: (x, rbase, rcount, work) where the
// ranges rng[rbase, rbase+rcount) are streamed against N+(x).
//   PLAIN:    one forward list per x over rng_all (the textbook edge
//             iterator: N+(u) is the set, every N+(v) is streamed).
//   DIRECTED: per x its forward list and its reverse list over rng_dir
//             (every edge streams the shorter of its two lists).
//   WEDGE:    per x one list over rng_sfx: for every arc u -> x the part of
//             N+(u) after x (relabelled graphs, where N+(x) holds only ids > x).
// Lists are grouped into buckets ("row-blocks" of one shape): the owner's
// outdeg bin (bin edges include 32, 256 and 2048, the set capacities) times
// the bin of the mean streamed-range length. A plan maps every bucket to a
// slot = (kernel kind, lane-group width); one launch per slot. The static
// plans map by outdeg alone (the density-threshold heuristic); the learned
// plan picks, per bucket, the menu slot of lowest predicted cost.

constexpr int kDBins = 25;
constexpr int kLBins = 32;
constexpr int kBuckets = TRIBIT_BUCKETS;
constexpr int kSlots = TRIBIT_MENU_SLOTS;
constexpr int kFeatures = TRIBIT_FEATURES;
static_assert(kBuckets == kDBins * kLBins, "bucket grid");
enum SlotKind {
  kKindWave64 = 0,
  kKindWave512 = 1,
  kKindBlock = 2,
  kKindTile = 3,
  kKindWaveMulti = 4
};

__host__ __device__ inline bool kind_is_block(int kind) { return kind == kKindBlock || kind == kKindTile; }

struct Plan {
  int nslots;
  int kind[kSlots];
  int g[kSlots];
  uint8_t table[kBuckets];  // bucket -> slot
  bool menu = false;        // the learned schedule's menu
  // how make_schedule resolved it (what ran): layout, count flags, the top
  // core's cut and the dense-block view
  bool resolved = false;
  int layout = 0, flags = 0;
  uint32_t core_d0 = 0xFFFFFFFFu, core_cut = 0;
  bool blocks = false;
};

struct BucketStats {
  uint64_t lists[kBuckets], sum_d[kBuckets], sum_k[kBuckets], sum_w[kBuckets], eligible[kBuckets], sum_t[kBuckets];
};

__host__ __device__ inline uint32_t d_bin_top(int b) {
  constexpr uint32_t top[kDBins - 1] = {1,   2,   3,   4,   6,   8,   12,   16,   24,   32,   48,   64,
                                        96,  128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 4096, 8192};
  return b < kDBins - 1 ? top[b] : 0xFFFFFFFFu;
}

__host__ __device__ inline int d_bin(uint32_t d) {
  int i = 0;
  while (i < kDBins - 1 && d > d_bin_top(i)) ++i;
  return i;
}

// floor(2 * log2(1 + mean range length)), capped
__host__ __device__ inline int l_bin(uint64_t w, uint32_t k) {
  // 32-bit division when it fits (a 64-bit one is a long software sequence on the GPU)
  uint64_t x = 1 + (k ? (w <= 0xFFFFFFFFull ? (uint64_t)((uint32_t)w / k) : w / k) : 0);
  if (x > (1ull << 20)) x = 1ull << 20;
  const int f = 63 - __builtin_clzll(x);
  const int b = 2 * f + (x * x >= (1ull << (2 * f + 1)) ? 1 : 0);
  return b < kLBins ? b : kLBins - 1;
}

__host__ __device__ inline int bucket_of(uint32_t d, uint64_t w, uint32_t k) { return d_bin(d) * kLBins + l_bin(w, k); }

__host__ __device__ inline uint32_t kind_max_d(int kind) {
  return (kind == kKindWave64 || kind == kKindWaveMulti) ? 32u : kind == kKindWave512 ? 256u : 0xFFFFFFFFu;
}

struct SchedView {
  const uint32_t* row_ptr;
  const uint32_t* fw_ptr;  // list 0 offsets (PLAIN: row_ptr, WEDGE: sfx_ptr)
  const uint32_t* rv_ptr;  // list 1 offsets (PLAIN, WEDGE: unused)
  const uint64_t* w_fw;    // list 0 work (PLAIN: wrow, WEDGE: w_sfx)
  const uint64_t* w_rv;
  int nlists;              // 1 (PLAIN, WEDGE) or 2 (DIRECTED)
  uint32_t min_d0;         // smallest owner outdeg for which list 0 can hold a triangle
  int layout;              // TRIBIT_SCHED_PLAIN / DIRECTED / WEDGE
  // dense core [core_d0, n) (counted on the matrix cores): entries whose
  // smallest vertex is in the core are dropped; core_d0 = UINT32_MAX: none
  const uint2* rng;
  uint32_t core_d0;
  uint32_t core_cut;       // entries sorted by source: kept iff rng.x < core_cut
};

struct PlanDev {
  int kind[kSlots];
  const uint8_t* table;  // bucket -> slot
};

struct Cnt8 {
  uint32_t v[kSlots];
};
struct Cnt8Plus {
  __host__ __device__ Cnt8 operator()(const Cnt8& a, const Cnt8& b) const {
    Cnt8 r;
    for (int i = 0; i < kSlots; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
  }
};

__device__ __forceinline__ void owner_list(const SchedView& v, uint64_t x, int l, uint32_t* base, uint32_t* count,
                                           uint64_t* work) {
  const uint32_t* p = (l == 0) ? v.fw_ptr : v.rv_ptr;
  *base = p[x];
  *count = p[x + 1] - p[x];
  *work = (l == 0) ? v.w_fw[x] : v.w_rv[x];
  if (v.core_d0 == 0xFFFFFFFFu || x < v.core_d0 || *count == 0) return;
  // owner in the core: lists whose triangles all have their smallest vertex
  // in the core go to the core kernel; source-sorted lists keep their prefix
  // of sources below the core
  uint32_t keep = 0;
  if (v.layout == TRIBIT_SCHED_WEDGE || (v.layout == TRIBIT_SCHED_DIRECTED && l == 1)) {
    uint32_t lo = 0, hi = *count;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (v.rng[*base + mid].x < v.core_cut)
        lo = mid + 1;
      else
        hi = mid;
    }
    keep = lo;
  }
  *work = keep ? (uint64_t)((double)*work * keep / *count) : 0;
  *count = keep;
}

__device__ __forceinline__ bool list_useful(const SchedView& v, int l, uint32_t d, uint32_t count) {
  // a forward list needs |N+(x)| >= 2 (its partner v is in N+(x) but not in N+(v))
  return count > 0 && d >= (l == 0 ? v.min_d0 : 1u);
}

// Cursor tasks: block lists (not tiles) of owners with more than kChunk ids,
// outside the wedge layout (whose narrow ranges go through the 16-bit path).
__device__ __forceinline__ bool cursor_list(int kind, uint32_t d, int layout) {
  return kind == kKindBlock && d > kChunk && layout != TRIBIT_SCHED_WEDGE;
}

__device__ __forceinline__ uint32_t list_splits(int kind, uint32_t count, uint64_t w, bool cursor = false) {
  if (cursor) {  // at most kChunk lists per task, about kCursorCap elements
    uint64_t s = (w + kCursorCap - 1) / kCursorCap;
    const uint64_t sl = (count + kChunk - 1) / kChunk;
    if (sl > s) s = sl;
    if (s < 1) s = 1;
    if (s > count) s = count;
    return (uint32_t)s;
  }
  const uint64_t cap = kind_is_block(kind) ? kBlockCap : kWaveCap;
  uint64_t s = (w + cap - 1) / cap;
  // wave tasks stage 64 lists per round trip: hubs with many short lists
  // (the top ids have tiny out-lists and huge in-degree) split by list count
  if (!kind_is_block(kind)) {
    const uint64_t sl = (count + kWaveListCap - 1) / kWaveListCap;
    if (sl > s) s = sl;
  }
  if (s < 1) s = 1;
  if (s > count) s = count;
  return (uint32_t)s;
}

__device__ __forceinline__ uint32_t list_chunks(int kind, uint32_t d, bool cursor = false) {
  return (!cursor && kind_is_block(kind) && d > kChunk) ? (d + kChunk - 1) / kChunk : 1u;
}

// Per owner: tasks per slot.
// Tasks of owner x per slot.
__device__ __forceinline__ Cnt8 owner_counts(const SchedView& sv, const PlanDev& pd, uint64_t x) {
  Cnt8 c;
  for (int i = 0; i < kSlots; ++i) c.v[i] = 0;
  const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
  for (int l = 0; l < sv.nlists; ++l) {
    uint32_t b, k;
    uint64_t w;
    owner_list(sv, x, l, &b, &k, &w);
    if (!list_useful(sv, l, d, k)) continue;
    const int s = pd.table[bucket_of(d, w, k)];
    const int kind = pd.kind[s];
    const bool cur = cursor_list(kind, d, sv.layout);
    c.v[s] += list_splits(kind, k, w, cur) * list_chunks(kind, d, cur);
  }
  return c;
}

// Task lists in owner order per slot (up to the order inside a block), in
// two passes over blocks of kSchedBlock owners: per block the per-slot task
// counts (LDS counters; scanned over the blocks on the device), then each
// owner takes its place in its block's region by an LDS atomic. Only
// per-block sums go through HBM.
constexpr unsigned kSchedBlock = 256;

__global__ __launch_bounds__(kSchedBlock) void k_sched_count(SchedView sv, PlanDev pd, uint64_t n,
                                                             Cnt8* __restrict__ blocksum) {
  __shared__ uint32_t sc[kSlots];
  if (threadIdx.x < kSlots) sc[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * kSchedBlock + threadIdx.x;
  if (x < n) {
    const Cnt8 c = owner_counts(sv, pd, x);
    for (int i = 0; i < kSlots; ++i)
      if (c.v[i]) atomicAdd(&sc[i], c.v[i]);
  }
  __syncthreads();
  if (threadIdx.x < kSlots) blocksum[blockIdx.x].v[threadIdx.x] = sc[threadIdx.x];
}

// Writes each owner's tasks in its block's region of each slot; inside a
// block (kSchedBlock consecutive owners) the order is that of LDS atomics.
__global__ __launch_bounds__(kSchedBlock) void k_sched_fill(SchedView sv, PlanDev pd, uint64_t n,
                                                            const Cnt8* __restrict__ blockoff, Cnt8 base,
                                                            uint4* __restrict__ tasks) {
  __shared__ uint32_t sc[kSlots];
  if (threadIdx.x < kSlots) sc[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * kSchedBlock + threadIdx.x;
  if (x >= n) return;
  const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
  for (int l = 0; l < sv.nlists; ++l) {
    uint32_t b, k;
    uint64_t w;
    owner_list(sv, x, l, &b, &k, &w);
    if (!list_useful(sv, l, d, k)) continue;
    const uint32_t bucket = (uint32_t)bucket_of(d, w, k);
    const int s = pd.table[bucket];
    const int kind = pd.kind[s];
    const bool cur = cursor_list(kind, d, sv.layout);
    const uint32_t sp = list_splits(kind, k, w, cur), nch = list_chunks(kind, d, cur);
    const uint32_t q = k / sp, r = k % sp;  // split i: q lists, one more for i < r
    uint4* out = tasks + base.v[s] + blockoff[blockIdx.x].v[s] + atomicAdd(&sc[s], sp * nch);
    uint32_t j = 0;
    for (uint32_t c = 0; c < nch; ++c)
      for (uint32_t i = 0; i < sp; ++i) {
        const uint32_t k0 = i * q + min(i, r);
        out[j++] = make_uint4((uint32_t)x, b + k0, q + (i < r ? 1u : 0u),
                              (cur ? kAllChunks : c) | (bucket << kBucketShift));
      }
  }
}

// Per bucket: lists, sum of outdeg, ranges, streamed elements, owners whose
// set fits a block bitmap window, 64-column tiles of the owners' rows.
// Privatised in LDS per block.
constexpr int kStat = 6;
__global__ __launch_bounds__(256) void k_bucket_stats(SchedView sv, const uint32_t* __restrict__ col, uint64_t n,
                                                      unsigned long long* __restrict__ st) {
  __shared__ unsigned long long acc[kStat * kBuckets];
  for (int i = threadIdx.x; i < kStat * kBuckets; i += blockDim.x) acc[i] = 0ull;
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) {
    const uint32_t r0 = sv.row_ptr[x], d = sv.row_ptr[x + 1] - r0;
    for (int l = 0; l < sv.nlists; ++l) {
      uint32_t b, k;
      uint64_t w;
      owner_list(sv, x, l, &b, &k, &w);
      if (!list_useful(sv, l, d, k)) continue;
      const int bk = bucket_of(d, w, k);
      const uint32_t span = col[r0 + d - 1] - (col[r0] & ~63u);
      uint32_t nt = 0, prev = 0xFFFFFFFFu;
      for (uint32_t e = r0; e < r0 + d; ++e) {
        const uint32_t tb = col[e] >> 6;
        nt += tb != prev;
        prev = tb;
      }
      atomicAdd(&acc[5 * kBuckets + bk], (unsigned long long)nt);
      atomicAdd(&acc[bk], 1ull);
      atomicAdd(&acc[kBuckets + bk], (unsigned long long)d);
      atomicAdd(&acc[2 * kBuckets + bk], (unsigned long long)k);
      atomicAdd(&acc[3 * kBuckets + bk], (unsigned long long)w);
      if (span < (64u << kC2Logs)) atomicAdd(&acc[4 * kBuckets + bk], 1ull);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kStat * kBuckets; i += blockDim.x)
    if (acc[i]) atomicAdd(&st[i], acc[i]);
}

// Per slot of a wedge plan: elements, ranges and two-window ranges of narrow
// wedge ranges (the lists' kept ranges whose end carries the narrow flag).
// A thread per owner walks lists of up to kNarrowLight ranges; longer lists
// (hubs have up to millions of in-arcs) go to `heavy` (at most m / kNarrowLight
// of them) and k_narrow_heavy walks each with a whole block.
constexpr uint32_t kNarrowLight = 256;

__device__ __forceinline__ void narrow_add(const uint2& r, unsigned long long& ne, unsigned long long& nr,
                                           unsigned long long& n2) {
  if (r.y & kNarrowFlag) {
    ne += (r.y & 0x3FFFFFFFu) - r.x;  // flags exist only when m < 2^30
    ++nr;
    n2 += (r.y & kTwoWinFlag) ? 1u : 0u;
  }
}

__global__ __launch_bounds__(256) void k_narrow_stats(SchedView sv, uint64_t n, const uint8_t* __restrict__ table,
                                                      unsigned long long* __restrict__ out,
                                                      uint4* __restrict__ heavy, uint32_t* __restrict__ nheavy) {
  __shared__ unsigned long long acc[3 * kSlots];
  if (threadIdx.x < 3 * kSlots) acc[threadIdx.x] = 0ull;
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool big = false;
  uint32_t b = 0, k = 0;
  int slot = 0;
  if (x < n) {
    const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
    uint64_t w;
    owner_list(sv, x, 0, &b, &k, &w);
    if (list_useful(sv, 0, d, k)) {
      slot = table[bucket_of(d, w, k)];
      big = k > kNarrowLight;
      if (!big) {
        unsigned long long ne = 0, nr = 0, n2 = 0;
        for (uint32_t i = b; i < b + k; ++i) narrow_add(sv.rng[i], ne, nr, n2);
        if (nr) {
          atomicAdd(&acc[slot], ne);
          atomicAdd(&acc[kSlots + slot], nr);
          if (n2) atomicAdd(&acc[2 * kSlots + slot], n2);
        }
      }
    }
  }
  // heavy owners: one append per wave
  const uint64_t hm = __ballot(big);
  if (hm) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(nheavy, (uint32_t)__popcll(hm));
    base = __shfl(base, 0);
    if (big) heavy[base + __popcll(hm & ((1ull << lane_id()) - 1))] = make_uint4(b, k, (uint32_t)slot, 0u);
  }
  __syncthreads();
  if (threadIdx.x < 3 * kSlots && acc[threadIdx.x]) atomicAdd(&out[threadIdx.x], acc[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_narrow_heavy(SchedView sv, const uint4* __restrict__ heavy,
                                                      const uint32_t* __restrict__ nheavy,
                                                      unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[3][256 / 64];
  const uint32_t nh = *nheavy;
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint4 t = heavy[h];
    unsigned long long ne = 0, nr = 0, n2 = 0;
    for (uint32_t i = t.x + threadIdx.x; i < t.x + t.y; i += blockDim.x) narrow_add(sv.rng[i], ne, nr, n2);
    ne = wave_sum_u64(ne);
    nr = wave_sum_u64(nr);
    n2 = wave_sum_u64(n2);
    if (lane_id() == 0) {
      part[0][threadIdx.x >> 6] = ne;
      part[1][threadIdx.x >> 6] = nr;
      part[2][threadIdx.x >> 6] = n2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < 256 / 64; ++i) {
        ne += part[0][i];
        nr += part[1][i];
        n2 += part[2][i];
      }
      if (nr) {
        atomicAdd(&out[t.z], ne);
        atomicAdd(&out[kSlots + t.z], nr);
        if (n2) atomicAdd(&out[2 * kSlots + t.z], n2);
      }
    }
    __syncthreads();
  }
}

// Timing logs: per (bucket) sum of the durations of a slot's tasks, from the
// per-task durations the counting kernels wrote (a block histogram in LDS, so
// the counting launches themselves do no atomics on shared counters).
__global__ __launch_bounds__(256) void k_tdur_reduce(const uint4* __restrict__ tasks, const uint32_t* __restrict__ tdur,
                                                     uint32_t ntasks, unsigned long long* __restrict__ tlog) {
  __shared__ unsigned long long h[kBuckets];
  for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0ull;
  __syncthreads();
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntasks; t += gridDim.x * blockDim.x)
    atomicAdd(&h[tasks[t].w >> kBucketShift], (unsigned long long)tdur[t]);
  __syncthreads();
  for (int i = threadIdx.x; i < kBuckets; i += blockDim.x)
    if (h[i]) atomicAdd(&tlog[i], h[i]);
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// A plan's bucket -> slot table, passed by value as a kernel argument.
struct PlanTable {
  uint8_t t[kBuckets];
};
__global__ void k_put_table(PlanTable pt, uint8_t* out) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kBuckets; i += blockDim.x) out[i] = pt.t[i];
}

// With `blocks` (wedge layout, dense diagonal blocks on the matrix cores) the
// wedge ranges are the blocks' copy with in-block arcs cut.
static SchedView sched_view(const tribit_graph* g, int sched, uint32_t core_d0 = 0xFFFFFFFFu,
                            uint32_t core_cut = 0, bool blocks = false) {
  SchedView v;
  v.row_ptr = g->row_ptr;
  v.min_d0 = 2;
  v.layout = sched;
  v.rng = sched == TRIBIT_SCHED_PLAIN ? g->rng_all : sched == TRIBIT_SCHED_WEDGE ? g->rng_sfx : g->rng_dir;
  v.core_d0 = core_d0;
  v.core_cut = core_cut;
  if (sched == TRIBIT_SCHED_PLAIN) {
    v.fw_ptr = g->row_ptr;
    v.rv_ptr = g->row_ptr;
    v.w_fw = g->wrow;
    v.w_rv = g->wrow;
    v.nlists = 1;
  } else if (sched == TRIBIT_SCHED_WEDGE) {
    v.fw_ptr = g->sfx_ptr;
    v.rv_ptr = g->sfx_ptr;
    v.w_fw = g->w_sfx;
    v.w_rv = g->w_sfx;
    v.nlists = 1;
    v.min_d0 = 1;
  } else {
    v.fw_ptr = g->fw_ptr;
    v.rv_ptr = g->rv_ptr;
    v.w_fw = g->w_fw;
    v.w_rv = g->w_rv;
    v.nlists = 2;
  }
  if (blocks && sched == TRIBIT_SCHED_WEDGE && g->blocks.rng) {
    v.rng = g->blocks.rng;
    v.fw_ptr = v.rv_ptr = g->blocks.ptr;
    v.w_fw = v.w_rv = g->blocks.w;
  }
  return v;
}

static const uint2* sched_ranges(const tribit_graph* g, int sched, bool blocks = false) {
  if (blocks && sched == TRIBIT_SCHED_WEDGE && g->blocks.rng) return g->blocks.rng;
  return sched == TRIBIT_SCHED_PLAIN ? g->rng_all : sched == TRIBIT_SCHED_WEDGE ? g->rng_sfx : g->rng_dir;
}

// ------------------------------------------------------------------ plans
// The learned scheduler's menu: (kind, lane-group width) per slot.
static const int kMenuKind[kSlots] = {kKindWave64, kKindWaveMulti, kKindWave512, kKindWave512,
                                      kKindBlock,  kKindBlock,     kKindBlock,   kKindTile};
static const int kMenuG[kSlots] = {1, 1, 2, 4, 4, 8, 16, 8};

static int kind_of_dbin(int db) {
  const uint32_t top = d_bin_top(db);
  return top <= 32 ? kKindWave64 : top <= 256 ? kKindWave512 : kKindBlock;
}

static bool slot_feasible(int slot, int bucket) {
  return d_bin_top(bucket / kLBins) <= kind_max_d(kMenuKind[slot]);
}

// Static plan of a fixed schedule: four slots by outdeg (<= 32, <= 256,
// <= 2048, > 2048) with the given widths.
static Plan static_plan(const int gw[4]) {
  Plan p;
  p.nslots = 4;
  const int kinds[4] = {kKindWave64, kKindWave512, kKindBlock, kKindBlock};
  for (int i = 0; i < 4; ++i) {
    p.kind[i] = kinds[i];
    p.g[i] = gw[i];
  }
  for (int b = 0; b < kBuckets; ++b) {
    const uint32_t top = d_bin_top(b / kLBins);
    p.table[b] = (uint8_t)(top <= 32 ? 0 : top <= 256 ? 1 : top <= kChunk ? 2 : 3);
  }
  return p;
}

static Plan menu_plan() {
  Plan p;
  p.menu = true;
  p.nslots = kSlots;
  for (int i = 0; i < kSlots; ++i) {
    p.kind[i] = kMenuKind[i];
    p.g[i] = kMenuG[i];
  }
  return p;
}

// The density-threshold heuristic over the menu (fallback of LEARNED): block
// tasks stream 8 lanes per list on the wedge layout (narrow ranges: two ids
// per load), 16 on the others.
static void heuristic_table(uint8_t* table, int layout) {
  const uint8_t blk = layout == TRIBIT_SCHED_WEDGE ? 5 : 6;
  for (int b = 0; b < kBuckets; ++b) {
    const int kind = kind_of_dbin(b / kLBins);
    table[b] = kind == kKindWave64 ? 0 : kind == kKindWave512 ? 3 : blk;
  }
}

// Default lane-group widths per variant and schedule (swept on R-MAT scale
// 22 and 25 over 1..32 lanes per list for wave tasks, 8..64 for block tasks).
// Wedge lists of small owners are short: one lane per list keeps 64 lists in
// flight per wave.
static void default_widths(int sched, int gw[4]) {
  if (sched == TRIBIT_SCHED_PLAIN) {
    gw[0] = 32, gw[1] = 16, gw[2] = 32, gw[3] = 16;
  } else if (sched == TRIBIT_SCHED_WEDGE) {
    gw[0] = 1, gw[1] = 4, gw[2] = 8, gw[3] = 8;
  } else {
    gw[0] = 1, gw[1] = 8, gw[2] = 16, gw[3] = 16;
  }
}

// Bucket statistics depend only on the graph, the range layout, the core and
// the dense-block view: computed once per graph and kept (the last
// combination the scheduler asked for). Also kept per graph: the resolved
// plan of the last count of each schedule (what tribit_variant_stats reports).
struct BucketCache {
  int layout = -1;
  uint32_t core_d0 = 0;
  bool blocks = false;
  BucketStats st;
  Plan last[4];  // by tribit_sched
};

static BucketCache* cache_of(tribit_graph* g) {
  if (!g->bcache) g->bcache = new BucketCache;
  return g->bcache;
}

void bucket_cache_free(tribit_graph* g) {
  delete g->bcache;
  g->bcache = nullptr;
}

// With keep = false the statistics are computed without replacing the kept ones.
void bucket_stats(tribit_ctx* ctx, tribit_graph* g, int sched, BucketStats* out, uint32_t core_d0 = 0xFFFFFFFFu,
                  uint32_t core_cut = 0, bool blocks = false, bool keep = true) {
  if (g->bcache && g->bcache->layout == sched && g->bcache->core_d0 == core_d0 && g->bcache->blocks == blocks) {
    *out = g->bcache->st;
    return;
  }
  ensure_layout(ctx, g, sched);
  memset(out, 0, sizeof(*out));
  DevBuf<unsigned long long> st(ctx, kStat * kBuckets);
  k_zero_u64<<<grid_for(kStat * kBuckets, kBlock), kBlock, 0, ctx->stream>>>(st.p, kStat * kBuckets);
  if (g->n && g->m)
    k_bucket_stats<<<grid_for(g->n, 256), 256, 0, ctx->stream>>>(sched_view(g, sched, core_d0, core_cut, blocks),
                                                                 g->col, g->n, st.p);
  std::vector<unsigned long long> h(kStat * kBuckets);
  TB_HIP(hipMemcpyAsync(h.data(), st.p, 8 * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  for (int b = 0; b < kBuckets; ++b) {
    out->lists[b] = h[b];
    out->sum_d[b] = h[kBuckets + b];
    out->sum_k[b] = h[2 * kBuckets + b];
    out->sum_w[b] = h[3 * kBuckets + b];
    out->eligible[b] = h[4 * kBuckets + b];
    out->sum_t[b] = h[5 * kBuckets + b];
  }
  if (!keep) return;
  BucketCache* c = cache_of(g);
  c->layout = sched;
  c->core_d0 = core_d0;
  c->blocks = blocks;
  c->st = *out;
}

// The heuristic's owner-outdeg classes (<= 32, <= 256, larger). The learned
// plan runs one menu slot per class (one launch per class).
constexpr int kClasses = TRIBIT_CLASSES;
static int class_of_bucket(int b) {
  const uint32_t top = d_bin_top(b / kLBins);
  return top <= 32 ? 0 : top <= 256 ? 1 : 2;
}

static uint32_t class_top(int c) { return c == 0 ? 32u : c == 1 ? 256u : 0xFFFFFFFFu; }

// A slot every bucket of the class fits (its sets hold the class's largest outdeg).
static bool class_slot_feasible(int c, int slot) { return class_top(c) <= kind_max_d(kMenuKind[slot]); }

static int heuristic_slot(int c, int layout) {
  uint8_t t[kBuckets];
  heuristic_table(t, layout);
  for (int b = 0; b < kBuckets; ++b)
    if (class_of_bucket(b) == c) return t[b];
  return 0;
}

// Feature vector of (outdeg class, menu slot) over a graph's bucket
// statistics; the same function builds training samples and scores plans.
// The cost model's target: log2 of the count's device time with the class on
// this slot and the other classes on the heuristic's slots, over the time of
// the heuristic plan.
void class_features(const BucketStats& st, int c, int slot, float* f) {
  double L = 0, D = 0, K = 0, W = 0, E = 0, T = 0, Lg = 0, Wg = 0;
  double hist[kDBins] = {};
  for (int b = 0; b < kBuckets; ++b) {
    if (!st.lists[b]) continue;
    Lg += (double)st.lists[b];
    Wg += (double)st.sum_w[b] + (double)st.sum_k[b];
    if (class_of_bucket(b) != c) continue;
    L += (double)st.lists[b];
    D += (double)st.sum_d[b];
    K += (double)st.sum_k[b];
    W += (double)st.sum_w[b];
    E += (double)st.eligible[b];
    T += (double)st.sum_t[b];
    hist[b / kLBins] += (double)st.lists[b];
  }
  // owner outdeg min / median / max over the class's lists, by outdeg bin
  // (the bin's top; the open last bin counts as 16384)
  auto top_of = [](int db) { return db < kDBins - 1 ? (double)d_bin_top(db) : 16384.0; };
  double dmin = 0, dmed = 0, dmax = 0, acc = 0;
  bool seen = false;
  for (int db = 0; db < kDBins; ++db) {
    if (hist[db] <= 0) continue;
    if (!seen) dmin = top_of(db);
    seen = true;
    dmax = top_of(db);
    if (acc < 0.5 * L && acc + hist[db] >= 0.5 * L) dmed = top_of(db);
    acc += hist[db];
  }
  // log2 magnitudes scaled to a few units (the perceptron takes raw inputs)
  const double Lc = L > 0 ? L : 1.0;
  f[0] = (float)(std::log2(1.0 + L) / 8.0);
  f[1] = (float)(std::log2(1.0 + D / Lc) / 4.0);                 // mean owner outdeg
  f[2] = (float)(std::log2(1.0 + K / Lc) / 4.0);                 // ranges per list
  f[3] = (float)(std::log2(1.0 + W / (K > 0 ? K : 1.0)) / 4.0);  // elements per range
  f[4] = (float)(std::log2(1.0 + W + K) / 8.0);                  // the class's work
  f[5] = (float)(E / Lc);                                        // bitmap-window eligible share
  f[6] = (float)(std::log2((1.0 + D) / (1.0 + T)) / 4.0);        // owner-row elements per 64-column tile
  f[7] = (float)(std::log2(1.0 + dmin) / 4.0);
  f[8] = (float)(std::log2(1.0 + dmed) / 4.0);
  f[9] = (float)(std::log2(1.0 + dmax) / 4.0);
  f[10] = (float)(std::log2(1.0 + L / 256.0) / 8.0);             // lists per CU
  f[11] = (float)((W + K) / (Wg > 0 ? Wg : 1.0));                // the class's share of the graph's work
  f[12] = c == 0;
  f[13] = c == 1;
  f[14] = c == 2;
  const int kind = kMenuKind[slot], g = kMenuG[slot];
  f[15] = kind == kKindWave64;
  f[16] = kind == kKindWaveMulti;
  f[17] = kind == kKindWave512;
  f[18] = kind == kKindBlock;
  f[19] = kind == kKindTile;
  f[20] = (float)(std::log2((double)g) / 4.0);
  f[21] = slot == heuristic_slot(c, TRIBIT_SCHED_WEDGE);
  f[22] = (float)(std::log2(1.0 + Lg) / 8.0);
  f[23] = (float)(std::log2(1.0 + Wg) / 8.0);
  static_assert(kFeatures == 24, "feature count");
}

static void stats_of_profile(const tribit_profile* p, BucketStats* st) {
  memcpy(st->lists, p->lists, sizeof(st->lists));
  memcpy(st->sum_d, p->sum_d, sizeof(st->sum_d));
  memcpy(st->sum_k, p->sum_k, sizeof(st->sum_k));
  memcpy(st->sum_w, p->sum_w, sizeof(st->sum_w));
  memcpy(st->eligible, p->eligible, sizeof(st->eligible));
  memcpy(st->sum_t, p->sum_t, sizeof(st->sum_t));
}

void profile_class_features(const tribit_profile* p, int cls, int slot, float* f) {
  BucketStats st;
  stats_of_profile(p, &st);
  class_features(st, cls, slot, f);
}

bool class_slot_ok(int cls, int slot) { return class_slot_feasible(cls, slot); }

void menu_slot(int slot, int* kind, int* g, uint32_t* max_d) {
  if (kind) *kind = kMenuKind[slot];
  if (g) *g = kMenuG[slot];
  if (max_d) *max_d = kind_max_d(kMenuKind[slot]);
}

uint32_t bucket_max_outdeg(int bucket) { return d_bin_top(bucket / kLBins); }

// LEARNED: one menu slot per outdeg class of the heuristic (<= 32, <= 256,
// larger). The heuristic's slot stays unless the cost model predicts, for the
// whole count, a feasible slot of the class at least kSwitch faster (predicted
// log2 time ratios, each against the heuristic plan, compared with the
// heuristic slot's own prediction). Classes are judged one at a time against
// the heuristic plan and the switches combined. Without a model, a class
// without lists or a prediction that is not finite: the heuristic's slot.
constexpr double kSwitch = 0.04;  // log2 units: a predicted gain of about 3%

static void learned_table(tribit_ctx* ctx, const BucketStats& st, int layout, int flags, uint8_t* table) {
  heuristic_table(table, layout);
  if (!ctx->model || layout != TRIBIT_SCHED_WEDGE) return;  // the model is trained on wedge-schedule timings
  bool any[kClasses] = {};
  for (int b = 0; b < kBuckets; ++b)
    if (st.lists[b]) any[class_of_bucket(b)] = true;
  std::vector<float> X;
  std::vector<std::pair<int, int>> ids;
  for (int c = 0; c < kClasses; ++c) {
    if (!any[c]) continue;
    for (int s = 0; s < kSlots; ++s) {
      if (!class_slot_feasible(c, s) || ((flags & TRIBIT_COUNT_NO_TILES) && kMenuKind[s] == kKindTile)) continue;
      float f[kFeatures];
      class_features(st, c, s, f);
      X.insert(X.end(), f, f + kFeatures);
      ids.push_back({c, s});
    }
  }
  if (ids.empty()) return;
  std::vector<float> y(ids.size());
  mlp_infer(ctx->model, X.data(), ids.size(), y.data());
  for (int c = 0; c < kClasses; ++c) {
    const int h = heuristic_slot(c, layout);
    double yh = NAN;
    for (size_t i = 0; i < ids.size(); ++i)
      if (ids[i].first == c && ids[i].second == h) yh = y[i];
    if (!std::isfinite(yh)) continue;
    int best = h;
    double yb = yh;
    for (size_t i = 0; i < ids.size(); ++i)
      if (ids[i].first == c && std::isfinite(y[i]) && y[i] < yb) {
        yb = y[i];
        best = ids[i].second;
      }
    if (best == h || yb > yh - kSwitch) continue;
    for (int b = 0; b < kBuckets; ++b)
      if (class_of_bucket(b) == c) table[b] = (uint8_t)best;
  }
}

// ----------------------------------------------------------------- launches
// Per-count launch arguments (no allocation or host sync between them, so a
// whole count can be captured into a hipGraph).
struct LaunchArgs {
  int att = 0;                           // mode word (attempts | flags)
  unsigned long long* result = nullptr;  // [0] triangles, [1] fallback tasks
  uint32_t* queue = nullptr;             // per slot kQueues task counters, kQStride apart
  uint32_t* tdur = nullptr;              // per-task durations, indexed like the schedule's tasks, or null
  const uint4* task0 = nullptr;          // the schedule's first task (tdur's origin)
  const uint32_t* gid = nullptr;         // graph of each vertex (batches), or null
  unsigned long long* gcount = nullptr;  // per-graph counts (batches), or null
  int waves[kSlots] = {0};               // workgroups per CU cap per slot (0: occupancy limit)
};

// Tasks per claim of a wave launch of `ntasks` tasks on `waves` waves: a
// multiple of `unit` (the tasks a wave takes at once), at most kWaveBatch,
// and small enough that a wave makes about kClaimsPerWave claims. A claim
// takes consecutive tasks of its sub-queue, which at the list's end are often
// the chunks of one hub's range list, so it stays small against a wave's share.
static uint32_t claim_size(uint32_t ntasks, uint64_t waves, uint32_t unit) {
  uint64_t b = ntasks / (waves * kClaimsPerWave);
  if (b > kWaveBatch) b = kWaveBatch;
  b = b / unit * unit;
  return (uint32_t)(b < unit ? unit : b);
}

template <int LOGS, int KMAX, int G, int U, bool BUF>
static void launch_wave_t(const LaunchArgs& a, int slot, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g,
                          const uint2* rng, const uint4* tasks, uint64_t* blocks) {
  int per_cu = 0;
  TB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count_wave<LOGS, KMAX, G, U, BUF>, kBlock, 0));
  if (a.waves[slot] > 0 && per_cu > a.waves[slot]) per_cu = a.waves[slot];
  if (per_cu < 1) per_cu = 1;
  // waves for about kWaveBatch tasks each (4 for the 512-slot sets, whose
  // short launches on small graphs gained from more waves; the 64-slot
  // kernel's did not)
  const uint64_t tpw = LOGS >= kC1Logs ? 4 : kWaveBatch;
  const uint64_t want = (ntasks + 4 * tpw - 1) / (4 * tpw);
  const unsigned grid = (unsigned)std::min<uint64_t>(want, (uint64_t)cus * per_cu);
  const uint32_t wb = claim_size(ntasks, 4ull * grid, 1);
  k_count_wave<LOGS, KMAX, G, U, BUF><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, g->m, rng, tasks, ntasks,
                                                             a.queue + (size_t)slot * kQueues * kQStride, a.result, a.att & 0xFF,
                                                             a.tdur ? a.tdur + (tasks - a.task0) : nullptr,
                                                             a.gid, a.gcount, wb);
  *blocks += grid;
}

template <int LOGS, int KMAX, int G>
static void launch_wave_g(const LaunchArgs& a, int slot, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g,
                          const uint2* rng, const uint4* tasks, uint64_t* blocks) {
  if (g->m * 4 <= kBufMax && !(a.att & kForceGlobal))
    launch_wave_t<LOGS, KMAX, G, kU, true>(a, slot, ntasks, cus, s, g, rng, tasks, blocks);
  else
    launch_wave_t<LOGS, KMAX, G, kU, false>(a, slot, ntasks, cus, s, g, rng, tasks, blocks);
}

template <int LOGS, int KMAX>
static void launch_wave(const LaunchArgs& a, int slot, int gw, uint32_t ntasks, int cus, hipStream_t s,
                        tribit_graph* g, const uint2* rng, const uint4* tasks, uint64_t* blocks) {
  if (!ntasks) return;
  switch (gw) {
    case 1: launch_wave_g<LOGS, KMAX, 1>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
    case 2: launch_wave_g<LOGS, KMAX, 2>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
    case 4: launch_wave_g<LOGS, KMAX, 4>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
    case 8: launch_wave_g<LOGS, KMAX, 8>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
    case 32: launch_wave_g<LOGS, KMAX, 32>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
    default: launch_wave_g<LOGS, KMAX, 16>(a, slot, ntasks, cus, s, g, rng, tasks, blocks); break;
  }
}

constexpr int kMultiTasks = 8;  // tasks per multi-task wave batch
constexpr uint32_t kMultiMinTasksPerCU = 4096;  // menu wave-64 launches switch to multi-task waves above this
constexpr uint32_t kBlockG4MaxTasksPerCU = 1536;  // menu 8-lane block launches use 4 lanes below this

template <int G, bool BUF>
static void launch_multi_t(const LaunchArgs& a, int slot, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g,
                           const uint2* rng, const uint4* tasks, uint64_t* blocks) {
  auto k = k_count_wave_multi<kC0Logs, kC0K, kMultiTasks, G, kU, BUF>;
  int per_cu = 0;
  TB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0));
  if (a.waves[slot] > 0 && per_cu > a.waves[slot]) per_cu = a.waves[slot];
  if (per_cu < 1) per_cu = 1;
  const uint64_t want = (ntasks + 4 * kWaveBatch - 1) / (4 * kWaveBatch);
  const unsigned grid = (unsigned)std::min<uint64_t>(want, (uint64_t)cus * per_cu);
  const uint32_t wb = claim_size(ntasks, 4ull * grid, kMultiTasks);
  k<<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, g->m, rng, tasks, ntasks, a.queue + (size_t)slot * kQueues * kQStride, a.result, a.att & 0xFF,
                            a.tdur ? a.tdur + (tasks - a.task0) : nullptr, wb, a.gid, a.gcount);
  *blocks += grid;
}

static void launch_multi(const LaunchArgs& a, int slot, int gw, uint32_t ntasks, int cus, hipStream_t s,
                         tribit_graph* g, const uint2* rng, const uint4* tasks, uint64_t* blocks) {
  if (!ntasks) return;
  const bool buf = g->m * 4 <= kBufMax && !(a.att & kForceGlobal);
  switch (gw) {
    case 2:
      buf ? launch_multi_t<2, true>(a, slot, ntasks, cus, s, g, rng, tasks, blocks)
          : launch_multi_t<2, false>(a, slot, ntasks, cus, s, g, rng, tasks, blocks);
      break;
    case 4:
      buf ? launch_multi_t<4, true>(a, slot, ntasks, cus, s, g, rng, tasks, blocks)
          : launch_multi_t<4, false>(a, slot, ntasks, cus, s, g, rng, tasks, blocks);
      break;
    default:
      buf ? launch_multi_t<1, true>(a, slot, ntasks, cus, s, g, rng, tasks, blocks)
          : launch_multi_t<1, false>(a, slot, ntasks, cus, s, g, rng, tasks, blocks);
      break;
  }
}

template <int LOGS, int G, int U, bool BUF, int TW, bool N16>
static void launch_block_t(const LaunchArgs& a, int slot, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g,
                           const uint2* rng, const TileSrc& ts, const uint4* tasks, uint64_t* blocks) {
  int per_cu = 0;
  TB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count_block<LOGS, G, U, BUF, TW, N16>,
                                                      kBlockThreads, 0));
  if (a.waves[slot] > 0 && per_cu > a.waves[slot]) per_cu = a.waves[slot];
  if (per_cu < 1) per_cu = 1;
  const unsigned grid = (unsigned)std::min<uint64_t>(ntasks, (uint64_t)cus * per_cu);
  k_count_block<LOGS, G, U, BUF, TW, N16><<<grid, kBlockThreads, 0, s>>>(
      g->row_ptr, g->col, g->m, g->col16, g->sfx_dh, g->sfx_split, rng, ts, tasks, ntasks, a.queue + (size_t)slot * kQueues * kQStride, a.result,
      a.att & ~kForceGlobal,
      a.tdur ? a.tdur + (tasks - a.task0) : nullptr, a.gid, a.gcount);
  *blocks += grid;
}

template <int LOGS, int G>
static void launch_block_g(const LaunchArgs& a, int slot, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g,
                           const uint2* rng, const TileSrc& ts, const uint4* tasks, uint64_t* blocks) {
  const bool buf = g->m * 4 <= kBufMax && !(a.att & kForceGlobal);
  // narrow ranges exist only in the wedge layout (col16 is built with it); the
  // dense blocks' compacted copy carries no narrow flags
  const bool n16 = buf && g->col16 && g->sfx_dh && g->sfx_split && rng == g->rng_sfx && !(a.att & kWideBit);
  if (ts.trng && ts.mask32) {
    if (n16)
      launch_block_t<LOGS, G, kU, true, 32, true>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
    else if (buf)
      launch_block_t<LOGS, G, kU, true, 32, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
    else
      launch_block_t<LOGS, G, kU, false, 32, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
  } else if (ts.trng) {
    if (n16)
      launch_block_t<LOGS, G, kU, true, 64, true>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
    else if (buf)
      launch_block_t<LOGS, G, kU, true, 64, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
    else
      launch_block_t<LOGS, G, kU, false, 64, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
  } else if (n16) {
    launch_block_t<LOGS, G, kU, true, 0, true>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
  } else if (buf) {
    launch_block_t<LOGS, G, kU, true, 0, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
  } else {
    launch_block_t<LOGS, G, kU, false, 0, false>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks);
  }
}

template <int LOGS>
static void launch_block(const LaunchArgs& a, int slot, int gw, uint32_t ntasks, int cus, hipStream_t s,
                         tribit_graph* g, const uint2* rng, const TileSrc& ts, const uint4* tasks, uint64_t* blocks) {
  if (!ntasks) return;
  switch (gw) {
    case 4: launch_block_g<LOGS, 4>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks); break;
    case 8: launch_block_g<LOGS, 8>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks); break;
    case 32: launch_block_g<LOGS, 32>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks); break;
    case 64: launch_block_g<LOGS, 64>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks); break;
    default: launch_block_g<LOGS, 16>(a, slot, ntasks, cus, s, g, rng, ts, tasks, blocks); break;
  }
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  TB_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

static int mode_word(int flags) {
  return ((flags & TRIBIT_COUNT_FORCE_LINEAR_PROBE) ? 0 : kCuckooSeeds) |
         ((flags & TRIBIT_COUNT_FORCE_GLOBAL_LOADS) ? kForceGlobal : 0) |
         ((flags & TRIBIT_COUNT_NO_BITMAP) ? kNoBitmapBit : 0) | ((flags & TRIBIT_COUNT_WIDE_IDS) ? kWideBit : 0);
}

// Tasks of one count under a plan, ready to launch.
struct Schedule {
  Plan plan;
  int gdef[kSlots] = {0};  // the plan's own widths (a batch's tuned widths override them)
  int kdef[kSlots] = {0};  // the plan's own kernel kinds (a batch's tuned tile width overrides block kinds)
  uint64_t core_c = 0;     // dense-core vertices counted on the matrix cores (0: none)
  bool blocks = false;     // dense diagonal blocks counted on the matrix cores, ranges cut to match
  int layout = 0;
  const uint2* rng = nullptr;
  TileSrc ts{nullptr, nullptr, nullptr};
  DevBuf<uint4> tasks;
  Cnt8 tot{}, base{};
  uint64_t ntask = 0;
};

// Bump allocation from one caller-held device buffer (the no-allocation entry
// point's scratch array): 256-byte aligned slices, released all at once by
// resetting `used`. A request beyond `size` fails with TRIBIT_ERR_TOO_LARGE.
struct Arena {
  char* base = nullptr;
  uint64_t size = 0, used = 0;
  template <typename T>
  T* take(uint64_t count) {
    uint64_t bytes = (count ? count : 1) * sizeof(T);
    bytes = (bytes + 255) & ~uint64_t(255);
    if (used + bytes > size)
      fail(TRIBIT_ERR_TOO_LARGE, "the schedule needs " + std::to_string((used + bytes) >> 20) +
                                     " MiB of scratch, more than the no-allocation entry point's " +
                                     std::to_string(size >> 20) + " MiB");
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

// A device buffer from the arena when there is one, else from the context's pool.
template <typename T>
static T* scratch_buf(tribit_ctx* ctx, Arena* arena, DevBuf<T>& own, uint64_t count) {
  if (arena) return arena->take<T>(count);
  own = DevBuf<T>(ctx, count ? count : 1);
  return own.p;
}

// Resolves the plan (learned table, tile opt-in/out) and builds the task
// lists (one host sync for the per-slot totals). With `arena`, every device
// buffer comes from it (the caller rules out what would allocate: learned
// tables, tiles, a dense core).
void make_schedule(tribit_ctx* ctx, tribit_graph* g, int layout, int flags, bool learned, Schedule& S,
                   uint64_t core_c, Arena* arena = nullptr) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n;
  Plan& plan = S.plan;
  S.layout = layout;
  ensure_layout(ctx, g, layout);  // no-op when built (count() builds it before its timing starts)
  uint32_t core_d0 = 0xFFFFFFFFu, core_cut = 0;
  if (core_c > g->n || !g->relabeled) core_c = 0;
  // dense diagonal blocks replace the top core (a block inside the core would
  // count its triangles twice)
  S.blocks = (flags & TRIBIT_COUNT_DENSE_BLOCKS) && layout == TRIBIT_SCHED_WEDGE && blocks_prepare(ctx, g);
  if (S.blocks) core_c = 0;
  S.core_c = core_c;
  if (core_c) {
    core_prepare(ctx, g, core_c);
    core_d0 = (uint32_t)(n - core_c);
    const uint32_t e0 = g->core_e0;
    // wedge entries start one past their arc; reverse entries at their row start
    core_cut = layout == TRIBIT_SCHED_WEDGE ? e0 + 1 : e0;
  }
  plan.resolved = true;
  plan.layout = layout;
  plan.flags = flags;
  plan.core_d0 = core_d0;
  plan.core_cut = core_cut;
  plan.blocks = S.blocks;
  if (learned) {
    // the model's choice is part of the scheduler's time
    BucketStats st;
    // statistics of the whole lists, as in the timing profiles the model is
    // trained on (counted without a dense core); the plan then runs with the core
    bucket_stats(ctx, g, layout, &st);
    learned_table(ctx, st, layout, flags, plan.table);
  }
  // test hooks / opt-outs: every block launch streams tiles, or none does. With
  // dense blocks none does: tile ranges follow the uncut wedge ranges.
  for (int i = 0; i < plan.nslots; ++i) {
    if ((flags & TRIBIT_COUNT_FORCE_TILES) && plan.kind[i] == kKindBlock) plan.kind[i] = kKindTile;
    if (((flags & TRIBIT_COUNT_NO_TILES) || S.blocks) && plan.kind[i] == kKindTile) plan.kind[i] = kKindBlock;
  }
  bool any_tiles = false;
  for (int i = 0; i < plan.nslots; ++i) any_tiles |= plan.kind[i] == kKindTile;
  if (any_tiles && g->m) {
    const int width = (flags & TRIBIT_COUNT_TILES32) ? 32 : 64;
    build_tiles(ctx, g, layout, width);  // once per graph, width and layout
    const tribit_graph::TileSet& T = g->tiles[width == 32 ? 0 : 1];
    S.ts.blk = T.blk;
    S.ts.mask = width == 64 ? reinterpret_cast<const unsigned long long*>(T.mask) : nullptr;
    S.ts.mask32 = width == 32 ? reinterpret_cast<const uint32_t*>(T.mask) : nullptr;
    S.ts.trng = T.trng[layout & 3];
  }
  DevBuf<uint8_t> dtable_own;
  uint8_t* dtable = scratch_buf(ctx, arena, dtable_own, kBuckets);
  {
    PlanTable pt;
    memcpy(pt.t, plan.table, kBuckets);
    k_put_table<<<1, kBlock, 0, s>>>(pt, dtable);  // a kernel argument: no host staging
  }
  PlanDev pd;
  for (int i = 0; i < kSlots; ++i) pd.kind[i] = i < plan.nslots ? plan.kind[i] : kKindBlock;
  pd.table = dtable;

  // per-block task counts (block nb covers no owner: its sum is zero), scanned
  const uint64_t nb = (n + kSchedBlock - 1) / kSchedBlock;
  DevBuf<Cnt8> cnt_own, off_own;
  Cnt8* cnt = scratch_buf(ctx, arena, cnt_own, nb + 1);
  Cnt8* off = scratch_buf(ctx, arena, off_own, nb + 1);
  const SchedView sv = sched_view(g, layout, core_d0, core_cut, S.blocks);
  S.rng = sched_ranges(g, layout, S.blocks);
  k_sched_count<<<(unsigned)(nb + 1), kSchedBlock, 0, s>>>(sv, pd, n, cnt);
  {
    Cnt8 zero;
    for (int i = 0; i < kSlots; ++i) zero.v[i] = 0;
    size_t bytes = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt, off, zero, (size_t)(nb + 1), Cnt8Plus(), s));
    DevBuf<uint8_t> tmp_own;
    uint8_t* tmp = scratch_buf(ctx, arena, tmp_own, bytes ? bytes : 8);
    TB_HIP(rocprim::exclusive_scan(tmp, bytes, cnt, off, zero, (size_t)(nb + 1), Cnt8Plus(), s));
  }
  // off[nb] = totals
  TB_HIP(hipMemcpyAsync(&S.tot, off + nb, sizeof(Cnt8), hipMemcpyDeviceToHost, s));
  ctx->sync();
  S.ntask = 0;
  for (int i = 0; i < kSlots; ++i) {
    S.base.v[i] = (uint32_t)S.ntask;
    S.ntask += S.tot.v[i];
  }
  if (S.ntask > 0xFFFFFFFFull) fail(TRIBIT_ERR_TOO_LARGE, "more than 2^32-1 tasks");
  if (arena)
    S.tasks.p = arena->take<uint4>(S.ntask);  // S.tasks.ctx stays null: the arena is never freed here
  else
    S.tasks = DevBuf<uint4>(ctx, S.ntask ? S.ntask : 1);
  if (nb) k_sched_fill<<<(unsigned)nb, kSchedBlock, 0, s>>>(sv, pd, n, off, S.base, S.tasks.p);
  TB_HIP(hipGetLastError());
  // every plan but the textbook one: a one-task wave-64 launch with thousands
  // of tasks per CU runs as multi-task waves (same tasks, eight per wave at
  // once); with fewer, the one-task kernel was faster (R-MAT s25: 1.40 -> 1.03
  // ms at 2.1M tasks; s22, 0.39M tasks: 0.37 -> 0.57 ms)
  const int cus = ctx->cus > 0 ? ctx->cus : 256;
  const bool adapt = plan.menu || layout != TRIBIT_SCHED_PLAIN;
  for (int i = 0; i < plan.nslots; ++i)
    if (plan.kind[i] == kKindWave64 && (plan.g[i] == 1 || plan.g[i] == 2 || plan.g[i] == 4) &&
        ((flags & TRIBIT_COUNT_FORCE_MULTI) || (adapt && S.tot.v[i] >= (uint64_t)kMultiMinTasksPerCU * cus)))
      plan.kind[i] = kKindWaveMulti;
  // menu plans: an 8-lane block launch of fewer than 1,536 tasks per CU
  // streams 4 lanes per list (R-MAT s20-s24, 8K-283K tasks: 3-12% faster;
  // s25, 540K tasks: 4% slower)
  if (plan.menu)
    for (int i = 0; i < plan.nslots; ++i)
      if (plan.kind[i] == kKindBlock && plan.g[i] == 8 && S.tot.v[i] < (uint64_t)kBlockG4MaxTasksPerCU * cus)
        plan.g[i] = 4;
}

// The counting launches of a schedule (capturable: no allocation, no sync).
// The caller zeroes a.result and a.queue. With `ev`, an event is recorded
// after each slot's launch (ev[i] for slot i).
void launch_schedule(tribit_ctx* ctx, tribit_graph* g, const Schedule& S, const LaunchArgs& a, hipEvent_t* ev,
                     uint64_t* blocks) {
  hipStream_t s = ctx->stream;
  const int cus = ctx->cus > 0 ? ctx->cus : 256;
  const Plan& plan = S.plan;
  for (int i = 0; i < plan.nslots; ++i) {
    const uint4* tk = S.tasks.p + S.base.v[i];
    const uint32_t nt = S.tot.v[i];
    if (plan.kind[i] == kKindWave64)
      launch_wave<kC0Logs, kC0K>(a, i, plan.g[i], nt, cus, s, g, S.rng, tk, blocks);
    else if (plan.kind[i] == kKindWaveMulti)
      launch_multi(a, i, plan.g[i], nt, cus, s, g, S.rng, tk, blocks);
    else if (plan.kind[i] == kKindWave512)
      launch_wave<kC1Logs, kC1K>(a, i, plan.g[i], nt, cus, s, g, S.rng, tk, blocks);
    else if (plan.kind[i] == kKindTile && S.ts.trng)
      launch_block<kC2Logs>(a, i, plan.g[i], nt, cus, s, g, S.rng, S.ts, tk, blocks);
    else
      launch_block<kC2Logs>(a, i, plan.g[i], nt, cus, s, g, S.rng, TileSrc{nullptr, nullptr, nullptr}, tk, blocks);
    if (ev) TB_HIP(hipEventRecord(ev[i], s));
  }
  if (S.core_c) {
    core_launch(ctx, g, a.result);
    if (ev) TB_HIP(hipEventRecord(ev[plan.nslots], s));
  }
  if (S.blocks) {
    blocks_launch(ctx, g, a.result);
    if (ev) TB_HIP(hipEventRecord(ev[plan.nslots + (S.core_c ? 1 : 0)], s));
  }
  TB_HIP(hipGetLastError());
}

// One scheduled count under `plan`. Returns the triangle count; fills
// per-phase times. With `tlog` (kSlots x kBuckets, zeroed by the caller) every
// task's wall-clock duration is added to tlog[slot][bucket] (recorded per task
// during the launches, summed after them).
static uint64_t count_once(tribit_ctx* ctx, tribit_graph* g, int sched, int flags, Plan& plan, bool learned,
                           double* ms_total, double* ms_sched, double* ms_kernel, double* ms_var,
                           tribit_count_info* info, unsigned long long* tlog, uint64_t core_c = 0,
                           double* ms_core = nullptr, Arena* arena = nullptr) {
  hipStream_t s = ctx->stream;
  TB_HIP(hipEventRecord(ctx->ev[2], s));
  Schedule S;
  S.plan = plan;
  DevBuf<unsigned long long> counters_own;
  // counters first: an arena that is too small for the tasks fails before any launch
  unsigned long long* counters = scratch_buf(ctx, arena, counters_own, kCountScratch);
  make_schedule(ctx, g, sched, flags, learned, S, core_c, arena);
  plan = S.plan;
  k_zero_u64<<<grid_for(kCountScratch, kBlock), kBlock, 0, s>>>(counters, kCountScratch);
  LaunchArgs a;
  a.att = mode_word(flags);
  a.result = counters;
  a.queue = reinterpret_cast<uint32_t*>(counters + kQueueOff);
  DevBuf<uint32_t> tdur;
  if (tlog) {
    tdur = DevBuf<uint32_t>(ctx, S.ntask ? S.ntask : 1);
    a.tdur = tdur.p;
    a.task0 = S.tasks.p;
  }
  TB_HIP(hipEventRecord(ctx->ev[3], s));
  uint64_t blocks = 0;
  launch_schedule(ctx, g, S, a, ctx->ev + 4, &blocks);
  hipEvent_t last = ctx->ev[3 + plan.nslots + (S.core_c ? 1 : 0) + (S.blocks ? 1 : 0)];
  if (tlog)
    for (int v = 0; v < plan.nslots; ++v)
      if (S.tot.v[v])
        k_tdur_reduce<<<(unsigned)std::min<uint64_t>((S.tot.v[v] + 255) / 256, 1024), 256, 0, s>>>(
            S.tasks.p + S.base.v[v], tdur.p + S.base.v[v], S.tot.v[v], tlog + (size_t)v * kBuckets);
  unsigned long long res[2] = {0, 0};
  TB_HIP(hipMemcpyAsync(res, a.result, sizeof(res), hipMemcpyDeviceToHost, s));
  TB_HIP(hipEventSynchronize(last));
  ctx->sync();
  *ms_sched = elapsed(ctx->ev[2], ctx->ev[3]);
  *ms_kernel = elapsed(ctx->ev[3], last);
  *ms_total = elapsed(ctx->ev[2], last);
  for (int v = 0; v < plan.nslots; ++v) ms_var[v] = elapsed(ctx->ev[3 + v], ctx->ev[4 + v]);
  if (ms_core) *ms_core = (S.core_c || S.blocks) ? elapsed(ctx->ev[3 + plan.nslots], last) : 0.0;
  if (info) {
    info->launch_blocks = blocks;
    info->tasks = S.ntask;
    for (int i = 0; i < 8; ++i) info->tasks_by_variant[i] = i < plan.nslots ? S.tot.v[i] : 0;
    info->nvariants = plan.nslots;
    info->fallback_tasks = res[1];
    for (int i = 0; i < 8; ++i) info->group_width[i] = i < plan.nslots ? plan.g[i] : 0;
    for (int i = 0; i < 8; ++i) info->slot_kind[i] = i < plan.nslots ? plan.kind[i] : 0;
    info->dense_core = S.core_c ? S.core_c : S.blocks ? g->blocks.vertices : 0;
    info->dense_blocks = S.blocks ? g->blocks.nblocks : 0;
  }
  return res[0];
}

static double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t k = v.size();
  return k ? ((k & 1) ? v[k / 2] : 0.5 * (v[k / 2 - 1] + v[k / 2])) : 0.0;
}

// The plan a schedule runs with: static widths for PLAIN/DIRECTED/WEDGE
// (overridable per class), the menu for LEARNED (table filled at count time).
static Plan plan_for(int sched, const int32_t* override_gw, int* layout) {
  if (sched == TRIBIT_SCHED_LEARNED) {
    *layout = -1;  // resolved by the caller
    Plan p = menu_plan();
    heuristic_table(p.table, TRIBIT_SCHED_WEDGE);  // re-filled for the layout that runs
    return p;
  }
  *layout = sched;
  int gw[4];
  default_widths(sched, gw);
  for (int i = 0; i < 4; ++i)
    if (override_gw && override_gw[i]) gw[i] = override_gw[i];
  return static_plan(gw);
}

// LEARNED runs one of these candidates; the first count of a graph (or after
// the model changes) times each once and keeps the fastest, so the heuristic
// plans judge the model's plan by measurement:
//   0: the cost model's plan on the preferred layout (+ dense core)
//   1: the density-threshold heuristic on the preferred layout (+ dense core)
//   2: the heuristic on the DIRECTED layout (+ dense core), for relabelled graphs
//   3: the textbook PLAIN plan
//   4: the heuristic with dense diagonal blocks on the matrix cores
// (preferred layout: WEDGE for relabelled graphs, else DIRECTED).
enum { kCandModel = 0, kCandHeur = 1, kCandDirected = 2, kCandPlain = 3, kCandBlocks = 4, kCands = 5 };
static_assert(kCands <= TRIBIT_CANDIDATES, "candidate times");

struct CandidateRun {
  int layout;
  Plan plan;
  bool model;
  uint64_t core_c;
  int flags = 0;  // count flags the candidate adds (dense diagonal blocks)
};

static bool candidate(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts& opts, int c, CandidateRun* out) {
  const int pref = g->rng_sfx ? TRIBIT_SCHED_WEDGE : TRIBIT_SCHED_DIRECTED;
  uint64_t core = 0;
  if (opts.dense_core > 0)
    core = std::min<uint64_t>((uint64_t)opts.dense_core, g->n);
  else if (opts.dense_core == 0 && g->relabeled) {
    if (g->core_pick < 0) g->core_pick = (int64_t)core_choose(ctx, g);  // the graph does not change
    core = (uint64_t)g->core_pick;
  }
  int dummy = 0;
  switch (c) {
    case kCandModel:
      if (!ctx->model || pref != TRIBIT_SCHED_WEDGE) return false;  // the model is trained on wedge timings
      *out = {pref, plan_for(TRIBIT_SCHED_LEARNED, nullptr, &dummy), true, core};
      return true;
    case kCandHeur:
      *out = {pref, plan_for(TRIBIT_SCHED_LEARNED, nullptr, &dummy), false, core};
      heuristic_table(out->plan.table, pref);
      return true;
    case kCandDirected:
      if (pref != TRIBIT_SCHED_WEDGE) return false;  // already candidate 1
      *out = {TRIBIT_SCHED_DIRECTED, plan_for(TRIBIT_SCHED_LEARNED, nullptr, &dummy), false, core};
      heuristic_table(out->plan.table, TRIBIT_SCHED_DIRECTED);
      return true;
    case kCandBlocks:
      // the heuristic plan with dense diagonal blocks on the matrix cores, when
      // the graph has any (detection runs once per graph, outside the timing)
      if (pref != TRIBIT_SCHED_WEDGE || opts.dense_core > 0 || !blocks_prepare(ctx, g)) return false;
      *out = {pref, plan_for(TRIBIT_SCHED_LEARNED, nullptr, &dummy), false, 0, TRIBIT_COUNT_DENSE_BLOCKS};
      heuristic_table(out->plan.table, pref);
      return true;
    default:
      // the textbook plan has no core of its own; a core the caller asks for
      // (dense_core > 0) is kept, as in every other candidate
      *out = {TRIBIT_SCHED_PLAIN, plan_for(TRIBIT_SCHED_PLAIN, nullptr, &dummy), false,
              opts.dense_core > 0 ? core : 0};
      return true;
  }
}

static const tribit_ctx::Judged* judged_lookup(const tribit_ctx* ctx, const tribit_graph* g,
                                               const tribit_count_opts& opts) {
  for (const auto& j : ctx->judged)
    if (j.n == g->n && j.m == g->m && j.wplus == g->wplus && j.relabeled == g->relabeled &&
        j.flags == opts.flags && j.core == opts.dense_core && j.model == ctx->model_version)
      return &j;
  return nullptr;
}

static void judged_store(tribit_ctx* ctx, const tribit_graph* g, const tribit_count_opts& opts) {
  constexpr size_t kKeep = 64;  // most recent shapes kept
  if (ctx->judged.size() >= kKeep) ctx->judged.erase(ctx->judged.begin());
  tribit_ctx::Judged j;
  j.n = g->n;
  j.m = g->m;
  j.wplus = g->wplus;
  j.model = ctx->model_version;
  j.relabeled = g->relabeled;
  j.flags = opts.flags;
  j.core = opts.dense_core;
  j.choice = g->judge_choice;
  for (int c = 0; c < kCands; ++c) j.ms[c] = g->judge_ms[c];
  ctx->judged.push_back(j);
}

bool judged_known(const tribit_ctx* ctx, const tribit_graph* g, const tribit_count_opts& opts) {
  return judged_lookup(ctx, g, opts) != nullptr;
}

// ---- the no-allocation count (tribit_count_noalloc) ------------------------
// Its scratch: one file-scope device array per GPU, held until the process
// ends. Host code never passes the symbol itself (on the host it names a
// stand-in, not device memory): noalloc_init looks its device address up once
// per context, with the context's device current, and kernels get that.
constexpr uint64_t kNoAllocScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocScratchBytes];

void noalloc_init(tribit_ctx* ctx) {
  void* p = nullptr;
  TB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
  if (!p) fail(TRIBIT_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  ctx->noalloc_scratch = p;
}

uint64_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

void count(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts& opts, tribit_count_info* info, bool noalloc) {
  const int reps = opts.repeats > 0 ? opts.repeats : 1;
  std::vector<double> tt, ts, tk, tv[kSlots];
  uint64_t first = 0;
  bool agree = true;
  tribit_count_info local;
  memset(&local, 0, sizeof(local));
  const bool learned = opts.sched == TRIBIT_SCHED_LEARNED;
  // what would allocate: judging, learned tables, tiles, a dense core matrix
  if (noalloc && (learned || opts.dense_core > 0 || (opts.flags & TRIBIT_COUNT_FORCE_TILES)))
    fail(TRIBIT_ERR_ARG, "the no-allocation count runs PLAIN, DIRECTED or WEDGE without a dense core or tiles");
  CandidateRun run;
  int chosen = -1;
  double ms_judge = 0.0;
  if (learned && (opts.flags & TRIBIT_COUNT_NO_JUDGE)) {
    chosen = kCandHeur;
    candidate(ctx, g, opts, chosen, &run);
  } else if (learned) {
    bool stale = g->judge_choice < 0 || g->judge_model != ctx->model_version ||
                 g->judge_flags != opts.flags || g->judge_core != opts.dense_core;
    if (stale) {
      // the same input judged before in this context (e.g. a warm-up pass)
      const tribit_ctx::Judged* hit = judged_lookup(ctx, g, opts);
      if (hit) {
        g->judge_choice = hit->choice;
        for (int c = 0; c < kCands; ++c) g->judge_ms[c] = hit->ms[c];
        g->judge_model = ctx->model_version;
        g->judge_flags = opts.flags;
        g->judge_core = opts.dense_core;
        stale = false;
      }
    }
    if (stale) {
      for (int c = 0; c < kCands; ++c) g->judge_ms[c] = 0.0;
      double best = INFINITY;
      uint64_t ref = 0;
      bool have = false;
      for (int c = 0; c < kCands; ++c) {
        CandidateRun cr;
        if (!candidate(ctx, g, opts, c, &cr)) continue;
        ensure_layout(ctx, g, cr.layout);  // a layout's first use: built before the timed runs
        if (c == kCandModel) {
          // the model keeps the heuristic's slot for every class it predicts
          // no gain for: then it is the heuristic candidate, not timed twice
          BucketStats st;
          bucket_stats(ctx, g, cr.layout, &st);
          uint8_t mt[kBuckets], ht[kBuckets];
          learned_table(ctx, st, cr.layout, opts.flags | cr.flags, mt);
          heuristic_table(ht, cr.layout);
          if (!memcmp(mt, ht, sizeof(mt))) continue;
        }
        // twice: the first run also builds what the candidate caches (core
        // matrix, bucket statistics, tiles); the second is the one compared,
        // unless the first already took more than twice the best time
        double a = 0.0, b, k, v[kSlots] = {0};
        for (int rep = 0; rep < 2; ++rep) {
          if (rep == 1 && a > 2.0 * best) break;
          Plan pl = cr.plan;
          const uint64_t tri = count_once(ctx, g, cr.layout, opts.flags | cr.flags, pl, cr.model, &a, &b, &k, v,
                                          nullptr, nullptr, cr.core_c, nullptr);
          if (have && tri != ref) fail(TRIBIT_ERR_HIP, "learned candidates disagree");
          ref = tri;
          have = true;
          ms_judge += a;
        }
        g->judge_ms[c] = a;
        if (a < best) {
          best = a;
          chosen = c;
        }
      }
      g->judge_choice = chosen;
      g->judge_model = ctx->model_version;
      g->judge_flags = opts.flags;
      g->judge_core = opts.dense_core;
      judged_store(ctx, g, opts);
    }
    chosen = g->judge_choice;
    candidate(ctx, g, opts, chosen, &run);
  } else {
    int layout = 0;
    run.plan = plan_for(opts.sched, opts.group_width, &layout);
    run.layout = layout;
    run.model = false;
    // fixed schedules use a dense core only when asked for (opts.dense_core > 0)
    run.core_c = (opts.dense_core > 0 && g->relabeled) ? std::min<uint64_t>((uint64_t)opts.dense_core, g->n) : 0;
  }
  std::vector<double> tcore;
  Arena arena;
  if (noalloc) {
    if (!layout_built(g, run.layout))
      fail(TRIBIT_ERR_ARG, "the no-allocation count needs the schedule's range layout resident: count the graph "
                           "once with tribit_count and this schedule first");
    arena.base = static_cast<char*>(ctx->noalloc_scratch);
    arena.size = kNoAllocScratchBytes;
    if (!arena.base) fail(TRIBIT_ERR_HIP, "no scratch array address for this device");
  }
  if (!noalloc) ensure_layout(ctx, g, run.layout);  // first use of a layout: built outside the timing
  for (int r = 0; r < reps; ++r) {
    double a, b, c, v[kSlots] = {0}, mc = 0.0;
    Plan plan = run.plan;
    arena.used = 0;  // every repeat reuses the whole array
    const uint64_t tri = count_once(ctx, g, run.layout, opts.flags | run.flags, plan, run.model, &a, &b, &c, v,
                                    &local, nullptr, run.core_c, &mc, noalloc ? &arena : nullptr);
    if (r == reps - 1) {
      run.plan = plan;
      cache_of(g)->last[opts.sched & 3] = plan;  // what ran, for tribit_variant_stats
    }
    tcore.push_back(mc);
    if (r == 0)
      first = tri;
    else if (tri != first)
      agree = false;
    tt.push_back(a);
    ts.push_back(b);
    tk.push_back(c);
    for (int i = 0; i < kSlots; ++i) tv[i].push_back(v[i]);
  }
  local.triangles = first;
  local.ms_count = median(tt);
  local.ms_sched = median(ts);
  local.ms_kernel = median(tk);
  for (int i = 0; i < kSlots; ++i) local.ms_by_variant[i] = i < run.plan.nslots ? median(tv[i]) : 0.0;
  local.repeats_agree = agree ? 1 : 0;
  local.ms_dense = median(tcore);
  local.candidate = learned ? chosen : -1;
  local.layout = run.layout;
  local.ms_judge = ms_judge;
  if (learned && !(opts.flags & TRIBIT_COUNT_NO_JUDGE))
    for (int c = 0; c < kCands; ++c) local.ms_candidates[c] = g->judge_ms[c];
  if (info) *info = local;
}

// Timed count under a menu plan given per bucket (NULL: the heuristic), with
// task durations logged per (slot, bucket).
void count_profile(tribit_ctx* ctx, tribit_graph* g, int layout, const uint8_t* table, tribit_profile* out) {
  memset(out, 0, sizeof(*out));
  Plan plan = menu_plan();
  if (table) {
    for (int b = 0; b < kBuckets; ++b) {
      if (table[b] >= kSlots || !slot_feasible(table[b], b)) fail(TRIBIT_ERR_ARG, "plan slot not feasible for bucket");
      plan.table[b] = table[b];
    }
  } else {
    heuristic_table(plan.table, layout);
  }
  BucketStats st;
  bucket_stats(ctx, g, layout, &st);
  DevBuf<unsigned long long> tl(ctx, (size_t)kSlots * kBuckets);
  k_zero_u64<<<grid_for(kSlots * kBuckets, kBlock), kBlock, 0, ctx->stream>>>(tl.p, kSlots * kBuckets);
  double a, b, c, v[kSlots] = {0};
  out->triangles = count_once(ctx, g, layout, 0, plan, false, &a, &b, &c, v, nullptr, tl.p);
  out->ms_kernel = c;
  for (int s2 = 0; s2 < kSlots; ++s2) out->ms_slot[s2] = v[s2];
  std::vector<unsigned long long> h((size_t)kSlots * kBuckets);
  TB_HIP(hipMemcpyAsync(h.data(), tl.p, 8 * h.size(), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  for (int s2 = 0; s2 < kSlots; ++s2)
    for (int bk = 0; bk < kBuckets; ++bk) out->ticks[s2][bk] = h[(size_t)s2 * kBuckets + bk];
  for (int bk = 0; bk < kBuckets; ++bk) {
    out->lists[bk] = st.lists[bk];
    out->sum_d[bk] = st.sum_d[bk];
    out->sum_k[bk] = st.sum_k[bk];
    out->sum_w[bk] = st.sum_w[bk];
    out->eligible[bk] = st.eligible[bk];
    out->sum_t[bk] = st.sum_t[bk];
    out->slot[bk] = plan.table[bk];
  }
  int khz = 0;
  TB_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  out->tick_ns = khz > 0 ? 1e6 / khz : 10.0;
}

// Device time of the counting launches under a menu plan given per bucket
// (NULL: the heuristic), median over `repeats` counts, no task logging: the
// cost model's training target.
void count_plan(tribit_ctx* ctx, tribit_graph* g, int layout, const uint8_t* table, int repeats, double* ms_kernel,
                uint64_t* triangles) {
  Plan base = menu_plan();
  if (table) {
    for (int b = 0; b < kBuckets; ++b) {
      if (table[b] >= kSlots || !slot_feasible(table[b], b)) fail(TRIBIT_ERR_ARG, "plan slot not feasible for bucket");
      base.table[b] = table[b];
    }
  } else {
    heuristic_table(base.table, layout);
  }
  ensure_layout(ctx, g, layout);
  std::vector<double> tk;
  uint64_t tri = 0;
  for (int r = 0; r < std::max(repeats, 1); ++r) {
    Plan plan = base;
    double a, b, c, v[kSlots] = {0};
    const uint64_t t = count_once(ctx, g, layout, 0, plan, false, &a, &b, &c, v, nullptr, nullptr);
    if (r && t != tri) fail(TRIBIT_ERR_HIP, "repeated counts disagree");
    tri = t;
    tk.push_back(c);
  }
  std::sort(tk.begin(), tk.end());
  *ms_kernel = tk[tk.size() / 2];
  *triangles = tri;
}

void learned_plan(tribit_ctx* ctx, tribit_graph* g, uint8_t* table) {
  BucketStats st;
  const int layout = g->rng_sfx ? TRIBIT_SCHED_WEDGE : TRIBIT_SCHED_DIRECTED;
  bucket_stats(ctx, g, layout, &st);
  learned_table(ctx, st, layout, 0, table);
}

// Per-slot work of a schedule's plan (LEARNED: the table the model picks).
// Per-slot work of a schedule's plan. LEARNED: the judged candidate (the
// heuristic before the first count), streamed work outside its dense core.
void variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out) {
  memset(out, 0, sizeof(*out));
  int layout = 0, flags = 0;
  Plan plan = plan_for(sched, nullptr, &layout);
  BucketStats st;
  uint32_t d0 = 0xFFFFFFFFu, cut = 0;
  bool blocks = false;
  const Plan* ran = g->bcache ? &g->bcache->last[sched & 3] : nullptr;
  if (ran && ran->resolved) {
    // the plan the last count of this schedule ran, as make_schedule resolved
    // it (table, kernel kinds and widths after its launch-size adjustments,
    // core cut, dense-block view); statistics computed without evicting the
    // scheduler's kept ones
    plan = *ran;
    layout = ran->layout;
    flags = ran->flags;
    d0 = ran->core_d0;
    cut = ran->core_cut;
    blocks = ran->blocks;
    bucket_stats(ctx, g, layout, &st, d0, cut, blocks, false);
  } else if (sched == TRIBIT_SCHED_LEARNED) {
    tribit_count_opts o;
    memset(&o, 0, sizeof(o));
    o.sched = TRIBIT_SCHED_LEARNED;
    o.flags = g->judge_choice >= 0 ? g->judge_flags : 0;
    o.dense_core = g->judge_choice >= 0 ? g->judge_core : 0;
    CandidateRun cr;
    candidate(ctx, g, o, g->judge_choice >= 0 ? g->judge_choice : kCandHeur, &cr);
    layout = cr.layout;
    plan = cr.plan;
    if (cr.core_c) {
      d0 = (uint32_t)(g->n - cr.core_c);
      uint32_t e0 = 0;
      TB_HIP(hipMemcpyAsync(&e0, g->row_ptr + d0, 4, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      cut = layout == TRIBIT_SCHED_WEDGE ? e0 + 1 : e0;
    }
    bucket_stats(ctx, g, layout, &st, d0, cut, false, false);
    if (cr.model) learned_table(ctx, st, layout, 0, plan.table);
    flags = o.flags;
  } else {
    bucket_stats(ctx, g, layout, &st);
  }
  for (int b = 0; b < kBuckets; ++b) {
    const int s = plan.table[b];
    out->rows[s] += st.lists[b];
    out->edges[s] += st.sum_k[b];
    out->wplus[s] += st.sum_w[b];
  }
  // block launches stream narrow wedge ranges as 16-bit halves
  if (layout == TRIBIT_SCHED_WEDGE && g->col16 && g->m * 4 <= kBufMax && !(flags & TRIBIT_COUNT_WIDE_IDS) &&
      !(flags & TRIBIT_COUNT_FORCE_GLOBAL_LOADS) && g->n) {
    DevBuf<uint8_t> dt(ctx, kBuckets);
    DevBuf<unsigned long long> ns(ctx, 3 * kSlots + 1);  // the last word: heavy-owner count
    DevBuf<uint4> heavy(ctx, g->m / kNarrowLight + 1);
    TB_HIP(hipMemcpyAsync(dt.p, plan.table, kBuckets, hipMemcpyHostToDevice, ctx->stream));
    k_zero_u64<<<1, 64, 0, ctx->stream>>>(ns.p, 3 * kSlots + 1);
    uint32_t* nheavy = reinterpret_cast<uint32_t*>(ns.p + 3 * kSlots);
    const SchedView sv = sched_view(g, layout, d0, cut, blocks);
    k_narrow_stats<<<grid_for(g->n, 256), 256, 0, ctx->stream>>>(sv, g->n, dt.p, ns.p, heavy.p, nheavy);
    k_narrow_heavy<<<2048, 256, 0, ctx->stream>>>(sv, heavy.p, nheavy, ns.p);
    unsigned long long h[3 * kSlots];
    TB_HIP(hipMemcpyAsync(h, ns.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    for (int s = 0; s < plan.nslots; ++s)
      if (plan.kind[s] == kKindBlock) {
        out->narrow[s] = h[s];
        out->narrow_ranges[s] = h[kSlots + s];
        out->narrow2_ranges[s] = h[2 * kSlots + s];
      }
  }
}

// ------------------------------------------------------------------ batches
void schedule_free(Schedule* s) { delete s; }

void batch_drop_exec(tribit_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Graph-major task order for a batch part's wave launches: key = the owner's
// graph, value = the task's position; a stable sort keeps owner order inside a
// graph (hubs still first within it, the kernels taking tasks from the end).
__global__ void k_task_graph(const uint4* __restrict__ t, uint32_t nt, const uint32_t* __restrict__ gid,
                             uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  key[j] = gid[t[j].x];
  idx[j] = j;
}

__global__ void k_task_gather(const uint4* __restrict__ t, const uint32_t* __restrict__ idx, uint32_t nt,
                              uint4* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nt) out[j] = t[idx[j]];
}

// Wave launches of a batch part stream short ranges of rows of the owner's
// own graph: grouped by graph, the tasks in flight at once touch the rows of
// a few graphs, which then stay in the caches, instead of the rows of every
// graph of the part at once (tasks in owner order interleave all graphs of a
// degree). Block launches keep hub-first order.
static void order_tasks_by_graph(tribit_ctx* ctx, const tribit_batch::Part& p, Schedule& S) {
  hipStream_t s = ctx->stream;
  const uint32_t kbits = bits_for((uint64_t)p.first + p.count);
  for (int i = 0; i < S.plan.nslots; ++i) {
    const int kind = S.plan.kind[i];
    const uint32_t nt = S.tot.v[i];
    if (nt < 2 || kind_is_block(kind)) continue;
    uint4* t = S.tasks.p + S.base.v[i];
    DevBuf<uint32_t> key(ctx, nt), idx(ctx, nt), key2(ctx, nt), idx2(ctx, nt);
    DevBuf<uint4> out(ctx, nt);
    k_task_graph<<<grid_for(nt, kBlock), kBlock, 0, s>>>(t, nt, p.gid, key.p, idx.p);
    size_t bytes = 0;
    TB_HIP(rocprim::radix_sort_pairs(nullptr, bytes, key.p, key2.p, idx.p, idx2.p, (size_t)nt, 0u, kbits, s));
    DevBuf<uint8_t> tmp(ctx, bytes ? bytes : 8);
    TB_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, key.p, key2.p, idx.p, idx2.p, (size_t)nt, 0u, kbits, s));
    k_task_gather<<<grid_for(nt, kBlock), kBlock, 0, s>>>(t, idx2.p, nt, out.p);
    TB_HIP(hipMemcpyAsync(t, out.p, (size_t)nt * sizeof(uint4), hipMemcpyDeviceToDevice, s));
    ctx->sync();  // the temporaries are freed on return
  }
}

// Task lists of every part for (sched, flags); returns the device ms spent.
static double batch_schedules(tribit_ctx* ctx, tribit_batch* b, int sched, int flags) {
  // a learned batch's tables come from the cost model: a new model makes them stale
  bool ready = b->sched == sched && b->flags == flags &&
               (sched != TRIBIT_SCHED_LEARNED || b->model_version == ctx->model_version);
  for (auto& p : b->parts) ready = ready && p.S;
  if (ready) return 0.0;
  batch_drop_exec(b);
  TB_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  for (auto& p : b->parts) {
    schedule_free(p.S);
    p.S = nullptr;
    int layout = 0;
    Plan plan = plan_for(sched, nullptr, &layout);
    const bool learned = sched == TRIBIT_SCHED_LEARNED;
    if (learned) layout = p.g->rng_sfx ? TRIBIT_SCHED_WEDGE : TRIBIT_SCHED_DIRECTED;
    if (layout == TRIBIT_SCHED_WEDGE && !p.g->rng_sfx) fail(TRIBIT_ERR_ARG, "the wedge schedule needs relabelled graphs");
    p.S = new Schedule;
    p.S->plan = plan;
    // per-graph attribution: no dense core and no dense blocks (their launches
    // add to the batch total, not to each graph's count)
    make_schedule(ctx, p.g, layout, flags & ~TRIBIT_COUNT_DENSE_BLOCKS, learned, *p.S, 0);
    order_tasks_by_graph(ctx, p, *p.S);
    for (int i = 0; i < kSlots; ++i) {
      p.S->gdef[i] = p.S->plan.g[i];
      p.S->kdef[i] = p.S->plan.kind[i];
    }
  }
  TB_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  TB_HIP(hipEventSynchronize(ctx->ev[1]));
  b->sched = sched;
  b->flags = flags;
  b->model_version = ctx->model_version;
  return elapsed(ctx->ev[0], ctx->ev[1]);
}

// Captures the zeroing of the counters and the counting launches of every
// part into one hipGraph (counters are cleared by a kernel, not a memset
// node). Returns the host ms of capture + instantiation.
static double batch_capture(tribit_ctx* ctx, tribit_batch* b, int att, uint32_t* launches, uint64_t* tasks) {
  batch_drop_exec(b);
  hipStream_t s = ctx->stream;
  // the tuned tile width: block launches become tile launches over tiles of
  // that width (built here, before the capture: building allocates and syncs)
  for (auto& p : b->parts) {
    Schedule& S = *p.S;
    const int tw = b->tile_width;
    bool tiles = false;
    for (int k = 0; k < S.plan.nslots; ++k) {
      const bool blockish = S.kdef[k] == kKindBlock || S.kdef[k] == kKindTile;
      S.plan.kind[k] = (tw && blockish) ? kKindTile : S.kdef[k];
      tiles |= S.plan.kind[k] == kKindTile;
    }
    if (tiles && p.g->m) {
      const int width = tw ? tw : (b->flags & TRIBIT_COUNT_TILES32) ? 32 : 64;
      build_tiles(ctx, p.g, S.layout, width);
      const tribit_graph::TileSet& T = p.g->tiles[width == 32 ? 0 : 1];
      S.ts.blk = T.blk;
      S.ts.mask = width == 64 ? reinterpret_cast<const unsigned long long*>(T.mask) : nullptr;
      S.ts.mask32 = width == 32 ? reinterpret_cast<const uint32_t*>(T.mask) : nullptr;
      S.ts.trng = T.trng[S.layout & 3];
    }
  }
  TB_HIP(hipStreamSynchronize(s));
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t blocks = 0;
  uint32_t nl = 2;
  uint64_t nt = 0;
  TB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    k_zero_u64<<<grid_for(b->ngraphs ? b->ngraphs : 1, kBlock), kBlock, 0, s>>>(b->gcount, b->ngraphs);
    const uint32_t nw = (uint32_t)(kCountScratch * b->parts.size());
    k_zero_u64<<<grid_for(nw ? nw : 1, kBlock), kBlock, 0, s>>>(b->scratch, nw);
    for (size_t i = 0; i < b->parts.size(); ++i) {
      auto& p = b->parts[i];
      for (int k = 0; k < kSlots; ++k) p.S->plan.g[k] = b->gw[k] ? b->gw[k] : p.S->gdef[k];
      LaunchArgs a;
      a.att = att;
      a.result = b->scratch + kCountScratch * i;
      a.queue = reinterpret_cast<uint32_t*>(b->scratch + kCountScratch * i + kQueueOff);
      a.gid = p.gid;
      a.gcount = b->gcount;
      for (int k = 0; k < kSlots; ++k) a.waves[k] = b->waves[k];
      launch_schedule(ctx, p.g, *p.S, a, nullptr, &blocks);
      for (int k = 0; k < p.S->plan.nslots; ++k) nl += p.S->tot.v[k] ? 1 : 0;
      nt += p.S->ntask;
    }
  } catch (...) {
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(s, &tmp);
    if (tmp) (void)hipGraphDestroy(tmp);
    throw;
  }
  TB_HIP(hipStreamEndCapture(s, &b->graph));
  TB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  if (launches) *launches = nl;
  if (tasks) *tasks = nt;
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Replays the captured graph `reps` times; per-graph counts of the last
// replay into `counts`; returns the median device ms; *agree = every replay
// gave the same counts.
static double batch_replay(tribit_ctx* ctx, tribit_batch* b, int reps, std::vector<unsigned long long>& counts,
                           bool* agree) {
  hipStream_t s = ctx->stream;
  std::vector<double> t;
  std::vector<unsigned long long> cur(b->ngraphs);
  *agree = true;
  for (int r = 0; r < reps; ++r) {
    TB_HIP(hipEventRecord(ctx->ev[2], s));
    TB_HIP(hipGraphLaunch(b->exec, s));
    TB_HIP(hipEventRecord(ctx->ev[3], s));
    if (b->ngraphs) TB_HIP(hipMemcpyAsync(cur.data(), b->gcount, 8 * b->ngraphs, hipMemcpyDeviceToHost, s));
    TB_HIP(hipEventSynchronize(ctx->ev[3]));
    ctx->sync();
    t.push_back(elapsed(ctx->ev[2], ctx->ev[3]));
    if (r == 0)
      counts = cur;
    else if (cur != counts)
      *agree = false;
  }
  return median(t);
}

static void batch_fill_info(tribit_batch* b, double ms, double ms_sched, double ms_cap, uint32_t launches,
                            uint64_t tasks, bool agree, const std::vector<unsigned long long>& counts,
                            tribit_batch_count_info* info) {
  if (!info) return;
  memset(info, 0, sizeof(*info));
  info->ms_count = ms;
  info->ms_sched = ms_sched;
  info->ms_capture = ms_cap;
  uint64_t tri = 0;
  for (auto c : counts) tri += c;
  info->triangles = tri;
  info->tasks = tasks;
  info->parts = (uint32_t)b->parts.size();
  info->launches = launches;
  info->repeats_agree = agree ? 1 : 0;
  info->tile_width = b->tile_width;
  if (!b->parts.empty() && b->parts[0].S) {
    const Plan& pl = b->parts[0].S->plan;
    info->nslots = pl.nslots;
    for (int i = 0; i < pl.nslots && i < 8; ++i) {
      info->group_width[i] = pl.g[i];
      info->waves[i] = b->waves[i];
    }
  }
}

void batch_count(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts& opts, uint64_t* counts,
                 tribit_batch_count_info* info) {
  const int reps = opts.repeats > 0 ? opts.repeats : 1;
  const double ms_sched = batch_schedules(ctx, b, opts.sched, opts.flags);
  double ms_cap = 0.0;
  if (!b->exec) ms_cap = batch_capture(ctx, b, mode_word(opts.flags), &b->launches, &b->tasks);
  const uint32_t launches = b->launches;
  const uint64_t tasks = b->tasks;
  std::vector<unsigned long long> c;
  bool agree = true;
  const double ms = batch_replay(ctx, b, reps, c, &agree);
  for (uint32_t j = 0; j < b->ngraphs; ++j) counts[j] = c[j];
  batch_fill_info(b, ms, ms_sched, ms_cap, launches, tasks, agree, c, info);
}

// Coordinate search: for every slot with tasks, the lane-group width, then
// the workgroups per CU, each kept when a replay of the recaptured graph is
// faster. Every configuration must give the same per-graph counts.
void batch_autotune(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts& opts, tribit_batch_count_info* info) {
  const double ms_sched = batch_schedules(ctx, b, opts.sched, opts.flags);
  const int att = mode_word(opts.flags);
  const int reps = opts.repeats > 0 ? opts.repeats : 3;
  std::vector<unsigned long long> ref, c;
  bool agree = true;
  uint32_t launches = 0;
  uint64_t tasks = 0;
  auto trial = [&]() {
    batch_capture(ctx, b, att, &launches, &tasks);
    const double ms = batch_replay(ctx, b, reps, c, &agree);
    if (!agree || (!ref.empty() && c != ref)) fail(TRIBIT_ERR_HIP, "batch counts differ between configurations");
    if (ref.empty()) ref = c;
    return ms;
  };
  double best = trial();
  const Plan& pl = b->parts.empty() ? Plan{} : b->parts[0].S->plan;
  const int nslots = b->parts.empty() ? 0 : pl.nslots;
  for (int i = 0; i < nslots; ++i) {
    uint64_t nt = 0;
    for (auto& p : b->parts) nt += p.S->tot.v[i];
    if (!nt) continue;
    // widths the slot's kernels have: multi-task waves (a part's large wave-64
    // launch) 1, 2 or 4 lanes per list; other waves 1 to 32; blocks 8 to 64
    bool multi = false;
    for (auto& p : b->parts) multi |= p.S->plan.kind[i] == kKindWaveMulti;
    const bool wave = pl.kind[i] == kKindWave64 || pl.kind[i] == kKindWave512;
    static const int kMultiW[6] = {1, 2, 4, 0, 0, 0}, kWaveW[6] = {1, 2, 4, 8, 16, 32},
                     kBlockW[6] = {8, 16, 32, 64, 0, 0};
    const int* gws = multi ? kMultiW : wave ? kWaveW : kBlockW;
    const int g0 = b->gw[i] ? b->gw[i] : b->parts[0].S->gdef[i];
    int bestg = g0;
    for (int q = 0; q < 6; ++q) {
      const int gw = gws[q];
      if (gw == 0 || gw == g0) continue;
      b->gw[i] = gw;
      const double ms = trial();
      if (ms < best) {
        best = ms;
        bestg = gw;
      }
    }
    b->gw[i] = bestg;
    const int w0 = b->waves[i];
    int bestw = w0;
    for (int wv : {1, 2, 4, 0}) {
      if (wv == w0) continue;
      b->waves[i] = wv;
      const double ms = trial();
      if (ms < best) {
        best = ms;
        bestw = wv;
      }
    }
    b->waves[i] = bestw;
  }
  // tile width: the block launches' lists streamed as 32- or 64-column bit
  // tiles instead of ids (tried when a part has block tasks)
  bool blocks = false;
  for (auto& p : b->parts)
    for (int i = 0; i < p.S->plan.nslots; ++i)
      blocks |= p.S->tot.v[i] && (p.S->kdef[i] == kKindBlock || p.S->kdef[i] == kKindTile);
  if (blocks) {
    const int t0 = b->tile_width;
    int bestt = t0;
    for (int tw : {32, 64, 0}) {
      if (tw == t0) continue;
      b->tile_width = tw;
      const double ms = trial();
      if (ms < best) {
        best = ms;
        bestt = tw;
      }
    }
    b->tile_width = bestt;
  }
  const double ms_cap = batch_capture(ctx, b, att, &launches, &tasks);
  const double ms = batch_replay(ctx, b, reps, c, &agree);
  if (!agree || (!ref.empty() && c != ref)) fail(TRIBIT_ERR_HIP, "batch counts differ between configurations");
  batch_fill_info(b, ms, ms_sched, ms_cap, launches, tasks, agree, c, info);
}

}  // namespace tb
