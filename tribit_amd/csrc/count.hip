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
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kBlock = 256;
constexpr uint32_t kWaveBatch = 16;      // wave tasks taken per queue atomic
constexpr uint32_t kChunk = 2048;        // probe-set chunk of class-3 rows
constexpr uint64_t kWaveCap = 1u << 15;  // target streamed elements per wave task
constexpr uint64_t kBlockCap = 1u << 18; // target streamed elements per block task
constexpr int kVariants = 4;             // one kernel launch per row class

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
  const uint32_t* T;
  uint32_t a;
  __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
    const uint32_t h = x * a;
    return (uint32_t)((T[slot_a<LOGS>(h)] == x) | (T[slot_b<LOGS>(h)] == x));
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
  const uint32_t* T;
  __device__ __forceinline__ uint32_t operator()(uint32_t w) const {
    uint32_t h = hmul<LOGS2>(w, 0x9E3779B1u);
    for (;;) {
      const uint32_t k = T[h];
      if (k == w) return 1u;
      if (k == kEmpty) return 0u;
      h = (h + 1) & ((1u << LOGS2) - 1);
    }
  }
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

struct ColSrc {
  const uint32_t* col;
  __amdgpu_buffer_rsrc_t rsrc;  // raw buffer over col (valid when col bytes < 2^32)
};

__device__ __forceinline__ ColSrc make_col_src(const uint32_t* col, uint64_t m) {
  ColSrc c;
  c.col = col;
  const uint64_t bytes = m * 4 < 0xFFFFFFFFull ? m * 4 : 0xFFFFFFFFull;
  c.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(col), (short)0, (int)(uint32_t)bytes, 0x00020000);
  return c;
}

template <int G, int U, bool SHARED, bool BUF, typename Probe>
__device__ __forceinline__ uint32_t stream_ranges(const ColSrc& src, const uint2* R, uint32_t nk, uint32_t* s_next,
                                                  uint32_t lane, const Probe& probe) {
  constexpr uint64_t kLeads = leader_mask<G>();
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint32_t base = 0, end = 0, cnt = 0, next = 0;
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
      const uint32_t k = kb + (uint32_t)__popcll(want & below);
      if (k < nk) {
        const uint2 r = R[k];
        base = r.x;
        end = r.y;
      }
    }
    const uint32_t p0 = base + gl;
    const int rem = (int)(end - p0);  // elements left for this lane's first load (may be <= 0)
    uint32_t x[U];
    if (BUF) {
      // an invalid element gets an offset past the buffer: no memory access, reads 0
      const uint32_t off = p0 * 4u;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const bool ok = j * G < rem;
        x[j] = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, (int)(ok ? off + (uint32_t)(j * G * 4) : 0xFFFFFFFFu), 0,
                                                    0);
        x[j] = ok ? x[j] : kNoKey;
      }
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) x[j] = (j * G < rem) ? src.col[p0 + j * G] : kNoKey;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) cnt += probe(x[j]);
    base += (uint32_t)(G * U);
  }
  return cnt;
}

// ------------------------------------------------------ wave-level tasks
// A task (x, rbase, rcount, chunk) pairs the set N+(x) with the lists whose
// ranges are rng[rbase, rbase + rcount). A wave owns one task: it builds the
// set in its LDS slice once, then stages the ranges KMAX at a time.

template <int LOGS, int KMAX, int G, int U, bool BUF>
__global__ __launch_bounds__(256) void k_count_wave(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, uint64_t m,
                                                    const uint2* __restrict__ rng,
                                                    const uint4* __restrict__ tasks, uint32_t ntasks,
                                                    uint32_t* __restrict__ queue,
                                                    unsigned long long* __restrict__ result, int attempts) {
  constexpr uint32_t S = 1u << LOGS;
  const ColSrc src = make_col_src(col, m);
  __shared__ uint32_t sT[4][2 * S];
  __shared__ uint2 sR[4][KMAX];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t* T = sT[wv];
  uint2* R = sR[wv];
  uint64_t acc = 0;
  uint32_t fallbacks = 0;
  for (;;) {
    uint32_t tb = 0;
    if (lane == 0) tb = atomicAdd(queue, kWaveBatch);
    tb = __shfl(tb, 0);
    if (tb >= ntasks) break;
    const uint32_t tend = min(tb + kWaveBatch, ntasks);
    for (uint32_t t = tb; t < tend; ++t) {
      const uint4 tk = tasks[t];
      const uint32_t x = tk.x, rbase = tk.y, rcount = tk.z;
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
    }
  }
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

template <int LOGS, int G, int U, bool BUF>
__global__ __launch_bounds__(kBlockThreads) void k_count_block(const uint32_t* __restrict__ row_ptr,
                                                               const uint32_t* __restrict__ col, uint64_t m,
                                                               const uint2* __restrict__ rng,
                                                               const uint4* __restrict__ tasks, uint32_t ntasks,
                                                               uint32_t* __restrict__ queue,
                                                               unsigned long long* __restrict__ result,
                                                               int attempts) {
  constexpr uint32_t S = 1u << LOGS;
  __shared__ uint32_t T[2 * S];
  __shared__ uint2 R[kChunk];
  __shared__ uint32_t s_task, s_next, s_fail;
  const uint32_t lane = threadIdx.x & 63;
  const ColSrc src = make_col_src(col, m);
  uint64_t acc = 0;
  uint32_t fallbacks = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t t = s_task;
    if (t >= ntasks) break;
    const uint4 tk = tasks[t];
    const uint32_t x = tk.x, rbase = tk.y, rcount = tk.z, c = tk.w;
    const uint32_t r0 = row_ptr[x], d = row_ptr[x + 1] - r0;
    const uint32_t p0 = c * kChunk, p1 = min(p0 + kChunk, d);
    const bool ranged = (p0 > 0) || (p1 < d);
    const uint32_t lo = col[r0 + p0], hi = col[r0 + p1 - 1];
    uint32_t sa = 0;
    bool ok = false;
    for (int attempt = 0; attempt < attempts && !ok; ++attempt) {
      sa = kSeedA[attempt];
      for (uint32_t i = threadIdx.x; i < 2 * S; i += kBlockThreads) T[i] = kEmpty;
      if (threadIdx.x == 0) s_fail = 0;
      __syncthreads();
      bool fail = false;
      for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBlockThreads)
        fail |= !cuckoo_insert<LOGS>(T, col[r0 + i], sa);
      if (fail) s_fail = 1;
      __syncthreads();
      ok = (s_fail == 0);
      __syncthreads();
    }
    if (!ok) {
      ++fallbacks;
      for (uint32_t i = threadIdx.x; i < 2 * S; i += kBlockThreads) T[i] = kEmpty;
      __syncthreads();
      for (uint32_t i = p0 + threadIdx.x; i < p1; i += kBlockThreads) linear_insert<LOGS + 1>(T, col[r0 + i]);
    }
    uint32_t cnt = 0;
    for (uint32_t c0 = 0; c0 < rcount; c0 += kChunk) {
      const uint32_t nk = min((uint32_t)kChunk, rcount - c0);
      for (uint32_t k = threadIdx.x; k < nk; k += kBlockThreads) {
        uint2 r = rng[rbase + c0 + k];
        if (ranged) {
          r.x = lower_bound_u32(col, r.x, r.y, lo);
          r.y = upper_bound_u32(col, r.x, r.y, hi);
        }
        R[k] = r;
      }
      if (threadIdx.x == 0) s_next = 0;
      __syncthreads();
      cnt += ok ? stream_ranges<G, U, true, BUF>(src, R, nk, &s_next, lane, CuckooProbe<LOGS>{T, sa})
                : stream_ranges<G, U, true, BUF>(src, R, nk, &s_next, lane, LinearProbe<LOGS + 1>{T});
      __syncthreads();
    }
    acc += cnt;
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
  if (threadIdx.x == 0 && fallbacks) atomicAdd(result + 1, (unsigned long long)fallbacks);
}

// ------------------------------------------------------------- scheduler
// A schedule is a set of owner lists: (x, rbase, rcount, work) where the
// ranges rng[rbase, rbase+rcount) are streamed against N+(x).
//   PLAIN:    one forward list per x over rng_all (the textbook edge
//             iterator: N+(u) is the set, every N+(v) is streamed).
//   DIRECTED: per x its forward list and its reverse list over rng_dir
//             (every edge streams the shorter of its two lists).
// The owner's outdeg picks the class (kernel variant); its work (streamed
// elements) picks the number of tasks the list is split into.

__device__ __forceinline__ int row_class(uint32_t d) {
  if (d == 0) return -1;
  if (d <= 32) return 0;
  if (d <= 256) return 1;
  if (d <= kChunk) return 2;
  return 3;
}

__device__ __forceinline__ uint32_t list_splits(int cls, uint32_t count, uint64_t w) {
  const uint64_t cap = cls <= 1 ? kWaveCap : kBlockCap;
  uint64_t s = (w + cap - 1) / cap;
  if (s < 1) s = 1;
  if (s > count) s = count;
  return (uint32_t)s;
}

struct SchedView {
  const uint32_t* row_ptr;
  const uint32_t* fw_ptr;  // list 0 offsets (PLAIN: row_ptr)
  const uint32_t* rv_ptr;  // list 1 offsets (PLAIN: unused)
  const uint64_t* w_fw;    // list 0 work (PLAIN: wrow)
  const uint64_t* w_rv;
  int nlists;              // 1 (PLAIN) or 2 (DIRECTED)
};

__device__ __forceinline__ void owner_list(const SchedView& v, uint64_t x, int l, uint32_t* base, uint32_t* count,
                                           uint64_t* work) {
  const uint32_t* p = (l == 0) ? v.fw_ptr : v.rv_ptr;
  *base = p[x];
  *count = p[x + 1] - p[x];
  *work = (l == 0) ? v.w_fw[x] : v.w_rv[x];
}

__device__ __forceinline__ bool list_useful(int l, uint32_t d, uint32_t count) {
  // a forward list needs |N+(x)| >= 2 (its partner v is in N+(x) but not in N+(v))
  return count > 0 && d >= (l == 0 ? 2u : 1u);
}

// Per owner: number of tasks in its class slot (uint4 = 4 class counters).
__global__ void k_sched_count(SchedView sv, uint64_t n, uint4* __restrict__ cnt) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x > n) return;
  uint32_t c4[4] = {0, 0, 0, 0};
  if (x < n) {
    const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
    const int cls = row_class(d);
    if (cls >= 0) {
      const uint32_t nch = (cls == 3) ? (d + kChunk - 1) / kChunk : 1u;
      for (int l = 0; l < sv.nlists; ++l) {
        uint32_t b, k;
        uint64_t w;
        owner_list(sv, x, l, &b, &k, &w);
        if (list_useful(l, d, k)) c4[cls] += list_splits(cls, k, w) * nch;
      }
    }
  }
  cnt[x] = make_uint4(c4[0], c4[1], c4[2], c4[3]);
}

// Writes each owner's tasks at its scanned offsets: wave tasks (classes 0, 1)
// into wt (class 1 after all class-0 tasks), block tasks (2, 3) into bt.
__global__ void k_sched_fill(SchedView sv, uint64_t n, const uint4* __restrict__ off, uint4 tot,
                             uint4* __restrict__ wt, uint4* __restrict__ bt) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
  const int cls = row_class(d);
  if (cls < 0) return;
  const uint4 o = off[x];
  uint4* out = (cls == 0) ? wt + o.x : (cls == 1) ? wt + tot.x + o.y : (cls == 2) ? bt + o.z : bt + tot.z + o.w;
  const uint32_t nch = (cls == 3) ? (d + kChunk - 1) / kChunk : 1u;
  uint32_t j = 0;
  for (int l = 0; l < sv.nlists; ++l) {
    uint32_t b, k;
    uint64_t w;
    owner_list(sv, x, l, &b, &k, &w);
    if (!list_useful(l, d, k)) continue;
    const uint32_t s = list_splits(cls, k, w);
    for (uint32_t c = 0; c < nch; ++c)
      for (uint32_t i = 0; i < s; ++i) {
        const uint32_t k0 = (uint32_t)((uint64_t)k * i / s), k1 = (uint32_t)((uint64_t)k * (i + 1) / s);
        out[j++] = make_uint4((uint32_t)x, b + k0, k1 - k0, c);
      }
  }
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

struct Uint4Plus {
  __host__ __device__ uint4 operator()(const uint4& a, const uint4& b) const {
    return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
  }
};

static SchedView sched_view(const tribit_graph* g, int sched) {
  SchedView v;
  v.row_ptr = g->row_ptr;
  if (sched == TRIBIT_SCHED_PLAIN) {
    v.fw_ptr = g->row_ptr;
    v.rv_ptr = g->row_ptr;
    v.w_fw = g->wrow;
    v.w_rv = g->wrow;
    v.nlists = 1;
  } else {
    v.fw_ptr = g->fw_ptr;
    v.rv_ptr = g->rv_ptr;
    v.w_fw = g->w_fw;
    v.w_rv = g->w_rv;
    v.nlists = 2;
  }
  return v;
}

static const uint2* sched_ranges(const tribit_graph* g, int sched) {
  return sched == TRIBIT_SCHED_PLAIN ? g->rng_all : g->rng_dir;
}

// Kernel variants: LDS set size (LOGS), staged ranges per batch (KMAX), lane
// group width G and loads in flight per lane U. In the DIRECTED schedule a
// streamed list is never longer than its owner's list, so class 0 streams
// lists of at most 32 elements and uses narrow groups.
constexpr int kC0Logs = 6, kC0K = 64;    // class 0: d <= 32, 2x64 slots
constexpr int kC1Logs = 9, kC1K = 256;   // class 1: d <= 256, 2x512 slots
constexpr int kC2Logs = 12;              // classes 2,3: <= 2048 ids, 2x4096 slots
constexpr int kU = 8;
constexpr int kForceGlobal = 1 << 8;  // launch-time flag folded into `attempts` (stripped before the kernel)

template <int LOGS, int KMAX, int G, int U, bool BUF>
static void launch_wave_t(uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                          const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                          uint64_t* blocks) {
  int per_cu = 0;
  TB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count_wave<LOGS, KMAX, G, U, BUF>, kBlock, 0));
  if (per_cu < 1) per_cu = 1;
  const uint64_t want = (ntasks + 4 * kWaveBatch - 1) / (4 * kWaveBatch);
  const unsigned grid = (unsigned)std::min<uint64_t>(want, (uint64_t)cus * per_cu);
  k_count_wave<LOGS, KMAX, G, U, BUF><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, g->m, rng, tasks, ntasks, queue,
                                                             result, attempts & (kForceGlobal - 1));
  *blocks += grid;
}

template <int LOGS, int KMAX, int G>
static void launch_wave_g(uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                          const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                          uint64_t* blocks) {
  if (g->m * 4 < 0xFFFFFFFFull && !(attempts & kForceGlobal))
    launch_wave_t<LOGS, KMAX, G, kU, true>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks);
  else
    launch_wave_t<LOGS, KMAX, G, kU, false>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks);
}

template <int LOGS, int KMAX>
static void launch_wave(int gw, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                        const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                        uint64_t* blocks) {
  if (!ntasks) return;
  switch (gw) {
    case 4: launch_wave_g<LOGS, KMAX, 4>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    case 8: launch_wave_g<LOGS, KMAX, 8>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    case 32: launch_wave_g<LOGS, KMAX, 32>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    default: launch_wave_g<LOGS, KMAX, 16>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
  }
}

template <int LOGS, int G, int U, bool BUF>
static void launch_block_t(uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                           const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                           uint64_t* blocks) {
  int per_cu = 0;
  TB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_count_block<LOGS, G, U, BUF>, kBlockThreads, 0));
  if (per_cu < 1) per_cu = 1;
  const unsigned grid = (unsigned)std::min<uint64_t>(ntasks, (uint64_t)cus * per_cu);
  k_count_block<LOGS, G, U, BUF><<<grid, kBlockThreads, 0, s>>>(g->row_ptr, g->col, g->m, rng, tasks, ntasks, queue,
                                                               result, attempts & (kForceGlobal - 1));
  *blocks += grid;
}

template <int LOGS, int G>
static void launch_block_g(uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                           const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                           uint64_t* blocks) {
  if (g->m * 4 < 0xFFFFFFFFull && !(attempts & kForceGlobal))
    launch_block_t<LOGS, G, kU, true>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks);
  else
    launch_block_t<LOGS, G, kU, false>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks);
}

template <int LOGS>
static void launch_block(int gw, uint32_t ntasks, int cus, hipStream_t s, tribit_graph* g, const uint2* rng,
                         const uint4* tasks, uint32_t* queue, unsigned long long* result, int attempts,
                         uint64_t* blocks) {
  if (!ntasks) return;
  switch (gw) {
    case 4: launch_block_g<LOGS, 4>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    case 8: launch_block_g<LOGS, 8>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    case 32: launch_block_g<LOGS, 32>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    case 64: launch_block_g<LOGS, 64>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
    default: launch_block_g<LOGS, 16>(ntasks, cus, s, g, rng, tasks, queue, result, attempts, blocks); break;
  }
}

// Default lane-group widths per variant and schedule.
// (swept on R-MAT scale 22 and 25 over 4..32 / 8..64 lanes per list)
static void default_widths(int sched, int gw[4]) {
  if (sched == TRIBIT_SCHED_PLAIN) {
    gw[0] = 32, gw[1] = 16, gw[2] = 32, gw[3] = 16;
  } else {
    gw[0] = 4, gw[1] = 8, gw[2] = 16, gw[3] = 16;
  }
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  TB_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// One scheduled count. Returns the triangle count; fills per-phase times.
static uint64_t count_once(tribit_ctx* ctx, tribit_graph* g, int sched, int flags, const int gw[4],
                           double* ms_total, double* ms_sched, double* ms_kernel, double* ms_var,
                           tribit_count_info* info) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n;
  TB_HIP(hipEventRecord(ctx->ev[2], s));
  // scratch: [0] count, [1] fallback tasks, [4..7] four u32 queue counters
  DevBuf<unsigned long long> scratch(ctx, 8);
  k_zero_u64<<<1, 64, 0, s>>>(scratch.p, 8);
  unsigned long long* d_result = scratch.p;
  uint32_t* d_queue = reinterpret_cast<uint32_t*>(scratch.p + 4);

  DevBuf<uint4> cnt(ctx, n + 1), off(ctx, n + 1);
  const SchedView sv = sched_view(g, sched);
  const uint2* rng = sched_ranges(g, sched);
  k_sched_count<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(sv, n, cnt.p);
  {
    size_t bytes = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt.p, off.p, make_uint4(0, 0, 0, 0), (size_t)(n + 1),
                                   Uint4Plus(), s));
    DevBuf<uint8_t> tmp(ctx, bytes ? bytes : 8);
    TB_HIP(rocprim::exclusive_scan(tmp.p, bytes, cnt.p, off.p, make_uint4(0, 0, 0, 0), (size_t)(n + 1), Uint4Plus(),
                                   s));
  }
  uint4 tot;  // off[n] = totals (cnt[n] is zero)
  TB_HIP(hipMemcpyAsync(&tot, off.p + n, sizeof(uint4), hipMemcpyDeviceToHost, s));
  ctx->sync();
  const uint64_t nw = (uint64_t)tot.x + tot.y, nb = (uint64_t)tot.z + tot.w;
  DevBuf<uint4> wt(ctx, nw ? nw : 1);
  DevBuf<uint4> bt(ctx, nb ? nb : 1);
  k_sched_fill<<<grid_for(n, kBlock), kBlock, 0, s>>>(sv, n, off.p, tot, wt.p, bt.p);
  cnt.reset();
  off.reset();
  TB_HIP(hipEventRecord(ctx->ev[3], s));

  const int cus = ctx->cus > 0 ? ctx->cus : 256;
  uint64_t blocks = 0;
  // one launch per variant; an event after each gives its device time
  const int att = ((flags & TRIBIT_COUNT_FORCE_LINEAR_PROBE) ? 0 : kCuckooSeeds) |
                  ((flags & TRIBIT_COUNT_FORCE_GLOBAL_LOADS) ? kForceGlobal : 0);
  launch_wave<kC0Logs, kC0K>(gw[0], tot.x, cus, s, g, rng, wt.p, d_queue + 0, d_result, att, &blocks);
  TB_HIP(hipEventRecord(ctx->ev[4], s));
  launch_wave<kC1Logs, kC1K>(gw[1], tot.y, cus, s, g, rng, wt.p + tot.x, d_queue + 1, d_result, att, &blocks);
  TB_HIP(hipEventRecord(ctx->ev[5], s));
  launch_block<kC2Logs>(gw[2], tot.z, cus, s, g, rng, bt.p, d_queue + 2, d_result, att, &blocks);
  TB_HIP(hipEventRecord(ctx->ev[6], s));
  launch_block<kC2Logs>(gw[3], tot.w, cus, s, g, rng, bt.p + tot.z, d_queue + 3, d_result, att, &blocks);
  TB_HIP(hipGetLastError());
  TB_HIP(hipEventRecord(ctx->ev[7], s));
  unsigned long long res[2] = {0, 0};
  TB_HIP(hipMemcpyAsync(res, d_result, sizeof(res), hipMemcpyDeviceToHost, s));
  TB_HIP(hipEventSynchronize(ctx->ev[7]));
  ctx->sync();
  *ms_sched = elapsed(ctx->ev[2], ctx->ev[3]);
  *ms_kernel = elapsed(ctx->ev[3], ctx->ev[7]);
  *ms_total = elapsed(ctx->ev[2], ctx->ev[7]);
  for (int v = 0; v < kVariants; ++v) ms_var[v] = elapsed(ctx->ev[3 + v], ctx->ev[4 + v]);
  if (info) {
    info->launch_blocks = blocks;
    info->tasks = nw + nb;
    info->tasks_by_variant[0] = tot.x;
    info->tasks_by_variant[1] = tot.y;
    info->tasks_by_variant[2] = tot.z;
    info->tasks_by_variant[3] = tot.w;
    info->nvariants = kVariants;
    info->fallback_tasks = res[1];
    for (int i = 0; i < 4; ++i) info->group_width[i] = gw[i];
  }
  return res[0];
}

static double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t k = v.size();
  return k ? ((k & 1) ? v[k / 2] : 0.5 * (v[k / 2 - 1] + v[k / 2])) : 0.0;
}

void count(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts& opts, tribit_count_info* info) {
  const int reps = opts.repeats > 0 ? opts.repeats : 1;
  std::vector<double> tt, ts, tk, tv[kVariants];
  uint64_t first = 0;
  bool agree = true;
  tribit_count_info local;
  memset(&local, 0, sizeof(local));
  int gw[4];
  default_widths(opts.sched, gw);
  for (int i = 0; i < 4; ++i)
    if (opts.group_width[i]) gw[i] = opts.group_width[i];
  for (int r = 0; r < reps; ++r) {
    double a, b, c, v[kVariants];
    const uint64_t tri = count_once(ctx, g, opts.sched, opts.flags, gw, &a, &b, &c, v, &local);
    if (r == 0)
      first = tri;
    else if (tri != first)
      agree = false;
    tt.push_back(a);
    ts.push_back(b);
    tk.push_back(c);
    for (int i = 0; i < kVariants; ++i) tv[i].push_back(v[i]);
  }
  local.triangles = first;
  local.ms_count = median(tt);
  local.ms_sched = median(ts);
  local.ms_kernel = median(tk);
  for (int i = 0; i < kVariants; ++i) local.ms_by_variant[i] = median(tv[i]);
  local.repeats_agree = agree ? 1 : 0;
  if (info) *info = local;
}

// Per-variant work of a schedule: owner lists, partners (edges), streamed
// elements (wedge checks).
__global__ void k_variant_stats(SchedView sv, uint64_t n, unsigned long long* __restrict__ st) {
  __shared__ unsigned long long acc[3 * kVariants];
  if (threadIdx.x < 3 * kVariants) acc[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) {
    const uint32_t d = sv.row_ptr[x + 1] - sv.row_ptr[x];
    const int c = row_class(d);
    if (c >= 0) {
      for (int l = 0; l < sv.nlists; ++l) {
        uint32_t b, k;
        uint64_t w;
        owner_list(sv, x, l, &b, &k, &w);
        if (!list_useful(l, d, k)) continue;
        atomicAdd(&acc[c], 1ull);
        atomicAdd(&acc[kVariants + c], (unsigned long long)k);
        atomicAdd(&acc[2 * kVariants + c], (unsigned long long)w);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 * kVariants && acc[threadIdx.x]) atomicAdd(&st[threadIdx.x], acc[threadIdx.x]);
}

void variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out) {
  memset(out, 0, sizeof(*out));
  DevBuf<unsigned long long> st(ctx, 3 * kVariants);
  k_zero_u64<<<1, 64, 0, ctx->stream>>>(st.p, 3 * kVariants);
  if (g->n)
    k_variant_stats<<<grid_for(g->n, kBlock), kBlock, 0, ctx->stream>>>(sched_view(g, sched), g->n, st.p);
  unsigned long long h[3 * kVariants];
  TB_HIP(hipMemcpyAsync(h, st.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  for (int i = 0; i < kVariants; ++i) {
    out->rows[i] = h[i];
    out->edges[i] = h[kVariants + i];
    out->wplus[i] = h[2 * kVariants + i];
  }
}

}  // namespace tb
