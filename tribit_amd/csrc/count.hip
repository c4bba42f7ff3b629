// count.hip -- scheduler + counting kernels (gfx950, wave64).
//
// Replaces compute_binary_popcount_mma_9 and its launch,
// src/single_gpu/tribit.cu:519-586 and :765-802. The reference forms, per
// (U word, 32-neighbour window), one b1 tensor-core AND+POPC; here the
// count is the degree-oriented edge iterator
//     triangles = sum over u, v in N+(u) of |N+(u) & N+(v)|
// with N+(u) staged once per task in an LDS hash set and every N+(v)
// streamed from HBM with coalesced int32 loads, one LDS probe per element.
//
// Scheduler (plain): rows with outdeg d >= 2 become tasks; the class of a
// row picks the kernel variant (LDS table size, wave- or block-level task)
// and its wedge count W+(u) = sum of outdeg(v) picks the work split.
//   class 0: d <= 32      wave task, 128-slot table        (k_count_wave<7>)
//   class 1: d <= 256     wave task, 1024-slot table       (k_count_wave<10>)
//   class 2: d <= 2048    block task, 8192-slot table      (k_count_block<13>)
//   class 3: d >  2048    block task per 2048-wide chunk of N+(u); each N+(v)
//                         is cut to the chunk's id range by binary search.
#include <cstring>
#include <algorithm>
#include <rocprim/device/device_scan.hpp>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kBlock = 256;
constexpr uint32_t kHashMul = 0x9E3779B1u;
constexpr uint32_t kWaveBatch = 16;      // wave tasks taken per queue atomic
constexpr uint32_t kChunk = 2048;        // probe-set chunk of class-3 rows
constexpr uint64_t kWaveCap = 1u << 15;  // target streamed elements per wave task
constexpr uint64_t kBlockCap = 1u << 18; // target streamed elements per block task
constexpr int kVariants = 4;             // one kernel launch per row class

// ------------------------------------------------------------ LDS hash set

template <int LOGS>
__device__ __forceinline__ uint32_t hslot(uint32_t w) {
  return (w * kHashMul) >> (32 - LOGS);
}

template <int LOGS>
__device__ __forceinline__ void hash_insert(uint32_t* T, uint32_t w) {
  uint32_t h = hslot<LOGS>(w);
  while (atomicCAS(&T[h], kEmpty, w) != kEmpty) h = (h + 1) & ((1u << LOGS) - 1);
}

template <int LOGS>
__device__ __forceinline__ uint32_t hash_probe(const uint32_t* T, uint32_t w) {
  uint32_t h = hslot<LOGS>(w);
  for (;;) {
    const uint32_t k = T[h];
    if (k == w) return 1u;
    if (k == kEmpty) return 0u;
    h = (h + 1) & ((1u << LOGS) - 1);
  }
}

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

// ------------------------------------------------------ wave-level tasks
// A wave owns one task (u, k0, k1): table of N+(u), then the v = N+(u)[k]
// for k in [k0,k1) are streamed by 64/G lane groups of G lanes; a group that
// runs dry takes the next v (ballot rank among the dry groups).

template <int LOGS, int G>
__global__ __launch_bounds__(256) void k_count_wave(const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col,
                                                    const uint2* __restrict__ tasks, uint32_t ntasks,
                                                    uint32_t* __restrict__ queue,
                                                    unsigned long long* __restrict__ result) {
  constexpr uint32_t S = 1u << LOGS;
  constexpr uint64_t kLeads = leader_mask<G>();
  __shared__ uint32_t lds[4 * S];
  const uint32_t lane = threadIdx.x & 63;
  uint32_t* T = lds + (threadIdx.x >> 6) * S;
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint64_t acc = 0;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(queue, kWaveBatch);
    base = __shfl(base, 0);
    if (base >= ntasks) break;
    const uint32_t tend = min(base + kWaveBatch, ntasks);
    for (uint32_t t = base; t < tend; ++t) {
      const uint2 tk = tasks[t];
      const uint32_t u = tk.x, k0 = tk.y & 0xFFFFu, k1 = tk.y >> 16;
      const uint32_t r0 = row_ptr[u], d = row_ptr[u + 1] - r0;
      for (uint32_t i = lane; i < S; i += 64) T[i] = kEmpty;
      wave_lds_sync();
      for (uint32_t i = lane; i < d; i += 64) hash_insert<LOGS>(T, col[r0 + i]);
      wave_lds_sync();
      uint32_t pos = 0, end = 0, cnt = 0, next = k0;
      for (;;) {
        const bool need = pos >= end;  // uniform within a group
        const uint64_t want = __ballot(need && gl == 0);
        if (want == kLeads && next >= k1) break;
        if (need) {
          const uint32_t k = next + (uint32_t)__popcll(want & below);
          if (k < k1) {
            const uint32_t v = col[r0 + k];
            pos = row_ptr[v];
            end = row_ptr[v + 1];
          }
        }
        next = min(next + (uint32_t)__popcll(want), k1);
        const uint32_t p = pos + gl;
        if (p < end) cnt += hash_probe<LOGS>(T, col[p]);
        pos += G;
      }
      acc += cnt;
      wave_lds_sync();
    }
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
}

// ----------------------------------------------------- block-level tasks
// A 256-thread block owns one task (u, k0, k1, chunk c): the probe set is
// N+(u)[c*kChunk, ...) in one shared table; its 4 waves share the v range
// through an LDS counter. When the row has more than one chunk, each N+(v)
// is cut to the chunk's id range [lo, hi] (both lists are sorted).

template <int LOGS, int G>
__global__ __launch_bounds__(256) void k_count_block(const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col,
                                                     const uint4* __restrict__ tasks, uint32_t ntasks,
                                                     uint32_t* __restrict__ queue,
                                                     unsigned long long* __restrict__ result) {
  constexpr uint32_t S = 1u << LOGS;
  constexpr uint64_t kLeads = leader_mask<G>();
  __shared__ uint32_t T[S];
  __shared__ uint32_t s_task, s_next;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t g = lane / G, gl = lane % G;
  const uint64_t below = (g == 0) ? 0ull : ((1ull << (g * G)) - 1);
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t t = s_task;
    if (t >= ntasks) break;
    const uint4 tk = tasks[t];
    const uint32_t u = tk.x, k0 = tk.y, k1 = tk.z, c = tk.w;
    const uint32_t r0 = row_ptr[u], d = row_ptr[u + 1] - r0;
    const uint32_t p0 = c * kChunk, p1 = min(p0 + kChunk, d);
    const bool ranged = (p0 > 0) || (p1 < d);
    const uint32_t lo = col[r0 + p0], hi = col[r0 + p1 - 1];
    for (uint32_t i = threadIdx.x; i < S; i += 256) T[i] = kEmpty;
    if (threadIdx.x == 0) s_next = k0;
    __syncthreads();
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += 256) hash_insert<LOGS>(T, col[r0 + i]);
    __syncthreads();
    uint32_t pos = 0, end = 0, cnt = 0;
    for (;;) {
      const bool need = pos >= end;
      const uint64_t want = __ballot(need && gl == 0);
      const uint32_t nw = (uint32_t)__popcll(want);
      uint32_t kb = 0;
      if (nw) {
        if (lane == 0) kb = atomicAdd(&s_next, nw);
        kb = __shfl(kb, 0);
        if (want == kLeads && kb >= k1) break;
      }
      if (need) {
        const uint32_t k = kb + (uint32_t)__popcll(want & below);
        if (k < k1) {
          const uint32_t v = col[r0 + k];
          pos = row_ptr[v];
          end = row_ptr[v + 1];
          if (ranged) {
            pos = lower_bound_u32(col, pos, end, lo);
            end = upper_bound_u32(col, pos, end, hi);
          }
        }
      }
      const uint32_t p = pos + gl;
      if (p < end) cnt += hash_probe<LOGS>(T, col[p]);
      pos += G;
    }
    acc += cnt;
    __syncthreads();
  }
  acc = wave_sum_u64(acc);
  if (lane == 0 && acc) atomicAdd(result, (unsigned long long)acc);
}

// ------------------------------------------------------------- scheduler

__device__ __forceinline__ int row_class(uint32_t d) {
  if (d < 2) return -1;
  if (d <= 32) return 0;
  if (d <= 256) return 1;
  if (d <= kChunk) return 2;
  return 3;
}

__device__ __forceinline__ uint32_t row_splits(int cls, uint32_t d, uint64_t w) {
  const uint64_t cap = cls <= 1 ? kWaveCap : kBlockCap;
  uint64_t s = (w + cap - 1) / cap;
  if (s < 1) s = 1;
  if (s > d) s = d;
  return (uint32_t)s;
}

// Per row: number of tasks in its class slot (uint4 = 4 class counters).
__global__ void k_sched_count(const uint32_t* __restrict__ row_ptr, const uint64_t* __restrict__ wrow, uint64_t n,
                              uint4* __restrict__ cnt) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  uint4 c = make_uint4(0, 0, 0, 0);
  if (u < n) {
    const uint32_t d = row_ptr[u + 1] - row_ptr[u];
    const int cls = row_class(d);
    if (cls >= 0) {
      uint32_t t = row_splits(cls, d, wrow[u]);
      if (cls == 3) t *= (d + kChunk - 1) / kChunk;
      if (cls == 0) c.x = t;
      else if (cls == 1) c.y = t;
      else if (cls == 2) c.z = t;
      else c.w = t;
    }
  }
  cnt[u] = c;
}

// Writes each row's tasks at its scanned offsets. Wave tasks (classes 0,1)
// go to wtasks (class 1 after all class-0 tasks), block tasks (2,3) to btasks.
__global__ void k_sched_fill(const uint32_t* __restrict__ row_ptr, const uint64_t* __restrict__ wrow, uint64_t n,
                             const uint4* __restrict__ off, uint4 tot, uint2* __restrict__ wtasks,
                             uint4* __restrict__ btasks) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const uint32_t d = row_ptr[u + 1] - row_ptr[u];
  const int cls = row_class(d);
  if (cls < 0) return;
  const uint32_t s = row_splits(cls, d, wrow[u]);
  const uint4 o = off[u];
  if (cls <= 1) {
    const uint32_t b = (cls == 0) ? o.x : tot.x + o.y;
    for (uint32_t i = 0; i < s; ++i) {
      const uint32_t k0 = (uint32_t)((uint64_t)d * i / s), k1 = (uint32_t)((uint64_t)d * (i + 1) / s);
      wtasks[b + i] = make_uint2((uint32_t)u, k0 | (k1 << 16));
    }
  } else {
    const uint32_t nch = (cls == 3) ? (d + kChunk - 1) / kChunk : 1u;
    const uint32_t b = (cls == 2) ? o.z : tot.z + o.w;
    for (uint32_t c = 0; c < nch; ++c)
      for (uint32_t i = 0; i < s; ++i) {
        const uint32_t k0 = (uint32_t)((uint64_t)d * i / s), k1 = (uint32_t)((uint64_t)d * (i + 1) / s);
        btasks[b + c * s + i] = make_uint4((uint32_t)u, k0, k1, c);
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

static float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  TB_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// One scheduled count. Returns the triangle count; fills per-phase times.
static uint64_t count_once(tribit_ctx* ctx, tribit_graph* g, double* ms_total, double* ms_sched, double* ms_kernel,
                           double* ms_var, tribit_count_info* info) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n;
  TB_HIP(hipEventRecord(ctx->ev[2], s));
  // scratch: result + 4 queue counters live in one 64-byte block
  DevBuf<unsigned long long> scratch(ctx, 8);
  k_zero_u64<<<1, 64, 0, s>>>(scratch.p, 8);
  unsigned long long* d_result = scratch.p;
  uint32_t* d_queue = reinterpret_cast<uint32_t*>(scratch.p + 4);

  DevBuf<uint4> cnt(ctx, n + 1), off(ctx, n + 1);
  k_sched_count<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(g->row_ptr, g->wrow, n, cnt.p);
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
  DevBuf<uint2> wt(ctx, nw ? nw : 1);
  DevBuf<uint4> bt(ctx, nb ? nb : 1);
  k_sched_fill<<<grid_for(n, kBlock), kBlock, 0, s>>>(g->row_ptr, g->wrow, n, off.p, tot, wt.p, bt.p);
  cnt.reset();
  off.reset();
  TB_HIP(hipEventRecord(ctx->ev[3], s));

  const int cus = ctx->cus > 0 ? ctx->cus : 256;
  uint64_t blocks = 0;
  // one launch per variant; an event after each gives its device time
  if (tot.x) {
    const unsigned grid = (unsigned)std::min<uint64_t>((tot.x + 4 * kWaveBatch - 1) / (4 * kWaveBatch),
                                                       (uint64_t)cus * 8);
    k_count_wave<7, 16><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, wt.p, tot.x, d_queue + 0, d_result);
    blocks += grid;
  }
  TB_HIP(hipEventRecord(ctx->ev[4], s));
  if (tot.y) {
    const unsigned grid = (unsigned)std::min<uint64_t>((tot.y + 4 * kWaveBatch - 1) / (4 * kWaveBatch),
                                                       (uint64_t)cus * 8);
    k_count_wave<10, 16><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, wt.p + tot.x, tot.y, d_queue + 1, d_result);
    blocks += grid;
  }
  TB_HIP(hipEventRecord(ctx->ev[5], s));
  if (tot.z) {
    const unsigned grid = (unsigned)std::min<uint64_t>(tot.z, (uint64_t)cus * 4);
    k_count_block<13, 16><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, bt.p, tot.z, d_queue + 2, d_result);
    blocks += grid;
  }
  TB_HIP(hipEventRecord(ctx->ev[6], s));
  if (tot.w) {
    const unsigned grid = (unsigned)std::min<uint64_t>(tot.w, (uint64_t)cus * 4);
    k_count_block<13, 16><<<grid, kBlock, 0, s>>>(g->row_ptr, g->col, bt.p + tot.z, tot.w, d_queue + 3, d_result);
    blocks += grid;
  }
  TB_HIP(hipGetLastError());
  TB_HIP(hipEventRecord(ctx->ev[7], s));
  uint64_t tri = 0;
  TB_HIP(hipMemcpyAsync(&tri, d_result, 8, hipMemcpyDeviceToHost, s));
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
  }
  return tri;
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
  for (int r = 0; r < reps; ++r) {
    double a, b, c, v[kVariants];
    const uint64_t tri = count_once(ctx, g, &a, &b, &c, v, &local);
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

// Per-variant work of the plain schedule: rows, sum outdeg, sum W+.
__global__ void k_variant_stats(const uint32_t* __restrict__ row_ptr, const uint64_t* __restrict__ wrow, uint64_t n,
                                unsigned long long* __restrict__ st) {
  __shared__ unsigned long long acc[3 * kVariants];
  if (threadIdx.x < 3 * kVariants) acc[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) {
    const uint32_t d = row_ptr[u + 1] - row_ptr[u];
    const int c = row_class(d);
    if (c >= 0) {
      atomicAdd(&acc[c], 1ull);
      atomicAdd(&acc[kVariants + c], (unsigned long long)d);
      atomicAdd(&acc[2 * kVariants + c], (unsigned long long)wrow[u]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 * kVariants && acc[threadIdx.x]) atomicAdd(&st[threadIdx.x], acc[threadIdx.x]);
}

void variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out) {
  (void)sched;
  memset(out, 0, sizeof(*out));
  DevBuf<unsigned long long> st(ctx, 3 * kVariants);
  k_zero_u64<<<1, 64, 0, ctx->stream>>>(st.p, 3 * kVariants);
  if (g->n) k_variant_stats<<<grid_for(g->n, kBlock), kBlock, 0, ctx->stream>>>(g->row_ptr, g->wrow, g->n, st.p);
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
