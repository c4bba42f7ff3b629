// prep.hip -- on-GPU generators and preprocessing (edge keys in HBM ->
// degree-oriented int32 CSR, per-row wedge counts, streamed-range lists), gfx950.
//
// Replaces the reference's preprocessing chain, src/single_gpu/tribit.cu:
// 664-763 (radix sort + unique of packed keys, L/U bit structures, singleton
// purge, warp schedule). The product here is a degree-oriented CSR instead:
//   keys (min<<32|max) --sort--> --unique--> m undirected edges
//   deg[v] (wave-aggregated atomics)
//   orient {a<b}: a->b iff deg(a) <= deg(b)       (rank = (deg, id))
//   --sort--> row_ptr[n+1] (u32), col[m] (u32, ascending within a row)
//   W+ = sum over v of indeg(v) * outdeg(v); wedge ranges (relabelled graphs)
// Every launch is a 1-D grid of 256-thread blocks (4 waves).
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kBlock = 256;

// ---------------------------------------------------------------- generators

__global__ void k_gen_rmat(uint64_t* __restrict__ keys, uint64_t ne, uint32_t scale, uint32_t t0, uint32_t t1,
                           uint32_t t2, uint64_t key) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += stride) {
    uint64_t src = 0, dst = 0, w = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) w = rand_word(key, e * 16 + (l >> 1));
      const uint32_t r = (l & 1) ? (uint32_t)(w >> 32) : (uint32_t)w;
      const uint64_t sb = (r >= t1) ? 1u : 0u;                       // quadrants (1,0),(1,1)
      const uint64_t db = (r >= t0 && r < t1) || (r >= t2) ? 1u : 0u;  // quadrants (0,1),(1,1)
      src |= sb << (scale - 1 - l);
      dst |= db << (scale - 1 - l);
    }
    keys[e] = (src == dst) ? kSentinel : (src < dst ? ((src << 32) | dst) : ((dst << 32) | src));
  }
}

__global__ void k_gen_er(uint64_t* __restrict__ keys, uint64_t ne, uint64_t n, uint64_t key) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += stride) {
    const uint64_t w = rand_word(key, e);
    const uint64_t u = ((w & 0xFFFFFFFFull) * n) >> 32;
    const uint64_t v = ((w >> 32) * n) >> 32;
    keys[e] = (u == v) ? kSentinel : (u < v ? ((u << 32) | v) : ((v << 32) | u));
  }
}

void rmat_thresholds(double a, double b, double c, uint32_t t[3]) {
  const double two32 = 4294967296.0;
  const double s[3] = {a, a + b, a + b + c};
  for (int i = 0; i < 3; ++i) {
    double v = s[i] * two32;
    if (v < 0) v = 0;
    if (v > 4294967295.0) v = 4294967295.0;
    t[i] = (uint32_t)v;
  }
}

static unsigned gen_grid(tribit_ctx* ctx, uint64_t ne) {
  uint64_t g = (ne + kBlock - 1) / kBlock;
  const uint64_t cap = (uint64_t)(ctx->cus > 0 ? ctx->cus : 256) * 64;
  return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

void gen_rmat(tribit_ctx* ctx, uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
              uint64_t* d_keys) {
  uint32_t t[3];
  rmat_thresholds(a, b, c, t);
  const uint64_t ne = (uint64_t)ef << scale;
  k_gen_rmat<<<gen_grid(ctx, ne), kBlock, 0, ctx->stream>>>(d_keys, ne, scale, t[0], t[1], t[2],
                                                            stream_key(seed, kStreamRmat));
  TB_HIP(hipGetLastError());
}

void gen_er(tribit_ctx* ctx, uint64_t n, uint64_t nsamples, uint64_t seed, uint64_t* d_keys) {
  k_gen_er<<<gen_grid(ctx, nsamples), kBlock, 0, ctx->stream>>>(d_keys, nsamples, n, stream_key(seed, kStreamEr));
  TB_HIP(hipGetLastError());
}

// ------------------------------------------------------------- preprocessing

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Flags keys that are neither the sentinel nor (min < max < n).
__global__ void k_validate(const uint64_t* __restrict__ keys, uint64_t nk, uint64_t n, uint32_t* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const uint64_t k = keys[i];
  if (k == kSentinel) return;
  const uint64_t a = k >> 32, b = k & 0xFFFFFFFFull;
  if (!(a < b && b < n)) atomicOr(bad, 1u);
}

// Checks every key (min << 32) | max (min < max < n) and packs it in place to
// (min << kb) | max, so the sort covers 2 kb bits instead of 32 + kb. The
// sentinel stays all ones and still sorts last: a packed key has min < max,
// so its low 2 kb bits are never all ones.
__global__ void k_validate_pack(uint64_t* __restrict__ keys, uint64_t nk, uint64_t n, uint32_t kb,
                                uint32_t* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const uint64_t k = keys[i];
  if (k == kSentinel) return;
  const uint64_t a = k >> 32, b = k & 0xFFFFFFFFull;
  if (!(a < b && b < n)) atomicOr(bad, 1u);
  keys[i] = (a << kb) | b;
}

// pos[i] = 1 iff sorted key i is the first of its value and not the sentinel.
__global__ void k_head_flags(const uint64_t* __restrict__ k, uint64_t nk, uint32_t* __restrict__ pos) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const uint64_t x = k[i];
  pos[i] = (x != kSentinel && (i == 0 || k[i - 1] != x)) ? 1u : 0u;
}

// out[pos[i]] = k[i] for the heads (pos = exclusive scan of the head flags).
__global__ void k_scatter_heads(const uint64_t* __restrict__ k, uint64_t nk, const uint32_t* __restrict__ pos,
                                uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  if (pos[i + 1] != pos[i]) out[pos[i]] = k[i];
}

// Degrees of the sorted unique keys. Consecutive keys repeat their smaller
// endpoint, so a wave adds each run of it with one atomic; the larger
// endpoints are written to `hi`, sorted, and added the same way by
// k_add_runs (one atomic per key took 37 ms for the 523M keys of R-MAT
// scale 25).
__global__ void k_degree_lo(const uint64_t* __restrict__ keys, uint64_t m, uint32_t kb, uint32_t* __restrict__ deg,
                            uint32_t* __restrict__ hi) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < m;
  const uint64_t k = valid ? keys[i] : 0;
  const uint32_t a = valid ? (uint32_t)(k >> kb) : 0xFFFFFFFFu;
  if (valid) hi[i] = (uint32_t)(k & ((1ull << kb) - 1));
  const uint32_t run = wave_run_length(a, valid);
  if (run) atomicAdd(&deg[a], run);
}

// deg[v] += the number of v in the sorted ids[0, m), one atomic per run per wave.
__global__ void k_add_runs(const uint32_t* __restrict__ ids, uint64_t m, uint32_t* __restrict__ deg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < m;
  const uint32_t a = valid ? ids[i] : 0xFFFFFFFFu;
  const uint32_t run = wave_run_length(a, valid);
  if (run) atomicAdd(&deg[a], run);
}

// Oriented arcs u -> v are sorted as packed keys (u << kb) | v with kb =
// bits_for(n): 2*kb bits (50 at 2^25 vertices) instead of 32 + kb, one radix
// digit pass fewer than the input key layout.
__device__ __forceinline__ uint64_t arc_key(uint32_t u, uint32_t v, uint32_t kb) {
  return ((uint64_t)u << kb) | (uint64_t)v;
}

// Degree orientation: {a<b} becomes a->b iff deg(a) <= deg(b).
__global__ void k_orient(const uint64_t* __restrict__ keys, uint64_t m, const uint32_t* __restrict__ deg,
                         uint32_t kb, uint64_t* __restrict__ okeys) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t k = keys[i];
  const uint32_t a = (uint32_t)(k >> kb), b = (uint32_t)(k & ((1ull << kb) - 1));
  okeys[i] = (deg[a] <= deg[b]) ? arc_key(a, b, kb) : arc_key(b, a, kb);
}

// col[i] = dst, odeg[src] += 1 (sorted by src: aggregated per run).
// Degree relabelling: new id = rank of (deg, old id); lab[old] = new.
__global__ void k_iota_u32(uint32_t* __restrict__ p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// cls[v] = floor(log2(deg[v])) + 1 (0 for an isolated vertex): octave classes.
__global__ void k_octave(const uint32_t* __restrict__ deg, uint64_t n, uint32_t* __restrict__ cls) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cls[i] = deg[i] ? 32u - (uint32_t)__clz(deg[i]) : 0u;
}

// out[i] = src[idx[i]].
__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n,
                             uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

__global__ void k_scatter_labels(const uint32_t* __restrict__ old_of_new, uint64_t n, uint32_t* __restrict__ lab) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lab[old_of_new[i]] = (uint32_t)i;
}

// With relabelled ids the degree order is the id order: a -> b iff a < b.
__global__ void k_orient_relabel(const uint64_t* __restrict__ keys, uint64_t m, const uint32_t* __restrict__ lab,
                                 uint32_t kb, uint64_t* __restrict__ okeys) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint64_t k = keys[i];
  const uint32_t a = lab[(uint32_t)(k >> kb)], b = lab[(uint32_t)(k & ((1ull << kb) - 1))];
  okeys[i] = a < b ? arc_key(a, b, kb) : arc_key(b, a, kb);
}

__global__ void k_csr_split(const uint64_t* __restrict__ okeys, uint64_t m, uint32_t kb, uint32_t* __restrict__ col,
                            uint32_t* __restrict__ odeg) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < m;
  const uint64_t k = valid ? okeys[i] : 0;
  const uint32_t s = valid ? (uint32_t)(k >> kb) : 0xFFFFFFFFu;
  if (valid) col[i] = (uint32_t)(k & ((1ull << kb) - 1));
  const uint32_t run = wave_run_length(s, valid);
  if (run) atomicAdd(&odeg[s], run);
}

struct RowStats {
  unsigned long long wplus;
  unsigned long long rows_active;
  unsigned int max_outdeg;
  unsigned int pad;
};

// Totals over a bounded grid (grid-stride loops, one atomic per block and
// field): one atomic per wave on a single address serialises at about 13 ns
// each, 7 ms for 2^25 rows and 100 ms for 2^29 arcs.
constexpr unsigned kTotalsBlocks = 1024;
inline unsigned totals_grid(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(kTotalsBlocks, (n + kBlock - 1) / kBlock));
}

// Edges {a, b}, a < b, of the deduplicated keys (a << kb) | b that join ids
// less than kBandIds apart, summed into *out: two keys per 16-byte load over a
// grid of at most kBandBlocks blocks, one atomic per block. keys is 256-byte
// aligned (hipMallocAsync).
constexpr uint64_t kBandIds = 512;
constexpr unsigned kBandBlocks = 4096;
__global__ void __launch_bounds__(kBlock) k_band_edges(const uint64_t* __restrict__ keys, uint64_t m, uint32_t kb,
                                                       unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_part[kBlock / 64];
  const uint64_t lo = (1ull << kb) - 1;
  const ulonglong2* __restrict__ k2 = reinterpret_cast<const ulonglong2*>(keys);
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m / 2;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const ulonglong2 k = k2[i];
    acc += ((k.x & lo) - (k.x >> kb) < kBandIds) + ((k.y & lo) - (k.y >> kb) < kBandIds);
  }
  if ((m & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const uint64_t k = keys[m - 1];
    acc += (k & lo) - (k >> kb) < kBandIds;
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < kBlock / 64; ++i) acc += s_part[i];
    if (acc) atomicAdd(out, (unsigned long long)acc);
  }
}

// pos[i] as k_head_flags, over a grid of at most kBandBlocks blocks, and the
// heads {a, b} (packed (a << kb) | b) that join ids less than kBandIds apart
// summed into *band, one atomic per block: the band count rides on the pass
// that dedup makes over the sorted keys anyway.
__global__ void __launch_bounds__(kBlock) k_head_flags_band(const uint64_t* __restrict__ k, uint64_t nk, uint32_t kb,
                                                            uint32_t* __restrict__ pos,
                                                            unsigned long long* __restrict__ band) {
  __shared__ unsigned long long s_part[kBlock / 64];
  const uint64_t lo = (1ull << kb) - 1;
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = k[i];
    const bool head = x != kSentinel && (i == 0 || k[i - 1] != x);
    pos[i] = head ? 1u : 0u;
    acc += head && (x & lo) - (x >> kb) < kBandIds;
  }
  acc = wave_sum_u64(acc);
  if (lane_id() == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < kBlock / 64; ++i) acc += s_part[i];
    if (acc) atomicAdd(band, (unsigned long long)acc);
  }
}

// TRIBIT_OPT_RELABEL 3 (auto). Octave-class ids keep runs of consecutive input
// ids together, which dense diagonal blocks need; degree-rank ids stream a few
// percent faster on everything else. Octave classes are taken when the edges
// joining ids less than kBandIds apart fill at least 1/16 of such id pairs:
// a block model with 512-id blocks at p_in 0.3 fills 15% of them, one at
// p_in 0.05 2.5%, a cell-ordered 2-D geometric graph of average degree 24
// 1.3%, R-MAT, Chung-Lu and Erdős–Rényi ids under 0.1%.
constexpr uint64_t kAutoMinIds = 4096;
int auto_relabel(uint64_t n, uint64_t band_edges) {
  if (n < kAutoMinIds) return 1;
  const double w = (double)std::min<uint64_t>(kBandIds, n);
  const double pairs = (w - 1.0) * (double)n - w * (w - 1.0) / 2.0;  // id pairs 1..w-1 apart
  return (double)band_edges * 16.0 >= pairs ? 2 : 1;
}

// Rows with outdeg >= 2, the largest outdeg and, with deg (undirected
// degrees), W+ = sum over v of indeg(v) * outdeg(v): every arc u -> v streams
// N+(v) once, and v has deg(v) - outdeg(v) in-arcs.
__global__ void __launch_bounds__(kBlock) k_row_flags(const uint32_t* __restrict__ row_ptr,
                                                      const uint32_t* __restrict__ deg, uint64_t n,
                                                      RowStats* __restrict__ st) {
  __shared__ unsigned long long s_act[kBlock / 64], s_w[kBlock / 64];
  __shared__ uint32_t s_max[kBlock / 64];
  uint64_t act = 0, w = 0;
  uint32_t dm = 0;
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = row_ptr[u + 1] - row_ptr[u];
    act += d >= 2;
    dm = d > dm ? d : dm;
    if (deg) w += (uint64_t)(deg[u] - d) * d;
  }
  act = wave_sum_u64(act);
  w = wave_sum_u64(w);
  dm = wave_max_u32(dm);
  if (lane_id() == 0) {
    s_act[threadIdx.x >> 6] = act;
    s_w[threadIdx.x >> 6] = w;
    s_max[threadIdx.x >> 6] = dm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < kBlock / 64; ++i) {
      act += s_act[i];
      w += s_w[i];
      dm = s_max[i] > dm ? s_max[i] : dm;
    }
    if (act) atomicAdd(&st->rows_active, (unsigned long long)act);
    if (w) atomicAdd(&st->wplus, (unsigned long long)w);
    if (dm) atomicMax(&st->max_outdeg, dm);
  }
}

// Per row u of a relabelled graph: the position where the row's last 64K-id
// window begins (the first id whose high 16 bits equal the last id's), for
// the narrow records of its wedge ranges. One lower-bound search per row.
__global__ void k_row_split(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            uint32_t* __restrict__ split) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const uint32_t b = row_ptr[u], e = row_ptr[u + 1];
  uint32_t lo = b;
  if (b < e) {
    const uint32_t l = col[e - 1] >> 16;
    uint32_t hi = e - 1;  // col[hi] is in window l
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if ((col[mid] >> 16) >= l)
        hi = mid;
      else
        lo = mid + 1;
    }
  }
  split[u] = lo;
}

// Relabelled graphs, one pass over the arcs e = (u -> v) in CSR order (okeys
// sorted, packed (u << kb) | v), every lane one arc, into arec[e]:
//   .x = row_ptr[u + 1] (the end of arc e's wedge range [e + 1, end), the part
//        of N+(u) after v), with bits 31 / 30 of the narrow record;
//   .y (with nar, m < 2^30) = the rest of that range's narrow record: bit 31
//        of .x: its ids share their high 16 bits H; bit 30 (with 31): shorter
//        than 2^16 ids with high halves H then H + 1, .y bits 0..15 the offset
//        of the first id with H + 1 (from the row's split position); .y bits
//        16..23 H - (v >> 16), 255 when larger. In CSR order the range's ids
//        are the row's next ones, so the reads stay in cache.
// One 8-byte record per arc, so the wedge-range fill gathers once per range.
__global__ void k_arc_rec(const uint64_t* __restrict__ okeys, uint64_t m, uint32_t kb,
                          const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                          const uint32_t* __restrict__ rsplit, bool nar, uint2* __restrict__ arec) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const uint64_t k = okeys[e];
  const uint32_t u = (uint32_t)(k >> kb);
  const uint32_t v = (uint32_t)(k & ((1ull << kb) - 1));
  const uint32_t end = row_ptr[u + 1];
  uint32_t x = end, y = 0;
  const uint32_t b = (uint32_t)e + 1;
  if (nar && b < end) {
    const uint32_t f = col[b] >> 16, l = col[end - 1] >> 16;
    const bool two = l == f + 1 && end - b < 65536u;
    if (f == l || two) {
      uint32_t d = f - (v >> 16);  // ids after v: H >= v's high half
      d = d < 255u ? d : 255u;
      x |= 0x80000000u | (two ? 0x40000000u : 0u);
      y = (d << 16) | (two ? rsplit[u] - b : 0u);
    }
  }
  arec[e] = make_uint2(x, y);
}

// ------------------------------------------------------------ range lists
// For every oriented edge u->v the count streams one list: N+(v) against the
// set N+(u) ("forward", owner u) or N+(u) against the set N+(v) ("reverse",
// owner v). The streamed list is stored as its [begin, end) range in col, so
// the counting kernels read 8 coalesced bytes per edge instead of gathering
// row_ptr. Two layouts are built:
//   rng_all: all edges forward, in CSR order (owner u, row_ptr offsets);
//   rng_dir: an edge is reverse iff outdeg(u) < outdeg(v) (the shorter list
//            is streamed): forward edges per owner first (fw_ptr), then the
//            reverse edges grouped by owner v (rv_ptr, offsets past m_fwd).

__global__ void k_ranges_all(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t m,
                             uint2* __restrict__ rng) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const uint32_t v = col[e];
  rng[e] = make_uint2(row_ptr[v], row_ptr[v + 1]);
}

// Per row u: forward-edge count and forward work; reverse-edge count.
__global__ void k_dir_count(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            uint32_t* __restrict__ fw_cnt, uint32_t* __restrict__ rv_out, uint64_t* __restrict__ w_fw) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  uint32_t f = 0, r = 0;
  uint64_t w = 0;
  if (u < n) {
    const uint32_t r0 = row_ptr[u], r1 = row_ptr[u + 1], du = r1 - r0;
    for (uint32_t e = r0; e < r1; ++e) {
      const uint32_t v = col[e];
      const uint32_t dv = row_ptr[v + 1] - row_ptr[v];
      if (du < dv) {
        ++r;
      } else {
        ++f;
        w += dv;
      }
    }
    w_fw[u] = w;
  }
  fw_cnt[u] = f;
  rv_out[u] = r;
}

// Writes the forward ranges (and partner ids) of row u at fw_ptr[u] and its
// reverse keys (v << 32 | u) at rv_off[u].
__global__ void k_dir_fill(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                           const uint32_t* __restrict__ fw_ptr, const uint32_t* __restrict__ rv_off,
                           uint2* __restrict__ rng, uint32_t* __restrict__ pid, uint64_t* __restrict__ rv_keys) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const uint32_t r0 = row_ptr[u], r1 = row_ptr[u + 1], du = r1 - r0;
  uint32_t f = fw_ptr[u], r = rv_off[u];
  for (uint32_t e = r0; e < r1; ++e) {
    const uint32_t v = col[e];
    const uint32_t b = row_ptr[v], en = row_ptr[v + 1];
    if (du < en - b) {
      rv_keys[r++] = ((uint64_t)v << 32) | (uint64_t)u;
    } else {
      pid[f] = v;
      rng[f++] = make_uint2(b, en);
    }
  }
}

// Sorted reverse keys (v << 32 | u): range of N+(u) at rng[base + j];
// rv_cnt[v] += 1 and w_rv[v] += outdeg(u), aggregated per run of v in a wave.
__global__ void k_rv_fill(const uint64_t* __restrict__ keys, uint64_t mr, const uint32_t* __restrict__ row_ptr,
                          uint2* __restrict__ rng, uint32_t* __restrict__ pid, uint32_t* __restrict__ rv_cnt,
                          unsigned long long* __restrict__ w_rv) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < mr;
  const uint64_t k = valid ? keys[j] : 0;
  const uint32_t v = valid ? (uint32_t)(k >> 32) : 0xFFFFFFFFu;
  uint64_t du = 0;
  if (valid) {
    const uint32_t u = (uint32_t)k;
    const uint32_t b = row_ptr[u], e = row_ptr[u + 1];
    rng[j] = make_uint2(b, e);
    pid[j] = u;
    du = e - b;
  }
  const uint32_t lane = lane_id();
  // segmented inclusive sum of du over equal v (keys sorted: equal v are contiguous)
  uint64_t sum = du;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(sum, o);
    const uint32_t kv = __shfl_up(v, o);
    if (lane >= (uint32_t)o && kv == v) sum += t;
  }
  const uint32_t nv = __shfl_down(v, 1);
  const uint64_t vmask = __ballot(valid);
  const uint64_t above = (lane == 63) ? 0ull : (vmask >> (lane + 1));
  const bool last_valid = !(above & 1ull);
  const bool tail = valid && (last_valid || nv != v);
  const uint32_t run = wave_run_length(v, valid);
  if (run) atomicAdd(&rv_cnt[v], run);
  if (tail && sum) atomicAdd(&w_rv[v], (unsigned long long)sum);
}

// Arcs e = (u -> v) sorted by head v: the suffix of N+(u) after v,
// [e + 1, end), is streamed against N+(v). With ids in degree-rank order
// N+(v) holds only ids > v, so the elements of N+(u) up to v can never match.
// w_sfx[v] += suffix length, aggregated per run of v in a wave.
// With narrow records (arec of k_arc_rec), wedge range j = [b, e) is flagged
// narrow when its ids share their high 16 bits H (end bit 31), or, shorter
// than 2^16 ids, when they have high half H up to position b + split[j] and
// H + 1 from there (end bits 31 and 30); dh[j] = H - (v >> 16) (255: not
// representable, the counting kernel reads col[b] instead).
__global__ void k_sfx_fill(const uint32_t* __restrict__ head, const uint32_t* __restrict__ arc, uint64_t m,
                           const uint2* __restrict__ arec, uint2* __restrict__ rng, uint8_t* __restrict__ dh,
                           uint16_t* __restrict__ split, unsigned long long* __restrict__ w_sfx) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < m;
  const uint32_t v = valid ? head[j] : 0xFFFFFFFFu;
  uint64_t len = 0;
  if (valid) {
    const uint32_t e = arc[j];
    const uint2 r = arec[e];
    if (dh) {
      dh[j] = (uint8_t)(r.y >> 16);
      split[j] = (uint16_t)r.y;
    }
    rng[j] = make_uint2(e + 1, r.x);
    len = (r.x & 0x3FFFFFFFu) - e - 1;
  }
  uint64_t sum = 0;
  if (wave_segment_tail(v, valid, len, &sum) && sum) atomicAdd(&w_sfx[v], (unsigned long long)sum);
}

// In-arcs per vertex: deg(v) - outdeg(v) (cnt[n] = 0 for the exclusive scan).
__global__ void k_in_count(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ deg, uint64_t n,
                           uint32_t* __restrict__ cnt) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) cnt[v] = deg[v] - (row_ptr[v + 1] - row_ptr[v]);
  if (v == n) cnt[v] = 0;
}

// W+(u) = sum of outdeg(v) over v in N+(u), per row (the PLAIN layout's work).
__global__ void k_row_wplus(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            uint64_t* __restrict__ wrow) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint64_t w = 0;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t v = col[e];
    w += row_ptr[v + 1] - row_ptr[v];
  }
  wrow[u] = w;
}

// Low 16 bits of every id (narrow wedge ranges are streamed from these).
__global__ void k_col16(const uint32_t* __restrict__ col, uint64_t m, uint16_t* __restrict__ col16) {
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m) col16[e] = (uint16_t)(col[e] & 0xFFFFu);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

__global__ void k_add_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += v;
}

template <typename F>
static void with_temp(tribit_ctx* ctx, F&& f) {
  size_t bytes = 0;
  TB_HIP(f(nullptr, bytes));
  DevBuf<uint8_t> tmp(ctx, bytes ? bytes : 8);
  TB_HIP(f(tmp.p, bytes));
}

// 9-bit digits for the sorts of 32-bit ids (three passes instead of four for
// ids below 2^27). Measured on MI355X over 2^29 random 25-bit keys: pairs
// with 32-bit values 14.7 -> 11.5 ms, keys alone 9.6 -> 8.0 ms. The 64-bit
// key sorts keep rocPRIM's 8-bit digits (9 bits: 22.1 -> 23.4 ms; 10 bits
// slower still).
using IdSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;

static void sort_keys(tribit_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t nk, uint32_t end_bit) {
  with_temp(ctx, [&](void* t, size_t& b) {
    return rocprim::radix_sort_keys(t, b, in, out, (size_t)nk, 0u, end_bit, ctx->stream);
  });
}

void preprocess(tribit_ctx* ctx, uint64_t* d_keys_in, uint64_t nkeys, uint64_t n, tribit_graph* g, int relabel_opt) {
  if (relabel_opt < 0) relabel_opt = ctx->relabel;
  DevBuf<uint64_t> keys;
  keys.ctx = ctx;
  keys.p = d_keys_in;  // owned from here on
  keys.n = nkeys;
  if (n == 0 || n > 0xFFFFFFFEull) fail(TRIBIT_ERR_TOO_LARGE, "vertex count must be in [1, 2^32-2]");
  hipStream_t s = ctx->stream;
  g->ctx = ctx;
  g->n = n;
  g->entries = nkeys;
  TB_HIP(hipEventRecord(ctx->ev[0], s));

  // keys checked and packed to (min << kb) | max for the sort
  const uint32_t kb = bits_for(n);
  DevBuf<uint32_t> flag(ctx, 4);
  k_fill_u32<<<1, 64, 0, s>>>(flag.p, 4, 0u);
  if (nkeys) k_validate_pack<<<grid_for(nkeys, kBlock), kBlock, 0, s>>>(keys.p, nkeys, n, kb, flag.p);

  // 1-2. sort + unique (tribit.cu:666-667). The sentinel sorts last.
  const uint32_t end_bit = 2 * kb;
  uint64_t m = 0;
  // id-band edges for the automatic choice of ids (TRIBIT_OPT_RELABEL 3)
  const bool want_band = relabel_opt == 3 && n >= kAutoMinIds;
  uint64_t band_edges = 0;
  DevBuf<unsigned long long> band;
  band.ctx = ctx;
  DevBuf<uint64_t> uniq(ctx, nkeys ? nkeys : 1);
  if (nkeys) {
    DevBuf<uint64_t> sorted(ctx, nkeys);
    sort_keys(ctx, keys.p, sorted.p, nkeys, end_bit);
    keys.reset();
    // unique: head flags, their exclusive scan (the last entry is the count),
    // scatter of the heads; the sentinel is never a head. 32-bit positions
    // (rocprim::unique took 16 ms for the 537M keys of R-MAT scale 25; more
    // than 2^32 - 2 keys still go through it).
    uint32_t bad = 0;
    if (nkeys <= 0xFFFFFFFEull) {
      DevBuf<uint32_t> head(ctx, nkeys + 1), pos(ctx, nkeys + 1);
      if (want_band) {
        band = DevBuf<unsigned long long>(ctx, 1);
        k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(band.p), 2, 0u);
        const unsigned grid = (unsigned)std::min<uint64_t>(kBandBlocks, (nkeys + kBlock - 1) / kBlock);
        k_head_flags_band<<<grid, kBlock, 0, s>>>(sorted.p, nkeys, kb, head.p, band.p);
      } else {
        k_head_flags<<<grid_for(nkeys, kBlock), kBlock, 0, s>>>(sorted.p, nkeys, head.p);
      }
      k_fill_u32<<<1, 64, 0, s>>>(head.p + nkeys, 1, 0u);
      with_temp(ctx, [&](void* t, size_t& b) {
        return rocprim::exclusive_scan(t, b, head.p, pos.p, 0u, (size_t)(nkeys + 1), rocprim::plus<uint32_t>(), s);
      });
      k_scatter_heads<<<grid_for(nkeys, kBlock), kBlock, 0, s>>>(sorted.p, nkeys, pos.p, uniq.p);
      uint32_t cnt = 0;
      TB_HIP(hipMemcpyAsync(&cnt, pos.p + nkeys, 4, hipMemcpyDeviceToHost, s));
      TB_HIP(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
      if (want_band) TB_HIP(hipMemcpyAsync(&band_edges, band.p, 8, hipMemcpyDeviceToHost, s));
      ctx->sync();
      band.reset();
      m = cnt;
    } else {
      DevBuf<uint64_t> ucount(ctx, 1);
      with_temp(ctx, [&](void* t, size_t& b) {
        return rocprim::unique(t, b, sorted.p, uniq.p, ucount.p, (size_t)nkeys, rocprim::equal_to<uint64_t>(), s);
      });
      uint64_t h[2] = {0, 0};
      TB_HIP(hipMemcpyAsync(&h[0], ucount.p, 8, hipMemcpyDeviceToHost, s));
      ctx->sync();
      if (h[0]) TB_HIP(hipMemcpyAsync(&h[1], uniq.p + h[0] - 1, 8, hipMemcpyDeviceToHost, s));
      TB_HIP(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
      ctx->sync();
      m = h[0] - ((h[0] && h[1] == kSentinel) ? 1 : 0);
      if (want_band && m && !bad) {
        band = DevBuf<unsigned long long>(ctx, 1);
        k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(band.p), 2, 0u);
        const unsigned grid = (unsigned)std::min<uint64_t>(kBandBlocks, (m / 2 + kBlock) / kBlock);
        k_band_edges<<<grid, kBlock, 0, s>>>(uniq.p, m, kb, band.p);
        TB_HIP(hipMemcpyAsync(&band_edges, band.p, 8, hipMemcpyDeviceToHost, s));
        ctx->sync();
        band.reset();
      }
    }
    if (bad) fail(TRIBIT_ERR_RANGE, "edge key is not (min < max < n)");
  } else {
    keys.reset();
  }
  if (m > 0xFFFFFFFFull) fail(TRIBIT_ERR_TOO_LARGE, "more than 2^32-1 edges in one graph");
  g->m = m;

  // 3. degrees
  DevBuf<uint32_t> deg(ctx, n);
  k_fill_u32<<<grid_for(n, kBlock), kBlock, 0, s>>>(deg.p, n, 0u);
  if (m) {
    DevBuf<uint32_t> hi(ctx, m), hi_sorted(ctx, m);
    k_degree_lo<<<grid_for(m, kBlock), kBlock, 0, s>>>(uniq.p, m, kb, deg.p, hi.p);
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::radix_sort_keys<IdSortConfig>(t, b, hi.p, hi_sorted.p, (size_t)m, 0u, bits_for(n), s);
    });
    k_add_runs<<<grid_for(m, kBlock), kBlock, 0, s>>>(hi_sorted.p, m, deg.p);
  }

  // 3b. optional degree relabelling: internal id = rank of (deg, id) (mode 1)
  // or of (octave of deg, id) (mode 2; mode 3 picks 1 or 2 from the id-band
  // edges), so the orientation below becomes the id order and the
  // out-neighbours of high-degree vertices are packed into a narrow id range
  // at the top.
  const int relabel = relabel_opt == 3 ? auto_relabel(n, band_edges) : relabel_opt;
  DevBuf<uint32_t> labels;
  labels.ctx = ctx;
  DevBuf<uint64_t> okeys;  // oriented arcs, packed (u << kb) | v
  okeys.ctx = ctx;
  if (relabel) {
    DevBuf<uint32_t> ids(ctx, n), deg_sorted(ctx, n), lab(ctx, n);
    labels = DevBuf<uint32_t>(ctx, n);
    k_iota_u32<<<grid_for(n, kBlock), kBlock, 0, s>>>(ids.p, n);
    if (relabel == 2) {
      // octave classes: rank of (floor(log2(deg)), id); runs of consecutive
      // input ids with similar degrees (a block of a block model) stay runs
      DevBuf<uint32_t> cls(ctx, n);
      k_octave<<<grid_for(n, kBlock), kBlock, 0, s>>>(deg.p, n, cls.p);
      with_temp(ctx, [&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs(t, b, cls.p, deg_sorted.p, ids.p, labels.p, (size_t)n, 0u, 6u, s);
      });
      k_gather_u32<<<grid_for(n, kBlock), kBlock, 0, s>>>(deg.p, labels.p, n, deg_sorted.p);
    } else {
      with_temp(ctx, [&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs<IdSortConfig>(t, b, deg.p, deg_sorted.p, ids.p, labels.p, (size_t)n, 0u,
                                         bits_for(n), s);
      });
    }
    std::swap(deg.p, deg_sorted.p);  // deg now indexed by internal id
    k_scatter_labels<<<grid_for(n, kBlock), kBlock, 0, s>>>(labels.p, n, lab.p);
    if (m) {
      okeys = DevBuf<uint64_t>(ctx, m);
      k_orient_relabel<<<grid_for(m, kBlock), kBlock, 0, s>>>(uniq.p, m, lab.p, kb, okeys.p);
    }
  } else if (m) {
    okeys = DevBuf<uint64_t>(ctx, m);
    k_orient<<<grid_for(m, kBlock), kBlock, 0, s>>>(uniq.p, m, deg.p, kb, okeys.p);
  }

  // 4-6. sort the arcs by (u, v) into the CSR
  DevBuf<uint32_t> col(ctx, m ? m : 1);
  DevBuf<uint32_t> row_ptr(ctx, n + 1);
  {
    DevBuf<uint32_t> odeg(ctx, n + 1);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(odeg.p, n + 1, 0u);
    if (m) {
      sort_keys(ctx, okeys.p, uniq.p, m, 2 * kb);
      okeys.reset();
      k_csr_split<<<grid_for(m, kBlock), kBlock, 0, s>>>(uniq.p, m, kb, col.p, odeg.p);
    }
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, odeg.p, row_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
  }

  // 7. totals (W+, active rows, largest outdeg) from the rows; relabelled
  // graphs also get every arc's record (row end, narrow record of its wedge range)
  DevBuf<RowStats> st(ctx, 1);
  k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(st.p), sizeof(RowStats) / 4, 0u);
  k_row_flags<<<totals_grid(n), kBlock, 0, s>>>(row_ptr.p, deg.p, n, st.p);
  const bool narrow = relabel && m && m < (1ull << 30);
  DevBuf<uint2> arec;
  arec.ctx = ctx;
  if (relabel && m) {
    arec = DevBuf<uint2>(ctx, m);
    DevBuf<uint32_t> rsplit;
    rsplit.ctx = ctx;
    if (narrow) {
      rsplit = DevBuf<uint32_t>(ctx, n);
      k_row_split<<<grid_for(n, kBlock), kBlock, 0, s>>>(row_ptr.p, col.p, n, rsplit.p);
    }
    k_arc_rec<<<grid_for(m, kBlock), kBlock, 0, s>>>(uniq.p, m, kb, row_ptr.p, col.p, rsplit.p, narrow, arec.p);
  }
  uniq.reset();
  RowStats hst;
  TB_HIP(hipMemcpyAsync(&hst, st.p, sizeof(hst), hipMemcpyDeviceToHost, s));

  // 8. wedge ranges (relabelled graphs): in-arcs of every v, each with its
  // suffix range; narrow ranges flagged, and col16 (the low 16 bits of every
  // id, padded to whole dwords plus one) built for streaming them
  DevBuf<uint2> rng_sfx;
  DevBuf<uint32_t> sfx_ptr;
  DevBuf<uint64_t> w_sfx;
  DevBuf<uint32_t> col16;
  DevBuf<uint8_t> sfx_dh;
  DevBuf<uint16_t> sfx_split;
  rng_sfx.ctx = sfx_ptr.ctx = w_sfx.ctx = ctx;
  if (relabel) {
    rng_sfx = DevBuf<uint2>(ctx, m ? m : 1);
    sfx_ptr = DevBuf<uint32_t>(ctx, n + 1);
    w_sfx = DevBuf<uint64_t>(ctx, n);
    if (narrow) {
      sfx_dh = DevBuf<uint8_t>(ctx, m);
      sfx_split = DevBuf<uint16_t>(ctx, m);
    }
    k_fill_u64<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(w_sfx.p), n);
    DevBuf<uint32_t> cnt(ctx, n + 1);
    k_in_count<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(row_ptr.p, deg.p, n, cnt.p);
    if (m) {
      DevBuf<uint32_t> arcs(ctx, m), head(ctx, m), arc_sorted(ctx, m);
      k_iota_u32<<<grid_for(m, kBlock), kBlock, 0, s>>>(arcs.p, m);
      with_temp(ctx, [&](void* t, size_t& b) {
        return rocprim::radix_sort_pairs<IdSortConfig>(t, b, col.p, head.p, arcs.p, arc_sorted.p, (size_t)m, 0u,
                                                       bits_for(n), s);
      });
      arcs.reset();
      k_sfx_fill<<<grid_for(m, kBlock), kBlock, 0, s>>>(head.p, arc_sorted.p, m, arec.p, rng_sfx.p, sfx_dh.p,
                                                        sfx_split.p, reinterpret_cast<unsigned long long*>(w_sfx.p));
    }
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, cnt.p, sfx_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
    if (narrow) {
      const uint64_t words = (m + 1) / 2 + 1;
      col16 = DevBuf<uint32_t>(ctx, words);
      k_fill_u32<<<1, 64, 0, s>>>(col16.p + words - 2, 2, 0u);
      k_col16<<<grid_for(m, kBlock), kBlock, 0, s>>>(col.p, m, reinterpret_cast<uint16_t*>(col16.p));
    }
  }
  arec.reset();

  // the graph owns its buffers from here on (freed by tribit_graph_free, also on error)
  g->row_ptr = row_ptr.release();
  g->col = col.release();
  g->deg = deg.release();
  g->labels = labels.release();
  g->relabeled = relabel;
  g->rng_sfx = rng_sfx.release();
  g->col16 = reinterpret_cast<uint16_t*>(col16.release());
  g->sfx_dh = sfx_dh.release();
  g->sfx_split = sfx_split.release();
  g->sfx_ptr = sfx_ptr.release();
  g->w_sfx = w_sfx.release();
  g->bytes = 4 * (n + 1) + 4 * (m ? m : 1) + 4 * n + (g->labels ? 4 * n : 0) +
             (g->rng_sfx ? 8 * (m ? m : 1) + 4 * (n + 1) + 8 * n : 0) + (g->col16 ? 4 * ((m + 1) / 2 + 1) + 3 * m : 0);
  // a graph without wedge ranges counts on the direction-split layout by default: built here
  if (!g->rng_sfx) build_layout(ctx, g, TRIBIT_SCHED_DIRECTED);
  TB_HIP(hipEventRecord(ctx->ev[1], s));
  TB_HIP(hipEventSynchronize(ctx->ev[1]));
  TB_HIP(hipGetLastError());
  float ms = 0.f;
  TB_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  g->ms_prep = ms;
  g->wplus = hst.wplus;
  g->rows_active = hst.rows_active;
  g->max_outdeg = hst.max_outdeg;
}

// ------------------------------------------------------------ range layouts
// The layouts a schedule streams beside the wedge ranges: rng_all (PLAIN,
// every arc forward in CSR order) and rng_dir (DIRECTED). Preprocessing builds
// the graph's default one (wedge ranges for relabelled graphs, else DIRECTED);
// the others are built on their first use, outside the count's timing, and
// their device time is kept per layout (tribit_graph_info.ms_layout).
void build_layout(tribit_ctx* ctx, tribit_graph* g, int layout) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, m = g->m;
  if (layout == TRIBIT_SCHED_PLAIN) {
    DevBuf<uint2> rng_all(ctx, m ? m : 1);
    DevBuf<uint64_t> wrow(ctx, n ? n : 1);
    if (m) k_ranges_all<<<grid_for(m, kBlock), kBlock, 0, s>>>(g->row_ptr, g->col, m, rng_all.p);
    if (n) k_row_wplus<<<grid_for(n, kBlock), kBlock, 0, s>>>(g->row_ptr, g->col, n, wrow.p);
    TB_HIP(hipGetLastError());
    g->rng_all = rng_all.release();
    g->wrow = wrow.release();
    g->bytes += 8 * (m ? m : 1) + 8 * (n ? n : 1);
    return;
  }
  if (layout != TRIBIT_SCHED_DIRECTED) fail(TRIBIT_ERR_ARG, "no such range layout");
  const uint32_t end_bit = 32 + bits_for(n);
  DevBuf<uint2> rng_dir(ctx, m ? m : 1);
  DevBuf<uint32_t> pid_dir(ctx, m ? m : 1);
  DevBuf<uint32_t> fw_ptr(ctx, n + 1), rv_ptr(ctx, n + 1);
  DevBuf<uint64_t> w_fw(ctx, n), w_rv(ctx, n);
  uint32_t m_fwd = 0;
  k_fill_u64<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(w_rv.p), n);
  if (m) {
    DevBuf<uint32_t> fw_cnt(ctx, n + 1), rv_cnt(ctx, n + 1), rv_off(ctx, n + 1);
    k_dir_count<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(g->row_ptr, g->col, n, fw_cnt.p, rv_cnt.p, w_fw.p);
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, fw_cnt.p, fw_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, rv_cnt.p, rv_off.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
    TB_HIP(hipMemcpyAsync(&m_fwd, fw_ptr.p + n, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    const uint64_t m_rev = m - m_fwd;
    DevBuf<uint64_t> rv_keys(ctx, m_rev ? m_rev : 1);
    k_dir_fill<<<grid_for(n, kBlock), kBlock, 0, s>>>(g->row_ptr, g->col, n, fw_ptr.p, rv_off.p, rng_dir.p,
                                                      pid_dir.p, rv_keys.p);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(rv_cnt.p, n + 1, 0u);
    if (m_rev) {
      DevBuf<uint64_t> rv_sorted(ctx, m_rev);
      sort_keys(ctx, rv_keys.p, rv_sorted.p, m_rev, end_bit);
      k_rv_fill<<<grid_for(m_rev, kBlock), kBlock, 0, s>>>(rv_sorted.p, m_rev, g->row_ptr, rng_dir.p + m_fwd,
                                                           pid_dir.p + m_fwd, rv_cnt.p,
                                                           reinterpret_cast<unsigned long long*>(w_rv.p));
    }
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, rv_cnt.p, rv_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
    k_add_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(rv_ptr.p, n + 1, m_fwd);
  } else {
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(fw_ptr.p, n + 1, 0u);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(rv_ptr.p, n + 1, 0u);
    k_fill_u64<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(w_fw.p), n);
  }
  TB_HIP(hipGetLastError());
  g->m_fwd = m_fwd;
  g->rng_dir = rng_dir.release();
  g->pid_dir = pid_dir.release();
  g->fw_ptr = fw_ptr.release();
  g->rv_ptr = rv_ptr.release();
  g->w_fw = w_fw.release();
  g->w_rv = w_rv.release();
  g->bytes += 12 * (m ? m : 1) + 8 * (n + 1) + 16 * n;
}

bool layout_built(const tribit_graph* g, int layout) {
  if (layout == TRIBIT_SCHED_PLAIN) return g->rng_all != nullptr;
  if (layout == TRIBIT_SCHED_DIRECTED) return g->rng_dir != nullptr;
  if (layout == TRIBIT_SCHED_WEDGE) return g->rng_sfx != nullptr;
  return false;
}

void ensure_layout(tribit_ctx* ctx, tribit_graph* g, int layout) {
  if (layout_built(g, layout)) return;
  if (layout == TRIBIT_SCHED_WEDGE) fail(TRIBIT_ERR_ARG, "wedge ranges need a relabelled graph (TRIBIT_OPT_RELABEL)");
  hipStream_t s = ctx->stream;
  hipEvent_t a = nullptr, b = nullptr;
  TB_HIP(hipEventCreate(&a));
  TB_HIP(hipEventCreate(&b));
  TB_HIP(hipEventRecord(a, s));
  build_layout(ctx, g, layout);
  TB_HIP(hipEventRecord(b, s));
  TB_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  TB_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  g->ms_layout[layout] = ms;
}

// ------------------------------------------- one rank's off-diagonal task
// The reference's off-diagonal task (main6_multi.cu:595-751, task shapes
// main.rs:186-220) counts the triangles a < b < c with (a, b) in `cols` and
// (a, c), (b, c) in `idxs`. As a count of this library: the idxs edges x -> c
// (x the smaller id) form the CSR H; every cols edge {a, b} is one streamed
// list, H's row of one endpoint against the set of H's row of the other, the
// endpoint with the longer row owning the set. The lists go into the reverse
// part of a direction-split layout (no forward part), which the DIRECTED
// schedule counts unchanged.

// Owner-first keys (o << 32) | s of the cols edges: o the endpoint with the
// longer H row (ties: the larger id); sentinels stay sentinels.
__global__ void k_offdiag_keys(const uint64_t* __restrict__ keys, uint64_t nk, const uint32_t* __restrict__ hrp,
                               uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const uint64_t k = keys[i];
  if (k == kSentinel) {
    out[i] = kSentinel;
    return;
  }
  const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
  const uint32_t la = hrp[a + 1] - hrp[a], lb = hrp[b + 1] - hrp[b];
  out[i] = la > lb ? (((uint64_t)a << 32) | b) : (((uint64_t)b << 32) | a);
}

// Sum of v[0, n) into *out (launched on totals_grid(n) blocks of kBlock).
__global__ void __launch_bounds__(kBlock) k_sum_u64(const unsigned long long* __restrict__ v, uint64_t n,
                                                    unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_part[kBlock / 64];
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    acc += v[i];
  acc = wave_sum_u64(acc);
  if (lane_id() == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < kBlock / 64; ++i) acc += s_part[i];
    if (acc) atomicAdd(out, (unsigned long long)acc);
  }
}

// Sorted distinct keys of in[0, nk) (the sentinel dropped) into out; returns
// their number. Sets *bad when a key is not (min < max < n).
static uint64_t sort_unique(tribit_ctx* ctx, const uint64_t* in, uint64_t nk, uint64_t n, uint64_t* out, bool* bad) {
  hipStream_t s = ctx->stream;
  if (!nk) return 0;
  if (nk > 0xFFFFFFFEull) fail(TRIBIT_ERR_TOO_LARGE, "more than 2^32-2 keys in one task set");
  DevBuf<uint32_t> flag(ctx, 1);
  k_fill_u32<<<1, 64, 0, s>>>(flag.p, 1, 0u);
  k_validate<<<grid_for(nk, kBlock), kBlock, 0, s>>>(in, nk, n, flag.p);
  DevBuf<uint64_t> sorted(ctx, nk);
  sort_keys(ctx, in, sorted.p, nk, 32 + bits_for(n));
  DevBuf<uint32_t> head(ctx, nk + 1), pos(ctx, nk + 1);
  k_head_flags<<<grid_for(nk, kBlock), kBlock, 0, s>>>(sorted.p, nk, head.p);
  k_fill_u32<<<1, 64, 0, s>>>(head.p + nk, 1, 0u);
  with_temp(ctx, [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, head.p, pos.p, 0u, (size_t)(nk + 1), rocprim::plus<uint32_t>(), s);
  });
  k_scatter_heads<<<grid_for(nk, kBlock), kBlock, 0, s>>>(sorted.p, nk, pos.p, out);
  uint32_t cnt = 0, f = 0;
  TB_HIP(hipMemcpyAsync(&cnt, pos.p + nk, 4, hipMemcpyDeviceToHost, s));
  TB_HIP(hipMemcpyAsync(&f, flag.p, 4, hipMemcpyDeviceToHost, s));
  ctx->sync();
  *bad = *bad || f != 0;
  return cnt;
}

void build_offdiag(tribit_ctx* ctx, const uint64_t* d_cols, uint64_t ncols, const uint64_t* d_idxs, uint64_t nidxs,
                   uint64_t n, tribit_graph* g) {
  if (n == 0 || n > 0xFFFFFFFEull) fail(TRIBIT_ERR_TOO_LARGE, "vertex count must be in [1, 2^32-2]");
  hipStream_t s = ctx->stream;
  g->ctx = ctx;
  g->n = n;
  g->entries = ncols + nidxs;
  bool bad = false;
  // H: the idxs edges as a CSR by the smaller id
  DevBuf<uint64_t> hk(ctx, nidxs ? nidxs : 1);
  const uint64_t mh = sort_unique(ctx, d_idxs, nidxs, n, hk.p, &bad);
  if (mh > 0xFFFFFFFFull) fail(TRIBIT_ERR_TOO_LARGE, "more than 2^32-1 index edges in one task");
  DevBuf<uint32_t> col(ctx, mh ? mh : 1), row_ptr(ctx, n + 1);
  {
    DevBuf<uint32_t> odeg(ctx, n + 1);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(odeg.p, n + 1, 0u);
    if (mh) k_csr_split<<<grid_for(mh, kBlock), kBlock, 0, s>>>(hk.p, mh, 32, col.p, odeg.p);
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, odeg.p, row_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
  }
  hk.reset();
  // the cols edges, owner first, distinct, as streamed ranges of H rows grouped by owner
  uint64_t mc = 0;
  DevBuf<uint64_t> ck(ctx, ncols ? ncols : 1);
  if (ncols) {
    DevBuf<uint32_t> flag(ctx, 1);
    k_fill_u32<<<1, 64, 0, s>>>(flag.p, 1, 0u);
    k_validate<<<grid_for(ncols, kBlock), kBlock, 0, s>>>(d_cols, ncols, n, flag.p);
    uint32_t f = 0;
    TB_HIP(hipMemcpyAsync(&f, flag.p, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    if (f) fail(TRIBIT_ERR_RANGE, "edge key is not (min < max < n)");
    DevBuf<uint64_t> ok(ctx, ncols);
    k_offdiag_keys<<<grid_for(ncols, kBlock), kBlock, 0, s>>>(d_cols, ncols, row_ptr.p, ok.p);
    // owner-first keys are no longer (min < max): dedup without the range check
    DevBuf<uint64_t> sorted(ctx, ncols);
    sort_keys(ctx, ok.p, sorted.p, ncols, 32 + bits_for(n));
    ok.reset();
    DevBuf<uint32_t> head(ctx, ncols + 1), pos(ctx, ncols + 1);
    k_head_flags<<<grid_for(ncols, kBlock), kBlock, 0, s>>>(sorted.p, ncols, head.p);
    k_fill_u32<<<1, 64, 0, s>>>(head.p + ncols, 1, 0u);
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, head.p, pos.p, 0u, (size_t)(ncols + 1), rocprim::plus<uint32_t>(), s);
    });
    k_scatter_heads<<<grid_for(ncols, kBlock), kBlock, 0, s>>>(sorted.p, ncols, pos.p, ck.p);
    uint32_t cnt = 0;
    TB_HIP(hipMemcpyAsync(&cnt, pos.p + ncols, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    mc = cnt;
  }
  if (bad) fail(TRIBIT_ERR_RANGE, "edge key is not (min < max < n)");
  DevBuf<uint2> rng(ctx, mc ? mc : 1);
  DevBuf<uint32_t> pid(ctx, mc ? mc : 1), fw_ptr(ctx, n + 1), rv_ptr(ctx, n + 1);
  DevBuf<uint64_t> w_fw(ctx, n), w_rv(ctx, n);
  {
    DevBuf<uint32_t> rv_cnt(ctx, n + 1);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(rv_cnt.p, n + 1, 0u);
    k_fill_u32<<<grid_for(n + 1, kBlock), kBlock, 0, s>>>(fw_ptr.p, n + 1, 0u);
    k_fill_u64<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(w_fw.p), n);
    k_fill_u64<<<grid_for(n, kBlock), kBlock, 0, s>>>(reinterpret_cast<unsigned long long*>(w_rv.p), n);
    if (mc)
      k_rv_fill<<<grid_for(mc, kBlock), kBlock, 0, s>>>(ck.p, mc, row_ptr.p, rng.p, pid.p, rv_cnt.p,
                                                        reinterpret_cast<unsigned long long*>(w_rv.p));
    with_temp(ctx, [&](void* t, size_t& b) {
      return rocprim::exclusive_scan(t, b, rv_cnt.p, rv_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s);
    });
  }
  ck.reset();
  // totals: streamed elements, owner rows
  DevBuf<RowStats> st(ctx, 1);
  k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t*>(st.p), sizeof(RowStats) / 4, 0u);
  k_sum_u64<<<totals_grid(n), kBlock, 0, s>>>(reinterpret_cast<const unsigned long long*>(w_rv.p), n,
                                              &st.p->wplus);
  k_row_flags<<<totals_grid(n), kBlock, 0, s>>>(row_ptr.p, nullptr, n, st.p);
  RowStats hst;
  TB_HIP(hipMemcpyAsync(&hst, st.p, sizeof(hst), hipMemcpyDeviceToHost, s));
  ctx->sync();
  g->m = mh;
  g->row_ptr = row_ptr.release();
  g->col = col.release();
  g->relabeled = 0;
  g->m_fwd = 0;
  g->rng_dir = rng.release();
  g->pid_dir = pid.release();
  g->fw_ptr = fw_ptr.release();
  g->rv_ptr = rv_ptr.release();
  g->w_fw = w_fw.release();
  g->w_rv = w_rv.release();
  g->wplus = hst.wplus;  // streamed elements of the task
  g->rows_active = hst.rows_active;
  g->max_outdeg = hst.max_outdeg;
  g->offdiag_lists = mc;
  g->bytes = 4 * (n + 1) * 3 + 4 * (mh ? mh : 1) + 12 * (mc ? mc : 1) + 16 * n;
}

// Releases every device buffer of a graph (the struct stays).
void free_graph_memory(tribit_graph* g) {
  if (!g || !g->ctx) return;
  free_tiles(g);
  core_release(g);
  blocks_release(g);
  bucket_cache_free(g);
  g->ctx->free(g->row_ptr);
  g->ctx->free(g->col);
  g->ctx->free(g->deg);
  g->ctx->free(g->labels);
  g->labels = nullptr;
  g->ctx->free(g->rng_sfx);
  g->ctx->free(g->col16);
  g->col16 = nullptr;
  g->ctx->free(g->sfx_dh);
  g->sfx_dh = nullptr;
  g->ctx->free(g->sfx_split);
  g->sfx_split = nullptr;
  g->ctx->free(g->sfx_ptr);
  g->ctx->free(g->w_sfx);
  g->rng_sfx = nullptr;
  g->sfx_ptr = nullptr;
  g->w_sfx = nullptr;
  g->ctx->free(g->wrow);
  g->ctx->free(g->rng_all);
  g->ctx->free(g->rng_dir);
  g->ctx->free(g->pid_dir);
  g->pid_dir = nullptr;
  g->ctx->free(g->fw_ptr);
  g->ctx->free(g->rv_ptr);
  g->ctx->free(g->w_fw);
  g->ctx->free(g->w_rv);
  g->row_ptr = g->col = g->deg = g->fw_ptr = g->rv_ptr = nullptr;
  g->wrow = g->w_fw = g->w_rv = nullptr;
  g->rng_all = g->rng_dir = nullptr;
}


void free_graph(tribit_graph* g) {
  if (!g) return;
  free_graph_memory(g);
  delete g;
}

}  // namespace tb
