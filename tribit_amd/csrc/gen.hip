// gen.hip -- on-GPU synthetic graph generators (Chung-Lu, 2-D random
// geometric, stochastic block model), gfx950. R-MAT and Erdos-Renyi live in
// prep.hip. Every generator writes packed keys (min<<32)|max with self-loops
// and rejected slots as the sentinel; preprocessing deduplicates.
//
// All randomness is the counter-based stream of device.hpp, so the CPU
// oracle (oracle/oracle.c) regenerates the identical key set; floating point
// appears only in host-side set-up (Chung-Lu CDF, RGG radius, thresholds),
// computed the same way on both sides.
#include <cmath>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <thread>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kGenBlock = 256;
constexpr uint64_t kStreamCl = 0x43484C55ull;    // "CHLU"
constexpr uint64_t kStreamClp = 0x43484C50ull;   // "CHLP"
constexpr uint64_t kStreamRgg = 0x52474732ull;   // "RGG2"
constexpr uint64_t kStreamSbmIn = 0x53424D49ull;   // "SBMI"
constexpr uint64_t kStreamSbmOut = 0x53424D4Full;  // "SBMO"

static unsigned grid_stride(tribit_ctx* ctx, uint64_t items) {
  uint64_t g = (items + kGenBlock - 1) / kGenBlock;
  const uint64_t cap = (uint64_t)(ctx->cus > 0 ? ctx->cus : 256) * 64;
  return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

// ------------------------------------------------------------- Chung-Lu

uint64_t chung_lu_samples(uint64_t n, double avg_deg) { return (uint64_t)((double)n * avg_deg / 2.0 + 0.5); }

// cdf_i = floor(sum_{j<=i} w_j / sum w * 2^62), w_j = (j+1)^(-1/(gamma-1)); last = 2^62.
void chung_lu_cdf(uint64_t n, double gamma, uint64_t* cdf) {
  const double e = -1.0 / (gamma - 1.0);
  std::vector<double> cum(n);
  double acc = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    acc += std::pow((double)(i + 1), e);
    cum[i] = acc;
  }
  const double scale = 4611686018427387904.0 / acc;
  for (uint64_t i = 0; i + 1 < n; ++i) {
    double v = cum[i] * scale;
    if (v > 4611686018427387904.0) v = 4611686018427387904.0;
    cdf[i] = (uint64_t)v;
  }
  cdf[n - 1] = 1ull << 62;
}

void chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg) {
  const uint64_t key = stream_key(seed, kStreamClp);
  auto unit = [](uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); };
  const double u0 = unit(rand_word(key, 3 * index));
  const double u1 = unit(rand_word(key, 3 * index + 1));
  const double u2 = unit(rand_word(key, 3 * index + 2));
  *gamma = 2.1 + 0.7 * u0;
  *n = (uint64_t)(1024.0 * std::exp2(8.0 * u1) + 0.5);
  *avg_deg = 8.0 + 56.0 * u2;
}

__device__ __forceinline__ uint64_t cdf_pick(const uint64_t* __restrict__ cdf, uint64_t n, uint64_t r) {
  uint64_t lo = 0, hi = n;  // first i with cdf[i] > r
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (cdf[mid] <= r)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n ? lo : n - 1;
}

__global__ void k_gen_chung_lu(const uint64_t* __restrict__ cdf, uint64_t n, uint64_t ne, uint64_t key,
                               uint64_t* __restrict__ keys) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += stride) {
    const uint64_t u = cdf_pick(cdf, n, rand_word(key, 2 * e) >> 2);
    const uint64_t v = cdf_pick(cdf, n, rand_word(key, 2 * e + 1) >> 2);
    keys[e] = (u == v) ? kSentinel : (u < v ? ((u << 32) | v) : ((v << 32) | u));
  }
}

void gen_chung_lu(tribit_ctx* ctx, uint64_t n, double gamma, double avg_deg, uint64_t seed, uint64_t* d_keys) {
  std::vector<uint64_t> cdf(n);
  chung_lu_cdf(n, gamma, cdf.data());
  DevBuf<uint64_t> dcdf(ctx, n);
  TB_HIP(hipMemcpyAsync(dcdf.p, cdf.data(), n * 8, hipMemcpyHostToDevice, ctx->stream));
  const uint64_t ne = chung_lu_samples(n, avg_deg);
  if (ne)
    k_gen_chung_lu<<<grid_stride(ctx, ne), kGenBlock, 0, ctx->stream>>>(dcdf.p, n, ne, stream_key(seed, kStreamCl),
                                                                        d_keys);
  TB_HIP(hipGetLastError());
  ctx->sync();  // the host CDF buffer goes out of scope
}

// ------------------------------------------------------- random geometric

struct RggSpec {
  uint64_t R, C;  // radius and cells per side, in 2^-30 units of the unit square
};

RggSpec rgg_spec(uint64_t n, double avg_deg) {
  RggSpec s;
  s.R = (uint64_t)(std::sqrt(avg_deg / (3.14159265358979323846 * (double)n)) * 1073741824.0);
  s.C = s.R ? ((1ull << 30) + s.R - 1) / s.R : 0;
  return s;
}

__global__ void k_rgg_points(uint64_t n, uint64_t key, uint64_t R, uint64_t C, uint64_t* __restrict__ ord) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t w = rand_word(key, i);
  const uint64_t x = w & 0x3FFFFFFFull, y = (w >> 32) & 0x3FFFFFFFull;
  const uint64_t cell = (y / R) * C + (x / R);
  ord[i] = (cell << 32) | i;
}

// Sorted order -> coordinates by new id, cell counts (sorted: run-aggregated).
__global__ void k_rgg_place(const uint64_t* __restrict__ ord, uint64_t n, uint64_t key, uint32_t* __restrict__ sx,
                            uint32_t* __restrict__ sy, uint32_t* __restrict__ ccount) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < n;
  uint32_t cell = 0xFFFFFFFFu;
  if (valid) {
    const uint64_t o = ord[k];
    const uint64_t w = rand_word(key, o & 0xFFFFFFFFull);
    sx[k] = (uint32_t)(w & 0x3FFFFFFFull);
    sy[k] = (uint32_t)((w >> 32) & 0x3FFFFFFFull);
    cell = (uint32_t)(o >> 32);
  }
  const uint32_t run = wave_run_length(cell, valid);
  if (run) atomicAdd(&ccount[cell], run);
}

// Pass 0 counts, pass 1 writes the edges (k, j), j > k, of point k.
template <bool WRITE>
__global__ void k_rgg_edges(const uint32_t* __restrict__ sx, const uint32_t* __restrict__ sy,
                            const uint32_t* __restrict__ cstart, uint64_t n, uint64_t R, uint64_t C,
                            uint32_t* __restrict__ ecount, const uint64_t* __restrict__ eoff,
                            uint64_t* __restrict__ keys) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t x = sx[k], y = sy[k], R2 = R * R;
  const int64_t cx = (int64_t)(x / R), cy = (int64_t)(y / R);
  uint64_t cnt = 0, o = WRITE ? eoff[k] : 0;
  for (int64_t dy = -1; dy <= 1; ++dy)
    for (int64_t dx = -1; dx <= 1; ++dx) {
      const int64_t nx = cx + dx, ny = cy + dy;
      if (nx < 0 || ny < 0 || nx >= (int64_t)C || ny >= (int64_t)C) continue;
      const uint64_t c = (uint64_t)ny * C + (uint64_t)nx;
      for (uint64_t j = cstart[c]; j < cstart[c + 1]; ++j) {
        if (j <= k) continue;
        const uint64_t ddx = x > sx[j] ? x - sx[j] : sx[j] - x;
        const uint64_t ddy = y > sy[j] ? y - sy[j] : sy[j] - y;
        if (ddx * ddx + ddy * ddy <= R2) {
          if (WRITE)
            keys[o++] = (k << 32) | j;
          else
            ++cnt;
        }
      }
    }
  if (!WRITE) ecount[k] = (uint32_t)cnt;
}

// Returns the device keys and their number (exact, no sentinels).
uint64_t* gen_rgg(tribit_ctx* ctx, uint64_t n, double avg_deg, uint64_t seed, uint64_t* nkeys) {
  const RggSpec sp = rgg_spec(n, avg_deg);
  if (sp.R < 1 || sp.C * sp.C > 0xFFFFFFFEull) fail(TRIBIT_ERR_ARG, "random geometric graph radius out of range");
  hipStream_t s = ctx->stream;
  const uint64_t key = stream_key(seed, kStreamRgg), ncell = sp.C * sp.C;
  DevBuf<uint32_t> sx(ctx, n), sy(ctx, n), cstart(ctx, ncell + 1);
  {
    DevBuf<uint64_t> ord(ctx, n), ord2(ctx, n);
    k_rgg_points<<<grid_for(n, kGenBlock), kGenBlock, 0, s>>>(n, key, sp.R, sp.C, ord.p);
    uint32_t bits = 1;
    while ((1ull << bits) < ncell) ++bits;
    size_t tb = 0;
    TB_HIP(rocprim::radix_sort_keys(nullptr, tb, ord.p, ord2.p, (size_t)n, 0u, 32u + bits, s));
    DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
    TB_HIP(rocprim::radix_sort_keys(tmp.p, tb, ord.p, ord2.p, (size_t)n, 0u, 32u + bits, s));
    DevBuf<uint32_t> ccount(ctx, ncell + 1);
    TB_HIP(hipMemsetAsync(ccount.p, 0, (ncell + 1) * 4, s));
    k_rgg_place<<<grid_for(n, kGenBlock), kGenBlock, 0, s>>>(ord2.p, n, key, sx.p, sy.p, ccount.p);
    size_t sb = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, sb, ccount.p, cstart.p, 0u, (size_t)(ncell + 1),
                                   rocprim::plus<uint32_t>(), s));
    DevBuf<uint8_t> stmp(ctx, sb ? sb : 8);
    TB_HIP(rocprim::exclusive_scan(stmp.p, sb, ccount.p, cstart.p, 0u, (size_t)(ncell + 1),
                                   rocprim::plus<uint32_t>(), s));
  }
  DevBuf<uint32_t> ecount(ctx, n + 1);
  DevBuf<uint64_t> eoff(ctx, n + 1);
  TB_HIP(hipMemsetAsync(ecount.p + n, 0, 4, s));
  k_rgg_edges<false><<<grid_for(n, kGenBlock), kGenBlock, 0, s>>>(sx.p, sy.p, cstart.p, n, sp.R, sp.C, ecount.p,
                                                                  nullptr, nullptr);
  size_t sb = 0;
  TB_HIP(rocprim::exclusive_scan(nullptr, sb, ecount.p, eoff.p, (uint64_t)0, (size_t)(n + 1),
                                 rocprim::plus<uint64_t>(), s));
  DevBuf<uint8_t> stmp(ctx, sb ? sb : 8);
  TB_HIP(rocprim::exclusive_scan(stmp.p, sb, ecount.p, eoff.p, (uint64_t)0, (size_t)(n + 1),
                                 rocprim::plus<uint64_t>(), s));
  uint64_t ne = 0;
  TB_HIP(hipMemcpyAsync(&ne, eoff.p + n, 8, hipMemcpyDeviceToHost, s));
  ctx->sync();
  uint64_t* keys = ctx->alloc_n<uint64_t>(ne ? ne : 1);
  if (ne)
    k_rgg_edges<true><<<grid_for(n, kGenBlock), kGenBlock, 0, s>>>(sx.p, sy.p, cstart.p, n, sp.R, sp.C, nullptr,
                                                                   eoff.p, keys);
  TB_HIP(hipGetLastError());
  *nkeys = ne;
  return keys;
}

// ------------------------------------------------------ stochastic blocks

uint64_t sbm_out_samples(uint64_t blocks, uint64_t size, double p_out) {
  const double n = (double)(blocks * size);
  const double cross = n * (n - 1.0) / 2.0 - (double)blocks * (double)size * ((double)size - 1.0) / 2.0;
  return (uint64_t)(p_out * cross + 0.5);
}

static uint64_t prob_threshold(double p) {
  if (p <= 0.0) return 0;
  if (p >= 1.0) return UINT64_MAX;
  return (uint64_t)(p * 18446744073709551616.0);
}

__global__ void k_sbm_in(uint64_t blocks, uint64_t size, uint64_t key, uint64_t thr, uint64_t* __restrict__ keys) {
  const uint64_t ss = size * size, total = blocks * ss;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint64_t b = t / ss, r = t % ss, i = r / size, j = r % size;
    uint64_t k = kSentinel;
    if (i < j && rand_word(key, t) < thr) k = ((b * size + i) << 32) | (b * size + j);
    keys[t] = k;
  }
}

__global__ void k_sbm_out(uint64_t n, uint64_t size, uint64_t mo, uint64_t key, uint64_t* __restrict__ keys) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mo; e += stride) {
    const uint64_t w = rand_word(key, e);
    const uint64_t u = ((w & 0xFFFFFFFFull) * n) >> 32, v = ((w >> 32) * n) >> 32;
    uint64_t k = kSentinel;
    if (u / size != v / size) k = u < v ? ((u << 32) | v) : ((v << 32) | u);
    keys[e] = k;
  }
}

uint64_t sbm_slots(uint64_t blocks, uint64_t size, double p_out) {
  return blocks * size * size + sbm_out_samples(blocks, size, p_out);
}

void gen_sbm(tribit_ctx* ctx, uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed,
             uint64_t* d_keys) {
  const uint64_t n = blocks * size, inner = blocks * size * size, mo = sbm_out_samples(blocks, size, p_out);
  k_sbm_in<<<grid_stride(ctx, inner), kGenBlock, 0, ctx->stream>>>(blocks, size, stream_key(seed, kStreamSbmIn),
                                                                   prob_threshold(p_in), d_keys);
  if (mo)
    k_sbm_out<<<grid_stride(ctx, mo), kGenBlock, 0, ctx->stream>>>(n, size, mo, stream_key(seed, kStreamSbmOut),
                                                                   d_keys + inner);
  TB_HIP(hipGetLastError());
}

}  // namespace tb
