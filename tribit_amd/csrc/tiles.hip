// tiles.hip -- 64-column bit tiles of the oriented rows, built on demand
// (gfx950). A tile is (blk, mask): blk = column / 64, bit (column % 64) of
// the u64 mask set for every out-neighbour in that block. A row's tiles are
// consecutive and ascending by blk. Tiles replace the reference's 32-column
// U/L words (src/single_gpu/tribit.cu:588-602, 672-740) for the rows where
// they are dense; the counting kernel streams them and intersects with
// AND + popcount.
//
// Built from the oriented CSR: per-row tile counts, a scan (tile_ptr), a
// per-row fill, then the tile ranges of every streamed list in both range
// layouts (trng_all in CSR order, trng_dir in rng_dir order) and, per owner
// list, the total tiles streamed (the density heuristic compares it with
// the elements streamed).
#include <cstring>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kTBlock = 256;

// Tiles per row (distinct col >> 6 in a sorted row); totals per wave.
__global__ void k_tile_count(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                             uint32_t* __restrict__ tcnt) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  uint32_t t = 0;
  if (u < n) {
    uint32_t prev = 0xFFFFFFFFu;
    for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
      const uint32_t b = col[e] >> 6;
      t += (b != prev);
      prev = b;
    }
  }
  tcnt[u] = t;
}

__global__ void k_tile_fill(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            const uint32_t* __restrict__ tile_ptr, uint32_t* __restrict__ blk,
                            unsigned long long* __restrict__ mask) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint32_t t = tile_ptr[u], prev = 0xFFFFFFFFu;
  unsigned long long acc = 0;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t c = col[e], b = c >> 6;
    if (b != prev) {
      if (prev != 0xFFFFFFFFu) {
        blk[t] = prev;
        mask[t] = acc;
        ++t;
      }
      prev = b;
      acc = 0;
    }
    acc |= 1ull << (c & 63);
  }
  if (prev != 0xFFFFFFFFu) {
    blk[t] = prev;
    mask[t] = acc;
  }
}

// Tile range of the list each range entry streams: partner ids from `pid`
// (or, when pid is null, the CSR column itself).
__global__ void k_tile_ranges(const uint32_t* __restrict__ tile_ptr, const uint32_t* __restrict__ pid, uint64_t m,
                              uint2* __restrict__ trng) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t p = pid[k];
  trng[k] = make_uint2(tile_ptr[p], tile_ptr[p + 1]);
}

// Per owner list [ptr[x], ptr[x+1]) of a range layout: tiles streamed.
__global__ void k_tile_work(const uint32_t* __restrict__ ptr, const uint2* __restrict__ trng, uint64_t n,
                            uint64_t* __restrict__ tw) {
  const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  uint64_t w = 0;
  for (uint32_t k = ptr[x]; k < ptr[x + 1]; ++k) {
    const uint2 r = trng[k];
    w += r.y - r.x;
  }
  tw[x] = w;
}

// Total tiles of the graph (cheap: one pass over col), for graph info and the
// decision to build the tiles at all.
uint64_t count_tiles(tribit_ctx* ctx, tribit_graph* g) {
  DevBuf<uint32_t> tcnt(ctx, g->n + 1);
  k_tile_count<<<grid_for(g->n + 1, kTBlock), kTBlock, 0, ctx->stream>>>(g->row_ptr, g->col, g->n, tcnt.p);
  DevBuf<uint64_t> total(ctx, 1);
  size_t tb = 0;
  TB_HIP(rocprim::reduce(nullptr, tb, tcnt.p, total.p, (uint64_t)0, (size_t)(g->n + 1), rocprim::plus<uint64_t>(),
                         ctx->stream));
  DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
  TB_HIP(rocprim::reduce(tmp.p, tb, tcnt.p, total.p, (uint64_t)0, (size_t)(g->n + 1), rocprim::plus<uint64_t>(),
                         ctx->stream));
  uint64_t h = 0;
  TB_HIP(hipMemcpyAsync(&h, total.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  return h;
}

void build_tiles(tribit_ctx* ctx, tribit_graph* g) {
  if (g->tile_ptr) return;
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, m = g->m;
  DevBuf<uint32_t> tile_ptr(ctx, n + 1);
  {
    DevBuf<uint32_t> tcnt(ctx, n + 1);
    k_tile_count<<<grid_for(n + 1, kTBlock), kTBlock, 0, s>>>(g->row_ptr, g->col, n, tcnt.p);
    size_t tb = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, tb, tcnt.p, tile_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(),
                                   s));
    DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
    TB_HIP(rocprim::exclusive_scan(tmp.p, tb, tcnt.p, tile_ptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(),
                                   s));
  }
  uint32_t W = 0;
  TB_HIP(hipMemcpyAsync(&W, tile_ptr.p + n, 4, hipMemcpyDeviceToHost, s));
  ctx->sync();
  DevBuf<uint32_t> blk(ctx, W ? W : 1);
  DevBuf<uint64_t> mask(ctx, W ? W : 1);
  k_tile_fill<<<grid_for(n, kTBlock), kTBlock, 0, s>>>(g->row_ptr, g->col, n, tile_ptr.p, blk.p,
                                                     reinterpret_cast<unsigned long long*>(mask.p));
  DevBuf<uint2> trng_all(ctx, m ? m : 1), trng_dir(ctx, m ? m : 1);
  DevBuf<uint64_t> tw_all(ctx, n), tw_fw(ctx, n), tw_rv(ctx, n);
  if (m) {
    k_tile_ranges<<<grid_for(m, kTBlock), kTBlock, 0, s>>>(tile_ptr.p, g->col, m, trng_all.p);
    k_tile_ranges<<<grid_for(m, kTBlock), kTBlock, 0, s>>>(tile_ptr.p, g->pid_dir, m, trng_dir.p);
  }
  k_tile_work<<<grid_for(n, kTBlock), kTBlock, 0, s>>>(g->row_ptr, trng_all.p, n, tw_all.p);
  k_tile_work<<<grid_for(n, kTBlock), kTBlock, 0, s>>>(g->fw_ptr, trng_dir.p, n, tw_fw.p);
  k_tile_work<<<grid_for(n, kTBlock), kTBlock, 0, s>>>(g->rv_ptr, trng_dir.p, n, tw_rv.p);
  TB_HIP(hipGetLastError());
  ctx->sync();
  g->ntiles = W;
  g->tile_ptr = tile_ptr.release();
  g->tile_blk = blk.release();
  g->tile_mask = mask.release();
  g->trng_all = trng_all.release();
  g->trng_dir = trng_dir.release();
  g->tw_all = tw_all.release();
  g->tw_fw = tw_fw.release();
  g->tw_rv = tw_rv.release();
  g->bytes += 4 * (n + 1) + 12ull * W + 16 * m + 24 * n;
}

void free_tiles(tribit_graph* g) {
  tribit_ctx* c = g->ctx;
  c->free(g->tile_ptr);
  c->free(g->tile_blk);
  c->free(g->tile_mask);
  c->free(g->trng_all);
  c->free(g->trng_dir);
  c->free(g->tw_all);
  c->free(g->tw_fw);
  c->free(g->tw_rv);
  g->tile_ptr = g->tile_blk = nullptr;
  g->tile_mask = g->tw_all = g->tw_fw = g->tw_rv = nullptr;
  g->trng_all = g->trng_dir = nullptr;
}

}  // namespace tb
