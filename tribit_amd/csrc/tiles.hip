// tiles.hip -- 64-column bit tiles of the oriented rows, built on demand
// (gfx950). A tile is (blk, mask): blk = column / 64, bit (column % 64) of
// the u64 mask set for every out-neighbour in that block. A row's tiles are
// consecutive and ascending by blk. Tiles replace the reference's 32-column
// U/L words (src/single_gpu/tribit.cu:588-602, 672-740) for the rows where
// they are dense; the counting kernel streams them and intersects with
// AND + popcount.
//
// Built from the oriented CSR: per-row tile counts, a scan (tile_ptr), a
// per-row fill that also records the tile of every element, then, per range
// layout on first use, the tile range of every streamed element range.
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
                            unsigned long long* __restrict__ mask, uint32_t* __restrict__ tidx) {
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
    tidx[e] = t;  // tile holding element e
  }
  if (prev != 0xFFFFFFFFu) {
    blk[t] = prev;
    mask[t] = acc;
  }
}

// Tile range of an element range [b, e) of one row: the tiles holding its
// first and last element. Bits of the first and last tile outside [b, e)
// belong to the same row; the schedules that cut ranges (the wedge suffix
// after the owner's id, id-range chunks) cut at ids the owner's set does not
// hold, so those bits never match.
__global__ void k_tile_ranges(const uint2* __restrict__ rng, uint64_t m, const uint32_t* __restrict__ tidx,
                              uint2* __restrict__ trng) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint2 r = rng[k];
  const uint32_t e = r.y & range_end_mask(m);  // bits 31 and 30: narrow-range flags of wedge ranges
  trng[k] = r.x < e ? make_uint2(tidx[r.x], tidx[e - 1] + 1) : make_uint2(0u, 0u);
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

// Builds the tiles (once) and the tile ranges of range layout `layout`
// (PLAIN: rng_all, DIRECTED: rng_dir, WEDGE: rng_sfx) on first use.
void build_tiles(tribit_ctx* ctx, tribit_graph* g, int layout) {
  hipStream_t s = ctx->stream;
  const uint64_t n = g->n, m = g->m;
  if (!g->tile_ptr) {
    DevBuf<uint32_t> tile_ptr(ctx, n + 1);
    {
      DevBuf<uint32_t> tcnt(ctx, n + 1);
      k_tile_count<<<grid_for(n + 1, kTBlock), kTBlock, 0, s>>>(g->row_ptr, g->col, n, tcnt.p);
      size_t tb = 0;
      TB_HIP(rocprim::exclusive_scan(nullptr, tb, tcnt.p, tile_ptr.p, 0u, (size_t)(n + 1),
                                     rocprim::plus<uint32_t>(), s));
      DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
      TB_HIP(rocprim::exclusive_scan(tmp.p, tb, tcnt.p, tile_ptr.p, 0u, (size_t)(n + 1),
                                     rocprim::plus<uint32_t>(), s));
    }
    uint32_t W = 0;
    TB_HIP(hipMemcpyAsync(&W, tile_ptr.p + n, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    DevBuf<uint32_t> blk(ctx, W ? W : 1), tidx(ctx, m ? m : 1);
    DevBuf<uint64_t> mask(ctx, W ? W : 1);
    k_tile_fill<<<grid_for(n, kTBlock), kTBlock, 0, s>>>(g->row_ptr, g->col, n, tile_ptr.p, blk.p,
                                                       reinterpret_cast<unsigned long long*>(mask.p), tidx.p);
    TB_HIP(hipGetLastError());
    ctx->sync();
    g->ntiles = W;
    g->tile_ptr = tile_ptr.release();
    g->tile_blk = blk.release();
    g->tile_mask = mask.release();
    g->tile_idx = tidx.release();
    g->bytes += 4 * (n + 1) + 12ull * W + 4 * m;
  }
  uint2** dst = layout == TRIBIT_SCHED_PLAIN ? &g->trng_all : layout == TRIBIT_SCHED_WEDGE ? &g->trng_sfx
                                                                                             : &g->trng_dir;
  const uint2* src = layout == TRIBIT_SCHED_PLAIN ? g->rng_all : layout == TRIBIT_SCHED_WEDGE ? g->rng_sfx
                                                                                               : g->rng_dir;
  if (*dst || !src) return;
  DevBuf<uint2> tr(ctx, m ? m : 1);
  if (m) k_tile_ranges<<<grid_for(m, kTBlock), kTBlock, 0, s>>>(src, m, g->tile_idx, tr.p);
  TB_HIP(hipGetLastError());
  ctx->sync();
  *dst = tr.release();
  g->bytes += 8 * (m ? m : 1);
}

void free_tiles(tribit_graph* g) {
  tribit_ctx* c = g->ctx;
  c->free(g->tile_ptr);
  c->free(g->tile_blk);
  c->free(g->tile_mask);
  c->free(g->tile_idx);
  c->free(g->trng_all);
  c->free(g->trng_dir);
  c->free(g->trng_sfx);
  g->tile_ptr = g->tile_blk = g->tile_idx = nullptr;
  g->tile_mask = nullptr;
  g->trng_all = g->trng_dir = g->trng_sfx = nullptr;
  g->ntiles = 0;
}

}  // namespace tb
