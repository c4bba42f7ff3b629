// tiles.hip -- 32- and 64-column bit tiles of the oriented rows, built on
// demand (gfx950). A tile of width W is (blk, mask): blk = column / W, bit
// (column % W) of the mask (uint32 for W = 32, uint64 for W = 64) set for
// every out-neighbour in that block. A row's tiles are consecutive and
// ascending by blk. Tiles replace the reference's 32-column U/L words
// (src/single_gpu/tribit.cu:588-602, 672-740) for the rows where they are
// dense; the counting kernel streams them and intersects with AND + popcount
// against its LDS bitmap, one 32- or 64-bit word per tile.
//
// Built from the oriented CSR: per-row tile counts, a scan (ptr), a per-row
// fill that also records the tile of every element, then, per range layout on
// first use, the tile range of every streamed element range.
#include <cstring>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr unsigned kTBlock = 256;

// Tiles per row (distinct col >> SH in a sorted row).
template <int SH>
__global__ void k_tile_count(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                             uint32_t* __restrict__ tcnt) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u > n) return;
  uint32_t t = 0;
  if (u < n) {
    uint32_t prev = 0xFFFFFFFFu;
    for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
      const uint32_t b = col[e] >> SH;
      t += (b != prev);
      prev = b;
    }
  }
  tcnt[u] = t;
}

template <int SH, typename M>
__global__ void k_tile_fill(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            const uint32_t* __restrict__ tile_ptr, uint32_t* __restrict__ blk, M* __restrict__ mask,
                            uint32_t* __restrict__ tidx) {
  constexpr uint32_t kLow = (1u << SH) - 1;
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint32_t t = tile_ptr[u], prev = 0xFFFFFFFFu;
  M acc = 0;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t c = col[e], b = c >> SH;
    if (b != prev) {
      if (prev != 0xFFFFFFFFu) {
        blk[t] = prev;
        mask[t] = acc;
        ++t;
      }
      prev = b;
      acc = 0;
    }
    acc |= (M)1 << (c & kLow);
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

static int set_of(int width) {
  if (width != 32 && width != 64) fail(TRIBIT_ERR_ARG, "tile width must be 32 or 64");
  return width == 32 ? 0 : 1;
}

// Per-row tile counts of width `width` into tcnt[n + 1] (last entry 0).
static void tile_counts(tribit_ctx* ctx, const tribit_graph* g, int width, uint32_t* tcnt) {
  const unsigned grid = grid_for(g->n + 1, kTBlock);
  if (width == 32)
    k_tile_count<5><<<grid, kTBlock, 0, ctx->stream>>>(g->row_ptr, g->col, g->n, tcnt);
  else
    k_tile_count<6><<<grid, kTBlock, 0, ctx->stream>>>(g->row_ptr, g->col, g->n, tcnt);
}

// Total tiles of width `width` (one pass over col), for graph info.
uint64_t count_tiles(tribit_ctx* ctx, tribit_graph* g, int width) {
  set_of(width);
  DevBuf<uint32_t> tcnt(ctx, g->n + 1);
  tile_counts(ctx, g, width, tcnt.p);
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

// Builds the tiles of width `width` (once) and their ranges for range layout
// `layout` (PLAIN, DIRECTED or WEDGE) on first use.
void build_tiles(tribit_ctx* ctx, tribit_graph* g, int layout, int width) {
  hipStream_t s = ctx->stream;
  ensure_layout(ctx, g, layout);
  tribit_graph::TileSet& T = g->tiles[set_of(width)];
  const uint64_t n = g->n, m = g->m;
  if (!T.ptr) {
    DevBuf<uint32_t> tile_ptr(ctx, n + 1);
    {
      DevBuf<uint32_t> tcnt(ctx, n + 1);
      tile_counts(ctx, g, width, tcnt.p);
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
    const uint64_t mbytes = (uint64_t)(W ? W : 1) * (width / 8);
    DevBuf<uint32_t> blk(ctx, W ? W : 1), tidx(ctx, m ? m : 1);
    DevBuf<uint8_t> mask(ctx, mbytes);
    if (width == 32)
      k_tile_fill<5, uint32_t><<<grid_for(n, kTBlock), kTBlock, 0, s>>>(
          g->row_ptr, g->col, n, tile_ptr.p, blk.p, reinterpret_cast<uint32_t*>(mask.p), tidx.p);
    else
      k_tile_fill<6, unsigned long long><<<grid_for(n, kTBlock), kTBlock, 0, s>>>(
          g->row_ptr, g->col, n, tile_ptr.p, blk.p, reinterpret_cast<unsigned long long*>(mask.p), tidx.p);
    TB_HIP(hipGetLastError());
    ctx->sync();
    T.ntiles = W;
    T.ptr = tile_ptr.release();
    T.blk = blk.release();
    T.mask = mask.release();
    T.idx = tidx.release();
    g->bytes += 4 * (n + 1) + (4ull + width / 8) * W + 4 * m;
  }
  const uint2* src = layout == TRIBIT_SCHED_PLAIN ? g->rng_all : layout == TRIBIT_SCHED_WEDGE ? g->rng_sfx
                                                                                               : g->rng_dir;
  uint2** dst = &T.trng[layout & 3];
  if (*dst || !src) return;
  DevBuf<uint2> tr(ctx, m ? m : 1);
  if (m) k_tile_ranges<<<grid_for(m, kTBlock), kTBlock, 0, s>>>(src, m, T.idx, tr.p);
  TB_HIP(hipGetLastError());
  ctx->sync();
  *dst = tr.release();
  g->bytes += 8 * (m ? m : 1);
}

void free_tiles(tribit_graph* g) {
  tribit_ctx* c = g->ctx;
  for (auto& T : g->tiles) {
    c->free(T.ptr);
    c->free(T.blk);
    c->free(T.mask);
    c->free(T.idx);
    for (auto& r : T.trng) c->free(r);
    T = tribit_graph::TileSet{};
  }
}

}  // namespace tb
