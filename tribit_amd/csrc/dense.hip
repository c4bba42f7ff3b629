// dense.hip -- the dense core on the matrix cores (gfx950 int8 MFMA).
//
// No counterpart in the reference (whose b1 tensor-core MMA,
// src/single_gpu/tribit.cu:561-576, has no CDNA equivalent). With vertex ids
// in degree-rank order the highest-degree vertices form the top id range
// [d0, n) and induce a dense subgraph (the "core"). Every triangle a < b < c
// whose smallest vertex a lies in the core has all three vertices there, and
// those triangles are
//     sum over r < c of A[r][c] * (A A^T)[r][c]
// for the core's oriented 0/1 adjacency A (upper triangular). That is a
// masked int8 GEMM with exact int32 accumulation: 128 x 128 output tiles of
// the upper triangle, K restricted to columns >= the tile's first column
// (A[c][k] = 0 for k <= c), 32 x 32 x 32 MFMA blocks fed from LDS. The
// streaming kernels skip exactly these triangles (the scheduler drops every
// list entry whose smallest vertex is in the core), so the two sums add up.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr int kCoreTile = 128;  // output tile side
constexpr int kCoreK = 128;     // K bytes per LDS stage
constexpr int kCorePad = 16;    // LDS row padding (bytes)
constexpr int kCoreThreads = 256;

typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));

__global__ void k_core_zero(uint4* __restrict__ p, uint64_t n16) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// The top core as bits: row u - d0 of kp bits (kp / 32 words) holds bit v - d0
// for every arc u -> v with u >= d0. A thread owns its row's words, so no
// atomics are needed.
__global__ void k_core_fill_bits(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                                 uint64_t d0, uint32_t kp, uint32_t* __restrict__ bits) {
  const uint64_t u = d0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  uint32_t* row = bits + (u - d0) * (uint64_t)(kp / 32);
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t c = col[e] - (uint32_t)d0;
    row[c >> 5] |= 1u << (c & 31);
  }
}

// Four bits to four bytes 0/1: the nibble's copies shifted by 0, 7, 14 and 21
// do not overlap, and bit k of the nibble lands on bit 8k.
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t nib) { return __umul24(nib, 0x204081u) & 0x01010101u; }

// 64 bits to 64 int8 0/1 values, bit k to byte k.
__device__ __forceinline__ void expand_bits64(uint2 b, int8_t* __restrict__ dst) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t h = ((q < 2) ? b.x : b.y) >> ((q & 1) * 16);
    uint4 o;
    o.x = nibble_bytes(h & 0xFu);
    o.y = nibble_bytes((h >> 4) & 0xFu);
    o.z = nibble_bytes((h >> 8) & 0xFu);
    o.w = nibble_bytes((h >> 12) & 0xFu);
    *reinterpret_cast<uint4*>(dst + 16 * q) = o;
  }
}

// Output row of accumulator element v of lane l in a 32 x 32 MFMA block
// (column = l % 32): rows come in four groups of 8, each lane half holding 4.
__device__ __forceinline__ int mfma_row(int v, int l) { return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3); }

// One 128 x 128 output tile (I <= J) of the masked product over a kp x kp 0/1
// upper-triangular matrix A: adds sum over rows r of tile block I and columns
// c of tile block J of A[r][c] * (A A^T)[r][c] to *result.
__device__ __forceinline__ void core_tile(const int8_t* __restrict__ A, uint32_t kp, uint32_t I, uint32_t J,
                                          unsigned long long* __restrict__ result) {
  // two LDS stages: stage s is written while the MFMAs of stage s - 1 read
  // the other buffer, so one barrier per stage orders both
  __shared__ __attribute__((aligned(16))) int8_t sA[2][kCoreTile][kCoreK + kCorePad];
  __shared__ __attribute__((aligned(16))) int8_t sB[2][kCoreTile][kCoreK + kCorePad];
  __shared__ unsigned long long sPart[kCoreThreads / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wr = w >> 1, wc = w & 1;
  const uint64_t rowA = (uint64_t)I * kCoreTile, rowB = (uint64_t)J * kCoreTile;
  v16i acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;
  // staging: thread tid moves 64 bytes of row tid / 2 of each operand per
  // stage; the next stage's loads are in flight while the MFMAs run
  const uint32_t lr = tid >> 1, lh = (tid & 1) * 64;
  const uint32_t k_begin = J * kCoreTile;
  const uint4* ga = reinterpret_cast<const uint4*>(A + (rowA + lr) * kp + k_begin + lh);
  const uint4* gb = reinterpret_cast<const uint4*>(A + (rowB + lr) * kp + k_begin + lh);
  uint4 pa0 = ga[0], pa1 = ga[1], pa2 = ga[2], pa3 = ga[3];
  uint4 pb0 = gb[0], pb1 = gb[1], pb2 = gb[2], pb3 = gb[3];
  // the epilogue's mask A[r][c] over this output tile is the first K stage of
  // the A operand (columns from J * 128): taken from LDS into one bit per
  // accumulator element instead of 64 byte reads per lane from global memory
  uint64_t msk = 0;
  uint32_t buf = 0;
  for (uint32_t k0 = k_begin; k0 < kp; k0 += kCoreK, buf ^= 1u) {
    // a wave stores stage s only after every wave passed the barrier of stage
    // s - 1, that is after the MFMAs of stage s - 2 read this buffer
    *reinterpret_cast<uint4*>(&sA[buf][lr][lh]) = pa0;
    *reinterpret_cast<uint4*>(&sA[buf][lr][lh + 16]) = pa1;
    *reinterpret_cast<uint4*>(&sA[buf][lr][lh + 32]) = pa2;
    *reinterpret_cast<uint4*>(&sA[buf][lr][lh + 48]) = pa3;
    *reinterpret_cast<uint4*>(&sB[buf][lr][lh]) = pb0;
    *reinterpret_cast<uint4*>(&sB[buf][lr][lh + 16]) = pb1;
    *reinterpret_cast<uint4*>(&sB[buf][lr][lh + 32]) = pb2;
    *reinterpret_cast<uint4*>(&sB[buf][lr][lh + 48]) = pb3;
    __syncthreads();
    if (k0 == k_begin) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int r = wr * 64 + i * 32 + mfma_row(v, lane), c = wc * 64 + j * 32 + (lane & 31);
            msk |= (uint64_t)(sA[0][r][c] != 0) << ((i * 2 + j) * 16 + v);
          }
    }
    if (k0 + kCoreK < kp) {  // next stage in flight during the MFMAs
      ga += kCoreK / 16;
      gb += kCoreK / 16;
      pa0 = ga[0], pa1 = ga[1], pa2 = ga[2], pa3 = ga[3];
      pb0 = gb[0], pb1 = gb[1], pb2 = gb[2], pb3 = gb[3];
    }
#pragma unroll
    for (int kk = 0; kk < kCoreK; kk += 32) {
      const int ko = kk + (lane >> 5) * 16;  // the same k slice for both operands
      v4i fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        fa[i] = *reinterpret_cast<const v4i*>(&sA[buf][wr * 64 + i * 32 + (lane & 31)][ko]);
        fb[i] = *reinterpret_cast<const v4i*>(&sB[buf][wc * 64 + i * 32 + (lane & 31)][ko]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }
  // masked sum: (A A^T)[r][c] counts where A[r][c] = 1
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if ((msk >> ((i * 2 + j) * 16 + v)) & 1) sum += (uint32_t)acc[i][j][v];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if (lane == 0) sPart[w] = sum;
  __syncthreads();
  if (tid == 0) {
    unsigned long long s = 0;
    for (int i = 0; i < kCoreThreads / 64; ++i) s += sPart[i];
    if (s) atomicAdd(result, s);
  }
}

// The top core's tile: 256 x 256 outputs per 512-thread block, each of its 8
// waves a 128 x 64 sub-tile of 4 x 2 MFMA blocks (128 accumulator registers
// per lane, two waves per SIMD). The core is kept as bits (kp^2 / 8 bytes:
// 32 MiB for 16,384 ids, resident in the last-level cache, where the int8
// matrix took 256 MiB and came from HBM); each 128-wide K stage reads 8 KiB
// of bits for 8.4M MACs and expands them to int8 0/1 in LDS (a multiply and
// a mask per four values). Two LDS stages (147 KiB: one block per CU). kp is
// a multiple of 256, so the output tile's mask, A over the tile's own 256
// columns, is the first two K stages of the A operand.
constexpr int kCoreTile2 = 256;
constexpr int kCoreThreads2 = 512;

__device__ __forceinline__ void core_tile256(const uint32_t* __restrict__ bits, uint32_t kp, uint32_t I, uint32_t J,
                                             unsigned long long* __restrict__ result) {
  constexpr int NW = kCoreThreads2 / 64;
  constexpr int WN = 64;                  // columns per wave
  constexpr int NJ = WN / 32;             // MFMA blocks per wave along the columns
  constexpr int WC = kCoreTile2 / WN;     // waves along the columns (two rows of waves)
  static_assert(WC * 2 == NW, "waves tile the 256 x 256 output");
  __shared__ __attribute__((aligned(16))) int8_t sA[2][kCoreTile2][kCoreK + kCorePad];
  __shared__ __attribute__((aligned(16))) int8_t sB[2][kCoreTile2][kCoreK + kCorePad];
  __shared__ unsigned long long sPart[NW];
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wr = w / WC, wc = w % WC;  // rows wr * 128, columns wc * WN of the tile
  const uint64_t rowA = (uint64_t)I * kCoreTile2, rowB = (uint64_t)J * kCoreTile2;
  const uint64_t kpw = kp / 32;  // words per bit row
  v16i acc[4][NJ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;
  // staging: two threads per row, 64 bits (= 64 int8 values) of each operand per stage
  const uint32_t lr = tid >> 1, lh = (tid & 1) * 64;
  const uint32_t k_begin = J * kCoreTile2;
  const uint2* ga = reinterpret_cast<const uint2*>(bits + (rowA + lr) * kpw + ((k_begin + lh) >> 5));
  const uint2* gb = reinterpret_cast<const uint2*>(bits + (rowB + lr) * kpw + ((k_begin + lh) >> 5));
  uint2 pa = *ga, pb = *gb;
  // this wave's columns lie in one K stage (the tile's columns 0..127, then 128..255)
  const uint32_t k_mask = k_begin + (wc * WN / kCoreK) * kCoreK, c_off = (wc * WN) % kCoreK;
  uint64_t msk[4] = {0ull, 0ull, 0ull, 0ull};  // bit j * 16 + v of msk[i]: A[r][c] of accumulator (i, j, v)
  uint32_t buf = 0;
  for (uint32_t k0 = k_begin; k0 < kp; k0 += kCoreK, buf ^= 1u) {
    // a wave stores stage s only after every wave passed the barrier of stage
    // s - 1, that is after the MFMAs of stage s - 2 read this buffer
    expand_bits64(pa, &sA[buf][lr][lh]);
    expand_bits64(pb, &sB[buf][lr][lh]);
    __syncthreads();
    if (k0 == k_mask) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            const int r = wr * 128 + i * 32 + mfma_row(v, lane), c = c_off + j * 32 + (lane & 31);
            msk[i] |= (uint64_t)(sA[buf][r][c] != 0) << (j * 16 + v);
          }
    }
    if (k0 + kCoreK < kp) {  // next stage in flight during the MFMAs
      ga += kCoreK / 64;
      gb += kCoreK / 64;
      pa = *ga;
      pb = *gb;
    }
#pragma unroll
    for (int kk = 0; kk < kCoreK; kk += 32) {
      const int ko = kk + (lane >> 5) * 16;  // the same k slice for both operands
      v4i fa[4], fb[NJ];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        fa[i] = *reinterpret_cast<const v4i*>(&sA[buf][wr * 128 + i * 32 + (lane & 31)][ko]);
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        fb[j] = *reinterpret_cast<const v4i*>(&sB[buf][wc * WN + j * 32 + (lane & 31)][ko]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }
  // masked sum: (A A^T)[r][c] counts where A[r][c] = 1
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if ((msk[i] >> (j * 16 + v)) & 1ull) sum += (uint32_t)acc[i][j][v];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
  if (lane == 0) sPart[w] = sum;
  __syncthreads();
  if (tid == 0) {
    unsigned long long s = 0;
    for (int i = 0; i < NW; ++i) s += sPart[i];
    if (s) atomicAdd(result, s);
  }
}

// tile pair (I <= J) of block t, the upper triangle enumerated column by column
__device__ __forceinline__ void tile_pair(uint64_t t, uint32_t* I, uint32_t* J) {
  uint32_t j = (uint32_t)((std::sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((uint64_t)(j + 1) * (j + 2) / 2 <= t) ++j;
  while ((uint64_t)j * (j + 1) / 2 > t) --j;
  *J = j;
  *I = (uint32_t)(t - (uint64_t)j * (j + 1) / 2);
}

__global__ __launch_bounds__(kCoreThreads2) void k_core_count(const uint32_t* __restrict__ bits, uint32_t kp,
                                                              unsigned long long* __restrict__ result) {
  uint32_t I, J;
  tile_pair(blockIdx.x, &I, &J);
  core_tile256(bits, kp, I, J, result);
}

// Builds the core matrix of the top `c` ids (cached per graph and size).
void core_prepare(tribit_ctx* ctx, tribit_graph* g, uint64_t c) {
  if (g->core_bits && g->core_c == c) return;
  core_release(g);
  if (c == 0) return;
  const uint64_t kp = (c + kCoreTile2 - 1) / kCoreTile2 * kCoreTile2;
  const uint64_t d0 = g->n - c;
  DevBuf<uint32_t> bits(ctx, kp * kp / 32);
  hipStream_t s = ctx->stream;
  k_core_zero<<<2048, 256, 0, s>>>(reinterpret_cast<uint4*>(bits.p), kp * kp / 128);
  k_core_fill_bits<<<grid_for(c, 256), 256, 0, s>>>(g->row_ptr, g->col, g->n, d0, (uint32_t)kp, bits.p);
  TB_HIP(hipGetLastError());
  uint32_t e0 = 0;
  TB_HIP(hipMemcpyAsync(&e0, g->row_ptr + d0, 4, hipMemcpyDeviceToHost, s));
  ctx->sync();
  g->core_bits = bits.release();
  g->core_c = c;
  g->core_kp = kp;
  g->core_e0 = e0;
}

void core_release(tribit_graph* g) {
  if (g->core_bits && g->ctx) g->ctx->free(g->core_bits);
  g->core_bits = nullptr;
  g->core_c = g->core_kp = 0;
}

// Tiles of the upper triangle, side `tile`.
uint64_t core_tiles(uint64_t kp, uint64_t tile) {
  const uint64_t nb = kp / tile;
  return nb * (nb + 1) / 2;
}

void core_launch(tribit_ctx* ctx, const tribit_graph* g, unsigned long long* result) {
  if (!g->core_bits) return;
  const uint64_t nt = core_tiles(g->core_kp, kCoreTile2);
  k_core_count<<<(unsigned)nt, kCoreThreads2, 0, ctx->stream>>>(g->core_bits, (uint32_t)g->core_kp, result);
  TB_HIP(hipGetLastError());
}

// int8 multiply-accumulates of a launch over the upper triangle in tiles of
// side `tile`, K from each tile's first column:
// sum over J of (J+1) * tile^2 * (kp - tile J).
double core_macs(uint64_t kp, uint64_t tile) {
  const uint64_t nb = kp / tile;
  double macs = 0.0;
  for (uint64_t J = 0; J < nb; ++J) macs += (double)(J + 1) * tile * tile * (double)(kp - tile * J);
  return macs;
}

static double kCoreMacPerMs = 1.3e12;

// Core size chosen for a relabelled graph: the candidate (multiples of 4096
// up to 65536 ids, at most n) with the largest estimated saving
//     streamed wedges of the core / wedge rate  -  MACs / MAC rate
// if that saving exceeds 0.5 ms; 0 = no core. The rates are the kernels'
// measured figures on MI355X (streaming kernels ~2e12 wedge checks/s, the
// core kernel ~1.3e15 int8 MAC/s including its staging: 7.7e11 MACs in
// 0.57 ms for 16,384 ids).
uint64_t core_choose(tribit_ctx* ctx, const tribit_graph* g) {
  if (!g->relabeled || g->n < 4096 || !g->m) return 0;
  const uint64_t top = std::min<uint64_t>(g->n, 65536);
  std::vector<uint32_t> rp(top + 1);
  TB_HIP(hipMemcpyAsync(rp.data(), g->row_ptr + (g->n - top), 4 * (top + 1), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  uint64_t best = 0;
  double best_gain = 0.5;
  double wedges = 0.0;  // of the top c ids, grown 4,096 ids at a time
  for (uint64_t c = 4096; c <= top; c += 4096) {
    for (uint64_t i = top - c; i < top - c + 4096; ++i) {
      const double d = (double)(rp[i + 1] - rp[i]);
      wedges += d * (d - 1.0) * 0.5;
    }
    const uint64_t kp = (c + kCoreTile2 - 1) / kCoreTile2 * kCoreTile2;
    const double sparse_ms = wedges / 2.0e9;
    const double dense_ms = core_macs(kp, kCoreTile2) / kCoreMacPerMs;
    if (sparse_ms - dense_ms > best_gain) {
      best_gain = sparse_ms - dense_ms;
      best = c;
    }
  }
  return best;
}

// ------------------------------------------------ dense diagonal blocks
// The same contraction on many small dense id ranges [s, e) instead of one
// top core (the blocks of a block model, once the octave-class relabelling
// keeps each block a run of consecutive ids). A triangle a < b < c with all
// three vertices in one block is counted by the block's masked product, one
// with a and b in a block and c after it by the pair kernel (k_blk_pairs),
// and the rest, whose arc a -> b crosses blocks, by the streaming kernels over
// a copy of the wedge ranges without those of in-block arcs (k_blk_trim).
//
// Detection: cross[p] = arcs a -> b with b - a <= kSpanMax and a <= p < b
// (short arcs over the gap between ids p and p + 1). A gap with fewer than
// kCrossMin such arcs ends a segment; segments of kBlockMin..kBlockMax ids
// whose induced density is at least kBlockDensity become blocks. The window is
// short so that sparse background edges seldom span a gap (a block model with
// p_out = 1e-4 puts ~0.2 background arcs over a block boundary) while inside
// a block of density rho even the gap after its first id has rho * kSpanMax
// of them (19 at rho = 0.3).
constexpr uint32_t kSpanMax = 64;
constexpr int32_t kCrossMin = 8;
constexpr uint32_t kBlockMin = 64, kBlockMax = 4096;
constexpr double kBlockDensity = 0.05;

// diff[u] += short out-arcs of u, diff[v] -= 1 for each (rows ascending:
// the short arcs are a prefix of the row).
__global__ void k_span_diff(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            int32_t* __restrict__ diff) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  int32_t c = 0;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t v = col[e];
    if (v - (uint32_t)u > kSpanMax) break;
    atomicSub(&diff[v], 1);
    ++c;
  }
  if (c) atomicAdd(&diff[u], c);
}

// cut[p] = 1 when the gap after id p ends a segment (the last id always does).
__global__ void k_cut_flags(const int32_t* __restrict__ cross, uint64_t n, uint32_t* __restrict__ cut) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) cut[p] = (p + 1 == n || cross[p] < kCrossMin) ? 1u : 0u;
}

// Segment k = ids with seg[v] == k (seg = exclusive scan of cut): its first id
// and length; intra[k] = arcs with both ends in it.
__global__ void k_seg_stats(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                            const uint32_t* __restrict__ seg, uint64_t n, uint32_t* __restrict__ start,
                            uint32_t* __restrict__ len, unsigned long long* __restrict__ intra) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = u < n;
  const uint32_t k = valid ? seg[u] : 0xFFFFFFFFu;
  if (valid && (u == 0 || seg[u - 1] != k)) start[k] = (uint32_t)u;
  const uint32_t run = wave_run_length(k, valid);
  if (run) atomicAdd(&len[k], run);
  if (!valid) return;
  uint32_t c = 0;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
    const uint32_t v = col[e];
    if (v - (uint32_t)u > kBlockMax || seg[v] != k) break;  // rows ascending: the segment's ids come first
    ++c;
  }
  if (c) atomicAdd(&intra[k], (unsigned long long)c);
}

// Segments dense enough to be blocks: (start, length) appended at out[*count].
__global__ void k_seg_select(const uint32_t* __restrict__ start, const uint32_t* __restrict__ len,
                             const unsigned long long* __restrict__ intra, uint32_t nseg, uint2* __restrict__ out,
                             uint32_t cap, uint32_t* __restrict__ count) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nseg) return;
  const uint32_t L = len[k];
  if (L < kBlockMin || L > kBlockMax) return;
  const double pairs = 0.5 * (double)L * (double)(L - 1);
  if ((double)intra[k] < kBlockDensity * pairs) return;
  const uint32_t at = atomicAdd(count, 1u);
  if (at < cap) out[at] = make_uint2(start[k], L);
}

__global__ void k_blk_mark(const uint2* __restrict__ blk, uint32_t nb, int32_t* __restrict__ blk_of) {
  const uint32_t i = blockIdx.x;
  if (i >= nb) return;
  for (uint32_t v = threadIdx.x; v < blk[i].y; v += blockDim.x) blk_of[blk[i].x + v] = (int32_t)i;
}

// A_i[(u - s) * kp + (v - s)] = 1 for every arc u -> v inside block i.
__global__ void k_blk_fill(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                           const uint2* __restrict__ blk, const uint64_t* __restrict__ aoff,
                           const uint32_t* __restrict__ kps, uint32_t nb, int8_t* __restrict__ A) {
  const uint32_t i = blockIdx.x;
  if (i >= nb) return;
  const uint32_t s = blk[i].x, e = s + blk[i].y, kp = kps[i];
  int8_t* M = A + aoff[i];
  for (uint32_t u = s + threadIdx.x; u < e; u += blockDim.x) {
    int8_t* row = M + (uint64_t)(u - s) * kp;
    for (uint32_t q = row_ptr[u]; q < row_ptr[u + 1]; ++q) {
      const uint32_t v = col[q];
      if (v >= e) break;
      row[v - s] = 1;
    }
  }
}

// A triangle u < v < w of a graph with dense diagonal blocks (contiguous id
// ranges) is counted once, by the first of these that applies:
//   all three in one block: the block's matrix on the matrix cores;
//   u and v in one block B, w after it: the pair kernel below, from the
//     sources u < v in B of w's in-arcs (a pair per w and block, rare unless
//     w has many neighbours in B) and B's matrix;
//   otherwise (u -> v crosses blocks): the streamed wedge of arc u -> v.
// (u and w in one block puts v in it too.) So the wedge ranges of in-block
// arcs are cut entirely (range copies were made by the caller) and w[v]
// reduced by their elements. Ranges of one owner are ordered by source, so
// the in-block sources are those whose arc index is >= row_ptr[s].
__global__ void k_blk_trim(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ sfx_ptr,
                           const uint2* __restrict__ blk, const int32_t* __restrict__ blk_of, uint64_t n,
                           uint32_t emask, uint2* __restrict__ rng, unsigned long long* __restrict__ w,
                           unsigned long long* __restrict__ total_cut) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int32_t i = blk_of[v];
  if (i < 0) return;
  const uint32_t first = row_ptr[blk[i].x];
  uint64_t cut = 0;
  for (uint32_t j = sfx_ptr[v]; j < sfx_ptr[v + 1]; ++j) {
    const uint2 r = rng[j];
    if (r.x <= first) continue;  // arc index r.x - 1 < row_ptr[s]: the source is below the block
    const uint32_t end = r.y & emask;
    cut += end - r.x;
    rng[j] = make_uint2(end, end);
  }
  if (cut) {
    w[v] -= cut;
    atomicAdd(total_cut, (unsigned long long)cut);
  }
}

// Arcs u -> w with u in a block and w after it, as (w << 32) | u, for the pair
// kernel; counted per source row first (pass 1: out == nullptr).
__global__ void k_blk_out_arcs(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                               const uint2* __restrict__ blk, const int32_t* __restrict__ blk_of, uint64_t n,
                               const uint64_t* __restrict__ off, uint32_t* __restrict__ cnt,
                               uint64_t* __restrict__ out) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int32_t i = blk_of[u];
  uint32_t k = 0;
  if (i >= 0) {
    const uint32_t e = blk[i].x + blk[i].y;
    // rows are sorted: the arcs leaving the block are a suffix of the row
    uint32_t lo = row_ptr[u], hi = row_ptr[u + 1];
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if (col[mid] < e)
        lo = mid + 1;
      else
        hi = mid;
    }
    k = row_ptr[u + 1] - lo;
    if (out) {
      uint64_t o = off[u];
      for (uint32_t q = lo; q < row_ptr[u + 1]; ++q) out[o++] = ((uint64_t)col[q] << 32) | u;
    }
  }
  if (!out) cnt[u] = k;
}

// Runs of the sorted (w, u) keys with one w and two or more sources in one
// block: a run starts where the previous key has another w or another block.
__global__ void k_blk_run_heads(const uint64_t* __restrict__ key, uint64_t nk, const int32_t* __restrict__ blk_of,
                                uint32_t* __restrict__ head) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nk) return;
  const uint64_t k = key[q];
  const int32_t b = blk_of[(uint32_t)k];
  const bool first = q == 0 || (key[q - 1] >> 32) != (k >> 32) || blk_of[(uint32_t)key[q - 1]] != b;
  const bool more = q + 1 < nk && (key[q + 1] >> 32) == (k >> 32) && blk_of[(uint32_t)key[q + 1]] == b;
  head[q] = (first && more) ? 1u : 0u;
}

// Length of every run (prepare time): keys from its first with one w and one block.
__global__ void k_blk_run_len(const uint64_t* __restrict__ key, uint64_t nk, const int32_t* __restrict__ blk_of,
                              const uint32_t* __restrict__ first, uint32_t nruns, uint2* __restrict__ runs) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nruns) return;
  const uint32_t q0 = first[t];
  const uint64_t w = key[q0] >> 32;
  const int32_t b = blk_of[(uint32_t)key[q0]];
  uint32_t q1 = q0 + 1;
  while (q1 < nk && (key[q1] >> 32) == w && blk_of[(uint32_t)key[q1]] == b) ++q1;
  runs[t] = make_uint2(q0, q1 - q0);
}

// Triangles u < v < w with u, v in one block and w after it: for every run
// (w, block i, sources u_1 < ... < u_k), the edges among its sources, read
// from block i's matrix. Short runs (at most kPairShort sources; almost all,
// most of them pairs): one thread per run.
constexpr uint32_t kPairShort = 32;

__device__ __forceinline__ uint64_t run_pairs(const uint64_t* __restrict__ key, uint2 r, uint32_t x0, uint32_t dx,
                                              const int32_t* __restrict__ blk_of, const uint2* __restrict__ blk,
                                              const uint64_t* __restrict__ aoff, const uint32_t* __restrict__ kps,
                                              const int8_t* __restrict__ A) {
  const int32_t b = blk_of[(uint32_t)key[r.x]];
  const uint32_t s = blk[b].x, kp = kps[b];
  const int8_t* M = A + aoff[b];
  uint64_t c = 0;
  for (uint32_t x = r.x + x0; x < r.x + r.y; x += dx) {
    const uint32_t ux = (uint32_t)key[x] - s;
    for (uint32_t y = x + 1; y < r.x + r.y; ++y) c += M[(uint64_t)ux * kp + ((uint32_t)key[y] - s)] != 0;
  }
  return c;
}

__global__ void k_blk_pairs(const uint64_t* __restrict__ key, const uint2* __restrict__ runs, uint32_t nruns,
                            const int32_t* __restrict__ blk_of, const uint2* __restrict__ blk,
                            const uint64_t* __restrict__ aoff, const uint32_t* __restrict__ kps,
                            const int8_t* __restrict__ A, unsigned long long* __restrict__ result) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t c = t < nruns ? run_pairs(key, runs[t], 0, 1, blk_of, blk, aoff, kps, A) : 0;
  c = wave_sum_u64(c);
  if (lane_id() == 0 && c) atomicAdd(result, (unsigned long long)c);
}

// Long runs (a vertex with many neighbours in one earlier block, e.g. a
// member of the block whose degree put its id above the block): one 256-thread
// workgroup per run, its threads striding over the first source of a pair.
__global__ void __launch_bounds__(256) k_blk_pairs_long(const uint64_t* __restrict__ key,
                                                        const uint2* __restrict__ runs,
                                                        const int32_t* __restrict__ blk_of,
                                                        const uint2* __restrict__ blk,
                                                        const uint64_t* __restrict__ aoff,
                                                        const uint32_t* __restrict__ kps,
                                                        const int8_t* __restrict__ A,
                                                        unsigned long long* __restrict__ result) {
  uint64_t c = run_pairs(key, runs[blockIdx.x], threadIdx.x, blockDim.x, blk_of, blk, aoff, kps, A);
  c = wave_sum_u64(c);
  if (lane_id() == 0 && c) atomicAdd(result, (unsigned long long)c);
}

// Owners' nonempty ranges of the cut copy, compacted per owner (pass 1:
// counts), flags cleared: the cut copy streams 32-bit ids.
__global__ void k_blk_compact(const uint32_t* __restrict__ sfx_ptr, const uint2* __restrict__ rng, uint64_t n,
                              uint32_t emask, const uint32_t* __restrict__ ptr, uint32_t* __restrict__ cnt,
                              uint2* __restrict__ out) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint32_t k = 0, o = out ? ptr[v] : 0;
  for (uint32_t j = sfx_ptr[v]; j < sfx_ptr[v + 1]; ++j) {
    const uint2 r = rng[j];
    const uint32_t end = r.y & emask;
    if (r.x < end) {
      if (out) out[o + k] = make_uint2(r.x, end);
      ++k;
    }
  }
  if (!out) cnt[v] = k;
}

__global__ __launch_bounds__(kCoreThreads) void k_blocks_count(const int8_t* __restrict__ A,
                                                               const uint4* __restrict__ work,
                                                               unsigned long long* __restrict__ result) {
  const uint4 w = work[blockIdx.x];
  core_tile(A + (((uint64_t)w.y << 32) | w.x), w.z, w.w & 0xFFFFu, w.w >> 16, result);
}

void blocks_release(tribit_graph* g) {
  tribit_graph::DenseBlocks& B = g->blocks;
  if (g->ctx) {
    g->ctx->free(B.A);
    g->ctx->free(B.work);
    g->ctx->free(B.rng);
    g->ctx->free(B.ptr);
    g->ctx->free(B.w);
    g->ctx->free(B.key);
    g->ctx->free(B.runs);
    g->ctx->free(B.runs_long);
    g->ctx->free(B.blk);
    g->ctx->free(B.blk_of);
    g->ctx->free(B.aoff);
    g->ctx->free(B.kps);
  }
  B = tribit_graph::DenseBlocks{};
}

// Positions of the set flags: out[pos[q]] = q where flag[q] (pos = exclusive scan of flag).
__global__ void k_flag_positions(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint64_t n,
                                 uint32_t* __restrict__ out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && flag[q]) out[pos[q]] = (uint32_t)q;
}

static bool blocks_build(tribit_ctx* ctx, tribit_graph* g);

// Detection and the blocks' structures, once per graph; their device time is
// kept (tribit_graph_info.ms_blocks), like a range layout's.
bool blocks_prepare(tribit_ctx* ctx, tribit_graph* g) {
  tribit_graph::DenseBlocks& B = g->blocks;
  if (B.found) return B.nblocks > 0;
  B.found = true;
  const uint64_t n = g->n, m = g->m;
  if (!g->relabeled || !g->rng_sfx || n < kBlockMin || !m) return false;
  hipEvent_t a = nullptr, b = nullptr;
  TB_HIP(hipEventCreate(&a));
  TB_HIP(hipEventCreate(&b));
  bool ok = false;
  try {
    TB_HIP(hipEventRecord(a, ctx->stream));
    ok = blocks_build(ctx, g);
    TB_HIP(hipEventRecord(b, ctx->stream));
    TB_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    TB_HIP(hipEventElapsedTime(&ms, a, b));
    B.ms_prepare = ms;
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ok;
}

static bool blocks_build(tribit_ctx* ctx, tribit_graph* g) {
  tribit_graph::DenseBlocks& B = g->blocks;
  const uint64_t n = g->n, m = g->m;
  hipStream_t s = ctx->stream;
  std::vector<uint2> hb;
  {
    DevBuf<int32_t> diff(ctx, n), cross(ctx, n);
    DevBuf<uint32_t> cut(ctx, n), seg(ctx, n);
    TB_HIP(hipMemsetAsync(diff.p, 0, 4 * n, s));
    k_span_diff<<<grid_for(n, 256), 256, 0, s>>>(g->row_ptr, g->col, n, diff.p);
    size_t tb = 0;
    TB_HIP(rocprim::inclusive_scan(nullptr, tb, diff.p, cross.p, (size_t)n, rocprim::plus<int32_t>(), s));
    DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
    TB_HIP(rocprim::inclusive_scan(tmp.p, tb, diff.p, cross.p, (size_t)n, rocprim::plus<int32_t>(), s));
    k_cut_flags<<<grid_for(n, 256), 256, 0, s>>>(cross.p, n, cut.p);
    size_t tb2 = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, tb2, cut.p, seg.p, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    DevBuf<uint8_t> tmp2(ctx, tb2 ? tb2 : 8);
    TB_HIP(rocprim::exclusive_scan(tmp2.p, tb2, cut.p, seg.p, 0u, (size_t)n, rocprim::plus<uint32_t>(), s));
    uint32_t nseg = 0;
    TB_HIP(hipMemcpyAsync(&nseg, seg.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    nseg += 1;  // the last id's cut flag is 1: its segment is the last
    DevBuf<uint32_t> start(ctx, nseg), len(ctx, nseg);
    DevBuf<unsigned long long> intra(ctx, nseg);
    TB_HIP(hipMemsetAsync(len.p, 0, 4ull * nseg, s));
    TB_HIP(hipMemsetAsync(intra.p, 0, 8ull * nseg, s));
    k_seg_stats<<<grid_for(n, 256), 256, 0, s>>>(g->row_ptr, g->col, seg.p, n, start.p, len.p, intra.p);
    constexpr uint32_t kCap = 1u << 20;
    const uint32_t cap = std::min<uint32_t>(nseg, kCap);
    DevBuf<uint2> sel(ctx, cap);
    DevBuf<uint32_t> cnt(ctx, 1);
    TB_HIP(hipMemsetAsync(cnt.p, 0, 4, s));
    k_seg_select<<<grid_for(nseg, 256), 256, 0, s>>>(start.p, len.p, intra.p, nseg, sel.p, cap, cnt.p);
    uint32_t nb = 0;
    TB_HIP(hipMemcpyAsync(&nb, cnt.p, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    nb = std::min(nb, cap);
    hb.resize(nb);
    if (nb) TB_HIP(hipMemcpyAsync(hb.data(), sel.p, 8ull * nb, hipMemcpyDeviceToHost, s));
    ctx->sync();
  }
  if (hb.empty()) return false;
  std::sort(hb.begin(), hb.end(), [](const uint2& a, const uint2& b) { return a.x < b.x; });
  const uint32_t nb = (uint32_t)hb.size();
  // matrices, padded to multiples of the output tile, and the workgroups' tile pairs
  std::vector<uint64_t> aoff(nb);
  std::vector<uint32_t> kps(nb);
  std::vector<uint4> work;
  uint64_t total = 0, verts = 0;
  double macs = 0.0;
  for (uint32_t i = 0; i < nb; ++i) {
    const uint32_t kp = (hb[i].y + kCoreTile - 1) / kCoreTile * kCoreTile;
    aoff[i] = total;
    kps[i] = kp;
    total += (uint64_t)kp * kp;
    verts += hb[i].y;
    macs += core_macs(kp, kCoreTile);
    for (uint32_t J = 0; J < kp / kCoreTile; ++J)
      for (uint32_t I = 0; I <= J; ++I)
        work.push_back(make_uint4((uint32_t)aoff[i], (uint32_t)(aoff[i] >> 32), kp, I | (J << 16)));
  }
  DevBuf<int8_t> A(ctx, total);
  DevBuf<uint4> dwork(ctx, work.size());
  DevBuf<uint2> dblk(ctx, nb);
  DevBuf<uint64_t> daoff(ctx, nb);
  DevBuf<uint32_t> dkp(ctx, nb);
  TB_HIP(hipMemcpyAsync(dwork.p, work.data(), 16ull * work.size(), hipMemcpyHostToDevice, s));
  TB_HIP(hipMemcpyAsync(dblk.p, hb.data(), 8ull * nb, hipMemcpyHostToDevice, s));
  TB_HIP(hipMemcpyAsync(daoff.p, aoff.data(), 8ull * nb, hipMemcpyHostToDevice, s));
  TB_HIP(hipMemcpyAsync(dkp.p, kps.data(), 4ull * nb, hipMemcpyHostToDevice, s));
  k_core_zero<<<2048, 256, 0, s>>>(reinterpret_cast<uint4*>(A.p), total / 16);
  k_blk_fill<<<nb, 256, 0, s>>>(g->row_ptr, g->col, dblk.p, daoff.p, dkp.p, nb, A.p);
  // the streamed ranges: a copy of the wedge ranges with in-block arcs cut
  DevBuf<int32_t> blk_of(ctx, n);
  TB_HIP(hipMemsetAsync(blk_of.p, 0xFF, 4 * n, s));
  k_blk_mark<<<nb, 256, 0, s>>>(dblk.p, nb, blk_of.p);
  DevBuf<uint2> rng(ctx, m);
  DevBuf<uint64_t> w(ctx, n);
  DevBuf<unsigned long long> cut(ctx, 1);
  TB_HIP(hipMemcpyAsync(rng.p, g->rng_sfx, 8 * m, hipMemcpyDeviceToDevice, s));
  TB_HIP(hipMemcpyAsync(w.p, g->w_sfx, 8 * n, hipMemcpyDeviceToDevice, s));
  TB_HIP(hipMemsetAsync(cut.p, 0, 8, s));
  k_blk_trim<<<grid_for(n, 256), 256, 0, s>>>(g->row_ptr, g->sfx_ptr, dblk.p, blk_of.p, n, range_end_mask(m),
                                              rng.p, reinterpret_cast<unsigned long long*>(w.p), cut.p);
  TB_HIP(hipGetLastError());
  unsigned long long hcut = 0;
  TB_HIP(hipMemcpyAsync(&hcut, cut.p, 8, hipMemcpyDeviceToHost, s));
  // the owners' lists without the empty ranges of in-block arcs
  DevBuf<uint32_t> bptr(ctx, n + 1);
  DevBuf<uint2> crng;
  crng.ctx = ctx;
  {
    DevBuf<uint32_t> cnt(ctx, n + 1);
    TB_HIP(hipMemsetAsync(cnt.p + n, 0, 4, s));
    k_blk_compact<<<grid_for(n, 256), 256, 0, s>>>(g->sfx_ptr, rng.p, n, range_end_mask(m), nullptr, cnt.p, nullptr);
    size_t tb = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, tb, cnt.p, bptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s));
    DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
    TB_HIP(rocprim::exclusive_scan(tmp.p, tb, cnt.p, bptr.p, 0u, (size_t)(n + 1), rocprim::plus<uint32_t>(), s));
    uint32_t kept = 0;
    TB_HIP(hipMemcpyAsync(&kept, bptr.p + n, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    crng = DevBuf<uint2>(ctx, kept ? kept : 1);
    k_blk_compact<<<grid_for(n, 256), 256, 0, s>>>(g->sfx_ptr, rng.p, n, range_end_mask(m), bptr.p, nullptr,
                                                   crng.p);
    TB_HIP(hipGetLastError());
    ctx->sync();
    rng.reset();
  }
  // the pair kernel's runs: arcs from a block to a later id, sorted by (w, u)
  DevBuf<uint64_t> key;
  key.ctx = ctx;
  DevBuf<uint2> runs, runs_long;
  runs.ctx = runs_long.ctx = ctx;
  uint64_t nkeys = 0;
  uint32_t nruns = 0, nruns_long = 0;
  {
    DevBuf<uint32_t> cnt(ctx, n);
    DevBuf<uint64_t> off(ctx, n);
    k_blk_out_arcs<<<grid_for(n, 256), 256, 0, s>>>(g->row_ptr, g->col, dblk.p, blk_of.p, n, nullptr, cnt.p,
                                                    nullptr);
    size_t tb = 0;
    TB_HIP(rocprim::exclusive_scan(nullptr, tb, cnt.p, off.p, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
    DevBuf<uint8_t> tmp(ctx, tb ? tb : 8);
    TB_HIP(rocprim::exclusive_scan(tmp.p, tb, cnt.p, off.p, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
    uint64_t last_off = 0;
    uint32_t last_cnt = 0;
    TB_HIP(hipMemcpyAsync(&last_off, off.p + n - 1, 8, hipMemcpyDeviceToHost, s));
    TB_HIP(hipMemcpyAsync(&last_cnt, cnt.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    ctx->sync();
    nkeys = last_off + last_cnt;
    if (nkeys) {
      DevBuf<uint64_t> raw(ctx, nkeys);
      key = DevBuf<uint64_t>(ctx, nkeys);
      k_blk_out_arcs<<<grid_for(n, 256), 256, 0, s>>>(g->row_ptr, g->col, dblk.p, blk_of.p, n, off.p, nullptr,
                                                      raw.p);
      size_t tb2 = 0;
      TB_HIP(rocprim::radix_sort_keys(nullptr, tb2, raw.p, key.p, (size_t)nkeys, 0u, 32u + bits_for(n), s));
      DevBuf<uint8_t> tmp2(ctx, tb2 ? tb2 : 8);
      TB_HIP(rocprim::radix_sort_keys(tmp2.p, tb2, raw.p, key.p, (size_t)nkeys, 0u, 32u + bits_for(n), s));
      DevBuf<uint32_t> head(ctx, nkeys + 1), pos(ctx, nkeys + 1);
      k_blk_run_heads<<<grid_for(nkeys, 256), 256, 0, s>>>(key.p, nkeys, blk_of.p, head.p);
      TB_HIP(hipMemsetAsync(head.p + nkeys, 0, 4, s));
      size_t tb3 = 0;
      TB_HIP(rocprim::exclusive_scan(nullptr, tb3, head.p, pos.p, 0u, (size_t)(nkeys + 1), rocprim::plus<uint32_t>(),
                                     s));
      DevBuf<uint8_t> tmp3(ctx, tb3 ? tb3 : 8);
      TB_HIP(rocprim::exclusive_scan(tmp3.p, tb3, head.p, pos.p, 0u, (size_t)(nkeys + 1), rocprim::plus<uint32_t>(),
                                     s));
      TB_HIP(hipMemcpyAsync(&nruns, pos.p + nkeys, 4, hipMemcpyDeviceToHost, s));
      ctx->sync();
      if (nruns) {
        // every run's first key and length; short and long runs apart
        DevBuf<uint32_t> first(ctx, nruns);
        DevBuf<uint2> all(ctx, nruns);
        k_flag_positions<<<grid_for(nkeys, 256), 256, 0, s>>>(head.p, pos.p, nkeys, first.p);
        k_blk_run_len<<<grid_for(nruns, 256), 256, 0, s>>>(key.p, nkeys, blk_of.p, first.p, nruns, all.p);
        std::vector<uint2> h(nruns), hs, hl;
        TB_HIP(hipMemcpyAsync(h.data(), all.p, 8ull * nruns, hipMemcpyDeviceToHost, s));
        ctx->sync();
        for (const uint2& r : h) (r.y <= kPairShort ? hs : hl).push_back(r);
        nruns = (uint32_t)hs.size();
        nruns_long = (uint32_t)hl.size();
        if (nruns) {
          runs = DevBuf<uint2>(ctx, nruns);
          TB_HIP(hipMemcpyAsync(runs.p, hs.data(), 8ull * nruns, hipMemcpyHostToDevice, s));
        }
        if (nruns_long) {
          runs_long = DevBuf<uint2>(ctx, nruns_long);
          TB_HIP(hipMemcpyAsync(runs_long.p, hl.data(), 8ull * nruns_long, hipMemcpyHostToDevice, s));
        }
      }
    }
    TB_HIP(hipGetLastError());
    ctx->sync();
  }
  B.nblocks = nb;
  B.vertices = verts;
  B.macs = macs;
  B.A = A.release();
  B.work = dwork.release();
  B.nwork = (uint32_t)work.size();
  const uint64_t kept = crng.n;
  B.rng = crng.release();
  B.ptr = bptr.release();
  B.w = w.release();
  B.cut = hcut;
  B.key = key.release();
  B.nkeys = nkeys;
  B.runs = runs.release();
  B.nruns = nruns;
  B.runs_long = runs_long.release();
  B.nruns_long = nruns_long;
  B.blk = dblk.release();
  B.blk_of = blk_of.release();
  B.aoff = daoff.release();
  B.kps = dkp.release();
  g->bytes += total + 16ull * work.size() + 8 * kept + 16 * n + 8 * nkeys + 8ull * (nruns + nruns_long) + 16ull * nb;
  return true;
}

void blocks_launch(tribit_ctx* ctx, const tribit_graph* g, unsigned long long* result) {
  const tribit_graph::DenseBlocks& B = g->blocks;
  if (!B.A || !B.nwork) return;
  k_blocks_count<<<B.nwork, kCoreThreads, 0, ctx->stream>>>(B.A, B.work, result);
  if (B.nruns)
    k_blk_pairs<<<(B.nruns + 255) / 256, 256, 0, ctx->stream>>>(B.key, B.runs, B.nruns, B.blk_of, B.blk, B.aoff,
                                                                B.kps, B.A, result);
  if (B.nruns_long)
    k_blk_pairs_long<<<B.nruns_long, 256, 0, ctx->stream>>>(B.key, B.runs_long, B.blk_of, B.blk, B.aoff, B.kps, B.A,
                                                              result);
  TB_HIP(hipGetLastError());
}

}  // namespace tb
