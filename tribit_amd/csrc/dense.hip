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
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

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

// A[(u - d0) * kp + (v - d0)] = 1 for every arc u -> v with u >= d0 (then v > u >= d0).
__global__ void k_core_fill(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col, uint64_t n,
                            uint64_t d0, uint32_t kp, int8_t* __restrict__ A) {
  const uint64_t u = d0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  int8_t* row = A + (u - d0) * (uint64_t)kp;
  for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) row[col[e] - d0] = 1;
}

// Output row of accumulator element v of lane l in a 32 x 32 MFMA block
// (column = l % 32): rows come in four groups of 8, each lane half holding 4.
__device__ __forceinline__ int mfma_row(int v, int l) { return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3); }

__global__ __launch_bounds__(kCoreThreads) void k_core_count(const int8_t* __restrict__ A, uint32_t kp,
                                                             unsigned long long* __restrict__ result) {
  __shared__ __attribute__((aligned(16))) int8_t sA[kCoreTile][kCoreK + kCorePad];
  __shared__ __attribute__((aligned(16))) int8_t sB[kCoreTile][kCoreK + kCorePad];
  __shared__ unsigned long long sPart[kCoreThreads / 64];
  // tile pair (I <= J) from the linear index over the upper triangle, column by column
  const uint64_t t = blockIdx.x;
  uint32_t J = (uint32_t)((std::sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((uint64_t)(J + 1) * (J + 2) / 2 <= t) ++J;
  while ((uint64_t)J * (J + 1) / 2 > t) --J;
  const uint32_t I = (uint32_t)(t - (uint64_t)J * (J + 1) / 2);
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
  for (uint32_t k0 = k_begin; k0 < kp; k0 += kCoreK) {
    *reinterpret_cast<uint4*>(&sA[lr][lh]) = pa0;
    *reinterpret_cast<uint4*>(&sA[lr][lh + 16]) = pa1;
    *reinterpret_cast<uint4*>(&sA[lr][lh + 32]) = pa2;
    *reinterpret_cast<uint4*>(&sA[lr][lh + 48]) = pa3;
    *reinterpret_cast<uint4*>(&sB[lr][lh]) = pb0;
    *reinterpret_cast<uint4*>(&sB[lr][lh + 16]) = pb1;
    *reinterpret_cast<uint4*>(&sB[lr][lh + 32]) = pb2;
    *reinterpret_cast<uint4*>(&sB[lr][lh + 48]) = pb3;
    __syncthreads();
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
        fa[i] = *reinterpret_cast<const v4i*>(&sA[wr * 64 + i * 32 + (lane & 31)][ko]);
        fb[i] = *reinterpret_cast<const v4i*>(&sB[wc * 64 + i * 32 + (lane & 31)][ko]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // masked sum: (A A^T)[r][c] counts where A[r][c] = 1
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint64_t c = rowB + wc * 64 + j * 32 + (lane & 31);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const uint64_t r = rowA + wr * 64 + i * 32 + mfma_row(v, lane);
        if (A[r * kp + c]) sum += (uint32_t)acc[i][j][v];
      }
    }
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

// Builds the core matrix of the top `c` ids (cached per graph and size).
void core_prepare(tribit_ctx* ctx, tribit_graph* g, uint64_t c) {
  if (g->core_A && g->core_c == c) return;
  core_release(g);
  if (c == 0) return;
  const uint64_t kp = (c + kCoreTile - 1) / kCoreTile * kCoreTile;
  const uint64_t d0 = g->n - c;
  DevBuf<int8_t> A(ctx, kp * kp);
  hipStream_t s = ctx->stream;
  k_core_zero<<<2048, 256, 0, s>>>(reinterpret_cast<uint4*>(A.p), kp * kp / 16);
  k_core_fill<<<grid_for(c, 256), 256, 0, s>>>(g->row_ptr, g->col, g->n, d0, (uint32_t)kp, A.p);
  TB_HIP(hipGetLastError());
  uint32_t e0 = 0;
  TB_HIP(hipMemcpyAsync(&e0, g->row_ptr + d0, 4, hipMemcpyDeviceToHost, s));
  ctx->sync();
  g->core_A = A.release();
  g->core_c = c;
  g->core_kp = kp;
  g->core_e0 = e0;
}

void core_release(tribit_graph* g) {
  if (g->core_A && g->ctx) g->ctx->free(g->core_A);
  g->core_A = nullptr;
  g->core_c = g->core_kp = 0;
}

// Tiles of the upper triangle.
uint64_t core_tiles(uint64_t kp) {
  const uint64_t nb = kp / kCoreTile;
  return nb * (nb + 1) / 2;
}

void core_launch(tribit_ctx* ctx, const tribit_graph* g, unsigned long long* result) {
  if (!g->core_A) return;
  const uint64_t nt = core_tiles(g->core_kp);
  k_core_count<<<(unsigned)nt, kCoreThreads, 0, ctx->stream>>>(g->core_A, (uint32_t)g->core_kp, result);
  TB_HIP(hipGetLastError());
}

// int8 multiply-accumulates of the core launch (the upper triangle, K from
// each tile's first column): sum over J of (J+1) * 128^2 * (kp - 128 J).
double core_macs(uint64_t kp) {
  const uint64_t nb = kp / kCoreTile;
  double macs = 0.0;
  for (uint64_t J = 0; J < nb; ++J) macs += (double)(J + 1) * kCoreTile * kCoreTile * (double)(kp - kCoreTile * J);
  return macs;
}

static double kCoreMacPerMs = 5.5e11;

// Core size chosen for a relabelled graph: the candidate (4096 .. 65536
// ids, at most n) with the largest estimated saving
//     streamed wedges of the core / wedge rate  -  MACs / MAC rate
// if that saving exceeds 0.5 ms; 0 = no core. The rates are the kernels'
// measured figures on MI355X (streaming kernels ~2e12 wedge checks/s, the
// core kernel ~5.5e14 int8 MAC/s including its staging).
uint64_t core_choose(tribit_ctx* ctx, const tribit_graph* g) {
  if (!g->relabeled || g->n < 4096 || !g->m) return 0;
  const uint64_t top = std::min<uint64_t>(g->n, 65536);
  std::vector<uint32_t> rp(top + 1);
  TB_HIP(hipMemcpyAsync(rp.data(), g->row_ptr + (g->n - top), 4 * (top + 1), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  uint64_t best = 0;
  double best_gain = 0.5;
  for (uint64_t c = 4096; c <= top; c *= 2) {
    double wedges = 0.0;
    for (uint64_t i = top - c; i < top; ++i) {
      const double d = (double)(rp[i + 1] - rp[i]);
      wedges += d * (d - 1.0) * 0.5;
    }
    const uint64_t kp = (c + kCoreTile - 1) / kCoreTile * kCoreTile;
    const double sparse_ms = wedges / 2.0e9;
    const double dense_ms = core_macs(kp) / kCoreMacPerMs;
    if (sparse_ms - dense_ms > best_gain) {
      best_gain = sparse_ms - dense_ms;
      best = c;
    }
  }
  return best;
}

}  // namespace tb
