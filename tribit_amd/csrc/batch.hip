// batch.hip -- building batches of graphs (gfx950). No counterpart in the
// reference, which counts one graph per process.
//
// A batch packs its graphs into parts: each part is the disjoint union of a
// run of graphs, graph j's vertex ids offset by the vertices of the graphs
// before it in the part. A part is preprocessed exactly like one graph
// (prep.hip); a triangle never spans two graphs, so the union's triangles are
// the graphs' triangles, and each vertex keeps the index of its graph for the
// per-graph attribution done by the counting kernels (count.hip).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace tb {

constexpr uint64_t kPartVerts = 1ull << 31;    // vertices per part
constexpr uint64_t kSentinelKey = 0xFFFFFFFFFFFFFFFFull;

// Index j of the segment holding i: off[j] <= i < off[j + 1] (off[0] = 0,
// off[count] = total, non-decreasing).
__device__ __forceinline__ uint32_t segment_of(const uint64_t* off, uint32_t count, uint64_t i) {
  uint32_t lo = 0, hi = count;  // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Offsets graph-local keys into the part's id space; validates min < max < n_j.
__global__ void k_offset_keys(uint64_t* __restrict__ keys, uint64_t nk, const uint64_t* __restrict__ koff,
                              const uint64_t* __restrict__ voff, const uint64_t* __restrict__ nv, uint32_t count,
                              uint32_t* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  const uint64_t k = keys[i];
  if (k == kSentinelKey) return;
  const uint32_t j = segment_of(koff, count, i);
  const uint64_t a = k >> 32, b = k & 0xFFFFFFFFull;
  if (!(a < b && b < nv[j])) {
    *bad = 1u;
    return;
  }
  keys[i] = ((a + voff[j]) << 32) | (b + voff[j]);
}

// Graph index of every internal vertex (through the relabelling, if any).
__global__ void k_vertex_gid(const uint32_t* __restrict__ labels, uint64_t n, const uint64_t* __restrict__ voff,
                             uint32_t count, uint32_t first, uint32_t* __restrict__ gid) {
  const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t old = labels ? labels[v] : v;
  gid[v] = first + segment_of(voff, count, old);
}

// Edges of every graph: each oriented row adds its length to its graph.
__global__ void k_graph_edges(const uint32_t* __restrict__ row_ptr, uint64_t n, const uint32_t* __restrict__ gid,
                              unsigned long long* __restrict__ m) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const uint32_t d = row_ptr[u + 1] - row_ptr[u];
  if (d) atomicAdd(&m[gid[u]], (unsigned long long)d);
}

__global__ void k_zero_words(unsigned long long* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

namespace {

// Runs of graphs per part: by input keys and vertices.
std::vector<std::pair<uint32_t, uint32_t>> split_parts(const std::vector<uint64_t>& nkeys,
                                                       const std::vector<uint64_t>& nv, uint64_t part_keys) {
  std::vector<std::pair<uint32_t, uint32_t>> runs;
  uint32_t first = 0;
  uint64_t k = 0, v = 0;
  for (uint32_t j = 0; j < nkeys.size(); ++j) {
    if (j > first && (k + nkeys[j] > part_keys || v + nv[j] > kPartVerts)) {
      runs.push_back({first, j - first});
      first = j;
      k = v = 0;
    }
    k += nkeys[j];
    v += nv[j];
  }
  if (nkeys.size() > first) runs.push_back({first, (uint32_t)nkeys.size() - first});
  return runs;
}

// Builds one part from keys already written to d_keys (graph-local ids, graph
// j of the run at koff[j]); takes ownership of d_keys.
void build_part(tribit_ctx* ctx, tribit_batch* b, uint32_t first, uint32_t count, uint64_t* d_keys,
                const std::vector<uint64_t>& koff_h, const std::vector<uint64_t>& nv_h) {
  hipStream_t s = ctx->stream;
  DevBuf<uint64_t> keys;
  keys.ctx = ctx;
  keys.p = d_keys;
  const uint64_t nk = koff_h[count];
  std::vector<uint64_t> voff_h(count + 1, 0);
  for (uint32_t j = 0; j < count; ++j) voff_h[j + 1] = voff_h[j] + nv_h[j];
  const uint64_t n = voff_h[count];
  if (n > 0xFFFFFFFEull) fail(TRIBIT_ERR_TOO_LARGE, "part has too many vertices");
  DevBuf<uint64_t> koff(ctx, count + 1), voff(ctx, count + 1), nv(ctx, count);
  DevBuf<uint32_t> bad(ctx, 1);
  TB_HIP(hipMemcpyAsync(koff.p, koff_h.data(), 8 * (count + 1), hipMemcpyHostToDevice, s));
  TB_HIP(hipMemcpyAsync(voff.p, voff_h.data(), 8 * (count + 1), hipMemcpyHostToDevice, s));
  TB_HIP(hipMemcpyAsync(nv.p, nv_h.data(), 8 * count, hipMemcpyHostToDevice, s));
  TB_HIP(hipMemsetAsync(bad.p, 0, 4, s));
  if (nk) k_offset_keys<<<grid_for(nk, 256), 256, 0, s>>>(keys.p, nk, koff.p, voff.p, nv.p, count, bad.p);
  uint32_t hbad = 0;
  TB_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
  ctx->sync();
  if (hbad) fail(TRIBIT_ERR_RANGE, "batch edge key is not (min < max < n) of its graph");
  tribit_batch::Part part;
  part.first = first;
  part.count = count;
  part.g = new tribit_graph;
  try {
    // a part's replay never counts dense diagonal blocks, so the automatic id
    // order resolves to degree rank, which streams faster
    preprocess(ctx, keys.release(), nk, n, part.g, ctx->relabel == 3 ? 1 : ctx->relabel);
    DevBuf<uint32_t> gid(ctx, n);
    k_vertex_gid<<<grid_for(n, 256), 256, 0, s>>>(part.g->labels, n, voff.p, count, first, gid.p);
    DevBuf<unsigned long long> m(ctx, b->ngraphs);
    k_zero_words<<<grid_for(b->ngraphs, 256), 256, 0, s>>>(m.p, b->ngraphs);
    k_graph_edges<<<grid_for(n, 256), 256, 0, s>>>(part.g->row_ptr, n, gid.p, m.p);
    std::vector<unsigned long long> mh(b->ngraphs);
    TB_HIP(hipMemcpyAsync(mh.data(), m.p, 8 * b->ngraphs, hipMemcpyDeviceToHost, s));
    ctx->sync();
    for (uint32_t j = 0; j < count; ++j) b->m[first + j] = mh[first + j];
    part.gid = gid.release();
    b->ms_prep += part.g->ms_prep;
  } catch (...) {
    free_graph(part.g);
    throw;
  }
  b->parts.push_back(part);
}

tribit_batch* new_batch(tribit_ctx* ctx, uint32_t ngraphs) {
  tribit_batch* b = new tribit_batch;
  b->ctx = ctx;
  b->ngraphs = ngraphs;
  b->n.assign(ngraphs, 0);
  b->m.assign(ngraphs, 0);
  return b;
}

void finish_batch(tribit_ctx* ctx, tribit_batch* b) {
  b->gcount = ctx->alloc_n<unsigned long long>(b->ngraphs ? b->ngraphs : 1);
  b->scratch = ctx->alloc_n<unsigned long long>(tb::kCountScratch * (b->parts.size() ? b->parts.size() : 1));
}

}  // namespace

tribit_batch* batch_from_host(tribit_ctx* ctx, const uint64_t* keys, const uint64_t* key_off, const uint64_t* nverts,
                              uint32_t ngraphs) {
  tribit_batch* b = new_batch(ctx, ngraphs);
  try {
    std::vector<uint64_t> nk(ngraphs), nv(ngraphs);
    for (uint32_t j = 0; j < ngraphs; ++j) {
      if (key_off[j + 1] < key_off[j] || nverts[j] < 1 || nverts[j] > 0xFFFFFFFEull)
        fail(TRIBIT_ERR_ARG, "batch offsets must be non-decreasing and every graph needs 1..2^32-2 vertices");
      nk[j] = key_off[j + 1] - key_off[j];
      nv[j] = nverts[j];
      b->n[j] = nverts[j];
    }
    for (auto [first, count] : split_parts(nk, nv, ctx->part_keys)) {
      std::vector<uint64_t> koff(count + 1, 0), nvr(nv.begin() + first, nv.begin() + first + count);
      for (uint32_t j = 0; j < count; ++j) koff[j + 1] = koff[j] + nk[first + j];
      uint64_t* d = ctx->alloc_n<uint64_t>(koff[count] ? koff[count] : 1);
      if (koff[count])
        TB_HIP(hipMemcpyAsync(d, keys + key_off[first], 8 * koff[count], hipMemcpyHostToDevice, ctx->stream));
      build_part(ctx, b, first, count, d, koff, nvr);
    }
    finish_batch(ctx, b);
  } catch (...) {
    batch_free(b);
    throw;
  }
  return b;
}

tribit_batch* batch_gen_chung_lu(tribit_ctx* ctx, uint64_t seed, uint64_t first_index, uint32_t ngraphs) {
  tribit_batch* b = new_batch(ctx, ngraphs);
  try {
    std::vector<uint64_t> nk(ngraphs), nv(ngraphs);
    std::vector<double> gam(ngraphs), avg(ngraphs);
    for (uint32_t j = 0; j < ngraphs; ++j) {
      chung_lu_batch_params(seed, first_index + j, &gam[j], &nv[j], &avg[j]);
      nk[j] = chung_lu_samples(nv[j], avg[j]);
      b->n[j] = nv[j];
    }
    for (auto [first, count] : split_parts(nk, nv, ctx->part_keys)) {
      std::vector<uint64_t> koff(count + 1, 0), nvr(nv.begin() + first, nv.begin() + first + count);
      for (uint32_t j = 0; j < count; ++j) koff[j + 1] = koff[j] + nk[first + j];
      uint64_t* d = ctx->alloc_n<uint64_t>(koff[count] ? koff[count] : 1);
      try {
        for (uint32_t j = 0; j < count; ++j) {
          const uint32_t gj = first + j;
          gen_chung_lu(ctx, nv[gj], gam[gj], avg[gj], (seed << 32) + first_index + gj, d + koff[j]);
        }
      } catch (...) {
        ctx->free(d);
        throw;
      }
      build_part(ctx, b, first, count, d, koff, nvr);
    }
    finish_batch(ctx, b);
  } catch (...) {
    batch_free(b);
    throw;
  }
  return b;
}

void batch_free(tribit_batch* b) {
  if (!b) return;
  batch_drop_exec(b);
  tribit_ctx* ctx = b->ctx;
  for (auto& p : b->parts) {
    schedule_free(p.S);
    p.S = nullptr;
    if (ctx) ctx->free(p.gid);
    free_graph(p.g);
  }
  if (ctx) {
    ctx->free(b->gcount);
    ctx->free(b->scratch);
  }
  delete b;
}

}  // namespace tb
