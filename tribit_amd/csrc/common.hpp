// common.hpp -- shared host/device helpers of libtribit_hip (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/tribit.h"

namespace tb {

// Internal error carrier: thrown inside the library, converted to a
// tribit_status at every extern "C" boundary (nothing escapes the ABI).
struct Error {
  int status;
  std::string msg;
};

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error{status, msg}; }

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
    fail(e == hipErrorOutOfMemory ? TRIBIT_ERR_NOMEM : TRIBIT_ERR_HIP, buf);
  }
}
#define TB_HIP(x) ::tb::hip_check((x), #x, __FILE__, __LINE__)

constexpr uint64_t kSentinel = 0xFFFFFFFFFFFFFFFFull;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

inline uint32_t bits_for(uint64_t n) {  // bits needed to hold ids 0..n-1 (>=1)
  uint32_t b = 1;
  while (b < 64 && (1ull << b) < n) ++b;
  return b;
}

inline unsigned grid_for(uint64_t items, unsigned block) {
  uint64_t g = (items + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)g;
}

}  // namespace tb

// Per-device context: one stream, stream-ordered pool allocations with
// byte accounting (cap for tests, peak for reporting), timing events.
struct tribit_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipMemPool_t pool = nullptr;
  int cus = 0;
  int lds_per_block = 0;
  char name[256] = {0};
  uint64_t cap = 0;    // 0 = unlimited
  uint64_t held = 0;   // bytes currently allocated through this context
  uint64_t peak = 0;
  std::unordered_map<void*, uint64_t> sizes;
  hipEvent_t ev[16] = {};
  int relabel = 3;     // TRIBIT_OPT_RELABEL: 0 input ids, 1 degree rank, 2 octave classes, 3 auto
  tribit_mlp* model = nullptr;  // cost model of the learned scheduler (not owned)
  uint64_t model_version = 1;   // bumped whenever the model is set (judged choices go stale)
  uint64_t part_keys = 1ull << 30;  // TRIBIT_OPT_BATCH_PART_KEYS: input keys per batch part
  // Judged choices of the learned schedule, kept per graph shape so that a
  // graph rebuilt from the same input (a warm-up pass, then the measured one)
  // runs the choice without judging again. Key: (n, m, W+, relabelled, flags,
  // core request, model version); a collision can only cost speed, never
  // exactness (every candidate counts the same triangles).
  struct Judged {
    uint64_t n, m, wplus, model;
    int relabeled, flags, core, choice;
    double ms[TRIBIT_CANDIDATES];
  };
  std::vector<Judged> judged;
  // device address of the no-allocation count's scratch array on this
  // context's device (count.hip; looked up once at tribit_init)
  void* noalloc_scratch = nullptr;
  std::string last_error;

  void* alloc(uint64_t bytes) {
    if (bytes == 0) bytes = 8;
    bytes = (bytes + 255) & ~uint64_t(255);
    if (cap && held + bytes > cap) tb::fail(TRIBIT_ERR_NOMEM, "allocation exceeds the configured cap");
    void* p = nullptr;
    hipError_t e = hipMallocAsync(&p, bytes, stream);
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      tb::fail(TRIBIT_ERR_NOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    sizes[p] = bytes;
    held += bytes;
    if (held > peak) peak = held;
    return p;
  }
  template <typename T>
  T* alloc_n(uint64_t n) {
    return static_cast<T*>(alloc(n * sizeof(T)));
  }
  void free(void* p) {
    if (!p) return;
    auto it = sizes.find(p);
    if (it != sizes.end()) {
      held -= it->second;
      sizes.erase(it);
    }
    (void)hipFreeAsync(p, stream);
  }
  void sync() { TB_HIP(hipStreamSynchronize(stream)); }
};

// Scoped device buffer owned by a context (freed on scope exit, also on throw).
template <typename T>
struct DevBuf {
  tribit_ctx* ctx = nullptr;
  T* p = nullptr;
  uint64_t n = 0;
  DevBuf() = default;
  DevBuf(tribit_ctx* c, uint64_t count) : ctx(c), p(c->alloc_n<T>(count)), n(count) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ctx(o.ctx), p(o.p), n(o.n) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    reset();
    ctx = o.ctx;
    p = o.p;
    n = o.n;
    o.p = nullptr;
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p && ctx) ctx->free(p);
    p = nullptr;
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

namespace tb {
struct BucketCache;

// Device counters of one count: u64 [0] triangles, [1] fallback tasks; from
// u64 index kQueueOff, the task queues: per menu slot, kQueues sub-queues of
// one u32 each, kQStride u32 apart so that no two share a cache line.
constexpr uint32_t kQueues = 8;
constexpr uint32_t kQStride = 64;
constexpr size_t kQueueOff = 32;
constexpr size_t kCountScratch = kQueueOff + (size_t)TRIBIT_MENU_SLOTS * kQueues * kQStride / 2;
}  // namespace tb

// Count-ready graph resident in HBM.
struct tribit_graph {
  tribit_ctx* ctx = nullptr;
  uint64_t n = 0, m = 0, entries = 0, wplus = 0, rows_active = 0;
  uint32_t max_outdeg = 0;
  uint32_t* row_ptr = nullptr;  // n+1, oriented CSR
  uint32_t* col = nullptr;      // m, ascending ids within a row
  uint32_t* deg = nullptr;      // n, undirected degree (by internal id)
  uint32_t* labels = nullptr;   // n, input id of each internal id (null: ids kept)
  int relabeled = 0;  // ids: 0 input, 1 degree rank, 2 octave classes
  uint64_t* wrow = nullptr;     // n, W+(u) = sum of outdeg(v) over v in N+(u)
  // streamed-range lists (see prep.hip): forward-only and direction-split
  uint2* rng_all = nullptr;     // m, in CSR order; owner u's ranges at row_ptr[u]
  uint2* rng_dir = nullptr;     // m: forward part [0, m_fwd), reverse part [m_fwd, m)
  uint32_t* pid_dir = nullptr;  // m: the vertex whose list each rng_dir entry streams
  uint32_t* fw_ptr = nullptr;   // n+1, owner's forward ranges in rng_dir
  uint32_t* rv_ptr = nullptr;   // n+1, owner's reverse ranges in rng_dir (offsets include m_fwd)
  uint64_t* w_fw = nullptr;     // n, streamed elements of the owner's forward ranges
  uint64_t* w_rv = nullptr;     // n, streamed elements of the owner's reverse ranges
  uint64_t m_fwd = 0;
  // wedge ranges (relabelled graphs only): arcs u->v grouped by v, each
  // streaming the suffix of N+(u) after v
  uint2* rng_sfx = nullptr;     // m
  uint32_t* sfx_ptr = nullptr;  // n+1, owner v's ranges in rng_sfx
  uint64_t* w_sfx = nullptr;    // n, streamed elements of v's ranges
  uint16_t* col16 = nullptr;    // low 16 bits of every id, for narrow wedge ranges (relabelled, m < 2^30)
  uint8_t* sfx_dh = nullptr;    // per wedge range: its ids' high half minus its owner's (255: read col)
  uint16_t* sfx_split = nullptr; // per wedge range crossing one 64K-id boundary: offset of its first id past it
  // bit tiles of the oriented rows (tiles.hip), one set per width (index 0:
  // 32 columns, uint32 masks; 1: 64 columns, uint64 masks), built on first use
  struct TileSet {
    uint64_t ntiles = 0;       // tiles built (0 until built)
    uint32_t* ptr = nullptr;   // n+1, each row's first tile
    uint32_t* blk = nullptr;   // ntiles, column / width
    void* mask = nullptr;      // ntiles, bit (column % width), uint32 or uint64
    uint32_t* idx = nullptr;   // m, tile holding each element
    uint2* trng[4] = {nullptr, nullptr, nullptr, nullptr};  // m, tile ranges aligned with each layout's ranges
  } tiles[2];
  uint64_t tiles32 = 0, tiles64 = 0;  // tile counts (computed on the first tribit_graph_get_info)
  // dense core (dense.hip): bit adjacency of the top core_c ids, padded to core_kp
  uint32_t* core_bits = nullptr;  // kp rows of kp / 32 words
  uint64_t core_c = 0, core_kp = 0;
  uint32_t core_e0 = 0;    // row_ptr[n - core_c]: first arc of the core's rows
  int64_t core_pick = -1;  // the learned schedule's core size from core_choose (-1: not chosen yet)
  // dense diagonal blocks (dense.hip): id ranges [s, e) whose induced subgraph
  // is dense, counted on the matrix cores; found on first use
  struct DenseBlocks {
    bool found = false;           // detection ran
    uint32_t nblocks = 0;
    uint64_t vertices = 0;        // ids inside blocks
    double macs = 0.0;            // int8 multiply-accumulates of one launch
    int8_t* A = nullptr;          // the blocks' 0/1 upper-triangular matrices, each padded to kp x kp
    uint4* work = nullptr;        // per workgroup: matrix offset (lo, hi), kp, tile pair I | J << 16
    uint32_t nwork = 0;
    uint2* rng = nullptr;         // the owners' wedge ranges without those of in-block arcs (32-bit ids)
    uint32_t* ptr = nullptr;      // n + 1: each owner's ranges in rng
    uint64_t* w = nullptr;        // n: streamed elements per owner over rng
    uint64_t cut = 0;             // streamed elements the cut removed
    // the pair kernel's input: arcs u -> w from a block to a later id, sorted
    // as (w << 32) | u, and the runs (first key, keys) of one w with two or
    // more sources in one block, short and long; the blocks' extents, offsets and sides
    uint64_t* key = nullptr;
    uint64_t nkeys = 0;
    uint2* runs = nullptr;
    uint32_t nruns = 0;
    uint2* runs_long = nullptr;
    uint32_t nruns_long = 0;
    uint2* blk = nullptr;
    int32_t* blk_of = nullptr;    // n: block of each id, -1 outside blocks
    uint64_t* aoff = nullptr;
    uint32_t* kps = nullptr;
    double ms_prepare = 0.0;      // device time of detection and these structures
  } blocks;
  tb::BucketCache* bcache = nullptr;  // scheduler statistics (count.hip)
  // the learned schedule's judged candidate (count.hip), valid for the model
  // version, flags and core request it was judged under
  int judge_choice = -1;
  uint64_t judge_model = 0;
  int judge_flags = 0, judge_core = 0;
  double judge_ms[TRIBIT_CANDIDATES] = {};
  double ms_prep = 0.0;
  double ms_layout[4] = {0, 0, 0, 0};  // device time of range layouts built after preprocessing, by schedule id
  bool tiles_counted = false;          // tiles32 / tiles64 computed (on the first tribit_graph_get_info)
  uint64_t offdiag_lists = 0;          // an off-diagonal task's graph: its streamed lists (cols edges)
  uint64_t bytes = 0;
};

namespace tb {
struct Schedule;
void schedule_free(Schedule* s);
}  // namespace tb

// Many graphs counted together: each part is the disjoint union of a run of
// graphs (vertex ids offset), preprocessed as one graph; every vertex carries
// its graph index, so each task's count goes to the owner's graph.
struct tribit_batch {
  struct Part {
    tribit_graph* g = nullptr;
    uint32_t first = 0, count = 0;  // graphs [first, first + count)
    uint32_t* gid = nullptr;        // device: graph index of each internal vertex
    tb::Schedule* S = nullptr;      // tasks (built on the first count)
  };
  tribit_ctx* ctx = nullptr;
  uint32_t ngraphs = 0;
  std::vector<Part> parts;
  std::vector<uint64_t> n, m;             // per graph: vertices, simple undirected edges
  unsigned long long* gcount = nullptr;   // device: per-graph triangle counts
  unsigned long long* scratch = nullptr;  // device: 8 words per part (count, fallbacks, queues)
  hipGraph_t graph = nullptr;             // captured counting launches of all parts
  hipGraphExec_t exec = nullptr;
  int sched = -1, flags = 0;              // what the schedules / capture were built for
  uint64_t model_version = 0;             // the context's cost model they were built with
  int gw[8] = {0};                        // tuned lane-group width per slot (0: the plan's)
  int waves[8] = {0};                     // tuned workgroups per CU per slot (0: occupancy limit)
  int tile_width = 0;                     // tuned: 0 the plans' kernels; 32 / 64 every block launch streams
                                          // bit tiles of that width
  uint32_t launches = 0;                  // kernels in the captured graph
  uint64_t tasks = 0;
  double ms_prep = 0.0;
};

namespace tb {
// prep.hip
// relabel: a TRIBIT_OPT_RELABEL mode, or -1 for the context's option.
void preprocess(tribit_ctx* ctx, uint64_t* d_keys, uint64_t nkeys, uint64_t n, tribit_graph* g, int relabel = -1);
void build_layout(tribit_ctx* ctx, tribit_graph* g, int layout);
void build_offdiag(tribit_ctx* ctx, const uint64_t* d_cols, uint64_t ncols, const uint64_t* d_idxs, uint64_t nidxs,
                   uint64_t n, tribit_graph* g);
bool layout_built(const tribit_graph* g, int layout);
void ensure_layout(tribit_ctx* ctx, tribit_graph* g, int layout);
void free_graph_memory(tribit_graph* g);
void bucket_cache_free(tribit_graph* g);
void free_graph(tribit_graph* g);
void gen_rmat(tribit_ctx* ctx, uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
              uint64_t* d_keys);
void gen_er(tribit_ctx* ctx, uint64_t n, uint64_t nsamples, uint64_t seed, uint64_t* d_keys);
void rmat_thresholds(double a, double b, double c, uint32_t t[3]);
// gen.hip
uint64_t chung_lu_samples(uint64_t n, double avg_deg);
void chung_lu_cdf(uint64_t n, double gamma, uint64_t* cdf);
void chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg);
void gen_chung_lu(tribit_ctx* ctx, uint64_t n, double gamma, double avg_deg, uint64_t seed, uint64_t* d_keys);
uint64_t* gen_rgg(tribit_ctx* ctx, uint64_t n, double avg_deg, uint64_t seed, uint64_t* nkeys);
uint64_t sbm_slots(uint64_t blocks, uint64_t size, double p_out);
void gen_sbm(tribit_ctx* ctx, uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed,
             uint64_t* d_keys);
// count.hip
void count(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts& opts, tribit_count_info* info,
           bool noalloc = false);
void noalloc_init(tribit_ctx* ctx);
uint64_t noalloc_scratch_bytes();
bool judged_known(const tribit_ctx* ctx, const tribit_graph* g, const tribit_count_opts& opts);
void variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out);

// Mask of a range end: wedge range ends carry flags in bits 31 (narrow) and
// 30 (two 64K-id windows) when m < 2^30; larger graphs carry none.
__host__ __device__ inline uint32_t range_end_mask(uint64_t m) {
  return m < (1ull << 30) ? 0x3FFFFFFFu : 0xFFFFFFFFu;
}
void count_profile(tribit_ctx* ctx, tribit_graph* g, int layout, const uint8_t* table, tribit_profile* out);
void profile_class_features(const tribit_profile* p, int cls, int slot, float* f);
bool class_slot_ok(int cls, int slot);
void count_plan(tribit_ctx* ctx, tribit_graph* g, int layout, const uint8_t* table, int repeats, double* ms_kernel,
                uint64_t* triangles);
void learned_plan(tribit_ctx* ctx, tribit_graph* g, uint8_t* table);
void menu_slot(int slot, int* kind, int* g, uint32_t* max_d);
uint32_t bucket_max_outdeg(int bucket);
void batch_count(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts& opts, uint64_t* counts,
                 tribit_batch_count_info* info);
void batch_autotune(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts& opts, tribit_batch_count_info* info);
void batch_drop_exec(tribit_batch* b);
// batch.hip
tribit_batch* batch_from_host(tribit_ctx* ctx, const uint64_t* keys, const uint64_t* key_off, const uint64_t* nverts,
                              uint32_t ngraphs);
tribit_batch* batch_gen_chung_lu(tribit_ctx* ctx, uint64_t seed, uint64_t first, uint32_t ngraphs);
void batch_free(tribit_batch* b);
// mlp.hip
tribit_mlp* mlp_create(tribit_ctx* ctx, int nfeat, uint64_t seed);
void mlp_free(tribit_mlp* m);
size_t mlp_num_params(const tribit_mlp* m);
tribit_ctx* mlp_context(const tribit_mlp* m);
int mlp_num_features(const tribit_mlp* m);
void mlp_get(tribit_mlp* m, float* h);
void mlp_set(tribit_mlp* m, const float* h);
double mlp_grad(tribit_mlp* m, const float* hX, const float* ht, int B, float* hgrad);
double mlp_train(tribit_mlp* m, const float* hX, const float* ht, uint64_t n, int epochs, int batch, float lr,
                 uint64_t seed);
void mlp_infer(tribit_mlp* m, const float* hX, uint64_t n, float* hout);
void mlp_infer_dev(tribit_mlp* m, const float* dX, uint64_t n, float* dout);
// dense.hip
void core_prepare(tribit_ctx* ctx, tribit_graph* g, uint64_t c);
void core_release(tribit_graph* g);
void core_launch(tribit_ctx* ctx, const tribit_graph* g, unsigned long long* result);
uint64_t core_choose(tribit_ctx* ctx, const tribit_graph* g);
double core_macs(uint64_t kp, uint64_t tile);
bool blocks_prepare(tribit_ctx* ctx, tribit_graph* g);  // false: no dense block found
void blocks_release(tribit_graph* g);
void blocks_launch(tribit_ctx* ctx, const tribit_graph* g, unsigned long long* result);
// tiles.hip
uint64_t count_tiles(tribit_ctx* ctx, tribit_graph* g, int width);
void build_tiles(tribit_ctx* ctx, tribit_graph* g, int layout, int width);
void free_tiles(tribit_graph* g);
// count.hip
// mtx.cpp
int read_mtx(const char* path, uint64_t** keys, uint64_t* nkeys, uint64_t* n, uint64_t* entries);
// webgraph.cpp
int read_webgraph(const char* basename, uint64_t** keys, uint64_t* nkeys, uint64_t* n);
}  // namespace tb
