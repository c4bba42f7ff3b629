/*
 * tribit.h -- C ABI of the MI355X (gfx950) exact triangle-counting library
 * libtribit_hip.so. Plain C: no C++ types, exceptions or torch types cross it.
 *
 * Reference interfaces replaced (sixt99/TriBIT):
 *   read_mtx                          src/single_gpu/tribit.cu:228-269
 *   count_triangles (preprocessing)   src/single_gpu/tribit.cu:656-763
 *   count_triangles (counting kernel) src/single_gpu/tribit.cu:765-802
 *   init_gpu_kernel5 / 6              src/multi_gpu/cuda/main5_multi.cu:578-587,
 *                                     src/multi_gpu/cuda/main6_multi.cu:585-593
 *   allocate_memory (pool warm-up)    src/single_gpu/tribit.cu:825-838
 *
 * Conventions.
 *   - Every entry point returns a tribit_status (0 = success) and never exits
 *     the process (the reference prints and exit(1)s, tribit.cu:62-65).
 *   - Edge keys are the reference's packed form: (uint64(min) << 32) | max with
 *     min < max (tribit.cu:258-267). Key 0xFFFFFFFFFFFFFFFF is a "no edge"
 *     sentinel and is dropped by preprocessing.
 *   - Host buffers passed in are borrowed for the call. The library owns all
 *     device memory; a tribit_graph keeps its count-ready structures resident
 *     in HBM until tribit_graph_free.
 *   - Calls are synchronous: they return after the result is on the host.
 *   - One context per device; use one host thread per context.
 *   - Counts are exact uint64.
 */
#ifndef TRIBIT_H
#define TRIBIT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TRIBIT_OK = 0,
  TRIBIT_ERR_IO = 1,        /* cannot open or read the input file */
  TRIBIT_ERR_FORMAT = 2,    /* malformed Matrix Market header or entry */
  TRIBIT_ERR_SHAPE = 3,     /* non-square matrix (tribit.cu:241-245) */
  TRIBIT_ERR_RANGE = 4,     /* vertex id outside [1, n] or key not (min<max<n) */
  TRIBIT_ERR_NOMEM = 5,     /* device allocation failed or exceeds the cap */
  TRIBIT_ERR_ARG = 6,       /* invalid argument */
  TRIBIT_ERR_HIP = 7,       /* HIP runtime error */
  TRIBIT_ERR_NODEVICE = 8,  /* no usable gfx950 device */
  TRIBIT_ERR_TOO_LARGE = 9  /* graph exceeds a 32-bit layout limit */
} tribit_status;

typedef struct tribit_ctx tribit_ctx;
typedef struct tribit_graph tribit_graph;
typedef struct tribit_mlp tribit_mlp;
typedef struct tribit_batch tribit_batch;

/* Count-ready graph description (filled by tribit_graph_get_info). */
typedef struct {
  uint64_t n;              /* vertices */
  uint64_t entries;        /* input keys (including duplicates and sentinels) */
  uint64_t m;              /* undirected simple edges == oriented edges */
  uint64_t wplus;          /* W+ = sum over oriented u->v of outdeg(v) */
  uint64_t rows_active;    /* rows with outdeg >= 2 */
  uint32_t max_outdeg;     /* max oriented out-degree */
  uint32_t reserved0;
  uint64_t tiles32;        /* 32-column bit tiles of the oriented rows */
  uint64_t tiles64;        /* 64-column bit tiles of the oriented rows */
  double ms_preprocess;    /* device time: keys in HBM -> count-ready (hipEvents): CSR, W+ and the default
                              range layout (wedge ranges of relabelled graphs, else the directed layout) */
  uint64_t device_bytes;   /* bytes held by the resident structures */
  double ms_layout[4];     /* device time of each other range layout, indexed by tribit_sched (PLAIN,
                              DIRECTED), built on its first count outside that count's timing; 0: not built */
  double ms_blocks;        /* device time of the dense diagonal blocks' detection and structures, built on the
                              first count that considers them (relabelled graphs, learned or DENSE_BLOCKS),
                              outside that count's timing; 0: not built */
} tribit_graph_info;

/* Scheduler choice for tribit_count. */
typedef enum {
  TRIBIT_SCHED_PLAIN = 0,    /* static degree thresholds; textbook direction: N+(u) is the set, N+(v) streamed */
  TRIBIT_SCHED_LEARNED = 1,  /* per-bucket kernel variant from the cost model, dense core by cost estimate;
                                the first count of a graph times the model's plan against the heuristic plans
                                and keeps the fastest (see tribit_count_info.candidate); the context keeps
                                the choice per input shape (n, m, W+), so a graph rebuilt from the same
                                input runs it without judging again */
  TRIBIT_SCHED_DIRECTED = 2, /* static thresholds; every edge streams the shorter of its two lists */
  TRIBIT_SCHED_WEDGE = 3     /* relabelled graphs: every arc u->v streams the part of N+(u) after v
                                against N+(v) (each out-out wedge of u checked once) */
} tribit_sched;

/* tribit_count_opts.flags */
#define TRIBIT_CANDIDATES 5               /* plans the learned schedule judges (tribit_count_info.candidate) */
#define TRIBIT_COUNT_FORCE_LINEAR_PROBE 1 /* test hook: skip the cuckoo sets, use the fallback tables */
#define TRIBIT_COUNT_FORCE_GLOBAL_LOADS 2 /* test hook: stream with global loads (the path for col >= 4 GiB) */
#define TRIBIT_COUNT_NO_TILES 4           /* never stream bit tiles (block launches stream elements) */
#define TRIBIT_COUNT_FORCE_TILES 8        /* every block launch streams 64-column bit tiles for owners whose
                                             set fits the LDS bitmap (test hook) */
#define TRIBIT_COUNT_NO_BITMAP 16         /* test hook: block tasks keep hashed sets even for narrow id windows */
#define TRIBIT_COUNT_WIDE_IDS 32          /* test hook: wedge ranges whose ids share their high 16 bits are
                                             streamed as 32-bit ids too (by default block tasks stream them
                                             from the 16-bit low halves, two ids per load) */
#define TRIBIT_COUNT_FORCE_MULTI 128      /* test hook: every launch of the 64-slot wave kernel runs as multi-task
                                             waves, whatever its task count */
#define TRIBIT_COUNT_DENSE_BLOCKS 512     /* WEDGE: dense diagonal blocks (id ranges whose induced subgraph is
                                             dense, e.g. the blocks of a block model under TRIBIT_OPT_RELABEL 2)
                                             are counted on the matrix cores, the rest streamed; no top core */
#define TRIBIT_COUNT_TILES32 256          /* launches that stream bit tiles stream 32-column tiles (uint32
                                             masks) instead of 64-column ones */
#define TRIBIT_COUNT_NO_JUDGE 64          /* LEARNED: run the density-threshold heuristic's candidate (preferred
                                             layout + dense core) without timing the candidates first; for
                                             graphs counted once, where judging would cost several counts */

typedef struct {
  int32_t sched;           /* tribit_sched */
  int32_t repeats;         /* timed repetitions (>=1); the count of every repeat must agree */
  int32_t flags;           /* TRIBIT_COUNT_* bits */
  int32_t group_width[4];  /* per kernel variant: lanes per streamed list (wave variants 1..32, block
                              variants 4..64, powers of two; 0 = default) */
  int32_t dense_core;      /* relabelled graphs: vertices of the top-id core counted on the matrix cores;
                              0 = chosen by the cost estimate for LEARNED, none otherwise; -1 = never */
} tribit_count_opts;

typedef struct {
  uint64_t triangles;      /* exact count */
  double ms_count;         /* scheduler + counting kernels, median over repeats (hipEvents) */
  double ms_sched;         /* scheduler part of ms_count (median) */
  double ms_kernel;        /* counting kernels part of ms_count (median) */
  uint64_t launch_blocks;  /* workgroups launched by the counting kernels */
  uint64_t tasks;          /* work items the scheduler produced */
  uint64_t tasks_by_variant[8];
  int32_t repeats_agree;   /* 1 if every repeat gave the same count */
  int32_t nvariants;       /* kernel variants (one launch each) */
  double ms_by_variant[8]; /* device time of each variant's launch (median) */
  uint64_t fallback_tasks; /* tasks whose cuckoo set could not be built (linear-probing fallback) */
  int32_t group_width[8];  /* lanes per streamed list used by each variant */
  uint64_t dense_core;     /* core vertices counted by the int8 MFMA kernel (0: none) */
  double ms_dense;         /* device time of the core launch (median) */
  int32_t candidate;       /* LEARNED: plan that ran (0 cost model, 1 heuristic, 2 heuristic on the
                              directed layout, 3 textbook plain, 4 heuristic with dense diagonal blocks on the
                              matrix cores); -1 for the fixed schedules */
  int32_t layout;          /* range layout that ran (TRIBIT_SCHED_PLAIN / DIRECTED / WEDGE) */
  double ms_judge;         /* LEARNED: device time of the candidate trials this call ran (first call) */
  double ms_candidates[TRIBIT_CANDIDATES]; /* LEARNED: judged time of each candidate (0 = not a candidate) */
  int32_t slot_kind[8];    /* kernel of each launch: 0 wave/64-slot set, 1 wave/512-slot set, 2 block, 3 block
                              streaming bit tiles, 4 multi-task wave (eight 64-slot sets) */
  uint32_t dense_blocks;   /* dense diagonal blocks counted on the matrix cores (their vertices in dense_core) */
  uint32_t reserved1;
} tribit_count_info;

/* Work assigned to each kernel variant by a scheduler, in the units of the
 * algorithmic byte count 4*rows + 4*edges + 4*wplus (4 bytes per vertex, per
 * edge and per wedge check). */
typedef struct {
  uint64_t rows[8];        /* owner lists (a vertex and the lists streamed against N+(vertex)) */
  uint64_t edges[8];       /* edges of those lists */
  uint64_t wplus[8];       /* streamed elements = wedge checks */
  uint64_t narrow[8];      /* of those, elements read as 16-bit halves (2 bytes each: narrow wedge ranges
                              streamed by block tasks) */
  uint64_t narrow_ranges[8]; /* ranges streamed that way (each also reads one byte: its ids' high half
                                relative to its owner's) */
  uint64_t narrow2_ranges[8]; /* of those, ranges across one 64K-id boundary (two more bytes each: the
                                 position of the crossing) */
} tribit_variant_work;

/* ---- context ------------------------------------------------------------ */
/* Selects `device`, creates a stream and sets the default memory pool to keep
 * freed memory (release threshold = max), as init_gpu_kernel5
 * (main5_multi.cu:578-587) and allocate_memory (tribit.cu:825-838) do. */
int tribit_init(int device, tribit_ctx** out);
int tribit_shutdown(tribit_ctx* ctx);
/* Test hook: fail device allocations beyond `bytes` in total (0 = no cap). */
int tribit_set_alloc_cap(tribit_ctx* ctx, uint64_t bytes);
/* Context options (tribit_set_option). */
#define TRIBIT_OPT_RELABEL 1 /* 1: preprocessing renumbers vertices by (degree, id) rank; 2: by
                                (octave of the degree, id) rank, so runs of consecutive input ids
                                of similar degree (the blocks of a block model) stay runs and can be
                                counted as dense diagonal blocks; 3 (default): 2 when the edges
                                joining input ids less than 512 apart fill at least 1/16 of such
                                id pairs (and n >= 4096), else 1; 0: input ids kept. Counts are the
                                same either way. */
#define TRIBIT_OPT_BATCH_PART_KEYS 2 /* input keys per batch part (default 2^30) */
int tribit_set_option(tribit_ctx* ctx, int option, int64_t value);
/* Peak bytes held from the memory pool since tribit_init. */
int tribit_pool_peak(tribit_ctx* ctx, uint64_t* bytes);
const char* tribit_status_str(int status);
/* Build description: library version, rocPRIM and HIP versions, target. */
const char* tribit_build_info(void);
/* Name and compute units of the context's device. */
int tribit_device_info(tribit_ctx* ctx, char* name, size_t name_len, int* cus, double* hbm_peak_gbs);

/* ---- input ----------------------------------------------------------------- */
/* Matrix Market coordinate reader with the semantics of read_mtx
 * (tribit.cu:228-269): banner skipped, '%' lines skipped, square check,
 * 1-based ids, value columns ignored, self-loops dropped, keys packed
 * (min<<32)|max. Validates ids (reference: undefined behaviour).
 * *keys is allocated by the library; release with tribit_free_host. */
int tribit_read_mtx(const char* path, uint64_t** keys, uint64_t* nkeys, uint64_t* n, uint64_t* entries);
/* BVGraph (WebGraph) reader with the semantics of the reference's multi-GPU
 * ingest (src/multi_gpu/src/main.rs:88, :150-161: BvGraphSeq over
 * <basename>.properties + <basename>.graph, every arc (u, v), u != v, as
 * (min<<32)|max, both directions of a symmetric graph kept). Sequential
 * decode of the published format (big-endian bit order; gamma, delta, unary
 * and zeta_k codes per the compressionflags). *keys is allocated by the
 * library; release with tribit_free_host. */
int tribit_read_webgraph(const char* basename, uint64_t** keys, uint64_t* nkeys, uint64_t* n);
void tribit_free_host(void* p);

/* Generate R-MAT edge keys in HBM (2^scale vertices, ef<<scale samples,
 * probabilities a,b,c, d = 1-a-b-c). Returned device buffer is owned by the
 * context and is consumed by tribit_graph_from_device_keys. */
int tribit_gen_rmat(tribit_ctx* ctx, uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
                    uint64_t** d_keys, uint64_t* nkeys);
/* Erdos-Renyi G(n, M): M samples of uniform vertex pairs. */
int tribit_gen_er(tribit_ctx* ctx, uint64_t n, uint64_t nsamples, uint64_t seed, uint64_t** d_keys, uint64_t* nkeys);
/* Copies generated keys back to a host buffer of nkeys (e.g. to partition
 * them by the reference's multi-GPU layout, main.rs:150-161), and releases a
 * generated key buffer that is not handed to tribit_graph_from_device_keys. */
int tribit_device_keys_to_host(tribit_ctx* ctx, const uint64_t* d_keys, uint64_t nkeys, uint64_t* host);
int tribit_free_device_keys(tribit_ctx* ctx, uint64_t* d_keys);

/* Chung-Lu by edge sampling: weights w_i = (i+1)^(-1/(gamma-1)), round(n*avg/2)
 * samples with both endpoints drawn from the weight distribution. */
int tribit_gen_chung_lu(tribit_ctx* ctx, uint64_t n, double gamma, double avg_deg, uint64_t seed,
                        uint64_t** d_keys, uint64_t* nkeys);
/* Parameters of graph `index` of a Chung-Lu batch drawn from `seed`:
 * gamma ~ U[2.1, 2.8], n log-uniform in [1024, 262144], avg_deg ~ U[8, 64]. */
void tribit_chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg);
/* 2-D random geometric graph: n points uniform in the unit square (2^30 grid),
 * radius sqrt(avg_deg / (pi n)); vertex ids in spatial (cell-major) order. */
int tribit_gen_rgg(tribit_ctx* ctx, uint64_t n, double avg_deg, uint64_t seed, uint64_t** d_keys, uint64_t* nkeys);
/* Stochastic block model: `blocks` blocks of `size` consecutive ids, pairs
 * inside a block kept with probability p_in, plus round(p_out * cross pairs)
 * uniform cross-block pairs. */
int tribit_gen_sbm(tribit_ctx* ctx, uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed,
                   uint64_t** d_keys, uint64_t* nkeys);

/* ---- preprocessing: keys -> count-ready graph (tribit.cu:656-763) --------- */
/* Copies host keys to HBM (untimed, tribit.cu:657-659) then preprocesses. */
int tribit_graph_from_host_keys(tribit_ctx* ctx, const uint64_t* keys, uint64_t nkeys, uint64_t n,
                                tribit_graph** out);
/* Preprocesses keys already in HBM; takes ownership of d_keys (frees it). */
int tribit_graph_from_device_keys(tribit_ctx* ctx, uint64_t* d_keys, uint64_t nkeys, uint64_t n,
                                  tribit_graph** out);
int tribit_graph_get_info(const tribit_graph* g, tribit_graph_info* info);
/* Copies the oriented CSR to host buffers (row_ptr[n+1], col[m], deg[n]);
 * any pointer may be NULL. Ids are internal ids (see tribit_graph_labels). */
int tribit_graph_download(const tribit_graph* g, uint32_t* row_ptr, uint32_t* col, uint32_t* deg);
/* Input id of every internal id (input_id[n]; may be NULL) and the ids the
 * graph got (*relabeled: 0 input ids, 1 degree rank, 2 octave classes; see
 * TRIBIT_OPT_RELABEL, whose mode 3 resolves to 1 or 2 per graph). */
int tribit_graph_labels(const tribit_graph* g, uint32_t* input_id, int* relabeled);
/* Bit tiles of width 32 or 64 of the oriented rows (built on first use; the
 * counting kernels stream them in tile launches): *ntiles = their number;
 * with non-NULL buffers also copies tile_ptr[n+1] (each row's first tile),
 * blk[ntiles] (column / width) and mask[ntiles] (uint32 for width 32, uint64
 * for 64; bit column % width). Replaces the reference's 32-column U/L words
 * (tribit.cu:588-602, 672-740) for the rows whose tiles are dense. */
int tribit_graph_tiles(tribit_ctx* ctx, tribit_graph* g, int width, uint64_t* ntiles, uint32_t* tile_ptr,
                       uint32_t* blk, void* mask);
int tribit_graph_free(tribit_graph* g);

/* ---- counting (tribit.cu:765-802) ------------------------------------------ */
int tribit_count(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts* opts, tribit_count_info* info);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). tribit_count with no allocation and no free: the schedule's
 * task lists, its counters and the result live in one scratch array of
 * tribit_noalloc_scratch_bytes() (128 MiB) per GPU, file-scope device memory
 * held until the process ends. opts->sched must be PLAIN, DIRECTED or WEDGE
 * (LEARNED judges candidates and builds tables: TRIBIT_ERR_ARG), with no dense
 * core and no TRIBIT_COUNT_FORCE_TILES. A schedule needing more scratch than
 * the array fails with TRIBIT_ERR_TOO_LARGE (R-MAT scale 25, 5.1M tasks, uses
 * about 90 MiB). Every call on one GPU shares that GPU's array and clears
 * what it needs at the start: calls that overlap on one GPU (two threads or
 * contexts on the same device) race and must be serialised by the caller. */
int tribit_count_noalloc(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts* opts, tribit_count_info* info);
uint64_t tribit_noalloc_scratch_bytes(void);
/* Per-variant work of the scheduler `sched` on this graph (not timed). */
int tribit_variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out);

/* ---- learned scheduler: row-block buckets, menu, timing logs --------------
 * Owner lists are grouped into TRIBIT_BUCKETS buckets (outdeg bin x bin of the
 * mean streamed-range length) and the buckets into TRIBIT_CLASSES classes by
 * owner outdeg (<= 32, <= 256, larger: the density-threshold heuristic's
 * split). A plan maps each bucket to a menu slot (kernel kind + lane-group
 * width); LEARNED maps each class to one slot: the heuristic's, unless the
 * cost model (tribit_set_cost_model) predicts a feasible slot at least about
 * 3% faster for the whole count. The model's inputs are the class's features
 * (tribit_class_features) and its target the log2 ratio of the count's device
 * time with the class moved to that slot over the heuristic plan's
 * (tribit_count_plan). */
#define TRIBIT_BUCKETS 800
#define TRIBIT_MENU_SLOTS 8
#define TRIBIT_FEATURES 24
#define TRIBIT_CLASSES 3
typedef struct {
  uint64_t lists[TRIBIT_BUCKETS];     /* owner lists in the bucket */
  uint64_t sum_d[TRIBIT_BUCKETS];     /* sum of owner outdegrees */
  uint64_t sum_k[TRIBIT_BUCKETS];     /* streamed ranges */
  uint64_t sum_w[TRIBIT_BUCKETS];     /* streamed elements */
  uint64_t eligible[TRIBIT_BUCKETS];  /* owners whose ids fit a 2^18-id bitmap window */
  uint64_t sum_t[TRIBIT_BUCKETS];     /* 64-column tiles of the owners' rows */
  uint64_t ticks[TRIBIT_MENU_SLOTS][TRIBIT_BUCKETS]; /* summed task durations (wall-clock ticks) */
  uint8_t slot[TRIBIT_BUCKETS];       /* the plan that ran */
  double tick_ns;                     /* nanoseconds per tick */
  double ms_kernel;                   /* device time of the counting launches */
  uint64_t triangles;
  double ms_slot[TRIBIT_MENU_SLOTS];  /* device time of each slot's launch */
} tribit_profile;
/* Counts once with the menu plan `table` (slot per bucket; NULL = heuristic)
 * over the range layout `sched` (PLAIN, DIRECTED or WEDGE) and logs every
 * task's duration into out->ticks[slot][bucket]. */
int tribit_count_profile(tribit_ctx* ctx, tribit_graph* g, int sched, const uint8_t* table, tribit_profile* out);
/* Device time of the counting launches (median over `repeats` >= 1 counts, no
 * task logging) under the menu plan `table` (slot per bucket; NULL = the
 * heuristic) over the range layout `sched`; the cost model's training target. */
int tribit_count_plan(tribit_ctx* ctx, tribit_graph* g, int sched, const uint8_t* table, int repeats,
                      double* ms_kernel, uint64_t* triangles);
/* Feature vector (TRIBIT_FEATURES floats) of (outdeg class, slot) from the
 * bucket statistics of a profile: the inputs the cost model is trained on and
 * scored with. The slot must be feasible for the whole class. */
int tribit_class_features(const tribit_profile* p, int cls, int slot, float* features);
/* Menu slot description: kind (0 wave/64-slot set, 1 wave/512-slot set,
 * 2 block), lane-group width and the largest owner outdeg it accepts. */
int tribit_menu_slot(int slot, int* kind, int* group_width, uint32_t* max_outdeg);
/* Largest owner outdeg of a bucket (UINT32_MAX for the last outdeg bin). */
uint32_t tribit_bucket_max_outdeg(int bucket);
/* Cost model used by TRIBIT_SCHED_LEARNED (NULL: the heuristic). Not owned;
 * keep it alive while it is set. */
int tribit_set_cost_model(tribit_ctx* ctx, tribit_mlp* m);
/* The slot per bucket LEARNED would use on this graph. */
int tribit_learned_plan(tribit_ctx* ctx, tribit_graph* g, uint8_t* table);

/* One-shot path with the shape of count_triangles (tribit.cu:656): host keys
 * in, exact count out; preprocessing and counting device times reported.
 * LEARNED does not judge here (the graph is counted once): it runs the choice
 * this context judged for the same input before (tribit_count on a graph of
 * equal n, m and W+), else the heuristic candidate (TRIBIT_COUNT_NO_JUDGE). */
int tribit_count_keys(tribit_ctx* ctx, const uint64_t* keys, uint64_t nkeys, uint64_t n,
                      const tribit_count_opts* opts, uint64_t* triangles, double* ms_preprocess,
                      tribit_count_info* info);

/* ---- batches: many graphs in one counting launch (no reference counterpart)
 * Graphs are packed into parts (disjoint unions, vertex ids offset, each part
 * under 2^30 input keys), preprocessed once per part; the counting launches
 * of all parts are captured into one hipGraph and replayed, and every task
 * adds its count to its owner's graph. */
typedef struct {
  double ms_count;         /* replay of the captured counting launches (hipEvents), median of repeats */
  double ms_sched;         /* building the task lists (first count of a configuration only) */
  double ms_capture;       /* stream capture + instantiation (first count of a configuration only) */
  uint64_t triangles;      /* sum over the batch */
  uint64_t tasks;
  uint32_t parts;
  uint32_t launches;       /* kernel launches in the captured graph */
  int32_t repeats_agree;
  int32_t nslots;
  int32_t group_width[8];  /* lanes per streamed list per slot */
  int32_t waves[8];        /* workgroups per CU per slot (0: occupancy limit) */
  int32_t tile_width;      /* 0: the plans' kernels; 32 or 64: block launches stream bit tiles of that width */
} tribit_batch_count_info;
/* Graph j has keys[key_offsets[j] .. key_offsets[j+1]) ((min<<32)|max of local
 * ids in [0, nverts[j]); duplicates and UINT64_MAX sentinels allowed). */
int tribit_batch_from_host(tribit_ctx* ctx, const uint64_t* keys, const uint64_t* key_offsets, const uint64_t* nverts,
                           uint32_t ngraphs, tribit_batch** out);
/* Graphs first .. first+ngraphs-1 of a Chung-Lu batch: graph i has the
 * parameters of tribit_chung_lu_batch_params(seed, i) and generator seed
 * (seed << 32) + i. */
int tribit_batch_gen_chung_lu(tribit_ctx* ctx, uint64_t seed, uint64_t first, uint32_t ngraphs, tribit_batch** out);
/* Sizes: graphs, parts, total vertices and edges, device preprocessing ms. */
int tribit_batch_info(const tribit_batch* b, uint32_t* ngraphs, uint32_t* nparts, uint64_t* total_n,
                      uint64_t* total_m, double* ms_preprocess);
/* Per graph vertices and simple undirected edges (either may be NULL). */
int tribit_batch_graph_sizes(const tribit_batch* b, uint64_t* n, uint64_t* m);
/* Exact per-graph counts (counts[ngraphs]); opts->sched PLAIN, DIRECTED, WEDGE
 * or LEARNED, opts->repeats replays of the captured launches. */
int tribit_batch_count(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts* opts, uint64_t* counts,
                       tribit_batch_count_info* info);
/* Autotuner: coordinate search over lane-group width and workgroups per CU
 * of every launch slot, then the tile width (block launches streaming ids, or
 * 32- or 64-column bit tiles), timing replays of the captured graph; every
 * configuration must give the same per-graph counts; the best stays with the
 * batch for later counts and is returned (tile_width, group_width, waves). */
int tribit_batch_autotune(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts* opts,
                          tribit_batch_count_info* best);
/* Work of every part's schedule `sched` summed over the batch (the units of
 * tribit_variant_stats; no reference counterpart: the algorithmic bytes of a
 * whole replay). */
int tribit_batch_variant_stats(tribit_ctx* ctx, tribit_batch* b, int sched, tribit_variant_work* out);
int tribit_batch_free(tribit_batch* b);

/* ---- one rank of the reference's multi-GPU decomposition (main.rs:97-220) ---
 * Keys are packed (min<<32)|max over global ids < num_nodes; chunk i holds
 * ptr[i+1] - ptr[i] keys (ptr cumulative, num_ptrs + 1 entries); the caller
 * owns the host chunks, the library frees its device memory before return.
 * Diagonal task (main5_multi.cu:589-728): the chunks hold every edge whose
 * larger id is in partition p; returns the triangles whose two larger
 * vertices are in p (the triangles of those edges).
 * Off-diagonal task (main6_multi.cu:595-751): cols hold the edges whose larger
 * id is in partition pc, idxs those whose larger id is in later partitions S
 * (and smaller id in partitions up to pc); returns the triangles a < b < c
 * with b in pc and c in S, i.e. those with (a, b) in cols and (a, c), (b, c)
 * in idxs: every cols edge streams one endpoint's idxs neighbours against
 * the other's (one count, no union). */
int tribit_count_partial_diagonal(tribit_ctx* ctx, const uint64_t* const* chunks, const uint64_t* nnz_ptr,
                                  uint32_t num_ptrs, uint64_t num_nodes, uint64_t* triangles);
int tribit_count_partial_off_diagonal(tribit_ctx* ctx, const uint64_t* const* cols, const uint64_t* nnz_cols_ptr,
                                      uint32_t num_ptrs_cols, const uint64_t* const* idxs,
                                      const uint64_t* nnz_idxs_ptr, uint32_t num_ptrs_idxs, uint64_t num_nodes,
                                      uint64_t* triangles);
/* The reference-named symbols, same signatures (main5_multi.cu:578-589,
 * main6_multi.cu:585-598; Rust declarations main.rs:15-36): init selects the
 * device for the calling thread; a count returns UINT64_MAX on error (the
 * reference exits) with the status in tribit_partial_last_status(). */
void init_gpu_kernel5(int gpu_id);
void init_gpu_kernel6(int gpu_id);
uint64_t count_partial_triangles_diagonal(uint64_t** chunks, uint64_t* nnz_ptr, uint32_t num_ptrs, uint64_t num_nodes);
uint64_t count_partial_triangles_off_diagonal(uint64_t** cols, uint64_t* nnz_cols_ptr, uint32_t num_ptrs_cols,
                                              uint64_t** idxs, uint64_t* nnz_idxs_ptr, uint32_t num_ptrs_idxs,
                                              uint64_t num_nodes);
int tribit_partial_last_status(void);
/* Sizes and device times of the calling thread's last successful partial
 * count (either entry point, either naming). No reference counterpart: the
 * reference times tasks on the host only (main.rs:357-361). */
typedef struct {
  uint64_t keys;       /* input keys of the task (cols + idxs for an off-diagonal task) */
  uint64_t edges;      /* distinct edges it counts over (off-diagonal: idxs edges + cols edges) */
  uint64_t lists;      /* streamed lists (diagonal: oriented edges; off-diagonal: distinct cols edges) */
  uint64_t streamed;   /* streamed elements, i.e. membership checks (diagonal: W+) */
  uint64_t triangles;
  double ms_h2d;       /* host -> device copies of the chunks (wall clock) */
  double ms_build;     /* device: keys in HBM -> count-ready structures */
  double ms_count;     /* device: scheduler + counting kernels */
} tribit_partial_info;
int tribit_partial_last_info(tribit_partial_info* out);

/* ---- cost model of the learned scheduler (no reference counterpart) -------
 * A 3-layer perceptron, F inputs -> 64 -> 64 -> 1 (ReLU), trained in FP32 by
 * the library's own HIP kernels (forward, backward, Adam; loss = mean squared
 * error / 2) and evaluated in BF16. Parameters, flat FP32:
 * W1[F][64] b1[64] W2[64][64] b2[64] W3[64] b3. Samples are row-major X[n][F]. */
int tribit_mlp_create(tribit_ctx* ctx, int nfeat, uint64_t seed, tribit_mlp** out);
int tribit_mlp_free(tribit_mlp* m);
/* Number of parameters (floats) of the model. */
uint64_t tribit_mlp_num_params(const tribit_mlp* m);
int tribit_mlp_get_params(tribit_mlp* m, float* params);
/* Replaces the parameters (and resets the optimiser state). */
int tribit_mlp_set_params(tribit_mlp* m, const float* params);
/* Loss and gradient (FP32, device kernels) of one batch X[B][F], t[B]. */
int tribit_mlp_grad(tribit_mlp* m, const float* X, const float* t, uint32_t B, float* grad, double* loss);
/* Mini-batch Adam over n samples; *loss = mean loss of the last epoch. */
int tribit_mlp_train(tribit_mlp* m, const float* X, const float* t, uint64_t n, int epochs, int batch, float lr,
                     uint64_t seed, double* loss);
/* BF16 inference: out[n]. */
int tribit_mlp_infer(tribit_mlp* m, const float* X, uint64_t n, float* out);

#ifdef __cplusplus
}
#endif
#endif
