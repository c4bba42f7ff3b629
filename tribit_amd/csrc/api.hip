// api.hip -- extern "C" entry points of libtribit_hip.so (declared in
// include/tribit.h). Every entry point catches internal errors and returns a
// tribit_status; nothing throws across the ABI and nothing exits the process.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <new>
#include <stdexcept>

#include <rocprim/rocprim_version.hpp>

#include "common.hpp"

namespace {

int guard_status(tribit_ctx* ctx, const tb::Error& e) {
  if (ctx) ctx->last_error = e.msg;
  if (getenv("TRIBIT_VERBOSE")) fprintf(stderr, "tribit: %s\n", e.msg.c_str());
  return e.status;
}

template <typename F>
int guarded(tribit_ctx* ctx, F&& f) {
  try {
    f();
    return TRIBIT_OK;
  } catch (const tb::Error& e) {
    return guard_status(ctx, e);
  } catch (const std::bad_alloc&) {
    return TRIBIT_ERR_NOMEM;
  } catch (...) {
    return TRIBIT_ERR_HIP;
  }
}

}  // namespace

extern "C" {

const char* tribit_status_str(int s) {
  switch (s) {
    case TRIBIT_OK: return "ok";
    case TRIBIT_ERR_IO: return "cannot open or read input";
    case TRIBIT_ERR_FORMAT: return "malformed Matrix Market input";
    case TRIBIT_ERR_SHAPE: return "matrix is not square";
    case TRIBIT_ERR_RANGE: return "vertex id out of range";
    case TRIBIT_ERR_NOMEM: return "device memory exhausted";
    case TRIBIT_ERR_ARG: return "invalid argument";
    case TRIBIT_ERR_HIP: return "HIP runtime error";
    case TRIBIT_ERR_NODEVICE: return "no usable device";
    case TRIBIT_ERR_TOO_LARGE: return "graph exceeds a 32-bit layout limit";
    default: return "unknown status";
  }
}

const char* tribit_build_info(void) {
  static char buf[160];
  snprintf(buf, sizeof(buf), "tribit-mi355x 0.1; rocPRIM %d.%d.%d; HIP %d.%d; gfx950", ROCPRIM_VERSION / 100000,
           (ROCPRIM_VERSION / 100) % 1000, ROCPRIM_VERSION % 100, HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
  return buf;
}

int tribit_init(int device, tribit_ctx** out) {
  if (!out) return TRIBIT_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return TRIBIT_ERR_NODEVICE;
  }
  if (device < 0 || device >= ndev) return TRIBIT_ERR_ARG;
  tribit_ctx* ctx = new (std::nothrow) tribit_ctx();
  if (!ctx) return TRIBIT_ERR_NOMEM;
  const int st = guarded(ctx, [&] {
    ctx->device = device;
    TB_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TB_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      tb::fail(TRIBIT_ERR_NODEVICE, std::string("library built for gfx950, device is ") + prop.gcnArchName);
    ctx->cus = prop.multiProcessorCount;
    ctx->lds_per_block = (int)prop.sharedMemPerBlock;
    char dname[128] = {0};
    if (hipDeviceGetName(dname, sizeof(dname) - 1, device) != hipSuccess || !dname[0])
      snprintf(dname, sizeof(dname), "%s", prop.name);
    snprintf(ctx->name, sizeof(ctx->name), "%s (%s, %d CUs)", dname, prop.gcnArchName, prop.multiProcessorCount);
    TB_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    TB_HIP(hipDeviceGetDefaultMemPool(&ctx->pool, device));
    uint64_t thr = UINT64_MAX;  // keep freed blocks in the pool (tribit.cu:829-830)
    TB_HIP(hipMemPoolSetAttribute(ctx->pool, hipMemPoolAttrReleaseThreshold, &thr));
    for (auto& e : ctx->ev) TB_HIP(hipEventCreate(&e));
    tb::noalloc_init(ctx);  // the device is current: its copy of the scratch array
  });
  if (st != TRIBIT_OK) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return TRIBIT_OK;
}

int tribit_shutdown(tribit_ctx* ctx) {
  if (!ctx) return TRIBIT_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TRIBIT_OK;
}

int tribit_set_alloc_cap(tribit_ctx* ctx, uint64_t bytes) {
  if (!ctx) return TRIBIT_ERR_ARG;
  ctx->cap = bytes;
  return TRIBIT_OK;
}

int tribit_set_option(tribit_ctx* ctx, int option, int64_t value) {
  if (!ctx) return TRIBIT_ERR_ARG;
  switch (option) {
    case TRIBIT_OPT_RELABEL:
      if (value < 0 || value > 3) return TRIBIT_ERR_ARG;
      ctx->relabel = (int)value;
      return TRIBIT_OK;
    case TRIBIT_OPT_BATCH_PART_KEYS:
      if (value < 1) return TRIBIT_ERR_ARG;
      ctx->part_keys = (uint64_t)value;
      return TRIBIT_OK;
    default:
      return TRIBIT_ERR_ARG;
  }
}

int tribit_pool_peak(tribit_ctx* ctx, uint64_t* bytes) {
  if (!ctx || !bytes) return TRIBIT_ERR_ARG;
  *bytes = ctx->peak;
  return TRIBIT_OK;
}

int tribit_device_info(tribit_ctx* ctx, char* name, size_t name_len, int* cus, double* hbm_peak_gbs) {
  if (!ctx) return TRIBIT_ERR_ARG;
  if (name && name_len) {
    strncpy(name, ctx->name, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (cus) *cus = ctx->cus;
  if (hbm_peak_gbs) *hbm_peak_gbs = 8000.0;  // MI355X HBM3E, AMD-published
  return TRIBIT_OK;
}

int tribit_read_mtx(const char* path, uint64_t** keys, uint64_t* nkeys, uint64_t* n, uint64_t* entries) {
  if (!path || !keys || !nkeys || !n) return TRIBIT_ERR_ARG;
  try {
    return tb::read_mtx(path, keys, nkeys, n, entries);
  } catch (const std::bad_alloc&) {
    return TRIBIT_ERR_NOMEM;
  } catch (...) {  // e.g. a parse thread that could not start
    return TRIBIT_ERR_IO;
  }
}

int tribit_read_webgraph(const char* basename, uint64_t** keys, uint64_t* nkeys, uint64_t* n) {
  if (!basename || !keys || !nkeys || !n) return TRIBIT_ERR_ARG;
  try {
    return tb::read_webgraph(basename, keys, nkeys, n);
  } catch (const std::bad_alloc&) {
    return TRIBIT_ERR_NOMEM;
  } catch (const std::length_error&) {
    return TRIBIT_ERR_NOMEM;
  } catch (...) {
    return TRIBIT_ERR_IO;
  }
}

void tribit_free_host(void* p) { free(p); }

int tribit_gen_rmat(tribit_ctx* ctx, uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
                    uint64_t** d_keys, uint64_t* nkeys) {
  if (!ctx || !d_keys || !nkeys || scale < 1 || scale > 31 || ef < 1) return TRIBIT_ERR_ARG;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    const uint64_t ne = (uint64_t)ef << scale;
    uint64_t* p = ctx->alloc_n<uint64_t>(ne);
    try {
      tb::gen_rmat(ctx, scale, ef, a, b, c, seed, p);
    } catch (...) {
      ctx->free(p);
      throw;
    }
    *d_keys = p;
    *nkeys = ne;
  });
}

int tribit_device_keys_to_host(tribit_ctx* ctx, const uint64_t* d_keys, uint64_t nkeys, uint64_t* host) {
  if (!ctx || (nkeys && (!d_keys || !host))) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    if (nkeys) TB_HIP(hipMemcpyAsync(host, d_keys, 8 * nkeys, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int tribit_free_device_keys(tribit_ctx* ctx, uint64_t* d_keys) {
  if (!ctx) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    ctx->free(d_keys);
    ctx->sync();
  });
}

int tribit_gen_er(tribit_ctx* ctx, uint64_t n, uint64_t nsamples, uint64_t seed, uint64_t** d_keys,
                  uint64_t* nkeys) {
  if (!ctx || !d_keys || !nkeys || n < 1 || n > 0xFFFFFFFEull) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    uint64_t* p = ctx->alloc_n<uint64_t>(nsamples ? nsamples : 1);
    try {
      if (nsamples) tb::gen_er(ctx, n, nsamples, seed, p);
    } catch (...) {
      ctx->free(p);
      throw;
    }
    *d_keys = p;
    *nkeys = nsamples;
  });
}

int tribit_gen_chung_lu(tribit_ctx* ctx, uint64_t n, double gamma, double avg_deg, uint64_t seed,
                        uint64_t** d_keys, uint64_t* nkeys) {
  if (!ctx || !d_keys || !nkeys || n < 2 || n > 0xFFFFFFFEull || !(gamma > 1.0) || !(avg_deg >= 0.0))
    return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    const uint64_t ne = tb::chung_lu_samples(n, avg_deg);
    uint64_t* p = ctx->alloc_n<uint64_t>(ne ? ne : 1);
    try {
      tb::gen_chung_lu(ctx, n, gamma, avg_deg, seed, p);
    } catch (...) {
      ctx->free(p);
      throw;
    }
    *d_keys = p;
    *nkeys = ne;
  });
}

void tribit_chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg) {
  double g = 0, d = 0;
  uint64_t nn = 0;
  tb::chung_lu_batch_params(seed, index, &g, &nn, &d);
  if (gamma) *gamma = g;
  if (n) *n = nn;
  if (avg_deg) *avg_deg = d;
}

int tribit_gen_rgg(tribit_ctx* ctx, uint64_t n, double avg_deg, uint64_t seed, uint64_t** d_keys, uint64_t* nkeys) {
  if (!ctx || !d_keys || !nkeys || n < 2 || n > 0xFFFFFFFEull || !(avg_deg > 0.0)) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    uint64_t ne = 0;
    uint64_t* p = tb::gen_rgg(ctx, n, avg_deg, seed, &ne);
    *d_keys = p;
    *nkeys = ne;
  });
}

int tribit_gen_sbm(tribit_ctx* ctx, uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed,
                   uint64_t** d_keys, uint64_t* nkeys) {
  if (!ctx || !d_keys || !nkeys || blocks < 1 || size < 1 || blocks * size > 0xFFFFFFFEull ||
      !(p_in >= 0.0 && p_in <= 1.0) || !(p_out >= 0.0 && p_out <= 1.0))
    return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    const uint64_t slots = tb::sbm_slots(blocks, size, p_out);
    uint64_t* p = ctx->alloc_n<uint64_t>(slots ? slots : 1);
    try {
      tb::gen_sbm(ctx, blocks, size, p_in, p_out, seed, p);
    } catch (...) {
      ctx->free(p);
      throw;
    }
    *d_keys = p;
    *nkeys = slots;
  });
}

int tribit_graph_from_device_keys(tribit_ctx* ctx, uint64_t* d_keys, uint64_t nkeys, uint64_t n,
                                  tribit_graph** out) {
  if (!ctx || !out || (!d_keys && nkeys)) return TRIBIT_ERR_ARG;
  *out = nullptr;
  tribit_graph* g = new (std::nothrow) tribit_graph();
  if (!g) return TRIBIT_ERR_NOMEM;
  const int st = guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    uint64_t* keys = d_keys;
    if (!keys) keys = ctx->alloc_n<uint64_t>(1);
    tb::preprocess(ctx, keys, nkeys, n, g);
  });
  if (st != TRIBIT_OK) {
    tb::free_graph_memory(g);
    delete g;
    return st;
  }
  *out = g;
  return TRIBIT_OK;
}

int tribit_graph_from_host_keys(tribit_ctx* ctx, const uint64_t* keys, uint64_t nkeys, uint64_t n,
                                tribit_graph** out) {
  if (!ctx || !out || (!keys && nkeys)) return TRIBIT_ERR_ARG;
  uint64_t* d = nullptr;
  const int st = guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    d = ctx->alloc_n<uint64_t>(nkeys ? nkeys : 1);
    if (nkeys) TB_HIP(hipMemcpyAsync(d, keys, nkeys * 8, hipMemcpyHostToDevice, ctx->stream));
    ctx->sync();
  });
  if (st != TRIBIT_OK) {
    if (d) ctx->free(d);
    return st;
  }
  return tribit_graph_from_device_keys(ctx, d, nkeys, n, out);
}

int tribit_graph_get_info(const tribit_graph* g, tribit_graph_info* info) {
  if (!g || !info) return TRIBIT_ERR_ARG;
  memset(info, 0, sizeof(*info));
  info->n = g->n;
  info->entries = g->entries;
  info->m = g->m;
  info->wplus = g->wplus;
  info->rows_active = g->rows_active;
  info->max_outdeg = g->max_outdeg;
  info->ms_preprocess = g->ms_prep;
  for (int l = 0; l < 4; ++l) info->ms_layout[l] = g->ms_layout[l];
  info->ms_blocks = g->blocks.ms_prepare;
  if (!g->tiles_counted && g->ctx) {  // tile counts on first request (untimed; not part of preprocessing)
    tribit_graph* gm = const_cast<tribit_graph*>(g);
    const int st = guarded(g->ctx, [&] {
      TB_HIP(hipSetDevice(g->ctx->device));
      gm->tiles32 = gm->m ? tb::count_tiles(g->ctx, gm, 32) : 0;
      gm->tiles64 = gm->m ? tb::count_tiles(g->ctx, gm, 64) : 0;
      gm->tiles_counted = true;
    });
    if (st != TRIBIT_OK) return st;
  }
  info->device_bytes = g->bytes;
  info->tiles32 = g->tiles32;
  info->tiles64 = g->tiles64;
  return TRIBIT_OK;
}

int tribit_graph_tiles(tribit_ctx* ctx, tribit_graph* g, int width, uint64_t* ntiles, uint32_t* tile_ptr,
                       uint32_t* blk, void* mask) {
  if (!ctx || !g || g->ctx != ctx || !ntiles || (width != 32 && width != 64)) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    const int layout = g->rng_sfx ? TRIBIT_SCHED_WEDGE : TRIBIT_SCHED_DIRECTED;
    if (g->m) tb::build_tiles(ctx, g, layout, width);
    const tribit_graph::TileSet& T = g->tiles[width == 32 ? 0 : 1];
    *ntiles = T.ntiles;
    if (tile_ptr) {
      if (T.ptr)
        TB_HIP(hipMemcpyAsync(tile_ptr, T.ptr, 4 * (g->n + 1), hipMemcpyDeviceToHost, ctx->stream));
      else
        memset(tile_ptr, 0, 4 * (g->n + 1));
    }
    if (blk && T.ntiles) TB_HIP(hipMemcpyAsync(blk, T.blk, 4 * T.ntiles, hipMemcpyDeviceToHost, ctx->stream));
    if (mask && T.ntiles)
      TB_HIP(hipMemcpyAsync(mask, T.mask, (width / 8) * T.ntiles, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int tribit_graph_download(const tribit_graph* g, uint32_t* row_ptr, uint32_t* col, uint32_t* deg) {
  if (!g) return TRIBIT_ERR_ARG;
  tribit_ctx* ctx = g->ctx;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    if (row_ptr) TB_HIP(hipMemcpyAsync(row_ptr, g->row_ptr, 4 * (g->n + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (col && g->m) TB_HIP(hipMemcpyAsync(col, g->col, 4 * g->m, hipMemcpyDeviceToHost, ctx->stream));
    if (deg) TB_HIP(hipMemcpyAsync(deg, g->deg, 4 * g->n, hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
  });
}

int tribit_graph_labels(const tribit_graph* g, uint32_t* input_id, int* relabeled) {
  if (!g) return TRIBIT_ERR_ARG;
  if (relabeled) *relabeled = g->relabeled;
  if (!input_id) return TRIBIT_OK;
  tribit_ctx* ctx = g->ctx;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    if (g->labels) {
      TB_HIP(hipMemcpyAsync(input_id, g->labels, 4 * g->n, hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
    } else {
      for (uint64_t i = 0; i < g->n; ++i) input_id[i] = (uint32_t)i;
    }
  });
}

int tribit_graph_free(tribit_graph* g) {
  if (!g) return TRIBIT_ERR_ARG;
  if (g->ctx) {
    (void)hipSetDevice(g->ctx->device);
    tb::free_graph_memory(g);
    (void)hipStreamSynchronize(g->ctx->stream);
  }
  delete g;
  return TRIBIT_OK;
}

int tribit_count(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts* opts, tribit_count_info* info) {
  if (!ctx || !g || g->ctx != ctx) return TRIBIT_ERR_ARG;
  tribit_count_opts o;
  memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  if (o.sched != TRIBIT_SCHED_PLAIN && o.sched != TRIBIT_SCHED_LEARNED && o.sched != TRIBIT_SCHED_DIRECTED &&
      o.sched != TRIBIT_SCHED_WEDGE)
    return TRIBIT_ERR_ARG;
  if (o.sched == TRIBIT_SCHED_WEDGE && !g->rng_sfx) {
    ctx->last_error = "the wedge schedule needs a relabelled graph (TRIBIT_OPT_RELABEL)";
    return TRIBIT_ERR_ARG;
  }

  for (int i = 0; i < 4; ++i) {
    const int w = o.group_width[i];
    if (w != 0 && w != 1 && w != 2 && w != 4 && w != 8 && w != 16 && w != 32 && w != 64) return TRIBIT_ERR_ARG;
    if (i <= 1 && w == 64) return TRIBIT_ERR_ARG;  // wave variants: at most 32 lanes per list
    if (i >= 2 && w != 0 && w < 4) return TRIBIT_ERR_ARG;  // block variants: at least 4 lanes per list
  }
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tribit_count_info local;
    memset(&local, 0, sizeof(local));
    tb::count(ctx, g, o, &local);
    if (info) *info = local;
  });
}

int tribit_count_noalloc(tribit_ctx* ctx, tribit_graph* g, const tribit_count_opts* opts, tribit_count_info* info) {
  if (!ctx || !g || g->ctx != ctx || !opts) return TRIBIT_ERR_ARG;
  const tribit_count_opts o = *opts;
  if (o.sched != TRIBIT_SCHED_PLAIN && o.sched != TRIBIT_SCHED_DIRECTED && o.sched != TRIBIT_SCHED_WEDGE) {
    ctx->last_error = "tribit_count_noalloc runs PLAIN, DIRECTED or WEDGE (LEARNED judges and allocates)";
    return TRIBIT_ERR_ARG;
  }
  if (o.dense_core > 0 || (o.flags & TRIBIT_COUNT_FORCE_TILES)) {
    ctx->last_error = "tribit_count_noalloc counts without a dense core and without tiles";
    return TRIBIT_ERR_ARG;
  }
  if (o.sched == TRIBIT_SCHED_WEDGE && !g->rng_sfx) {
    ctx->last_error = "the wedge schedule needs a relabelled graph (TRIBIT_OPT_RELABEL)";
    return TRIBIT_ERR_ARG;
  }
  for (int i = 0; i < 4; ++i) {
    const int w = o.group_width[i];
    if (w != 0 && w != 1 && w != 2 && w != 4 && w != 8 && w != 16 && w != 32 && w != 64) return TRIBIT_ERR_ARG;
    if (i <= 1 && w == 64) return TRIBIT_ERR_ARG;
    if (i >= 2 && w != 0 && w < 4) return TRIBIT_ERR_ARG;
  }
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tribit_count_info local;
    memset(&local, 0, sizeof(local));
    tb::count(ctx, g, o, &local, /*noalloc=*/true);
    if (info) *info = local;
  });
}

uint64_t tribit_noalloc_scratch_bytes(void) { return tb::noalloc_scratch_bytes(); }

int tribit_variant_stats(tribit_ctx* ctx, tribit_graph* g, int sched, tribit_variant_work* out) {
  if (!ctx || !g || !out || g->ctx != ctx) return TRIBIT_ERR_ARG;
  if (sched < TRIBIT_SCHED_PLAIN || sched > TRIBIT_SCHED_WEDGE) return TRIBIT_ERR_ARG;
  if (sched == TRIBIT_SCHED_WEDGE && !g->rng_sfx) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::variant_stats(ctx, g, sched, out);
  });
}

int tribit_count_keys(tribit_ctx* ctx, const uint64_t* keys, uint64_t nkeys, uint64_t n,
                      const tribit_count_opts* opts, uint64_t* triangles, double* ms_preprocess,
                      tribit_count_info* info) {
  if (!triangles) return TRIBIT_ERR_ARG;
  tribit_graph* g = nullptr;
  int st = tribit_graph_from_host_keys(ctx, keys, nkeys, n, &g);
  if (st != TRIBIT_OK) return st;
  tribit_count_info local;
  memset(&local, 0, sizeof(local));
  tribit_count_opts o;
  memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  // a graph counted once is not judged (judging runs every candidate twice):
  // LEARNED runs the judged choice if this context has judged the same input
  // before, else the heuristic candidate
  if (o.sched == TRIBIT_SCHED_LEARNED && !(o.flags & TRIBIT_COUNT_NO_JUDGE) && !tb::judged_known(ctx, g, o))
    o.flags |= TRIBIT_COUNT_NO_JUDGE;
  st = tribit_count(ctx, g, &o, &local);
  if (st == TRIBIT_OK) {
    *triangles = local.triangles;
    if (ms_preprocess) *ms_preprocess = g->ms_prep;
    if (info) *info = local;
  }
  tribit_graph_free(g);
  return st;
}

int tribit_mlp_create(tribit_ctx* ctx, int nfeat, uint64_t seed, tribit_mlp** out) {
  if (!ctx || !out) return TRIBIT_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    *out = tb::mlp_create(ctx, nfeat, seed);
  });
}

int tribit_mlp_free(tribit_mlp* m) {
  if (!m) return TRIBIT_ERR_ARG;
  tribit_ctx* c = tb::mlp_context(m);
  if (c && c->model == m) {
    c->model = nullptr;
    ++c->model_version;
  }
  tb::mlp_free(m);
  return TRIBIT_OK;
}

uint64_t tribit_mlp_num_params(const tribit_mlp* m) { return m ? (uint64_t)tb::mlp_num_params(m) : 0; }

static tribit_ctx* mlp_ctx(tribit_mlp* m) {
  tribit_ctx* c = tb::mlp_context(m);
  (void)hipSetDevice(c->device);
  return c;
}

int tribit_mlp_get_params(tribit_mlp* m, float* params) {
  if (!m || !params) return TRIBIT_ERR_ARG;
  return guarded(mlp_ctx(m), [&] { tb::mlp_get(m, params); });
}

int tribit_mlp_set_params(tribit_mlp* m, const float* params) {
  if (!m || !params) return TRIBIT_ERR_ARG;
  return guarded(mlp_ctx(m), [&] { tb::mlp_set(m, params); });
}

int tribit_mlp_grad(tribit_mlp* m, const float* X, const float* t, uint32_t B, float* grad, double* loss) {
  if (!m || !X || !t || !grad || B == 0) return TRIBIT_ERR_ARG;
  return guarded(mlp_ctx(m), [&] {
    const double l = tb::mlp_grad(m, X, t, (int)B, grad);
    if (loss) *loss = l;
  });
}

int tribit_mlp_train(tribit_mlp* m, const float* X, const float* t, uint64_t n, int epochs, int batch, float lr,
                     uint64_t seed, double* loss) {
  if (!m || !X || !t || n == 0 || epochs < 0 || batch < 1) return TRIBIT_ERR_ARG;
  return guarded(mlp_ctx(m), [&] {
    const double l = tb::mlp_train(m, X, t, n, epochs, batch, lr, seed);
    if (loss) *loss = l;
  });
}

int tribit_mlp_infer(tribit_mlp* m, const float* X, uint64_t n, float* out) {
  if (!m || (n && (!X || !out))) return TRIBIT_ERR_ARG;
  return guarded(mlp_ctx(m), [&] {
    if (n) tb::mlp_infer(m, X, n, out);
  });
}

int tribit_count_profile(tribit_ctx* ctx, tribit_graph* g, int sched, const uint8_t* table, tribit_profile* out) {
  if (!ctx || !g || g->ctx != ctx || !out) return TRIBIT_ERR_ARG;
  if (sched != TRIBIT_SCHED_PLAIN && sched != TRIBIT_SCHED_DIRECTED && sched != TRIBIT_SCHED_WEDGE)
    return TRIBIT_ERR_ARG;
  if (sched == TRIBIT_SCHED_WEDGE && !g->rng_sfx) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::count_profile(ctx, g, sched, table, out);
  });
}

int tribit_count_plan(tribit_ctx* ctx, tribit_graph* g, int sched, const uint8_t* table, int repeats,
                      double* ms_kernel, uint64_t* triangles) {
  if (!ctx || !g || g->ctx != ctx || !ms_kernel || !triangles || repeats < 1) return TRIBIT_ERR_ARG;
  if (sched != TRIBIT_SCHED_PLAIN && sched != TRIBIT_SCHED_DIRECTED && sched != TRIBIT_SCHED_WEDGE)
    return TRIBIT_ERR_ARG;
  if (sched == TRIBIT_SCHED_WEDGE && !g->rng_sfx) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::count_plan(ctx, g, sched, table, repeats, ms_kernel, triangles);
  });
}

int tribit_class_features(const tribit_profile* p, int cls, int slot, float* features) {
  if (!p || !features || cls < 0 || cls >= TRIBIT_CLASSES || slot < 0 || slot >= TRIBIT_MENU_SLOTS ||
      !tb::class_slot_ok(cls, slot))
    return TRIBIT_ERR_ARG;
  tb::profile_class_features(p, cls, slot, features);
  return TRIBIT_OK;
}

int tribit_menu_slot(int slot, int* kind, int* group_width, uint32_t* max_outdeg) {
  if (slot < 0 || slot >= TRIBIT_MENU_SLOTS) return TRIBIT_ERR_ARG;
  tb::menu_slot(slot, kind, group_width, max_outdeg);
  return TRIBIT_OK;
}

uint32_t tribit_bucket_max_outdeg(int bucket) {
  return (bucket < 0 || bucket >= TRIBIT_BUCKETS) ? 0 : tb::bucket_max_outdeg(bucket);
}

int tribit_set_cost_model(tribit_ctx* ctx, tribit_mlp* m) {
  if (!ctx) return TRIBIT_ERR_ARG;
  if (m && (tb::mlp_context(m) != ctx || tb::mlp_num_features(m) != TRIBIT_FEATURES)) return TRIBIT_ERR_ARG;
  ctx->model = m;
  ++ctx->model_version;
  return TRIBIT_OK;
}

int tribit_learned_plan(tribit_ctx* ctx, tribit_graph* g, uint8_t* table) {
  if (!ctx || !g || g->ctx != ctx || !table) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::learned_plan(ctx, g, table);
  });
}

int tribit_batch_from_host(tribit_ctx* ctx, const uint64_t* keys, const uint64_t* key_offsets, const uint64_t* nverts,
                           uint32_t ngraphs, tribit_batch** out) {
  if (!ctx || !out || !key_offsets || !nverts || ngraphs == 0 || (!keys && key_offsets[ngraphs] > key_offsets[0]))
    return TRIBIT_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    *out = tb::batch_from_host(ctx, keys, key_offsets, nverts, ngraphs);
  });
}

int tribit_batch_gen_chung_lu(tribit_ctx* ctx, uint64_t seed, uint64_t first, uint32_t ngraphs, tribit_batch** out) {
  if (!ctx || !out || ngraphs == 0) return TRIBIT_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    *out = tb::batch_gen_chung_lu(ctx, seed, first, ngraphs);
  });
}

int tribit_batch_info(const tribit_batch* b, uint32_t* ngraphs, uint32_t* nparts, uint64_t* total_n,
                      uint64_t* total_m, double* ms_preprocess) {
  if (!b) return TRIBIT_ERR_ARG;
  uint64_t tn = 0, tm = 0;
  for (uint32_t j = 0; j < b->ngraphs; ++j) {
    tn += b->n[j];
    tm += b->m[j];
  }
  if (ngraphs) *ngraphs = b->ngraphs;
  if (nparts) *nparts = (uint32_t)b->parts.size();
  if (total_n) *total_n = tn;
  if (total_m) *total_m = tm;
  if (ms_preprocess) *ms_preprocess = b->ms_prep;
  return TRIBIT_OK;
}

int tribit_batch_graph_sizes(const tribit_batch* b, uint64_t* n, uint64_t* m) {
  if (!b) return TRIBIT_ERR_ARG;
  for (uint32_t j = 0; j < b->ngraphs; ++j) {
    if (n) n[j] = b->n[j];
    if (m) m[j] = b->m[j];
  }
  return TRIBIT_OK;
}

int tribit_batch_variant_stats(tribit_ctx* ctx, tribit_batch* b, int sched, tribit_variant_work* out) {
  if (!ctx || !b || !out || b->ctx != ctx) return TRIBIT_ERR_ARG;
  if (sched < TRIBIT_SCHED_PLAIN || sched > TRIBIT_SCHED_WEDGE) return TRIBIT_ERR_ARG;
  for (const auto& part : b->parts)
    if (sched == TRIBIT_SCHED_WEDGE && !part.g->rng_sfx) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    memset(out, 0, sizeof(*out));
    for (const auto& part : b->parts) {
      tribit_variant_work w;
      tb::variant_stats(ctx, part.g, sched, &w);
      for (int i = 0; i < 8; ++i) {
        out->rows[i] += w.rows[i];
        out->edges[i] += w.edges[i];
        out->wplus[i] += w.wplus[i];
        out->narrow[i] += w.narrow[i];
        out->narrow_ranges[i] += w.narrow_ranges[i];
        out->narrow2_ranges[i] += w.narrow2_ranges[i];
      }
    }
  });
}

static int batch_opts(const tribit_count_opts* opts, const tribit_batch* b, tribit_count_opts* o) {
  memset(o, 0, sizeof(*o));
  o->sched = TRIBIT_SCHED_LEARNED;
  if (opts) *o = *opts;
  if (o->sched < TRIBIT_SCHED_PLAIN || o->sched > TRIBIT_SCHED_WEDGE) return TRIBIT_ERR_ARG;
  if (o->sched == TRIBIT_SCHED_WEDGE)
    for (const auto& p : b->parts)
      if (!p.g->rng_sfx) return TRIBIT_ERR_ARG;
  return TRIBIT_OK;
}

int tribit_batch_count(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts* opts, uint64_t* counts,
                       tribit_batch_count_info* info) {
  if (!ctx || !b || b->ctx != ctx || !counts) return TRIBIT_ERR_ARG;
  tribit_count_opts o;
  const int st = batch_opts(opts, b, &o);
  if (st) return st;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::batch_count(ctx, b, o, counts, info);
  });
}

int tribit_batch_autotune(tribit_ctx* ctx, tribit_batch* b, const tribit_count_opts* opts,
                          tribit_batch_count_info* best) {
  if (!ctx || !b || b->ctx != ctx) return TRIBIT_ERR_ARG;
  tribit_count_opts o;
  const int st = batch_opts(opts, b, &o);
  if (st) return st;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    tb::batch_autotune(ctx, b, o, best);
  });
}

int tribit_batch_free(tribit_batch* b) {
  if (!b) return TRIBIT_ERR_ARG;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  tb::batch_free(b);
  return TRIBIT_OK;
}

// ---- one rank of the reference's multi-GPU decomposition ------------------
namespace {

thread_local tribit_partial_info t_partial_info;

// The chunks' keys copied into one device buffer (chunk by chunk, no host
// gather); the caller owns the result. Sets the key count and the copy time.
uint64_t* upload_chunks(tribit_ctx* ctx, const uint64_t* const* chunks, const uint64_t* ptr, uint32_t num,
                        uint64_t* nkeys, double* ms) {
  const uint64_t total = ptr[num] - ptr[0];
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t* d = ctx->alloc_n<uint64_t>(total ? total : 1);
  uint64_t at = 0;
  for (uint32_t i = 0; i < num; ++i) {
    const uint64_t k = ptr[i + 1] - ptr[i];
    if (k) TB_HIP(hipMemcpyAsync(d + at, chunks[i], 8 * k, hipMemcpyHostToDevice, ctx->stream));
    at += k;
  }
  ctx->sync();
  *ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *nkeys = total;
  return d;
}

// Diagonal task: the triangles of its edges (preprocessing + the learned
// schedule's heuristic candidate: a graph counted once is not judged).
uint64_t count_diagonal(tribit_ctx* ctx, const uint64_t* const* chunks, const uint64_t* ptr, uint32_t num,
                        uint64_t n) {
  tribit_partial_info pi;
  memset(&pi, 0, sizeof(pi));
  tribit_graph* g = new tribit_graph;
  try {
    uint64_t nk = 0;
    uint64_t* d = upload_chunks(ctx, chunks, ptr, num, &nk, &pi.ms_h2d);
    // counted once without judging (no dense diagonal blocks): the automatic id
    // order resolves to degree rank, which streams faster
    tb::preprocess(ctx, d, nk, n, g, ctx->relabel == 3 ? 1 : ctx->relabel);
    tribit_count_opts o;
    memset(&o, 0, sizeof(o));
    o.sched = TRIBIT_SCHED_LEARNED;
    o.flags = TRIBIT_COUNT_NO_JUDGE;
    o.repeats = 1;
    tribit_count_info ci;
    tb::count(ctx, g, o, &ci);
    pi.keys = nk;
    pi.edges = g->m;
    pi.lists = g->m;
    pi.streamed = g->wplus;
    pi.triangles = ci.triangles;
    pi.ms_build = g->ms_prep + g->ms_layout[ci.layout & 3];
    pi.ms_count = ci.ms_count;
    tb::free_graph(g);
    t_partial_info = pi;
    return ci.triangles;
  } catch (...) {
    tb::free_graph(g);
    throw;
  }
}

// Off-diagonal task: triangles a < b < c with (a, b) in cols and (a, c),
// (b, c) in idxs, counted directly (build_offdiag) on the DIRECTED schedule.
uint64_t count_off_diagonal(tribit_ctx* ctx, const uint64_t* const* cols, const uint64_t* cptr, uint32_t cnum,
                            const uint64_t* const* idxs, const uint64_t* iptr, uint32_t inum, uint64_t n) {
  tribit_partial_info pi;
  memset(&pi, 0, sizeof(pi));
  tribit_graph* g = new tribit_graph;
  DevBuf<uint64_t> dc, dx;
  try {
    uint64_t nc = 0, nx = 0;
    dc.ctx = dx.ctx = ctx;
    dc.p = upload_chunks(ctx, cols, cptr, cnum, &nc, &pi.ms_h2d);
    dx.p = upload_chunks(ctx, idxs, iptr, inum, &nx, &pi.ms_h2d);
    TB_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    tb::build_offdiag(ctx, dc.p, nc, dx.p, nx, n, g);
    TB_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    TB_HIP(hipEventSynchronize(ctx->ev[1]));
    float ms = 0.f;
    TB_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    dc.reset();
    dx.reset();
    tribit_count_opts o;
    memset(&o, 0, sizeof(o));
    o.sched = TRIBIT_SCHED_DIRECTED;
    o.repeats = 1;
    tribit_count_info ci;
    tb::count(ctx, g, o, &ci);
    pi.keys = nc + nx;
    pi.edges = g->m + g->offdiag_lists;
    pi.lists = g->offdiag_lists;
    pi.streamed = g->wplus;
    pi.triangles = ci.triangles;
    pi.ms_build = ms;
    pi.ms_count = ci.ms_count;
    tb::free_graph(g);
    t_partial_info = pi;
    return ci.triangles;
  } catch (...) {
    tb::free_graph(g);
    throw;
  }
}

bool chunks_ok(const uint64_t* const* chunks, const uint64_t* ptr, uint32_t num) {
  if (!ptr) return false;
  for (uint32_t i = 0; i < num; ++i)
    if (ptr[i + 1] < ptr[i] || (ptr[i + 1] > ptr[i] && !chunks)) return false;
  for (uint32_t i = 0; i < num; ++i)
    if (ptr[i + 1] > ptr[i] && !chunks[i]) return false;
  return true;
}

thread_local tribit_ctx* t_rank_ctx = nullptr;
thread_local int t_rank_status = TRIBIT_OK;

}  // namespace

int tribit_count_partial_diagonal(tribit_ctx* ctx, const uint64_t* const* chunks, const uint64_t* nnz_ptr,
                                  uint32_t num_ptrs, uint64_t num_nodes, uint64_t* triangles) {
  if (!ctx || !triangles || !chunks_ok(chunks, nnz_ptr, num_ptrs) || num_nodes == 0) return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    *triangles = count_diagonal(ctx, chunks, nnz_ptr, num_ptrs, num_nodes);
  });
}

int tribit_count_partial_off_diagonal(tribit_ctx* ctx, const uint64_t* const* cols, const uint64_t* nnz_cols_ptr,
                                      uint32_t num_ptrs_cols, const uint64_t* const* idxs,
                                      const uint64_t* nnz_idxs_ptr, uint32_t num_ptrs_idxs, uint64_t num_nodes,
                                      uint64_t* triangles) {
  if (!ctx || !triangles || !chunks_ok(cols, nnz_cols_ptr, num_ptrs_cols) ||
      !chunks_ok(idxs, nnz_idxs_ptr, num_ptrs_idxs) || num_nodes == 0)
    return TRIBIT_ERR_ARG;
  return guarded(ctx, [&] {
    TB_HIP(hipSetDevice(ctx->device));
    *triangles = count_off_diagonal(ctx, cols, nnz_cols_ptr, num_ptrs_cols, idxs, nnz_idxs_ptr, num_ptrs_idxs,
                                    num_nodes);
  });
}

int tribit_partial_last_info(tribit_partial_info* out) {
  if (!out) return TRIBIT_ERR_ARG;
  *out = t_partial_info;
  return TRIBIT_OK;
}

// Reference-named entry points (same signatures as main5_multi.cu:578-589 and
// main6_multi.cu:585-598): a thread-local context per calling thread; errors
// give UINT64_MAX and tribit_partial_last_status() (the reference exits).
void init_gpu_kernel5(int gpu_id) {
  if (t_rank_ctx) {
    tribit_shutdown(t_rank_ctx);
    t_rank_ctx = nullptr;
  }
  t_rank_status = tribit_init(gpu_id, &t_rank_ctx);
}

void init_gpu_kernel6(int gpu_id) { init_gpu_kernel5(gpu_id); }

uint64_t count_partial_triangles_diagonal(uint64_t** chunks, uint64_t* nnz_ptr, uint32_t num_ptrs,
                                          uint64_t num_nodes) {
  if (!t_rank_ctx) init_gpu_kernel5(0);
  uint64_t t = 0;
  t_rank_status = t_rank_ctx ? tribit_count_partial_diagonal(t_rank_ctx, chunks, nnz_ptr, num_ptrs, num_nodes, &t)
                             : t_rank_status;
  return t_rank_status == TRIBIT_OK ? t : UINT64_MAX;
}

uint64_t count_partial_triangles_off_diagonal(uint64_t** cols, uint64_t* nnz_cols_ptr, uint32_t num_ptrs_cols,
                                              uint64_t** idxs, uint64_t* nnz_idxs_ptr, uint32_t num_ptrs_idxs,
                                              uint64_t num_nodes) {
  if (!t_rank_ctx) init_gpu_kernel6(0);
  uint64_t t = 0;
  t_rank_status = t_rank_ctx ? tribit_count_partial_off_diagonal(t_rank_ctx, cols, nnz_cols_ptr, num_ptrs_cols, idxs,
                                                                 nnz_idxs_ptr, num_ptrs_idxs, num_nodes, &t)
                             : t_rank_status;
  return t_rank_status == TRIBIT_OK ? t : UINT64_MAX;
}

int tribit_partial_last_status(void) { return t_rank_status; }

}  // extern "C"
