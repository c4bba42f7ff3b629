/*
 * oracle.c -- CPU oracle for exact triangle counting (TEST INFRASTRUCTURE ONLY).
 * See oracle.h for what pins each function. Plain C99 + OpenMP, integer only.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <ctype.h>
#include <math.h>
#include <errno.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void or_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Ingest: restates read_mtx, tribit.cu:228-269.                             */
/* ------------------------------------------------------------------------ */

static int parse_u64(const char** s, uint64_t* out) {
  const char* p = *s;
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  if (!isdigit((unsigned char)*p)) return 0;
  uint64_t v = 0;
  while (isdigit((unsigned char)*p)) {
    uint64_t d = (uint64_t)(*p - '0');
    if (v > (UINT64_MAX - d) / 10) return 0;
    v = v * 10 + d;
    ++p;
  }
  *out = v;
  *s = p;
  return 1;
}

int or_read_mtx(const char* path, uint64_t** keys_out, uint64_t* nkeys_out, uint64_t* n_out,
                uint64_t* n_entries_out) {
  FILE* f = fopen(path, "rb");
  if (!f) return OR_ERR_IO;
  char* line = NULL;
  size_t cap = 0;
  ssize_t len;
  /* tribit.cu:236: the first line (banner) is skipped unconditionally. */
  if ((len = getline(&line, &cap, f)) < 0) {
    free(line);
    fclose(f);
    return OR_ERR_FORMAT;
  }
  /* tribit.cu:238-239: '%' lines skipped; the first other line is the size line. */
  int have_size = 0;
  uint64_t nr = 0, nc = 0, nnz = 0;
  while ((len = getline(&line, &cap, f)) >= 0) {
    if (line[0] == '%') continue;
    const char* p = line;
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
    if (*p == '\0') continue; /* divergence: blank line skipped (reference: UB) */
    if (!parse_u64(&p, &nr) || !parse_u64(&p, &nc) || !parse_u64(&p, &nnz)) {
      free(line);
      fclose(f);
      return OR_ERR_FORMAT;
    }
    have_size = 1;
    break;
  }
  if (!have_size) {
    free(line);
    fclose(f);
    return OR_ERR_FORMAT;
  }
  /* tribit.cu:241-245: non-square -> error. */
  if (nr != nc) {
    free(line);
    fclose(f);
    return OR_ERR_SHAPE;
  }
  if (nr > 0xFFFFFFFFull) {
    free(line);
    fclose(f);
    return OR_ERR_RANGE;
  }
  uint64_t cap_keys = nnz ? nnz : 16;
  uint64_t* keys = (uint64_t*)malloc(cap_keys * sizeof(uint64_t));
  if (!keys) {
    free(line);
    fclose(f);
    return OR_ERR_NOMEM;
  }
  uint64_t nk = 0, ne = 0;
  int st = OR_OK;
  /* tribit.cu:252-268: entries "r c [values...]", 1-based, self-loops dropped,
   * packed as (min<<32)|max. Blank lines are skipped; the loop ends at the
   * first other line that does not start with two integers (as
   * `file >> r >> c` skips whitespace and fails there). */
  while ((len = getline(&line, &cap, f)) >= 0) {
    const char* p = line;
    const char* q = line;
    while (*q == ' ' || *q == '\t' || *q == '\r') ++q;
    if (*q == '\n' || *q == '\0') continue; /* blank line: `file >> r >> c` skips whitespace */
    uint64_t r, c;
    if (!parse_u64(&p, &r) || !parse_u64(&p, &c)) break;
    if (r < 1 || c < 1 || r > nr || c > nr) {
      st = OR_ERR_RANGE;
      break;
    }
    ++ne;
    r -= 1;
    c -= 1;
    if (r == c) continue;
    uint64_t lo = r < c ? r : c, hi = r < c ? c : r;
    if (nk == cap_keys) {
      cap_keys *= 2;
      uint64_t* nkp = (uint64_t*)realloc(keys, cap_keys * sizeof(uint64_t));
      if (!nkp) {
        st = OR_ERR_NOMEM;
        break;
      }
      keys = nkp;
    }
    keys[nk++] = (lo << 32) | hi;
  }
  free(line);
  fclose(f);
  if (st != OR_OK) {
    free(keys);
    return st;
  }
  *keys_out = keys;
  *nkeys_out = nk;
  *n_out = nr;
  if (n_entries_out) *n_entries_out = ne;
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Counter-based generators. The spec (shared with the GPU generators):      */
/*   mix64      = SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)     */
/*   stream_key = mix64(seed ^ mix64(stream))                                */
/*   rand(k, i) = mix64(k + i * 0xD1342543DE82EF95)                          */
/* ------------------------------------------------------------------------ */

uint64_t or_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
uint64_t or_stream_key(uint64_t seed, uint64_t stream) { return or_mix64(seed ^ or_mix64(stream)); }
uint64_t or_rand(uint64_t key, uint64_t i) { return or_mix64(key + i * 0xD1342543DE82EF95ull); }

void or_rmat_thresholds(double a, double b, double c, uint64_t t[3]) {
  const double two32 = 4294967296.0;
  double s[3] = {a, a + b, a + b + c};
  for (int i = 0; i < 3; ++i) {
    double v = s[i] * two32;
    if (v < 0) v = 0;
    if (v > 4294967295.0) v = 4294967295.0;
    t[i] = (uint64_t)v;
  }
}

/* Stream ids of the generators. */
#define OR_STREAM_RMAT 0x524D4154ull /* "RMAT" */
#define OR_STREAM_ER 0x45524E59ull   /* "ERNY" */

/* R-MAT (Chakrabarti, Zhan & Faloutsos, SDM 2004), no noise, no permutation.
 * Edge e, level l (0 = most significant bit): r = 32-bit half (l&1) of
 * rand(key, e*16 + l/2); quadrant (0,0) if r<t0, (0,1) if r<t1, (1,0) if r<t2,
 * else (1,1), as (src bit, dst bit). */
int or_gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, uint64_t* keys) {
  if (scale < 1 || scale > 31) return OR_ERR_ARG;
  uint64_t t[3];
  or_rmat_thresholds(a, b, c, t);
  const uint64_t key = or_stream_key(seed, OR_STREAM_RMAT);
  const uint64_t ne = (uint64_t)ef << scale;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)ne; ++e) {
    uint64_t src = 0, dst = 0, w = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) w = or_rand(key, (uint64_t)e * 16 + (l >> 1));
      uint64_t r = (l & 1) ? (w >> 32) : (w & 0xFFFFFFFFull);
      uint64_t sb, db;
      if (r < t[0]) {
        sb = 0;
        db = 0;
      } else if (r < t[1]) {
        sb = 0;
        db = 1;
      } else if (r < t[2]) {
        sb = 1;
        db = 0;
      } else {
        sb = 1;
        db = 1;
      }
      src |= sb << (scale - 1 - l);
      dst |= db << (scale - 1 - l);
    }
    if (src == dst)
      keys[e] = OR_SENTINEL;
    else
      keys[e] = src < dst ? ((src << 32) | dst) : ((dst << 32) | src);
  }
  return OR_OK;
}

/* G(n, M) with replacement: pair e = (lo32(w)*n >> 32, hi32(w)*n >> 32), w = rand(key, e). */
int or_gen_er(uint64_t n, uint64_t nedges, uint64_t seed, uint64_t* keys) {
  if (n < 1 || n > 0xFFFFFFFFull) return OR_ERR_ARG;
  const uint64_t key = or_stream_key(seed, OR_STREAM_ER);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)nedges; ++e) {
    uint64_t w = or_rand(key, (uint64_t)e);
    uint64_t u = ((w & 0xFFFFFFFFull) * n) >> 32;
    uint64_t v = ((w >> 32) * n) >> 32;
    if (u == v)
      keys[e] = OR_SENTINEL;
    else
      keys[e] = u < v ? ((u << 32) | v) : ((v << 32) | u);
  }
  return OR_OK;
}

#define OR_STREAM_CL 0x43484C55ull     /* "CHLU" */
#define OR_STREAM_CLP 0x43484C50ull    /* "CHLP" */
#define OR_STREAM_RGG 0x52474732ull    /* "RGG2" */
#define OR_STREAM_SBMI 0x53424D49ull   /* "SBMI" */
#define OR_STREAM_SBMO 0x53424D4Full   /* "SBMO" */

static double or_unit(uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); }

uint64_t or_chung_lu_samples(uint64_t n, double avg_deg) { return (uint64_t)((double)n * avg_deg / 2.0 + 0.5); }

int or_chung_lu_cdf(uint64_t n, double gamma, uint64_t* cdf) {
  if (n < 1 || gamma <= 1.0) return OR_ERR_ARG;
  const double e = -1.0 / (gamma - 1.0);
  double* cum = (double*)malloc(n * sizeof(double));
  if (!cum) return OR_ERR_NOMEM;
  double acc = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    acc += pow((double)(i + 1), e);
    cum[i] = acc;
  }
  const double scale = 4611686018427387904.0 / acc; /* 2^62 / total */
  for (uint64_t i = 0; i + 1 < n; ++i) {
    double v = cum[i] * scale;
    if (v > 4611686018427387904.0) v = 4611686018427387904.0;
    cdf[i] = (uint64_t)v;
  }
  cdf[n - 1] = 1ull << 62;
  free(cum);
  return OR_OK;
}

static uint64_t cdf_pick(const uint64_t* cdf, uint64_t n, uint64_t r) {
  uint64_t lo = 0, hi = n; /* first i with cdf[i] > r */
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] <= r)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n ? lo : n - 1;
}

int or_gen_chung_lu(uint64_t n, double gamma, double avg_deg, uint64_t seed, uint64_t* keys) {
  if (n < 2 || n > 0xFFFFFFFEull) return OR_ERR_ARG;
  uint64_t* cdf = (uint64_t*)malloc(n * sizeof(uint64_t));
  if (!cdf) return OR_ERR_NOMEM;
  int st = or_chung_lu_cdf(n, gamma, cdf);
  if (st) {
    free(cdf);
    return st;
  }
  const uint64_t M = or_chung_lu_samples(n, avg_deg);
  const uint64_t key = or_stream_key(seed, OR_STREAM_CL);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)M; ++e) {
    uint64_t u = cdf_pick(cdf, n, or_rand(key, 2 * (uint64_t)e) >> 2);
    uint64_t v = cdf_pick(cdf, n, or_rand(key, 2 * (uint64_t)e + 1) >> 2);
    keys[e] = (u == v) ? OR_SENTINEL : (u < v ? ((u << 32) | v) : ((v << 32) | u));
  }
  free(cdf);
  return OR_OK;
}

void or_chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg) {
  const uint64_t key = or_stream_key(seed, OR_STREAM_CLP);
  const double u0 = or_unit(or_rand(key, 3 * index));
  const double u1 = or_unit(or_rand(key, 3 * index + 1));
  const double u2 = or_unit(or_rand(key, 3 * index + 2));
  *gamma = 2.1 + 0.7 * u0;
  *n = (uint64_t)(1024.0 * exp2(8.0 * u1) + 0.5);
  *avg_deg = 8.0 + 56.0 * u2;
}

uint64_t or_gen_rgg(uint64_t n, double avg_deg, uint64_t seed, uint64_t* keys, uint64_t cap) {
  const double two30 = 1073741824.0;
  const uint64_t R = (uint64_t)(sqrt(avg_deg / (3.14159265358979323846 * (double)n)) * two30);
  if (n < 2 || R < 1) return 0;
  const uint64_t C = ((1ull << 30) + R - 1) / R; /* cells per side */
  const uint64_t key = or_stream_key(seed, OR_STREAM_RGG);
  uint32_t* px = (uint32_t*)malloc(n * 4);
  uint32_t* py = (uint32_t*)malloc(n * 4);
  uint64_t* ord = (uint64_t*)malloc(n * 8);
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t w = or_rand(key, i);
    px[i] = (uint32_t)(w & 0x3FFFFFFFull);
    py[i] = (uint32_t)((w >> 32) & 0x3FFFFFFFull);
    uint64_t cell = (py[i] / R) * C + (px[i] / R);
    ord[i] = (cell << 32) | i; /* n < 2^32 */
  }
  or_sort_u64(ord, n);
  /* sorted position = new vertex id; cell_start over sorted order */
  uint64_t ncell = C * C;
  uint64_t* cstart = (uint64_t*)calloc(ncell + 1, 8);
  uint32_t* sx = (uint32_t*)malloc(n * 4);
  uint32_t* sy = (uint32_t*)malloc(n * 4);
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t i = ord[k] & 0xFFFFFFFFull;
    sx[k] = px[i];
    sy[k] = py[i];
    cstart[(ord[k] >> 32) + 1]++;
  }
  for (uint64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
  const uint64_t R2 = R * R;
  uint64_t out = 0;
  int overflow = 0;
  for (uint64_t k = 0; k < n && !overflow; ++k) {
    int64_t cx = (int64_t)(sx[k] / R), cy = (int64_t)(sy[k] / R);
    for (int64_t dy = -1; dy <= 1; ++dy)
      for (int64_t dx = -1; dx <= 1; ++dx) {
        int64_t nx = cx + dx, ny = cy + dy;
        if (nx < 0 || ny < 0 || nx >= (int64_t)C || ny >= (int64_t)C) continue;
        uint64_t c = (uint64_t)ny * C + (uint64_t)nx;
        for (uint64_t j = cstart[c]; j < cstart[c + 1]; ++j) {
          if (j <= k) continue;
          uint64_t ddx = sx[k] > sx[j] ? sx[k] - sx[j] : sx[j] - sx[k];
          uint64_t ddy = sy[k] > sy[j] ? sy[k] - sy[j] : sy[j] - sy[k];
          if (ddx * ddx + ddy * ddy <= R2) {
            if (out == cap) {
              overflow = 1;
              break;
            }
            keys[out++] = (k << 32) | j;
          }
        }
      }
  }
  free(px);
  free(py);
  free(ord);
  free(cstart);
  free(sx);
  free(sy);
  return overflow ? UINT64_MAX : out;
}

static uint64_t or_sbm_out_samples(uint64_t blocks, uint64_t size, double p_out) {
  const double n = (double)(blocks * size);
  const double cross = n * (n - 1.0) / 2.0 - (double)blocks * (double)size * ((double)size - 1.0) / 2.0;
  return (uint64_t)(p_out * cross + 0.5);
}

uint64_t or_sbm_slots(uint64_t blocks, uint64_t size, double p_out) {
  return blocks * size * size + or_sbm_out_samples(blocks, size, p_out);
}

static uint64_t or_prob_threshold(double p) {
  if (p <= 0.0) return 0;
  if (p >= 1.0) return UINT64_MAX;
  return (uint64_t)(p * 18446744073709551616.0);
}

int or_gen_sbm(uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed, uint64_t* keys) {
  const uint64_t n = blocks * size;
  if (blocks < 1 || size < 1 || n > 0xFFFFFFFEull) return OR_ERR_ARG;
  const uint64_t kin = or_stream_key(seed, OR_STREAM_SBMI), kout = or_stream_key(seed, OR_STREAM_SBMO);
  const uint64_t tin = or_prob_threshold(p_in);
  const uint64_t ss = size * size;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)(blocks * ss); ++t) {
    const uint64_t b = (uint64_t)t / ss, r = (uint64_t)t % ss, i = r / size, j = r % size;
    uint64_t k = OR_SENTINEL;
    if (i < j && or_rand(kin, (uint64_t)t) < tin) k = ((b * size + i) << 32) | (b * size + j);
    keys[t] = k;
  }
  const uint64_t mo = or_sbm_out_samples(blocks, size, p_out);
  uint64_t* ko = keys + blocks * ss;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)mo; ++e) {
    const uint64_t w = or_rand(kout, (uint64_t)e);
    const uint64_t u = ((w & 0xFFFFFFFFull) * n) >> 32, v = ((w >> 32) * n) >> 32;
    uint64_t k = OR_SENTINEL;
    if (u / size != v / size) k = u < v ? ((u << 32) | v) : ((v << 32) | u);
    ko[e] = k;
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Sort + dedup: restates radix_sort + unique, tribit.cu:666-667.            */
/* ------------------------------------------------------------------------ */

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y);
}

/* Parallel LSD radix sort, 8-bit digits; digit passes on which every key
 * agrees are skipped. */
void or_sort_u64(uint64_t* keys, uint64_t nkeys) {
  if (nkeys < 2) return;
  uint64_t* tmp = (uint64_t*)malloc(nkeys * sizeof(uint64_t));
  if (!tmp) {
    qsort(keys, nkeys, sizeof(uint64_t), cmp_u64);
    return;
  }
  int nt = omp_get_max_threads();
  if (nt > 64) nt = 64;
  uint64_t* hist = (uint64_t*)calloc((size_t)nt * 256, sizeof(uint64_t));
  uint64_t* src = keys;
  uint64_t* dst = tmp;
  for (int shift = 0; shift < 64; shift += 8) {
    memset(hist, 0, (size_t)nt * 256 * sizeof(uint64_t));
#pragma omp parallel num_threads(nt)
    {
      int t = omp_get_thread_num();
      int T = omp_get_num_threads();
      uint64_t lo = nkeys * (uint64_t)t / (uint64_t)T, hi = nkeys * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t* h = hist + (size_t)t * 256;
      for (uint64_t i = lo; i < hi; ++i) h[(src[i] >> shift) & 255]++;
    }
    /* skip a pass where all keys share the digit */
    int skip = 0;
    for (int d = 0; d < 256; ++d) {
      uint64_t tot = 0;
      for (int t = 0; t < nt; ++t) tot += hist[(size_t)t * 256 + d];
      if (tot == nkeys) skip = 1;
      if (tot) break;
    }
    if (skip) continue;
    uint64_t run = 0;
    for (int d = 0; d < 256; ++d)
      for (int t = 0; t < nt; ++t) {
        uint64_t c = hist[(size_t)t * 256 + d];
        hist[(size_t)t * 256 + d] = run;
        run += c;
      }
#pragma omp parallel num_threads(nt)
    {
      int t = omp_get_thread_num();
      int T = omp_get_num_threads();
      uint64_t lo = nkeys * (uint64_t)t / (uint64_t)T, hi = nkeys * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t* h = hist + (size_t)t * 256;
      for (uint64_t i = lo; i < hi; ++i) dst[h[(src[i] >> shift) & 255]++] = src[i];
    }
    uint64_t* sw = src;
    src = dst;
    dst = sw;
  }
  if (src != keys) memcpy(keys, src, nkeys * sizeof(uint64_t));
  free(hist);
  free(tmp);
}

uint64_t or_dedup(uint64_t* keys, uint64_t nkeys) {
  or_sort_u64(keys, nkeys);
  uint64_t m = 0;
  for (uint64_t i = 0; i < nkeys; ++i) {
    if (keys[i] == OR_SENTINEL) break;
    if (m == 0 || keys[m - 1] != keys[i]) keys[m++] = keys[i];
  }
  return m;
}

/* ------------------------------------------------------------------------ */
/* Degree orientation + CSR.                                                 */
/* ------------------------------------------------------------------------ */

void or_graph_free(or_graph* g) {
  if (!g) return;
  free(g->row_ptr);
  free(g->col);
  free(g->deg);
  memset(g, 0, sizeof(*g));
}

static int orient_csr(const uint64_t* ukeys, uint64_t m, uint64_t n, or_graph* g, int by_id) {
  memset(g, 0, sizeof(*g));
  g->n = n;
  g->m = m;
  g->row_ptr = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  g->col = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
  g->deg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint64_t* okeys = (uint64_t*)malloc((m ? m : 1) * sizeof(uint64_t));
  if (!g->row_ptr || !g->col || !g->deg || !okeys) {
    free(okeys);
    or_graph_free(g);
    return OR_ERR_NOMEM;
  }
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t a = ukeys[i] >> 32, b = ukeys[i] & 0xFFFFFFFFull;
    if (a >= n || b >= n || a >= b) {
      free(okeys);
      or_graph_free(g);
      return OR_ERR_RANGE;
    }
    g->deg[a]++;
    g->deg[b]++;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    uint64_t a = ukeys[i] >> 32, b = ukeys[i] & 0xFFFFFFFFull;
    /* a < b: a->b iff (deg a, a) < (deg b, b) iff deg a <= deg b; by_id: a->b */
    okeys[i] = (by_id || g->deg[a] <= g->deg[b]) ? ((a << 32) | b) : ((b << 32) | a);
  }
  or_sort_u64(okeys, m);
  for (uint64_t i = 0; i < m; ++i) {
    g->row_ptr[(okeys[i] >> 32) + 1]++;
    g->col[i] = (uint32_t)(okeys[i] & 0xFFFFFFFFull);
  }
  for (uint64_t v = 0; v < n; ++v) g->row_ptr[v + 1] += g->row_ptr[v];
  free(okeys);
  return OR_OK;
}

int or_orient_csr(const uint64_t* ukeys, uint64_t m, uint64_t n, or_graph* g) { return orient_csr(ukeys, m, n, g, 0); }

int or_id_csr(const uint64_t* ukeys, uint64_t m, uint64_t n, or_graph* g) { return orient_csr(ukeys, m, n, g, 1); }

int or_degree_labels(const uint64_t* ukeys, uint64_t m, uint64_t n, uint32_t* old_of_new, uint32_t* new_of_old) {
  /* counting sort of the vertices by degree, stable in id: new id = rank of (deg, id) */
  uint32_t* deg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint64_t* start = (uint64_t*)calloc(n + 1, sizeof(uint64_t));
  if (!deg || !start) {
    free(deg);
    free(start);
    return OR_ERR_NOMEM;
  }
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t a = ukeys[i] >> 32, b = ukeys[i] & 0xFFFFFFFFull;
    if (a >= n || b >= n || a >= b) {
      free(deg);
      free(start);
      return OR_ERR_RANGE;
    }
    deg[a]++;
    deg[b]++;
  }
  for (uint64_t v = 0; v < n; ++v) start[deg[v] + 1]++; /* deg <= n-1 */
  for (uint64_t d = 0; d < n; ++d) start[d + 1] += start[d];
  for (uint64_t v = 0; v < n; ++v) {
    uint64_t r = start[deg[v]]++;
    old_of_new[r] = (uint32_t)v;
    if (new_of_old) new_of_old[v] = (uint32_t)r;
  }
  free(deg);
  free(start);
  return OR_OK;
}

int or_octave_labels(const uint64_t* ukeys, uint64_t m, uint64_t n, uint32_t* old_of_new, uint32_t* new_of_old) {
  /* counting sort of the vertices by octave class, stable in id: class = number
   * of bits of deg (0 for an isolated vertex), new id = rank of (class, id) */
  uint32_t* deg = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint64_t start[34] = {0};
  if (!deg) return OR_ERR_NOMEM;
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t a = ukeys[i] >> 32, b = ukeys[i] & 0xFFFFFFFFull;
    if (a >= n || b >= n || a >= b) {
      free(deg);
      return OR_ERR_RANGE;
    }
    deg[a]++;
    deg[b]++;
  }
  for (uint64_t v = 0; v < n; ++v) {
    uint32_t c = 0;
    while (c < 32 && (deg[v] >> c)) ++c;
    deg[v] = c;
    start[c + 1]++;
  }
  for (int c = 0; c < 33; ++c) start[c + 1] += start[c];
  for (uint64_t v = 0; v < n; ++v) {
    uint64_t r = start[deg[v]]++;
    old_of_new[r] = (uint32_t)v;
    if (new_of_old) new_of_old[v] = (uint32_t)r;
  }
  free(deg);
  return OR_OK;
}

void or_relabel_keys(const uint64_t* ukeys, uint64_t m, const uint32_t* new_of_old, uint64_t* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)m; ++i) {
    uint64_t a = new_of_old[ukeys[i] >> 32], b = new_of_old[ukeys[i] & 0xFFFFFFFFull];
    out[i] = a < b ? ((a << 32) | b) : ((b << 32) | a);
  }
}

uint64_t or_wplus(const or_graph* g) {
  uint64_t w = 0;
#pragma omp parallel for reduction(+ : w) schedule(dynamic, 4096)
  for (int64_t u = 0; u < (int64_t)g->n; ++u)
    for (uint64_t k = g->row_ptr[u]; k < g->row_ptr[u + 1]; ++k) {
      uint32_t v = g->col[k];
      w += g->row_ptr[v + 1] - g->row_ptr[v];
    }
  return w;
}

uint64_t or_tiles(const or_graph* g, int width, uint64_t* tile_ptr, uint32_t* blk, uint64_t* mask) {
  if (width != 32 && width != 64) return 0;
  const int sh = width == 32 ? 5 : 6;
  uint64_t w = 0;
  for (uint64_t u = 0; u < g->n; ++u) {
    if (tile_ptr) tile_ptr[u] = w;
    uint64_t prev = UINT64_MAX;
    for (uint64_t k = g->row_ptr[u]; k < g->row_ptr[u + 1]; ++k) {
      uint64_t c = g->col[k];
      uint64_t b = c >> sh;
      if (b != prev) {
        if (blk) blk[w] = (uint32_t)b;
        if (mask) mask[w] = 0;
        prev = b;
        ++w;
      }
      if (mask) mask[w - 1] |= 1ull << (c & (uint64_t)(width - 1));
    }
  }
  if (tile_ptr) tile_ptr[g->n] = w;
  return w;
}

/* ------------------------------------------------------------------------ */
/* Counting.                                                                 */
/* ------------------------------------------------------------------------ */

static uint64_t merge_count(const uint32_t* a, uint64_t na, const uint32_t* b, uint64_t nb) {
  uint64_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j])
      ++i;
    else if (b[j] < a[i])
      ++j;
    else {
      ++c;
      ++i;
      ++j;
    }
  }
  return c;
}

uint64_t or_count_rows(const or_graph* g, uint64_t r0, uint64_t r1, int threads) {
  if (r1 > g->n) r1 = g->n;
  uint64_t total = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64) num_threads(threads)
  for (int64_t u = (int64_t)r0; u < (int64_t)r1; ++u) {
    const uint32_t* nu = g->col + g->row_ptr[u];
    uint64_t du = g->row_ptr[u + 1] - g->row_ptr[u];
    if (du < 2) continue;
    for (uint64_t k = 0; k < du; ++k) {
      uint32_t v = nu[k];
      total += merge_count(nu, du, g->col + g->row_ptr[v], g->row_ptr[v + 1] - g->row_ptr[v]);
    }
  }
  return total;
}

uint64_t or_count(const or_graph* g, int threads) { return or_count_rows(g, 0, g->n, threads); }

uint64_t or_count_strided(const or_graph* g, uint64_t start, uint64_t stride, int threads) {
  if (stride == 0) stride = 1;
  const int64_t cnt = start < g->n ? (int64_t)((g->n - start + stride - 1) / stride) : 0;
  uint64_t total = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64) num_threads(threads)
  for (int64_t i = 0; i < cnt; ++i) {
    const uint64_t u = start + (uint64_t)i * stride;
    const uint32_t* nu = g->col + g->row_ptr[u];
    uint64_t du = g->row_ptr[u + 1] - g->row_ptr[u];
    if (du < 2) continue;
    for (uint64_t k = 0; k < du; ++k) {
      uint32_t v = nu[k];
      total += merge_count(nu, du, g->col + g->row_ptr[v], g->row_ptr[v + 1] - g->row_ptr[v]);
    }
  }
  return total;
}

uint64_t or_count_bruteforce(uint64_t n, const uint64_t* keys, uint64_t nkeys) {
  if (n > 2048) return UINT64_MAX;
  uint8_t* adj = (uint8_t*)calloc((n * n) > 0 ? n * n : 1, 1);
  for (uint64_t i = 0; i < nkeys; ++i) {
    if (keys[i] == OR_SENTINEL) continue;
    uint64_t a = keys[i] >> 32, b = keys[i] & 0xFFFFFFFFull;
    if (a >= n || b >= n || a == b) continue;
    adj[a * n + b] = adj[b * n + a] = 1;
  }
  uint64_t t = 0;
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t j = i + 1; j < n; ++j) {
      if (!adj[i * n + j]) continue;
      for (uint64_t k = j + 1; k < n; ++k) t += adj[i * n + k] & adj[j * n + k];
    }
  free(adj);
  return t;
}

/* ------------------------------------------------------------------------ */
/* Restatement of the reference pipeline, tribit.cu:656-786, on              */
/* deduplicated keys sorted by (min, max). Single-threaded; for fixtures.    */
/* ------------------------------------------------------------------------ */

static uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

int or_ref_pipeline(const uint64_t* ukeys, uint64_t m, uint64_t n, or_ref_meta* out) {
  memset(out, 0, sizeof(*out));
  out->m = m;
  const uint64_t N = 32 * ((n + 31) / 32); /* tribit.cu:246 */
  const uint64_t NB = N / 32;
  if (m == 0) return OR_OK;
  /* ---- L: key = max + (min/32)*N, bit 1<<(31-min%32) (tribit.cu:317-324, 588-594),
   * sorted pairs then OR-reduced (tribit.cu:679-682). */
  uint64_t* kl = (uint64_t*)malloc(m * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t c = ukeys[i] >> 32, r = ukeys[i] & 0xFFFFFFFFull;
    kl[i] = r + (c / 32) * N;
  }
  /* the masks depend only on (key, min%32); sort (key<<5 | min%32) pairs */
  uint64_t* kp = (uint64_t*)malloc(m * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) kp[i] = (kl[i] << 5) | ((ukeys[i] >> 32) & 31);
  or_sort_u64(kp, m);
  uint64_t* lkey = (uint64_t*)malloc(m * sizeof(uint64_t));
  uint32_t* lmask = (uint32_t*)malloc(m * sizeof(uint32_t));
  uint64_t wl = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t key = kp[i] >> 5;
    uint32_t bit = 1u << (31 - (uint32_t)(kp[i] & 31));
    if (wl == 0 || lkey[wl - 1] != key) {
      lkey[wl] = key;
      lmask[wl] = 0;
      ++wl;
    }
    lmask[wl - 1] |= bit;
  }
  out->w_l = wl;
  /* row_id_l = key % N (tribit.cu:686); set_ptr_l = histogram of key/N + scan (:688-695) */
  uint32_t* row_id_l = (uint32_t*)malloc(wl * sizeof(uint32_t));
  uint64_t* set_ptr_l = (uint64_t*)calloc(NB + 1, sizeof(uint64_t));
  for (uint64_t k = 0; k < wl; ++k) {
    row_id_l[k] = (uint32_t)(lkey[k] % N);
    set_ptr_l[lkey[k] / N + 1]++;
  }
  for (uint64_t b = 0; b < NB; ++b) set_ptr_l[b + 1] += set_ptr_l[b];
  /* ---- purge (tribit.cu:604-615, 698-707): keep an edge iff a sorted
   * neighbour shares its min. With m == 1 the reference reads past the end;
   * here the lone edge is dropped. */
  uint64_t* pk = (uint64_t*)malloc(m * sizeof(uint64_t));
  uint64_t mp = 0;
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t c = ukeys[i] >> 32;
    int keep = (i > 0 && (ukeys[i - 1] >> 32) == c) || (i + 1 < m && (ukeys[i + 1] >> 32) == c);
    if (keep) pk[mp++] = ukeys[i];
  }
  out->m_purged = mp;
  /* ---- U: key = min*(N/32) + max/32, bit 1<<(31-max%32) (tribit.cu:326-333, 596-602);
   * keys are already sorted because pk is sorted by (min,max) (:722 sorts sorted data). */
  uint64_t* ukey = (uint64_t*)malloc((mp ? mp : 1) * sizeof(uint64_t));
  uint32_t* umask = (uint32_t*)malloc((mp ? mp : 1) * sizeof(uint32_t));
  uint64_t wu = 0;
  for (uint64_t i = 0; i < mp; ++i) {
    uint64_t c = pk[i] >> 32, r = pk[i] & 0xFFFFFFFFull;
    uint64_t key = c * NB + r / 32;
    uint32_t bit = 1u << (31 - (uint32_t)(r % 32));
    if (wu == 0 || ukey[wu - 1] != key) {
      ukey[wu] = key;
      umask[wu] = 0;
      ++wu;
    }
    umask[wu - 1] |= bit;
  }
  out->w_u = wu;
  /* col_id_u = key % (N/32) (:729); RLE of key/(N/32) -> set_ptr_u[R+1] (:734-739) */
  uint32_t* col_id_u = (uint32_t*)malloc((wu ? wu : 1) * sizeof(uint32_t));
  uint64_t* set_ptr_u = (uint64_t*)malloc((wu + 1) * sizeof(uint64_t));
  uint64_t R = 0;
  for (uint64_t k = 0; k < wu; ++k) {
    col_id_u[k] = (uint32_t)(ukey[k] % NB);
    if (k == 0 || ukey[k - 1] / NB != ukey[k] / NB) set_ptr_u[R++] = k;
  }
  set_ptr_u[R] = wu;
  out->r = R;
  /* d_idxs = max endpoints (:714-717); popc_long = exclusive popcount prefix (:742-744);
   * popc_array[i] = popc_long[set_ptr_u[i]] (:745-748) */
  uint64_t* popc_long = (uint64_t*)malloc((wu + 1) * sizeof(uint64_t));
  popc_long[0] = 0;
  for (uint64_t k = 0; k < wu; ++k) popc_long[k + 1] = popc_long[k] + popc32(umask[k]);
  uint64_t* popc_array = (uint64_t*)malloc((R + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i <= R; ++i) popc_array[i] = popc_long[set_ptr_u[i]];
  /* per word: nb_k = ceil((popc_array[i+1]-popc_long[k])/32) (:645-654); T = sum (:750-757) */
  uint64_t T = 0, tri = 0;
  uint64_t i = 0;
  for (uint64_t k = 0; k < wu; ++k) {
    while (set_ptr_u[i + 1] <= k) ++i;
    uint64_t rem = popc_array[i + 1] - popc_long[k];
    uint64_t nb = (rem + 31) / 32;
    T += nb;
    /* the nb warps of word k (kernel, tribit.cu:519-586) */
    uint64_t deg = popc_array[i + 1] - popc_array[i];
    uint64_t bdx = (deg + 31) / 32;
    uint64_t pad = (32 - (deg & 31)) & 31;
    uint64_t cb = col_id_u[k];
    for (uint64_t q = 1; q <= nb; ++q) {
      uint64_t bx = bdx - q;
      for (uint64_t lane = 0; lane < 32; ++lane) {
        uint64_t idxx = popc_array[i] + 32 * bx + lane - pad; /* may wrap below: checked */
        if (!(popc_array[i] <= idxx && idxx < popc_array[i + 1])) continue;
        uint32_t r2 = (uint32_t)(pk[idxx] & 0xFFFFFFFFull);
        /* exact-match search of r2 in L column block cb (:552-555) */
        uint64_t lo = set_ptr_l[cb], hi = set_ptr_l[cb + 1];
        uint32_t ml = 0;
        while (lo < hi) {
          uint64_t mid = lo + (hi - lo) / 2;
          if (row_id_l[mid] < r2)
            lo = mid + 1;
          else if (r2 < row_id_l[mid])
            hi = mid;
          else {
            ml = lmask[mid];
            break;
          }
        }
        tri += popc32(umask[k] & ml);
      }
    }
  }
  out->t = T;
  out->blocks = (T + 3) / 4;
  out->triangles = tri;
  free(kl);
  free(kp);
  free(lkey);
  free(lmask);
  free(row_id_l);
  free(set_ptr_l);
  free(pk);
  free(ukey);
  free(umask);
  free(col_id_u);
  free(set_ptr_u);
  free(popc_long);
  free(popc_array);
  return OR_OK;
}
