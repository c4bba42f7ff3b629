/*
 * oracle.h -- CPU oracle for exact triangle counting (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * check and bench.py's CPU-baseline leg load it. The HIP library
 * (tribit_amd/libtribit_hip.so) never links or calls anything here.
 *
 * Every function restates one piece of the reference's single-GPU path
 * (src/single_gpu/tribit.cu of sixt99/TriBIT) or a graph-theoretic
 * definition, and cites the lines it follows. Integer arithmetic only; counts
 * are uint64.
 *
 * Pinning (what checks each function independently of its own formula):
 *   or_read_mtx          hand-written .mtx fixtures with hand-derived results;
 *                        wiki-Vote shape (n=8297, 103,689 entries, m=100,762).
 *   or_gen_rmat, _er ... GPU generator bit-equality (two independent codes of
 *                        one spec) + distributional checks; the spec is ours.
 *   or_dedup             sortedness/uniqueness invariants; brute-force set.
 *   or_orient_csr        invariants (rank order, row sums, each undirected
 *                        edge once); brute force on tiny graphs.
 *   or_count             closed forms (K_n, wheels, windmills, cycles,
 *                        bipartite, K_{a,b,c}), O(n^3) brute force, networkx
 *                        golden counts, wiki-Vote = 608,389 (SNAP value).
 *   or_ref_meta/_count   restatement of tribit.cu:656-786; pinned by the
 *                        wiki-Vote metadata derived independently when the
 *                        reference was surveyed (W_L=70,956, m_purged=99,318,
 *                        W_U=54,880, R=2,964, T=113,795) and by equality of
 *                        its count with or_count.
 *   or_degree_labels     a permutation; degrees non-decreasing along new ids,
 *                        ties in input-id order; counts unchanged.
 *   or_octave_labels     equals a sort of (int.bit_length(deg), id) computed
 *                        in Python; counts unchanged.
 *   or_tiles             popcount(tile masks) == row lengths; decode of the
 *                        tiles == CSR rows.
 */
#ifndef TRIBIT_ORACLE_H
#define TRIBIT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared meaning with the HIP library's tribit_status). */
enum {
  OR_OK = 0,
  OR_ERR_IO = 1,       /* cannot open / read file */
  OR_ERR_FORMAT = 2,   /* malformed header or entry */
  OR_ERR_SHAPE = 3,    /* non-square matrix */
  OR_ERR_RANGE = 4,    /* vertex id out of [1, n] */
  OR_ERR_NOMEM = 5,
  OR_ERR_ARG = 6
};

#define OR_SENTINEL 0xFFFFFFFFFFFFFFFFull

/* ---- ingest ------------------------------------------------------------ */
/* Restates read_mtx, tribit.cu:228-269: banner line skipped, '%' lines skipped
 * up to the size line, square check, 1-based -> 0-based, value columns
 * ignored, self-loops dropped, each entry packed as (min<<32)|max.
 * Divergences on malformed input only (documented): blank lines before the
 * size line are skipped; an id outside [1,n] is an error (reference: UB);
 * reading stops at the first line that is not "int int ...", as the
 * reference's `file >> r >> c` loop does at a '%' line.
 * *keys is malloc'ed (caller frees with or_free). */
int or_read_mtx(const char* path, uint64_t** keys, uint64_t* nkeys, uint64_t* n, uint64_t* n_entries);

/* ---- generators (spec shared with the GPU generators) ------------------- */
uint64_t or_mix64(uint64_t z);
uint64_t or_stream_key(uint64_t seed, uint64_t stream);
uint64_t or_rand(uint64_t key, uint64_t i);
/* R-MAT thresholds: t[0]=floor(a*2^32), t[1]=floor((a+b)*2^32), t[2]=floor((a+b+c)*2^32) */
void or_rmat_thresholds(double a, double b, double c, uint64_t t[3]);
/* Writes ef<<scale packed keys; self-loops become OR_SENTINEL. */
int or_gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, uint64_t* keys);
/* Erdos-Renyi G(n, M): M = n*avg_deg/2 uniform pairs (then dedup). */
int or_gen_er(uint64_t n, uint64_t nedges, uint64_t seed, uint64_t* keys);

/* Chung-Lu by edge sampling: weights w_i = (i+1)^(-1/(gamma-1)); cdf_i =
 * floor(sum_{j<=i} w_j / sum w * 2^62) (u64, last = 2^62); M = round(n*avg/2)
 * samples, each endpoint = first i with cdf_i > (rand >> 2). */
uint64_t or_chung_lu_samples(uint64_t n, double avg_deg);
int or_chung_lu_cdf(uint64_t n, double gamma, uint64_t* cdf);
int or_gen_chung_lu(uint64_t n, double gamma, double avg_deg, uint64_t seed, uint64_t* keys);
/* Parameters of graph `index` of a Chung-Lu batch drawn from `seed`:
 * gamma ~ U[2.1, 2.8], n log-uniform in [1024, 262144], avg_deg ~ U[8, 64]. */
void or_chung_lu_batch_params(uint64_t seed, uint64_t index, double* gamma, uint64_t* n, double* avg_deg);
/* 2-D random geometric graph: points on a 2^30 x 2^30 integer grid
 * (x = low 30 bits, y = bits 32..61 of rand(key, i)); edge iff squared
 * distance <= R^2 with R = floor(sqrt(avg_deg / (pi n)) * 2^30). Vertex ids are
 * the points' ranks in (cell, index) order, cells of side R row-major.
 * Writes at most `cap` keys; returns the number written (or UINT64_MAX if cap
 * is too small). */
uint64_t or_gen_rgg(uint64_t n, double avg_deg, uint64_t seed, uint64_t* keys, uint64_t cap);
/* Stochastic block model: `blocks` blocks of `size` consecutive ids; pair
 * (i<j) in one block kept iff rand(key_in, b*size*size + i*size + j) <
 * floor(p_in*2^64); plus M_out = round(p_out * cross pairs) samples of uniform
 * pairs whose endpoints lie in different blocks. keys holds
 * blocks*size*size + M_out slots (self/duplicate/same-block slots become the
 * sentinel). */
uint64_t or_sbm_slots(uint64_t blocks, uint64_t size, double p_out);
int or_gen_sbm(uint64_t blocks, uint64_t size, double p_in, double p_out, uint64_t seed, uint64_t* keys);

/* ---- preprocessing ------------------------------------------------------ */
/* Sort ascending + unique + drop OR_SENTINEL, in place. Returns m.
 * Restates radix_sort + unique, tribit.cu:666-667 (CUB SortKeys/Unique). */
uint64_t or_dedup(uint64_t* keys, uint64_t nkeys);
void or_sort_u64(uint64_t* keys, uint64_t nkeys);

typedef struct {
  uint64_t n;         /* vertices */
  uint64_t m;         /* undirected simple edges == oriented edges */
  uint64_t* row_ptr;  /* n+1: oriented CSR row pointer */
  uint32_t* col;      /* m: out-neighbours, ascending id within a row */
  uint32_t* deg;      /* n: undirected degree */
} or_graph;

/* Degree orientation of deduplicated keys: edge {a<b} is a->b iff
 * deg(a) <= deg(b) (rank = (deg, id), ties by lower id), else b->a. */
int or_orient_csr(const uint64_t* ukeys, uint64_t m, uint64_t n, or_graph* g);
/* Id orientation: edge {a<b} is a->b (the orientation of octave-class ids,
 * whose order is not the degree order inside a class). */
int or_id_csr(const uint64_t* ukeys, uint64_t m, uint64_t n, or_graph* g);
void or_graph_free(or_graph* g);

/* Degree relabelling: new id = rank of (deg, id) (a counting sort by degree,
 * stable in id). old_of_new[n] and, if not NULL, new_of_old[n]. With these ids
 * the degree orientation above is the id order (a -> b iff a < b). */
int or_degree_labels(const uint64_t* ukeys, uint64_t m, uint64_t n, uint32_t* old_of_new, uint32_t* new_of_old);
/* Octave-class relabelling: new id = rank of (bit length of deg, id), a
 * counting sort over 33 classes, stable in id. No counterpart in the
 * reference: the library's block-preserving id order (TRIBIT_OPT_RELABEL 2). */
int or_octave_labels(const uint64_t* ukeys, uint64_t m, uint64_t n, uint32_t* old_of_new, uint32_t* new_of_old);
/* out[i] = (min<<32)|max of the relabelled pair of ukeys[i] (not re-sorted). */
void or_relabel_keys(const uint64_t* ukeys, uint64_t m, const uint32_t* new_of_old, uint64_t* out);

/* W+ = sum over oriented edges u->v of outdeg(v). */
uint64_t or_wplus(const or_graph* g);

/* Bit tiles of the oriented rows: one tile per (row, col/width) that holds at
 * least one neighbour. tile_ptr[n+1]; blk[k] = col/width; mask[k] bit
 * (col%width) set (LSB-first). width 32 or 64. Returns tile count; pass NULL
 * outputs to only count. */
uint64_t or_tiles(const or_graph* g, int width, uint64_t* tile_ptr, uint32_t* blk, uint64_t* mask);

/* ---- counting ------------------------------------------------------------ */
/* Edge iterator with sorted-merge intersection, OpenMP over rows. */
uint64_t or_count(const or_graph* g, int threads);
/* Same, restricted to rows [r0, r1) (bounded CPU baseline samples). */
uint64_t or_count_rows(const or_graph* g, uint64_t r0, uint64_t r1, int threads);
/* Same, over rows u = start, start+stride, ... (representative samples). */
uint64_t or_count_strided(const or_graph* g, uint64_t start, uint64_t stride, int threads);
/* O(n^3) brute force on an adjacency matrix; n <= 2048. */
uint64_t or_count_bruteforce(uint64_t n, const uint64_t* keys, uint64_t nkeys);

/* ---- restatement of the reference's own structures ----------------------- */
typedef struct {
  uint64_t m;            /* unique undirected edges */
  uint64_t w_l;          /* compressed_size_l, tribit.cu:682 */
  uint64_t m_purged;     /* after singleton purge, tribit.cu:704 */
  uint64_t w_u;          /* compressed_size_u, tribit.cu:725 */
  uint64_t r;            /* non_zero_rows, tribit.cu:734 */
  uint64_t t;            /* total_num_warps, tribit.cu:757 */
  uint64_t blocks;       /* total_num_blocks = ceil(T/4), tribit.cu:758 */
  uint64_t triangles;    /* sum of popc(U & L) as the kernel forms it */
} or_ref_meta;
/* Restates tribit.cu:656-786 on deduplicated id-oriented keys. */
int or_ref_pipeline(const uint64_t* ukeys, uint64_t m, uint64_t n, or_ref_meta* out);

void or_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
