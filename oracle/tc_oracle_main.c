/*
 * tc_oracle -- command-line front end of the CPU oracle (TEST INFRASTRUCTURE).
 *
 *   tc_oracle -i graph.mtx            [--threads T] [--ref-meta]
 *   tc_oracle --rmat S,EF,A,B,C,SEED  [--threads T]
 *   tc_oracle --er N,M,SEED           [--threads T]
 *
 * Prints n, m, W+ and the exact triangle count, with the CSR-build and
 * counting times measured separately (wall clock, seconds).
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double now(void) { return omp_get_wtime(); }

int main(int argc, char** argv) {
  const char* mtx = NULL;
  const char* rmat = NULL;
  const char* er = NULL;
  int threads = 0, ref_meta = 0;
  static struct option lo[] = {{"rmat", required_argument, 0, 'r'},
                               {"er", required_argument, 0, 'e'},
                               {"threads", required_argument, 0, 't'},
                               {"ref-meta", no_argument, 0, 'M'},
                               {0, 0, 0, 0}};
  int o;
  while ((o = getopt_long(argc, argv, "i:", lo, NULL)) != -1) {
    switch (o) {
      case 'i': mtx = optarg; break;
      case 'r': rmat = optarg; break;
      case 'e': er = optarg; break;
      case 't': threads = atoi(optarg); break;
      case 'M': ref_meta = 1; break;
      default: fprintf(stderr, "usage: %s -i f.mtx | --rmat S,EF,A,B,C,SEED | --er N,M,SEED\n", argv[0]); return 2;
    }
  }
  if (threads > 0) omp_set_num_threads(threads);
  uint64_t *keys = NULL, nkeys = 0, n = 0;
  double t0 = now();
  if (mtx) {
    int st = or_read_mtx(mtx, &keys, &nkeys, &n, NULL);
    if (st) {
      fprintf(stderr, "read error %d\n", st);
      return 1;
    }
  } else if (rmat) {
    unsigned s, ef;
    double a, b, c;
    unsigned long long seed;
    if (sscanf(rmat, "%u,%u,%lf,%lf,%lf,%llu", &s, &ef, &a, &b, &c, &seed) != 6) return 2;
    n = 1ull << s;
    nkeys = (uint64_t)ef << s;
    keys = (uint64_t*)malloc(nkeys * sizeof(uint64_t));
    or_gen_rmat(s, ef, a, b, c, seed, keys);
  } else if (er) {
    unsigned long long nn, mm, seed;
    if (sscanf(er, "%llu,%llu,%llu", &nn, &mm, &seed) != 3) return 2;
    n = nn;
    nkeys = mm;
    keys = (uint64_t*)malloc(nkeys * sizeof(uint64_t));
    or_gen_er(n, nkeys, seed, keys);
  } else {
    fprintf(stderr, "no input\n");
    return 2;
  }
  double t1 = now();
  uint64_t m = or_dedup(keys, nkeys);
  if (ref_meta) {
    or_ref_meta r;
    or_ref_pipeline(keys, m, n, &r);
    printf("ref_m %llu\nref_w_l %llu\nref_m_purged %llu\nref_w_u %llu\nref_r %llu\nref_t %llu\nref_blocks %llu\nref_triangles %llu\n",
           (unsigned long long)r.m, (unsigned long long)r.w_l, (unsigned long long)r.m_purged,
           (unsigned long long)r.w_u, (unsigned long long)r.r, (unsigned long long)r.t,
           (unsigned long long)r.blocks, (unsigned long long)r.triangles);
  }
  or_graph g;
  if (or_orient_csr(keys, m, n, &g)) return 1;
  double t2 = now();
  uint64_t wp = or_wplus(&g);
  double t3 = now();
  uint64_t tri = or_count(&g, threads);
  double t4 = now();
  printf("n %llu\nm %llu\nentries %llu\nwplus %llu\ntriangles %llu\n", (unsigned long long)n,
         (unsigned long long)m, (unsigned long long)nkeys, (unsigned long long)wp, (unsigned long long)tri);
  printf("t_input_s %.3f\nt_csr_s %.3f\nt_count_s %.3f\nthreads %d\n", t1 - t0, t2 - t1, t4 - t3,
         omp_get_max_threads());
  or_graph_free(&g);
  free(keys);
  return 0;
}
