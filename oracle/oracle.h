/* TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the trie matcher:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls into it.
 *
 * Two CPU restatements live here:
 *   oracle_count / oracle_enumerate  -- an independent depth-first
 *     backtracking matcher (its own static order, merge of sorted lists),
 *     used to pin exact counts and embedding sets;
 *   trie_cpu_run  -- an OpenMP port of the reference's level-by-level trie
 *     expansion (src/device_funcs.cu:267-407 driven as in
 *     src/host_funcs.cu:39-171), run with a given matching order; it reports
 *     per-level table sizes and the merge-model byte count, and is the CPU
 *     baseline that bench.py times.
 *   oracle_heuristic_order  -- a second, independent restatement of the
 *     score order of src/graph.cpp:62-93 / src/score.cpp:18-99.
 *
 * Graphs are CSR: off_out/adj_out (out-lists), off_in/adj_in (in-lists),
 * sig[2v] = in-degree, sig[2v+1] = out-degree.  A query is given as k and a
 * k*k 0/1 arc matrix qa[u*k+v] (1 <=> arc u->v), self-loops allowed.
 */
#ifndef CUTS_ORACLE_H
#define CUTS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t V;
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
} oracle_graph;

/* Number of maps f: Q -> D that are injective, map every arc u->v (u != v)
 * of Q to an arc of D and satisfy sig(f(u)) >= sig(u) componentwise.
 * nthreads <= 0 uses the OpenMP default. */
uint64_t oracle_count(const oracle_graph *d, uint32_t k, const uint8_t *qa, int nthreads);

/* Options of oracle_count_ex: explicit signature thresholds per query
 * vertex (NULL = the query's own degrees) and a restriction of the image
 * of query vertex root_q (-1 = none) to root_begin + i*root_stride <
 * root_end. */
typedef struct {
  const uint32_t *thr_in, *thr_out;
  int32_t root_q;
  uint32_t root_begin, root_end, root_stride;
} oracle_opts;
uint64_t oracle_count_ex(const oracle_graph *d, uint32_t k, const uint8_t *qa, const oracle_opts *o,
                         int nthreads);

/* Same set, enumerated: writes up to cap rows of k words (row[u] = f(u)),
 * returns the total number of embeddings (which may exceed cap). */
uint64_t oracle_enumerate(const oracle_graph *d, uint32_t k, const uint8_t *qa, uint32_t *rows,
                          uint64_t cap);
/* oracle_enumerate with oracle_count_ex's thresholds and root restriction. */
uint64_t oracle_enumerate_ex(const oracle_graph *d, uint32_t k, const uint8_t *qa, const oracle_opts *o,
                             uint32_t *rows, uint64_t cap);

/* Score order restated independently of the product's host code. Returns 0
 * on success, -1 if some step has no vertex scoring above (0,0,0). */
int oracle_heuristic_order(uint32_t k, const uint8_t *qa, uint32_t *order);

typedef struct {
  uint64_t count;          /* |T_{k-1}| */
  uint64_t level[32];      /* |T_l| for l < k (level[k-1] == count) */
  double bytes;            /* merge-model algorithmic bytes: 8 per data vertex
                              (signature scan); per prefix expanded, 8 + 4*len
                              for each distinct backward list (a symmetric
                              graph's in- and out-list are one list); 16 per
                              trie node of levels below k-1 (written, read) */
  double seconds;          /* wall time of the matching phase */
  int threads;
  int complete;            /* 0 if max_seconds stopped the run early */
} trie_cpu_result;

/* Level-by-level trie expansion with the given order (order[i] = query
 * vertex at position i).  sample_stride > 1 keeps only every stride-th
 * level-0 vertex (bounded CPU baseline).  max_rows bounds each level's
 * table (chunked depth-first over prefix groups). Returns 0 or -1. */
int trie_cpu_run(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order,
                 uint32_t sample_stride, uint64_t max_rows, int nthreads, trie_cpu_result *res);

/* As trie_cpu_run, but stops between row groups once max_seconds have
 * passed (0 = no limit); count/level then cover only the sub-tries that were
 * completed, and res->complete is 0. */
int trie_cpu_run_timed(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order,
                       uint32_t sample_stride, uint64_t max_rows, int nthreads, double max_seconds,
                       trie_cpu_result *res);
/* As trie_cpu_run_timed, with level 0 restricted to the data vertices
 * rb, rb+rs, ... below re (the matcher's root samples) before sampling. */
int trie_cpu_run_roots(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order, uint32_t rb,
                       uint32_t re, uint32_t rs, uint32_t sample_stride, uint64_t max_rows, int nthreads,
                       double max_seconds, trie_cpu_result *res);
/* omp_get_max_threads() of the oracle library (the CPU baseline's pool). */
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
