/* TEST INFRASTRUCTURE ONLY -- the CPU oracle and CPU trie baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library; the product path never does.
 *
 * Parity pins (see tests/test_oracle.py): counts against networkx
 * DiGraphMatcher monomorphism counts (tests/golden/nx_counts.json), brute
 * force over all injective maps on tiny graphs, and the closed forms
 * n!/(n-k)! (any loop-free k-vertex query into the complete digraph),
 * trace(A^3) (triangles), sum d(d-1) (2-paths), sum d!/(d-s)! (s-stars).
 * The score-order restatement is pinned against the reference's own host
 * code compiled under oracle/_ref when that is built, and against
 * hand-derived orders.  The trie baseline's counts are pinned to
 * oracle_count; its per-level sizes are pinned to the oracle's own prefix
 * counts (level l of the trie == embeddings of the query restricted to the
 * first l+1 positions with the full-query signature filter).
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAXK 32
#define MAXC 64

typedef struct {
  uint32_t k;
  uint32_t ord[MAXK];
  uint32_t nc[MAXK];
  uint32_t cj[MAXK][MAXC]; /* earlier position */
  uint8_t cd[MAXK][MAXC];  /* 0: candidate in out(f(j)); 1: candidate in in(f(j)) */
  uint32_t qin[MAXK], qout[MAXK];
  int root_pos; /* position whose image is restricted, -1 for none */
  uint32_t rb, re, rs;
} oplan;

static void degrees(uint32_t k, const uint8_t *qa, uint32_t *din, uint32_t *dout) {
  for (uint32_t u = 0; u < k; ++u) {
    din[u] = dout[u] = 0;
    for (uint32_t v = 0; v < k; ++v) {
      dout[u] += qa[u * k + v] != 0;
      din[u] += qa[v * k + u] != 0;
    }
  }
}

/* Fills the constraint lists of a plan for a given order. */
static void plan_from_order(uint32_t k, const uint8_t *qa, const uint32_t *ord, oplan *p,
                            const oracle_opts *o) {
  uint32_t din[MAXK], dout[MAXK];
  degrees(k, qa, din, dout);
  if (o && o->thr_in) memcpy(din, o->thr_in, k * sizeof(uint32_t));
  if (o && o->thr_out) memcpy(dout, o->thr_out, k * sizeof(uint32_t));
  memset(p, 0, sizeof(*p));
  p->root_pos = -1;
  if (o && o->root_q >= 0) {
    for (uint32_t i = 0; i < k; ++i)
      if (ord[i] == (uint32_t)o->root_q) p->root_pos = (int)i;
    p->rb = o->root_begin;
    p->re = o->root_end;
    p->rs = o->root_stride ? o->root_stride : 1;
  }
  p->k = k;
  for (uint32_t i = 0; i < k; ++i) {
    uint32_t v = ord[i];
    p->ord[i] = v;
    p->qin[i] = din[v];
    p->qout[i] = dout[v];
    for (uint32_t j = 0; j < i; ++j) {
      uint32_t t = ord[j];
      if (qa[t * k + v]) { p->cj[i][p->nc[i]] = j; p->cd[i][p->nc[i]++] = 0; }
      if (qa[v * k + t]) { p->cj[i][p->nc[i]] = j; p->cd[i][p->nc[i]++] = 1; }
    }
  }
}

/* The oracle's own static order: most total degree first, then the vertex
 * with most arcs into the placed set (ties: larger degree, lower id). */
static void oracle_order(uint32_t k, const uint8_t *qa, uint32_t *ord) {
  uint32_t din[MAXK], dout[MAXK];
  char placed[MAXK] = {0};
  degrees(k, qa, din, dout);
  for (uint32_t i = 0; i < k; ++i) {
    int best = -1;
    long bs = -1, bd = -1;
    for (uint32_t v = 0; v < k; ++v) {
      if (placed[v]) continue;
      long s = 0;
      for (uint32_t j = 0; j < i; ++j) s += qa[ord[j] * k + v] + qa[v * k + ord[j]];
      long dg = din[v] + dout[v];
      if (s > bs || (s == bs && dg > bd)) { best = (int)v; bs = s; bd = dg; }
    }
    ord[i] = (uint32_t)best;
    placed[best] = 1;
  }
}

/* binary search membership in a[lo,hi) */
static inline int member(const uint32_t *a, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    uint32_t y = a[mid];
    if (y == x) return 1;
    if (y < x) lo = mid + 1;
    else hi = mid;
  }
  return 0;
}

static inline void list_of(const oracle_graph *d, uint8_t dir, uint32_t v, const uint32_t **a,
                           uint32_t *lo, uint32_t *hi) {
  if (dir == 0) { *a = d->adj_out; *lo = d->off_out[v]; *hi = d->off_out[v + 1]; }
  else { *a = d->adj_in; *lo = d->off_in[v]; *hi = d->off_in[v + 1]; }
}

/* Does candidate x pass position i given f[0..i)? (all constraints checked) */
static inline int accept(const oracle_graph *d, const oplan *p, uint32_t i, const uint32_t *f, uint32_t x,
                         int skip) {
  if (d->sig[2 * (size_t)x] < p->qin[i] || d->sig[2 * (size_t)x + 1] < p->qout[i]) return 0;
  if ((int)i == p->root_pos && (x < p->rb || x >= p->re || (x - p->rb) % p->rs)) return 0;
  for (uint32_t j = 0; j < i; ++j)
    if (f[j] == x) return 0;
  for (uint32_t c = 0; c < p->nc[i]; ++c) {
    if ((int)c == skip) continue;
    const uint32_t *a;
    uint32_t lo, hi;
    list_of(d, p->cd[i][c], f[p->cj[i][c]], &a, &lo, &hi);
    if (!member(a, lo, hi, x)) return 0;
  }
  return 1;
}

/* Candidate source for position i: the shortest constraint list, or all
 * vertices when the position has no constraint. */
static inline int source(const oracle_graph *d, const oplan *p, uint32_t i, const uint32_t *f,
                         const uint32_t **a, uint32_t *lo, uint32_t *hi) {
  *a = NULL;
  *lo = *hi = 0;
  if (p->nc[i] == 0) { *hi = d->V; return -1; }
  int best = 0;
  uint32_t bl = UINT32_MAX;
  for (uint32_t c = 0; c < p->nc[i]; ++c) {
    const uint32_t *aa;
    uint32_t l, h;
    list_of(d, p->cd[i][c], f[p->cj[i][c]], &aa, &l, &h);
    if (h - l < bl) { bl = h - l; best = (int)c; *a = aa; *lo = l; *hi = h; }
  }
  return best;
}

static uint64_t dfs_count(const oracle_graph *d, const oplan *p, uint32_t i, uint32_t *f) {
  const uint32_t *a;
  uint32_t lo, hi;
  int skip = source(d, p, i, f, &a, &lo, &hi);
  uint64_t n = 0;
  for (uint32_t t = lo; t < hi; ++t) {
    uint32_t x = a ? a[t] : t;
    if (!accept(d, p, i, f, x, skip)) continue;
    if (i + 1 == p->k) { ++n; continue; }
    f[i] = x;
    n += dfs_count(d, p, i + 1, f);
  }
  return n;
}

uint64_t oracle_count(const oracle_graph *d, uint32_t k, const uint8_t *qa, int nthreads) {
  return oracle_count_ex(d, k, qa, NULL, nthreads);
}

uint64_t oracle_count_ex(const oracle_graph *d, uint32_t k, const uint8_t *qa, const oracle_opts *o,
                         int nthreads) {
  if (k == 0 || k > MAXK) return 0;
  uint32_t ord[MAXK];
  oplan p;
  oracle_order(k, qa, ord);
  plan_from_order(k, qa, ord, &p, o);
  for (uint32_t i = 0; i < k; ++i)
    if (p.nc[i] > MAXC) return 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  /* level 0 */
  uint32_t *l0 = (uint32_t *)malloc(sizeof(uint32_t) * (d->V + 1));
  uint64_t n0 = 0;
  uint32_t f0[MAXK] = {0};
  for (uint32_t x = 0; x < d->V; ++x)
    if (accept(d, &p, 0, f0, x, -1)) l0[n0++] = x;
  if (k == 1) { free(l0); return n0; }
  /* level 1 as explicit pairs, so that hub subtrees spread over threads */
  uint64_t n1 = 0, cap1 = 1024;
  uint32_t *l1 = (uint32_t *)malloc(sizeof(uint32_t) * 2 * cap1);
  for (uint64_t t = 0; t < n0; ++t) {
    uint32_t f[MAXK];
    const uint32_t *a;
    uint32_t lo, hi;
    f[0] = l0[t];
    int skip = source(d, &p, 1, f, &a, &lo, &hi);
    for (uint32_t x = lo; x < hi; ++x) {
      uint32_t v = a ? a[x] : x;
      if (!accept(d, &p, 1, f, v, skip)) continue;
      if (n1 == cap1) { cap1 *= 2; l1 = (uint32_t *)realloc(l1, sizeof(uint32_t) * 2 * cap1); }
      l1[2 * n1] = f[0];
      l1[2 * n1 + 1] = v;
      ++n1;
    }
  }
  uint64_t total = 0;
  if (k == 2) {
    total = n1;
  } else {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total) num_threads(nthreads)
    for (long long t = 0; t < (long long)n1; ++t) {
      uint32_t f[MAXK];
      f[0] = l1[2 * t];
      f[1] = l1[2 * t + 1];
      total += dfs_count(d, &p, 2, f);
    }
  }
  free(l1);
  free(l0);
  return total;
}

typedef struct {
  uint32_t *rows;
  uint64_t cap, n;
} sink;

static void dfs_enum(const oracle_graph *d, const oplan *p, uint32_t i, uint32_t *f, sink *s) {
  const uint32_t *a;
  uint32_t lo, hi;
  int skip = source(d, p, i, f, &a, &lo, &hi);
  for (uint32_t t = lo; t < hi; ++t) {
    uint32_t x = a ? a[t] : t;
    if (!accept(d, p, i, f, x, skip)) continue;
    f[i] = x;
    if (i + 1 == p->k) {
      if (s->n < s->cap)
        for (uint32_t j = 0; j < p->k; ++j) s->rows[s->n * p->k + p->ord[j]] = f[j];
      s->n++;
      continue;
    }
    dfs_enum(d, p, i + 1, f, s);
  }
}

uint64_t oracle_enumerate(const oracle_graph *d, uint32_t k, const uint8_t *qa, uint32_t *rows,
                          uint64_t cap) {
  return oracle_enumerate_ex(d, k, qa, NULL, rows, cap);
}

uint64_t oracle_enumerate_ex(const oracle_graph *d, uint32_t k, const uint8_t *qa, const oracle_opts *o,
                             uint32_t *rows, uint64_t cap) {
  if (k == 0 || k > MAXK) return 0;
  uint32_t ord[MAXK], f[MAXK];
  oplan p;
  oracle_order(k, qa, ord);
  plan_from_order(k, qa, ord, &p, o);
  sink s = {rows, cap, 0};
  dfs_enum(d, &p, 0, f, &s);
  return s.n;
}

int oracle_heuristic_order(uint32_t k, const uint8_t *qa, uint32_t *order) {
  /* out-neighbourhoods as bit sets; k <= 32 */
  if (k == 0 || k > MAXK) return -1;
  uint64_t out[MAXK];
  for (uint32_t u = 0; u < k; ++u) {
    out[u] = 0;
    for (uint32_t v = 0; v < k; ++v)
      if (qa[u * k + v]) out[u] |= 1ull << v;
  }
  uint32_t first = 0, best = 0;
  for (uint32_t v = 0; v < k; ++v) {
    uint32_t od = (uint32_t)__builtin_popcountll(out[v]);
    if (od > best) { best = od; first = v; }
  }
  uint64_t S = 1ull << first, NS = out[first];
  order[0] = first;
  for (uint32_t n = 1; n < k; ++n) {
    uint32_t b1 = 0, b2 = 0, b3 = 0;
    int pick = -1;
    for (uint32_t v = 0; v < k; ++v) {
      if (S >> v & 1) continue;
      uint32_t into = 0;
      for (uint32_t u = 0; u < k; ++u)
        if ((S >> u & 1) && (out[u] >> v & 1)) ++into;
      uint32_t s1 = into + (uint32_t)__builtin_popcountll(out[v] & S);
      uint32_t s2 = (uint32_t)__builtin_popcountll(out[v] & NS & ~S);
      uint32_t s3 = (uint32_t)__builtin_popcountll(out[v] & ~(S | NS));
      int gt = s1 != b1 ? s1 > b1 : (s2 != b2 ? s2 > b2 : s3 > b3);
      if (gt) { b1 = s1; b2 = s2; b3 = s3; pick = (int)v; }
    }
    if (pick < 0) return -1;
    order[n] = (uint32_t)pick;
    S |= 1ull << pick;
    NS |= out[pick];
  }
  return 0;
}

/* ---------------- CPU trie baseline ------------------------------------ */

typedef struct {
  const oracle_graph *d;
  oplan p;
  uint64_t max_rows;
  int nthreads;
  uint64_t level[MAXK];
  double bytes;
  double t0, max_seconds;
  int stopped;
  int symmetric; /* in-lists equal out-lists: one list per backward neighbour */
} trie_ctx;

/* Expands the n rows of R (each `len` words = positions 0..len-1) into
 * position len, recursing depth-first over row groups. */
static void trie_expand(trie_ctx *c, const uint32_t *R, uint64_t n, uint32_t len) {
  const oracle_graph *d = c->d;
  const oplan *p = &c->p;
  const uint32_t k = p->k;
  const int last = (len + 1 == k);
  /* candidate upper bounds for grouping */
  uint64_t start = 0;
  while (start < n) {
    if (c->max_seconds > 0 && omp_get_wtime() - c->t0 > c->max_seconds) {
      c->stopped = 1;
      return;
    }
    uint64_t end = start, acc = 0;
    while (end < n) {
      const uint32_t *f = R + end * len;
      const uint32_t *a;
      uint32_t lo, hi;
      source(d, p, len, f, &a, &lo, &hi);
      uint64_t L = hi - lo;
      /* a group holds at most max_rows candidates (rows of the next level),
       * or 64 x max_rows at the last level (embeddings examined, nothing
       * stored), so the time cap is checked every fraction of a second even
       * where one prefix level fans out into the last */
      if (end > start && acc + L > (last ? 64 * c->max_rows : c->max_rows)) break;
      acc += L;
      ++end;
    }
    const uint64_t g = end - start;
    const int nt = c->nthreads;
    uint32_t **buf = (uint32_t **)calloc(nt, sizeof(uint32_t *));
    uint64_t *bn = (uint64_t *)calloc(nt, sizeof(uint64_t));
    uint64_t *bc = (uint64_t *)calloc(nt, sizeof(uint64_t));
    double *by = (double *)calloc(nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      uint64_t cap = 0, m = 0;
      uint32_t *out = NULL;
      double bytes = 0;
#pragma omp for schedule(dynamic, 64)
      for (long long r = 0; r < (long long)g; ++r) {
        const uint32_t *f = R + (start + r) * len;
        const uint32_t *a;
        uint32_t lo, hi;
        int skip = source(d, p, len, f, &a, &lo, &hi);
        /* merge model: an offset pair and every element of each distinct
         * backward list (both arcs of a symmetric pair share one list;
         * plan_from_order puts a position's two constraints side by side) */
        for (uint32_t cc = 0; cc < p->nc[len]; ++cc) {
          if (c->symmetric && cc > 0 && p->cj[len][cc - 1] == p->cj[len][cc]) continue;
          const uint32_t *aa;
          uint32_t l2, h2;
          list_of(d, p->cd[len][cc], f[p->cj[len][cc]], &aa, &l2, &h2);
          bytes += 8.0 + 4.0 * (h2 - l2);
        }
        for (uint32_t x = lo; x < hi; ++x) {
          uint32_t v = a ? a[x] : x;
          if (!accept(d, p, len, f, v, skip)) continue;
          if (last) { ++m; continue; }
          if (m == cap) {
            cap = cap ? cap * 2 : 1024;
            out = (uint32_t *)realloc(out, cap * (len + 1) * sizeof(uint32_t));
          }
          memcpy(out + m * (len + 1), f, len * sizeof(uint32_t));
          out[m * (len + 1) + len] = v;
          ++m;
        }
      }
      buf[t] = out;
      bn[t] = last ? 0 : m;
      bc[t] = m;
      by[t] = bytes;
    }
    uint64_t tot = 0;
    for (int t = 0; t < nt; ++t) { tot += bc[t]; c->bytes += by[t]; }
    c->level[len] += tot;
    if (!last && tot) {
      c->bytes += 16.0 * tot; /* write + read back of the new trie level */
      uint32_t *cat = (uint32_t *)malloc(tot * (len + 1) * sizeof(uint32_t));
      uint64_t o = 0;
      for (int t = 0; t < nt; ++t) {
        if (bn[t]) memcpy(cat + o * (len + 1), buf[t], bn[t] * (len + 1) * sizeof(uint32_t));
        o += bn[t];
      }
      for (int t = 0; t < nt; ++t) free(buf[t]);
      free(buf);
      free(bn);
      free(bc);
      free(by);
      trie_expand(c, cat, tot, len + 1);
      free(cat);
    } else {
      for (int t = 0; t < nt; ++t) free(buf[t]);
      free(buf);
      free(bn);
      free(bc);
      free(by);
    }
    start = end;
  }
}

int trie_cpu_run(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order,
                 uint32_t sample_stride, uint64_t max_rows, int nthreads, trie_cpu_result *res) {
  return trie_cpu_run_timed(d, k, qa, order, sample_stride, max_rows, nthreads, 0.0, res);
}

int trie_cpu_run_timed(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order,
                       uint32_t sample_stride, uint64_t max_rows, int nthreads, double max_seconds,
                       trie_cpu_result *res) {
  return trie_cpu_run_roots(d, k, qa, order, 0, d->V, 1, sample_stride, max_rows, nthreads, max_seconds, res);
}

int trie_cpu_run_roots(const oracle_graph *d, uint32_t k, const uint8_t *qa, const uint32_t *order, uint32_t rb,
                       uint32_t re, uint32_t rs, uint32_t sample_stride, uint64_t max_rows, int nthreads,
                       double max_seconds, trie_cpu_result *res) {
  if (k == 0 || k > MAXK || rs == 0) return -1;
  trie_ctx *c = (trie_ctx *)calloc(1, sizeof(trie_ctx));
  c->d = d;
  plan_from_order(k, qa, order, &c->p, NULL);
  for (uint32_t i = 0; i < k; ++i)
    if (c->p.nc[i] > MAXC) { free(c); return -1; }
  c->max_rows = max_rows ? max_rows : (1ull << 26);
  c->symmetric = 1;
  for (uint32_t x = 0; x <= d->V && c->symmetric; ++x) c->symmetric = d->off_in[x] == d->off_out[x];
  for (uint64_t e = 0; c->symmetric && e < d->off_out[d->V]; ++e) c->symmetric = d->adj_in[e] == d->adj_out[e];
  c->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  if (sample_stride == 0) sample_stride = 1;
  double t0 = omp_get_wtime();
  c->t0 = t0;
  c->max_seconds = max_seconds;
  uint32_t *l0 = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)d->V + 1));
  uint64_t n0 = 0, seen = 0;
  uint32_t dummy[MAXK] = {0};
  for (uint32_t x = rb; x < d->V && x < re; x += rs)
    if (accept(d, &c->p, 0, dummy, x, -1)) {
      if (seen++ % sample_stride == 0) l0[n0++] = x;
    }
  c->bytes = 8.0 * d->V;
  c->level[0] = n0;
  if (k > 1 && n0) {
    c->bytes += 16.0 * n0;
    trie_expand(c, l0, n0, 1);
  }
  double t1 = omp_get_wtime();
  free(l0);
  memset(res, 0, sizeof(*res));
  for (uint32_t i = 0; i < k; ++i) res->level[i] = c->level[i];
  res->count = c->level[k - 1];
  res->bytes = c->bytes;
  res->seconds = t1 - t0;
  res->threads = c->nthreads;
  res->complete = !c->stopped;
  free(c);
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }
