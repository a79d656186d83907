/*
 * cuts_amd -- C ABI of the MI355X trie-based subgraph matcher.
 *
 * The reference (cuTS) exposes its path through three de-facto boundaries:
 * the `cuts <data_graph> <query_graph> [trunks]` process (main.cu:24-46),
 * the host functions `search` / `search_dfs_bfs_strategy`
 * (inc/host_funcs.h:21-25) and the kernels `initialize_searching` /
 * `search_kernel` (inc/device_funcs.h:54-58).  This header replaces the
 * second and third: plain pointers and sizes, status codes instead of
 * exit(-1), no C++ or torch types.  Every entry point returns CUTS_OK (0) or
 * a negative status; cuts_last_error() gives the message of the last
 * failure on the calling thread.
 */
#ifndef CUTS_AMD_H
#define CUTS_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CUTS_OK = 0,
  CUTS_E_ARG = -1,     /* invalid argument */
  CUTS_E_IO = -2,      /* file could not be read / parsed / written */
  CUTS_E_QUERY = -3,   /* query unsupported: k > 16, not weakly connected */
  CUTS_E_HIP = -4,     /* HIP runtime error */
  CUTS_E_NOMEM = -5,   /* device or host allocation failed */
  CUTS_E_NODEV = -6    /* no usable GPU */
};

typedef struct cuts_graph cuts_graph;               /* host CSR graph */
typedef struct cuts_plan cuts_plan;                 /* query + matching order */
typedef struct cuts_context cuts_context;           /* one GPU: stream + trie arena */
typedef struct cuts_device_graph cuts_device_graph; /* data graph resident in HBM */

const char *cuts_last_error(void);
const char *cuts_version(void);

/* ---- host graphs ------------------------------------------------------ */
/* Edge-list reader; replaces file_reader + Graph::Graph (src/util.cpp:19-52,
 * src/graph.cpp:18-61). Same format: first line V, then "s<TAB>t" arcs. */
int cuts_graph_load(const char *path, cuts_graph **out);
/* Builds a graph from m arcs src[i]->dst[i]; symmetrize adds dst->src. */
int cuts_graph_from_arcs(uint32_t V, const uint32_t *src, const uint32_t *dst, uint64_t m,
                         int symmetrize, cuts_graph **out);
int cuts_graph_save(const cuts_graph *g, const char *path);
typedef struct {
  uint32_t V;
  uint64_t arcs;
  int symmetric;
  const uint32_t *out_off, *out_adj, *in_off, *in_adj, *sig; /* sig[2v]=in, sig[2v+1]=out */
} cuts_graph_view;
int cuts_graph_view_of(const cuts_graph *g, cuts_graph_view *view);
void cuts_graph_free(cuts_graph *g);
/* Two-hop index (no reference counterpart): the table a device upload of a
 * symmetric graph builds when its wedges fit the budget.  For every vertex
 * u, the vertices v != u sharing a neighbour with u, ascending
 * (v[off[u] .. off[u+1]]), and |N(u) & N(v)| for each (c).  off holds V+1
 * entries; v and c hold cap entries.  *entries receives the table size;
 * CUTS_E_ARG when it exceeds cap (call with cap = 0 to size the buffers). */
int cuts_graph_two_hop(const cuts_graph *g, uint64_t *off, uint32_t *v, uint32_t *c, uint64_t cap,
                       uint64_t *entries);

/* ---- synthetic inputs (no reference counterpart) ---------------------- */
int cuts_gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint32_t max_deg, uint64_t seed,
                      cuts_graph **out);
int cuts_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                  cuts_graph **out);
/* One occurrence of an undirected pattern (k vertices, ne edges as pairs). */
int cuts_find_pattern(const cuts_graph *d, uint32_t k, const uint32_t *edges, uint32_t ne,
                      uint64_t seed, uint32_t *witness);
/* Random-walk query with k vertices and up to n_edges edges; writes
 * *n_out edge pairs into edges (capacity k*(k-1)/2 pairs) and the walk's
 * data vertices into witness[k]. */
int cuts_walk_query(const cuts_graph *d, uint32_t k, uint32_t n_edges, uint64_t seed, uint32_t *edges,
                    uint32_t *n_out, uint32_t *witness);

/* ---- query plans: Graph::sort_search_order (src/graph.cpp:62-134) ----- */
/* order == NULL computes the score order of src/score.cpp:18-99. */
int cuts_plan_create(const cuts_graph *query, const uint32_t *order, cuts_plan **out);
/* Copies k, order[k], P/C CSR (par_off[k+1], par[], chi_off[k+1], chi[]);
 * any output pointer may be NULL. */
int cuts_plan_info(const cuts_plan *p, uint32_t *k, uint32_t *order, uint32_t *par_off, uint32_t *par,
                   uint32_t *chi_off, uint32_t *chi);
void cuts_plan_free(cuts_plan *p);

/* ---- device ----------------------------------------------------------- */
int cuts_context_create(int device, cuts_context **out);
void cuts_context_destroy(cuts_context *ctx);
/* Copies the CSR + signatures to HBM (src/gpu_memory_allocation.cu:26-36). */
int cuts_device_graph_upload(cuts_context *ctx, const cuts_graph *g, cuts_device_graph **out);
void cuts_device_graph_free(cuts_device_graph *dg);
/* What an upload built (no reference counterpart). */
typedef struct {
  uint32_t V;
  uint64_t arcs;
  int symmetric;
  uint64_t self_loops;          /* vertices u with u in out(u) */
  uint32_t hubs, hub_min_deg;   /* hub bitmap rows and their degree threshold */
  uint64_t two_hop_entries;     /* 0: no two-hop index */
  uint32_t two_hop_dense_rows;
  int arc_common;               /* 1: per-arc common-neighbour counts present */
  double two_hop_build_s, arc_common_build_s;
  int two_hop_on_device;        /* 1: the index was built on the device (V <= 2^17) */
  uint32_t hub_rows;            /* graphs without the index: rows of |N(hub) & N(v)| (0: none) */
  double hub_rows_build_s;
} cuts_device_graph_info;
int cuts_device_graph_query(const cuts_device_graph *dg, cuts_device_graph_info *info);
/* Per-arc common-neighbour counts (no reference counterpart): built on the
 * device at upload for symmetric loop-free graphs; out[i] = |N(u) & N(w)|
 * for the arc (u, w) at position i of the out-lists.  n must equal the
 * number of arcs. */
int cuts_device_graph_arc_common(const cuts_device_graph *dg, uint32_t *out, uint64_t n);
/* The device graph's two-hop index read back (no reference counterpart):
 * the layout of cuts_graph_two_hop (off: V+1 entries, v and c: cap entries,
 * CUTS_E_ARG when the index holds more), plus its dense rows when dense_id
 * is non-null: dense_id[u] = row or -1 (V entries) and row r at
 * dense[r*V .. r*V+V) holding |N(u) & N(v)| as uint16 (dense_cap entries,
 * at least rows*V).  CUTS_E_ARG when the upload built no index. */
int cuts_device_graph_two_hop(const cuts_device_graph *dg, uint64_t *off, uint32_t *v, uint32_t *c, uint64_t cap,
                              int32_t *dense_id, uint16_t *dense, uint64_t dense_cap);
/* Hub rows read back (no reference counterpart): on symmetric loop-free
 * graphs too large for the two-hop index, the upload gives vertices of
 * degree >= 256 (CUTS_H2_HUB_MIN_DEG), longest lists first, up to 16384
 * (CUTS_H2_HUB_ROWS) a row of |N(u) & N(v)| for every v as uint16.
 * row_id[u] = row or -1 (V entries); row r at rows[r*V .. r*V+V) (cap
 * entries, at least hub_rows*V).  CUTS_E_ARG when the upload built none. */
int cuts_device_graph_hub_rows(const cuts_device_graph *dg, int32_t *row_id, uint16_t *rows, uint64_t cap);

enum {
  CUTS_FLAG_NO_SHORTCUT = 1u << 0, /* last level: always scan candidate lists */
  CUTS_FLAG_NO_FUSE = 1u << 1,     /* never fuse the last two levels */
  CUTS_FLAG_WRITE = 1u << 2,       /* materialise embeddings into emb_rows */
  CUTS_FLAG_NO_XCD = 1u << 3,      /* plain blockIdx -> tile mapping */
  CUTS_FLAG_PROFILE = 1u << 4,     /* time every kernel class with HIP events */
  CUTS_FLAG_NO_BITMAP = 1u << 5,   /* lower-bound searches even in hub lists */
  CUTS_FLAG_NO_PREFIX_BASE = 1u << 6, /* fused closed form fully per work item */
  CUTS_FLAG_NO_PAIR = 1u << 7,        /* never count the last two positions together */
  CUTS_FLAG_NO_CHAIN = 1u << 8,       /* not as a chain either (k-1 hanging on k-2) */
  CUTS_FLAG_NO_TWO_HOP = 1u << 9,     /* intersections by list walks even where the two-hop index exists */
  CUTS_FLAG_NO_SHARED = 1u << 10,     /* chain level: position k-3's candidates found per row, not per parent */
  CUTS_FLAG_NO_ARC_COMMON = 1u << 11, /* ignore the per-arc common-neighbour table */
  CUTS_FLAG_SYNC_WRITE = 1u << 12,    /* CUTS_FLAG_WRITE: copy each chunk's rows before the next chunk runs
                                         (default: copies stream on a second stream, overlapped) */
  CUTS_FLAG_SYNC_LEVELS = 1u << 13    /* read every level's work back before expanding it (default: a
                                         level that fits the buffers already allocated is expanded
                                         without that readback; results are identical) */
};

/* Kernel classes reported by CUTS_FLAG_PROFILE. */
enum {
  CUTS_K_LEVEL0 = 0,   /* level-0 signature filter */
  CUTS_K_PLAN = 1,     /* shortest-list selection per prefix */
  CUTS_K_SCAN = 2,     /* exclusive scan of the list lengths */
  CUTS_K_TILE = 3,     /* tile -> prefix index */
  CUTS_K_EXPAND = 4,   /* expansion writing the next trie level */
  CUTS_K_COUNT = 5,    /* expansion counting the last level */
  CUTS_K_FUSED = 6,    /* expansion + closed-form last level */
  CUTS_K_CLOSED = 7,   /* closed-form last level over stored rows */
  CUTS_K_PAIR = 8,     /* expansion + the last two positions counted as a pair */
  CUTS_K_COMMON = 9,   /* intersection sizes |N(u) & N(w)| by list walks */
  CUTS_K_CHAIN = 10,   /* the last two positions counted as a chain */
  CUTS_K_CLASSES = 11
};

typedef struct {
  uint32_t trunks;       /* level-0 chunks, as the reference's 3rd CLI arg; 0 = automatic */
  uint32_t flags;        /* CUTS_FLAG_* */
  uint64_t budget_bytes; /* trie-table budget in HBM; 0 = default */
  /* level-0 restriction: only data vertices root_begin + i*root_stride
   * (< root_end) may match position 0; all zero = every vertex. */
  uint32_t root_begin, root_end, root_stride;
  uint32_t blocks_per_cu; /* launch shape for the expansion kernels; 0 = default (6 blocks per CU) */
  /* CUTS_FLAG_WRITE: host buffer receiving up to emb_cap rows of k words,
   * row = data vertices in matching-order positions (ans.txt layout). */
  uint32_t *emb_rows;
  uint64_t emb_cap;
  /* > 0 (counting runs only): stop once this much wall time has passed,
   * checked between expansion chunks; a level-0 trunk cut short is rolled
   * back, so the result covers exactly the first roots_done level-0
   * vertices (level 0 is in ascending vertex order). */
  double time_limit_ms;
  /* > 0: give up once the expansion work (sum of the per-level scan totals,
   * i.e. candidates examined) exceeds this; deterministic, used to admit
   * queries into fixed workloads.  The result then has work_capped = 1 and
   * no valid count. */
  uint64_t max_work;
  /* threads per block of the counting expansions (the fused, pair, chain
   * and last-level kernels): 64, 128 or 256 (0 = 256).  A tile stays 1,024
   * work items, so each lane carries 16, 8 or 4 of them. */
  uint32_t expand_threads;
  /* blocks per CU of the tiled kernels of each class (index CUTS_K_*: the
   * expansions, the intersection walks, the chain level); 0 = blocks_per_cu.
   * Kernels that walk lists at random addresses often run faster with fewer
   * blocks (less L2 thrashing) than the counting ones. */
  uint8_t class_blocks_per_cu[16];
} cuts_match_options;

typedef struct {
  uint64_t count;           /* exact number of embeddings */
  uint64_t level_sizes[16]; /* |T_l| for every order position l < k */
  double device_ms;         /* first matcher kernel .. final count readback */
  uint64_t emb_written;     /* rows copied to emb_rows */
  uint32_t chunks;          /* expansion chunks run (automatic trunking) */
  uint32_t complete;        /* 0 if time_limit_ms stopped the run early */
  /* CUTS_FLAG_PROFILE: per kernel class, summed HIP-event time, launches and
   * algorithmic (demand) bytes as defined next to each kernel */
  double kernel_ms[11];
  uint64_t kernel_launches[11];
  double kernel_bytes[11];
  uint64_t roots_done;      /* level-0 vertices whose subtrees are counted */
  uint32_t last_root;       /* data vertex of the last of them */
  uint32_t work_capped;     /* 1 if max_work stopped the run (count invalid) */
  uint64_t work;            /* candidates examined (sum of scan totals) */
  /* CUTS_FLAG_PROFILE: merge-model algorithmic bytes per kernel class -- the
   * level-0 signature scan (8 per vertex) and level-0 nodes (16 each) in
   * class 0; for every expansion level, 8 per backward neighbour and 4 per
   * element of its list for each prefix expanded, plus 16 per trie node of
   * levels below k-1 (written and read back), in the class that performs the
   * expansion.  The sum does not depend on kernels or modes: it equals the
   * CPU trie walk's figure for the same graph, query and order. */
  double merge_bytes[11];
} cuts_match_result;

void cuts_match_options_default(cuts_match_options *o);
/* One query against a resident data graph; replaces the level loop of
 * search / search_dfs_bfs_strategy (src/host_funcs.cu:39-171). */
int cuts_match(cuts_context *ctx, const cuts_device_graph *dg, const cuts_plan *plan,
               const cuts_match_options *opt, cuts_match_result *res);

/* Many queries against one resident data graph, run as single-query
 * matches on `workers` host threads (0 = 4, at most 16), each with its own
 * stream and trie arena (created inside the context on first use and kept;
 * their trie budgets split the context's), so that the small kernels of
 * different queries share the GPU.  Queries are handed out dynamically in
 * order; res[i] is query i's result.  opts: n_opts = 0 (defaults), 1 (the
 * same options for every query) or n.  *wall_ms (may be NULL) receives the
 * host wall time of the whole set.  No reference counterpart: the reference
 * runs one query per process (src/host_funcs.cu:115-171). */
int cuts_match_many(cuts_context *ctx, const cuts_device_graph *dg, const cuts_plan *const *plans, uint32_t n,
                    const cuts_match_options *opts, uint32_t n_opts, uint32_t workers, cuts_match_result *res,
                    double *wall_ms);

/* No reference counterpart: cuts_match for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and result;
 * every buffer of the run (trie levels, scans, counters, degree tables the
 * device graph has not cached yet) is carved from a static scratch array of
 * cuts_noalloc_scratch_bytes() bytes that each GPU holds for the life of the
 * process; nothing is allocated or freed.  The trie levels take at most an
 * eighth of it (budget_bytes is clamped), so large outputs run in more
 * chunks.  Fails with CUTS_E_ARG when the data graph has more vertices than
 * the scratch allows (level 0: about 32 bytes per vertex within a quarter
 * of it) and with CUTS_E_NOMEM when a run outgrows it.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int cuts_match_noalloc(cuts_context *ctx, const cuts_device_graph *dg, const cuts_plan *plan,
                       const cuts_match_options *opt, cuts_match_result *res);
uint64_t cuts_noalloc_scratch_bytes(void);
/* Pinned (page-locked) host memory (no reference counterpart): an emb_rows
 * buffer allocated here receives CUTS_FLAG_WRITE rows by asynchronous copies
 * that overlap the expansion; pageable memory works but copies block. */
int cuts_host_alloc(uint64_t bytes, void **out);
void cuts_host_free(void *p);

/* Table log (no reference counterpart; the reference keeps the level sizes
 * in `lengths`, src/host_funcs.cu:95,153, and drops them).  Appends one JSON
 * line to `path`: {"label", "k", "order", "level_sizes", "count", "work",
 * "device_ms", "complete"} -- the training data of the table-size network.
 * The CLI writes it when the environment variable CUTS_TABLE_LOG names a
 * file. */
int cuts_log_tables(const char *path, const char *label, const cuts_plan *plan, const cuts_match_result *r);

/* ---- batched matching (no reference counterpart) --------------------- */
/* Many queries against one resident data graph.  A level-synchronous device
 * work queue of (query, level, row chunk) items: by default one
 * persistent-grid launch per trie level processes the items of every query
 * at that level and publishes the next level's items, and no block ever
 * waits for another; with `persistent`, one cooperative launch runs every
 * level, its blocks meeting at a bounded device-wide barrier between levels.
 * The launch sequence is captured once in a hipGraph and replayed.  The
 * reference runs one query per process with a host sync per level
 * (src/host_funcs.cu:141-157) and spreads level-0 trunks over MPI ranks
 * (main_mpi.cu:257-376). */
typedef struct cuts_batch cuts_batch;
typedef struct {
  uint32_t flags;         /* CUTS_FLAG_NO_SHORTCUT / CUTS_FLAG_NO_FUSE / CUTS_FLAG_NO_BITMAP */
  uint32_t blocks_per_cu; /* grid of each level launch = CUs x this, 1..8; 0 = 6 */
  uint32_t vchunk;        /* level-0 vertices per initial item; 0 = 4096 */
  uint32_t no_graph;      /* 1: plain launches instead of a hipGraph */
  uint64_t queue_cap;     /* work items per level; 0 = 4M */
  uint64_t arena_bytes;   /* trie-row arena, two halves alternating between levels; 0 = 8 GiB */
  uint32_t ticket_queue;  /* 1: the asynchronous ticket-queue kernel -- disabled, create fails */
  uint32_t item_work;     /* candidate work at which a new item is started; 0 = 16384 */
  const uint32_t *roots;  /* NULL, or 3 words per query: first, end, stride of the data
                             vertices order position 0 may take (cuts_match_options'
                             root_begin/root_end/root_stride, per query) */
  uint32_t persistent;    /* 1: one cooperative launch runs every level (the persistent
                             kernel); blocks meet at a device-wide barrier between levels
                             instead of a launch boundary.  blocks_per_cu is clamped to
                             what one CU holds at once. */
  uint32_t barrier_timeout_ms; /* persistent: a block that waits longer than this at a level
                                  barrier aborts the run; every block then leaves, the run
                                  fails with CUTS_E_HIP and the context stays usable;
                                  0 = 60000; 0xFFFFFFFF = no wait (the first barrier
                                  aborts: for tests of that path) */
} cuts_batch_options;
typedef struct {
  uint64_t count;
  uint64_t level_sizes[16];
  double latency_ms;      /* launch start .. this query's last item done */
} cuts_batch_query_result;
void cuts_batch_options_default(cuts_batch_options *o);
/* The context, the device graph and the plans must outlive the batch. */
int cuts_batch_create(cuts_context *ctx, const cuts_device_graph *dg, const cuts_plan *const *plans, uint32_t n,
                      const cuts_batch_options *opt, cuts_batch **out);
/* results: n entries.  Fails with CUTS_E_NOMEM if the queue or arena filled
 * up (counts are then invalid).  arena_bytes: the largest level's rows. */
int cuts_batch_run(cuts_batch *b, cuts_batch_query_result *results, double *device_ms, uint64_t *arena_bytes);
void cuts_batch_free(cuts_batch *b);

/* ---- table-size predictor (no reference counterpart) ------------------ */
/* A 3-layer message-passing network (hidden 64, BF16 weights, FP32
 * accumulation) mapping (query graph, prefix set S, data-graph features) to
 * y ~ log(1 + |T(S)|) / 10, where |T(S)| is the trie table size of a prefix
 * with vertex set S.  Forward, backward and Adam run in HIP kernels.
 * A sample is k <= 16 vertices: x[16][24] input features (rows >= k
 * ignored), adj[16] neighbour bit masks, meta = k | (S bit mask << 16). */
enum { CUTS_MPNN_NODES = 16, CUTS_MPNN_FEATS = 24, CUTS_MPNN_HIDDEN = 64 };
typedef struct cuts_mpnn cuts_mpnn;
uint32_t cuts_mpnn_num_params(void);
/* params: cuts_mpnn_num_params() floats (FP32 master copy). */
int cuts_mpnn_create(cuts_context *ctx, const float *params, cuts_mpnn **out);
int cuts_mpnn_get_params(cuts_mpnn *m, float *params);
int cuts_mpnn_set_params(cuts_mpnn *m, const float *params);
int cuts_mpnn_forward(cuts_mpnn *m, uint32_t n, const float *x, const uint16_t *adj, const uint32_t *meta, float *y);
/* Mean squared error over the n samples and its gradient w.r.t. the BF16
 * weights (grad may be NULL). */
int cuts_mpnn_gradient(cuts_mpnn *m, uint32_t n, const float *x, const uint16_t *adj, const uint32_t *meta,
                       const float *target, float *grad, double *loss);
/* gradient + one Adam step (beta 0.9/0.999, eps 1e-8). */
int cuts_mpnn_train_step(cuts_mpnn *m, uint32_t n, const float *x, const uint16_t *adj, const uint32_t *meta,
                         const float *target, float lr, double *loss);
void cuts_mpnn_free(cuts_mpnn *m);

/* Calibration (no reference counterpart): copies `bytes` bytes through a
 * plain streaming kernel `iters` times and reports read+write GB/s.  Its
 * dispatches are the known-size yardstick for HBM byte counters. */
int cuts_copy_probe(cuts_context *ctx, uint64_t bytes, uint32_t iters, double *gbps);

/* ---- reference-compatible entry points ------------------------------- */
/* search(query, data, write_to_disk) and search_dfs_bfs_strategy(...,
 * trunks) of inc/host_funcs.h:21-25: load both files, print the progress
 * lines and the "<data>,<query>,<ms>ms,<count>" report line on stdout,
 * write ans.txt in the current directory when write_to_disk. */
int cuts_search(const char *query_file, const char *data_file, int write_to_disk, uint64_t *count);
int cuts_search_dfs_bfs_strategy(const char *query_file, const char *data_file, int write_to_disk,
                                 uint32_t trunks, uint64_t *count);
/* ans.txt writer of src/util.cpp:53-74 for n rows of k words (positions). */
int cuts_write_matches(const char *path, const cuts_plan *plan, const uint32_t *rows, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
