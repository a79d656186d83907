// Internal interface between the C ABI layer (capi.cpp) and the HIP code
// (match.hip, batch.hip).  Not part of the public ABI.
#pragma once
#include <array>
#include <string>
#include <unordered_map>
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace cuts {

// Error carrying a cuts.h status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define HIPCHECK(x) ::cuts::hip_check((x), #x)

// Plan in the form the kernels read: per position the list of constraints
// (earlier position | direction<<7) and the signature thresholds, plus the
// closed-form last-level description.  Passed by value as a kernel argument.
struct KPlan {
  uint32_t k;
  uint32_t sig_in[kMaxQuery], sig_out[kMaxQuery];
  uint32_t nj[kMaxQuery];
  uint8_t jc[kMaxQuery][2 * kMaxQuery];
  // last position counted as |N(row[sc_pos])| minus prefix members
  uint32_t sc_ok;      // 1 if the closed form applies
  uint32_t sc_pos;     // position whose list is counted
  uint32_t sc_adjmask; // positions certainly inside that list
  uint32_t sc_thr;     // degree threshold of the last position
  // last two positions counted together (symmetric graphs): when neither is
  // a backward neighbour of the other and each has one backward neighbour,
  // at positions pr_a and pr_b older than position k-3, a prefix of
  // positions 0..k-3 has |A| * |B| - |A & B| completions, with
  // A = N(row[pr_a]) and B = N(row[pr_b]) minus the prefix, filtered by
  // degree thresholds
  uint32_t pr_ok;
  uint32_t pr_a, pr_b;
  uint32_t pr_ta, pr_tb, pr_tab;  // degree thresholds of A, B and A & B
  uint32_t pr_ma, pr_mb;          // positions < k-2 certainly inside N(row[pr_a]) / N(row[pr_b])
  // pair level whose position k-3 has exactly the constraints pr_a, pr_b
  // (thresholds <= 2): every survivor v of a prefix adds the same
  // (|A|-1)(|B|-1) - (|A & B|-1), and the survivors are N(r_a) & N(r_b) minus
  // the prefix, so the prefix's count is closed-form from one two-hop lookup
  uint32_t pr_collapse;
  // pr_a and pr_b adjacent in the query: |A & B|'s base is the per-arc
  // common-neighbour count of the arc (row[pr_a], row[pr_b])
  uint32_t pr_adj;
  // last two positions as a chain (symmetric graphs): k-1's one backward
  // neighbour is k-2, whose one backward neighbour ch_u is older than k-3.
  // A prefix P of positions 0..k-3 with u = row[ch_u] and
  // A = {x in N(u) : deg(x) >= ch_ta} - P has
  //   sum_{x in A} |{y in N(x) : deg(y) >= ch_tb} - P - {x}|
  // completions, which take the vertex tables H(u) = sum over N(u) of the
  // degree-filtered neighbour counts and one intersection N(u) & N(r_i) per
  // other prefix vertex -- those of positions < k-3 shared by all rows with
  // the same parent prefix
  uint32_t ch_ok;
  uint32_t ch_u;
  uint32_t ch_ta, ch_tb;
  // chain level whose position k-3 has no constraint on position k-4: its
  // candidates depend on the row's parent prefix only (shared by siblings)
  uint32_t ch_sib;
  uint32_t qadj[kMaxQuery];       // bit j of qadj[i]: positions i and j adjacent in the query
};

// loop_free: the data graph has no self-loop.  The pair, chain and collapsed
// closed forms assume u not in N(u); on a graph with self-loops they are
// off and those positions are expanded (the last-level closed form counts
// u in N(u) itself and stays on).
KPlan make_kplan(const Plan& p, bool symmetric_data, bool loop_free);

struct DeviceGraph {
  int device = 0;
  uint32_t V = 0;
  uint64_t A = 0;
  bool symmetric = false;
  uint32_t max_deg = 0;
  uint64_t loops = 0;  // vertices u with u in out(u)
  uint32_t *off_out = nullptr, *adj_out = nullptr, *off_in = nullptr, *adj_in = nullptr;
  uint32_t* sig = nullptr;                 // 2V: in, out
  // signature of every list entry, parallel to adj_out / adj_in (the in-side
  // aliases the out-side on symmetric graphs): a candidate's signature is
  // read at its list index, coalesced with the list word itself
  uint2 *nsig_out = nullptr, *nsig_in = nullptr;
  std::recursive_mutex tables_mu;          // guards the two lazily built table maps below
  std::map<uint32_t, uint32_t*> cnt_tables; // degree threshold -> #neighbours meeting it
  // (ta, tb) -> 2V uint64: per vertex u, sum over x in N(u) with deg(x) >= ta
  // of |{y in N(x) : deg(y) >= tb}|, then of deg(x)
  std::map<std::pair<uint32_t, uint32_t>, uint64_t*> chain_tables;
  // adjacency bitmaps of hub vertices (symmetric graphs): hub_id[v] = row of
  // v in `bits` or -1; a row is `words` uint64 covering all V vertices
  int32_t* hub_id = nullptr;
  uint64_t* bits = nullptr;
  uint32_t words = 0, hubs = 0, hub_min_deg = 0;
  // two-hop index (symmetric graphs whose wedges fit the budget): per vertex
  // u, the vertices v != u with a common neighbour, ascending (h2_v over
  // h2_off[u] .. h2_off[u+1]), and |N(u) & N(v)| for each (h2_c)
  uint64_t* h2_off = nullptr;
  uint32_t *h2_v = nullptr, *h2_c = nullptr;
  uint64_t h2_entries = 0;
  // dense rows of the index for vertices whose two-hop list holds at least
  // V/64 entries: h2_did[u] = row or -1, a row = V uint16 counts (a lookup
  // is one load instead of a lower-bound search)
  int32_t* h2_did = nullptr;
  uint16_t* h2_dense = nullptr;
  uint32_t h2_dense_rows = 0;
  double h2_build_s = 0;
  bool h2_on_gpu = false;  // built by k_two_hop (V <= 2^17) rather than on the host
  // per-arc common-neighbour counts (symmetric loop-free graphs): arc_cn[i]
  // = |N(u) & N(w)| for the arc (u, w) at adj_out[i]; built on the device at
  // upload
  uint32_t* arc_cn = nullptr;
  double arc_cn_build_s = 0;
  // hub rows (symmetric loop-free graphs too large for the two-hop index):
  // for up to 16,384 vertices of degree >= 256 (within 1/8 of free HBM), longest lists first,
  // |N(u) & N(v)| for every v as V uint16 (hr_id[u] = row or -1); built on
  // the device at upload.  They serve the intersections of the pair and
  // chain levels in which u or w has a row
  int32_t* hr_id = nullptr;
  uint16_t* hr = nullptr;
  uint32_t hr_rows = 0;
  double hr_build_s = 0;
  ~DeviceGraph();
};

// Per-level scratch of the trie: rows plus the arrays that plan the
// expansion of those rows.
struct LevelBuf {
  uint32_t* rows = nullptr;  // cap rows of (level+1) words
  uint64_t cap = 0;
  uint32_t* rows2 = nullptr;    // second output buffer of the written last level
  uint64_t rows2_cap = 0;
  uint32_t* minlen = nullptr;   // cap
  uint8_t* minsel = nullptr;    // cap
  uint64_t* work_off = nullptr; // cap+1
  uint32_t* cbase = nullptr;    // cap: per-prefix closed-form base (fused levels)
  // allocated only on the levels that use them (pbase_cap etc. rows each)
  uint4* pbase = nullptr;       // per-prefix |A|, |B|, |A & B| bases (pair level)
  uint32_t* nsurv = nullptr;    // per prefix, survivors (pair level), the shared
                                // intersections of its children (chain level) or
                                // the filled length of its block in `cand`
  uint32_t* parent = nullptr;   // parent prefix of every row (rows of the chain level)
  uint64_t pbase_cap = 0, nsurv_cap = 0, parent_cap = 0;
  // chain level with shared candidates: per prefix of this level, the start
  // of its block in `cand` (n+1; a block holds as many entries as the
  // prefix's shortest constraint list) and the candidates of position k-3
  uint64_t* coff = nullptr;
  uint32_t* cand = nullptr;
  uint64_t coff_cap = 0, cand_cap = 0;
  uint64_t* bsum = nullptr;     // scan partials
  uint64_t bsum_cap = 0;
  uint32_t* tile_p = nullptr;   // tiles+1
  uint64_t tile_cap = 0;
};

// Bump allocator over a fixed device region (the no-allocation entry
// point's static scratch array): carve() hands out 256-byte aligned pieces
// and never frees; it throws CUTS_E_NOMEM when the region is exhausted.
struct ScratchArena {
  char* base = nullptr;
  uint64_t size = 0, used = 0;
  void* carve(uint64_t bytes);
};

struct Context {
  int device = 0;
  int cus = 256;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t default_budget = 0;
  LevelBuf lv[kMaxQuery];
  // [0] embeddings, [1] level k-2 size of the fused or chain level (level
  // k-3 with the pair level), [2] level k-2 size of the pair level, [3] demand bytes, [4..7] eight
  // uint32 tile counters, [8] merge-model bytes, [10] rows written by one
  // expansion
  unsigned long long* counters = nullptr;
  hipEvent_t kev[2] = {nullptr, nullptr};
  // embedding output: rows of the last level are copied to the caller's
  // buffer on their own stream while the next chunk expands; copy_ev[b] marks
  // the end of the last copy out of output buffer b
  hipStream_t copy_stream = nullptr;
  hipEvent_t copy_ev[2] = {nullptr, nullptr};
  uint64_t* host_pinned = nullptr;        // small pinned staging
  // coherent host words that kernels write directly (system-scope stores):
  // a writing level's row count and status, level 0's size, the run's
  // counters, read after the stream synchronises with no copy; host_res_dev
  // is its device address
  uint64_t* host_res = nullptr;
  uint64_t* host_res_dev = nullptr;
  // row counts of levels whose successor took its size from the device
  // (16 words, written by k_plan_small), and the level sizes of the last
  // complete run of each query (by device graph, plan, roots and flags):
  // a writing level of at most one planning block's rows in that run hands
  // its size to the next level on the device
  unsigned long long* lvcnt = nullptr;
  std::unordered_map<std::string, std::array<uint64_t, kMaxQuery>> level_hints;
  char* scratch = nullptr;                // this device's no-allocation scratch array
  ~Context();
};

struct MatchOpts {
  uint32_t trunks = 0;
  uint32_t flags = 0;
  uint64_t budget = 0;
  uint32_t root_begin = 0, root_end = 0, root_stride = 0;
  uint32_t blocks_per_cu = 0;
  uint32_t* emb_rows = nullptr;
  uint64_t emb_cap = 0;
  double time_limit_ms = 0;
  uint64_t max_work = 0;
  uint32_t expand_threads = 0;
  uint8_t class_bpc[16] = {0};  // per kernel class blocks per CU (0: blocks_per_cu)
};
struct MatchResult {
  uint64_t count = 0;
  uint64_t level[kMaxQuery] = {0};
  double device_ms = 0;
  uint64_t emb_written = 0;
  uint32_t chunks = 0;
  uint32_t complete = 1;
  uint64_t roots_done = 0;
  uint32_t last_root = 0;
  uint32_t work_capped = 0;
  uint64_t work = 0;
  double kernel_ms[11] = {0};
  uint64_t kernel_launches[11] = {0};
  double kernel_bytes[11] = {0};
  double merge_bytes[11] = {0};
};

void context_init(Context& c, int device);
void upload_graph(Context& c, const Graph& g, DeviceGraph& dg);
void run_match(Context& c, DeviceGraph& dg, const Plan& p, const MatchOpts& o, MatchResult& r);
// Bytes of the static scratch array behind run_match_noalloc.
uint64_t noalloc_scratch_bytes();
// run_match with every buffer of the run carved from the device's static
// scratch array: no device allocation at all.  Degree tables the graph has
// not cached yet are built in the scratch too (and not cached).  Calls that
// overlap on one device share the array and race.
void run_match_noalloc(Context& c, DeviceGraph& dg, const Plan& p, const MatchOpts& o, MatchResult& r);
// Streams `bytes` through a plain copy kernel `iters` times; returns GB/s
// (read + write).
double copy_probe(Context& c, uint64_t bytes, uint32_t iters);

}  // namespace cuts
