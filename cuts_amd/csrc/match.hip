// Trie-based partial-match expansion on gfx950 (MI355X).
//
// Semantics follow the reference's kernels: level 0 is every data vertex
// whose signature dominates that of the first query vertex
// (src/device_funcs.cu:267-279); level l keeps, for each prefix, the
// candidates v that lie in out(pre[j]) for every parent position j and in
// in(pre[j]) for every child position j, are not already used by the prefix,
// and whose signature dominates that of the query vertex at position l
// (src/device_funcs.cu:280-407); the last level is counted, not stored
// (src/device_funcs.cu:363-391).
//
// Layout and schedule are MI355X-first rather than a translation:
//   * a trie level is a dense row table (row = the data vertices of all
//     positions so far), so a prefix is one contiguous read instead of the
//     reference's parent-pointer chase (src/device_funcs.cu:39-55);
//   * expansion work is flattened: per prefix the shortest constraint list
//     gives an exact upper bound L_p on its children; an exclusive scan of
//     L_p maps work items to (prefix, list index), so a hub's list is spread
//     over many wavefronts and every wave64 lane checks one candidate;
//   * the per-tile slice of the scan is staged in LDS and searched there;
//   * survivors are compacted with ballot/mbcnt inside the wave, an LDS
//     prefix over the block's waves, and one global atomic per block;
//   * the scan total bounds the next level exactly, so a chunk of the work
//     space is sized to the HBM budget before it runs: trunking is automatic
//     and no table can overflow (the reference has no overflow check);
//   * when the last position has a single constraint on a symmetric graph,
//     its candidates are counted in closed form (|N(u)| meeting the degree
//     threshold, minus prefix members of N(u)), and that count is fused into
//     the expansion of the level before, so the largest level is never
//     materialised;
//   * when the last two positions are two pendants on older vertices (a
//     pair) or a 2-edge tail on an older vertex (a chain), both are counted
//     in closed form inside the expansion of position k-3, from degree
//     tables and |N(u) & N(w)| looked up in a two-hop index built at upload
//     (or counted by flattened list walks without it); a pair whose position
//     k-3 is exactly the common neighbours of its two list owners is counted
//     per prefix of level k-4 without expanding position k-3 at all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <mutex>

#include "batch.hpp"

// Scratch of the no-allocation entry point (run_match_noalloc): every
// buffer of such a run -- trie levels, scans, counters, degree tables the
// graph has not cached -- is carved from this array.  Each device holds its
// own copy for the life of the process.  Host code never passes this name
// to a kernel: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress (scratch_address below).
constexpr uint64_t kScratchBytes = 256ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace cuts {

// The device address of this device's scratch array, looked up once per
// device with that device current (context_init calls it).
char* scratch_address(int device) {
  static std::mutex mu;
  static char* addr[64] = {};
  if (device < 0 || device >= 64) throw Error(-1, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(-4, "hipGetSymbolAddress returned a null scratch address");
    addr[device] = (char*)p;
  }
  return addr[device];
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(-4, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int BLOCK = 256;  // 4 wave64s
constexpr int IPT = 4;      // work items per thread per tile
constexpr int TILE = BLOCK * IPT;
constexpr int LDS_OFF = 2048;      // staged work offsets per tile
// Items in lock-step per thread in the fused last-two-levels expansion, for
// rows of up to 6 and of 7..10 words (more items keep more independent loads
// in flight, at the cost of registers; 8 for short rows spilled and ran 14%
// slower on the R-MAT 20 step, 4 for long rows ran 1% faster than 2), and in
// the counting and writing expansions for rows of 7..10 words (4 ran 1%
// faster than 2).  The pair, chain and collapsing expansions take 4 and 2.
#ifndef FUSED_G_SMALL
#define FUSED_G_SMALL 4
#endif
#ifndef FUSED_G_LARGE
#define FUSED_G_LARGE 4
#endif
#ifndef COUNT_G_LARGE
#define COUNT_G_LARGE 4
#endif
constexpr int SCAN_IPT = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_IPT;
constexpr uint64_t MAX_CHUNK = 1ull << 31;  // work items per launch sequence (written rows stay below 2^32)
// work items of one speculative counting launch (counts, offsets and tile
// numbers are 64-bit; the tile index of such a level is at most 1 GiB)
constexpr uint64_t MAX_SPEC_ITEMS = 1ull << 38;

struct GView {
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
  uint32_t V;
  const int32_t* hub_id;  // null when the graph has no hub bitmaps
  const uint2 *nsig_out, *nsig_in;  // signature of every list entry
  const uint64_t* bits;
  uint32_t words;
  const uint64_t* h2_off;  // two-hop index or null
  const uint32_t *h2_v, *h2_c;
  const int32_t* h2_did;   // dense rows of the index or null
  const uint16_t* h2_dense;
  const uint32_t* arc_cn;  // per arc (u, w) of adj_out: |N(u) & N(w)|, or null
  const int32_t* hr_id;    // graphs without the two-hop index: hub row of a vertex or -1, or null
  const uint16_t* hr;      // hub rows: row r = |N(hub_r) & N(v)| for every v (V uint16)
};

GView view_of(const DeviceGraph& dg) {
  return GView{dg.off_out, dg.adj_out, dg.off_in, dg.adj_in, dg.sig, dg.V, dg.hub_id, dg.nsig_out, dg.nsig_in, dg.bits,
               dg.words, dg.h2_off, dg.h2_v, dg.h2_c, dg.h2_did, dg.h2_dense, dg.arc_cn, dg.hr_id, dg.hr};
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
// readfirstlane returns int: both halves are taken as uint32 before they
// are widened, or a low word >= 2^31 would be sign-extended into the high one
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// lower-bound membership test of x in adj[lo, hi)
__device__ __forceinline__ bool member_safe(const uint32_t* __restrict__ adj, uint32_t lo, uint32_t hi,
                                            uint32_t x) {
  const uint32_t end = hi;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (adj[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < end && adj[lo] == x;
}

// Demand bytes of one lower-bound search over `len` words: the probes plus
// the final compare.
__device__ __forceinline__ uint32_t bs_bytes(uint32_t len) { return 4u * ((32u - __clz(len)) + 1u); }

// x in out-list(u) of a symmetric graph: one bit test when u is a hub,
// else a lower-bound search of u's sorted list.
__device__ __forceinline__ bool adj_has(const GView& g, uint32_t u, uint32_t x, uint64_t& bytes) {
  if (g.hub_id) {
    const int32_t h = g.hub_id[u];
    bytes += 4;
    if (h >= 0) {
      bytes += 8;
      return (g.bits[(uint64_t)h * g.words + (x >> 6)] >> (x & 63)) & 1ull;
    }
  }
  const uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
  bytes += 8 + bs_bytes(hi - lo);
  return member_safe(g.adj_out, lo, hi, x);
}

// |N(u) & N(w)| for u != w from the two-hop index: a lower-bound search of
// w in u's two-hop list (0 when w is not there).  Every common neighbour of
// two distinct vertices has degree >= 2, so the count serves degree
// thresholds up to 2.
__device__ __forceinline__ uint32_t two_hop(const GView& g, uint32_t u, uint32_t w, uint64_t& bytes) {
  if (g.h2_did) {
    const int32_t dr = g.h2_did[u];
    if (dr >= 0) {
      bytes += 6;
      return g.h2_dense[(uint64_t)dr * g.V + w];
    }
  }
  uint64_t lo = g.h2_off[u], hi = g.h2_off[u + 1];
  const uint64_t end = hi;
  bytes += 16 + 4 * ((64 - __clzll(hi - lo)) + 1);
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (g.h2_v[mid] < w) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && g.h2_v[lo] == w) ? g.h2_c[lo] : 0u;
}

// |N(u) & N(w)| for u != w from a hub row when u or w has one, else ~0u.
// Like the two-hop index it serves degree thresholds up to 2.
__device__ __forceinline__ uint32_t hub_common(const GView& g, uint32_t u, uint32_t w) {
  if (!g.hr_id) return ~0u;
  const int32_t ru = g.hr_id[u];
  if (ru >= 0) return g.hr[(uint64_t)ru * g.V + w];
  const int32_t rw = g.hr_id[w];
  if (rw >= 0) return g.hr[(uint64_t)rw * g.V + u];
  return ~0u;
}

// |N(u) & N(w)| for adjacent u, w from the per-arc table: w's index in u's
// sorted list (a lower-bound search), then one load.  Like the two-hop
// index it serves degree thresholds up to 2.
__device__ __forceinline__ uint32_t arc_common(const GView& g, uint32_t u, uint32_t w, uint64_t& bytes) {
  uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
  bytes += 8 + bs_bytes(hi - lo) + 4;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (g.adj_out[mid] < w) lo = mid + 1;
    else hi = mid;
  }
  return g.arc_cn[lo];
}

// Per-arc common-neighbour counts of a symmetric loop-free graph: for every
// arc (u, w) with u < w, |N(u) & N(w)| by walking the shorter list with a
// membership test in the longer one (a bit test when it belongs to a hub),
// stored at (u, w) and at the mirror arc (w, u).  One thread per arc; the
// arc's source is found by an upper-bound search of the offsets.
__global__ __launch_bounds__(256) void k_arc_common(GView g, uint64_t A, uint32_t* __restrict__ cn) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < A; i += (uint64_t)gridDim.x * 256) {
    uint32_t lo = 0, hi = g.V;  // largest u with off[u] <= i
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (g.off_out[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t u = lo, w = g.adj_out[i];
    if (u >= w) continue;
    const uint32_t du = g.off_out[u + 1] - g.off_out[u], dw = g.off_out[w + 1] - g.off_out[w];
    const uint32_t s = du <= dw ? u : w, o = du <= dw ? w : u;
    uint32_t c = 0;
    uint64_t dummy = 0;
    for (uint32_t j = g.off_out[s]; j < g.off_out[s + 1]; ++j) c += adj_has(g, o, g.adj_out[j], dummy);
    cn[i] = c;
    uint32_t a = g.off_out[w], b = g.off_out[w + 1];
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if (g.adj_out[mid] < u) a = mid + 1;
      else b = mid;
    }
    cn[a] = c;
  }
}

// One block per hub: sets the bits of its neighbours.
__global__ __launch_bounds__(256) void k_hub_bits(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                  const uint32_t* __restrict__ hubs, uint32_t nh, uint32_t words,
                                                  uint64_t* __restrict__ bits) {
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint32_t u = hubs[h];
    uint64_t* row = bits + (uint64_t)h * words;
    for (uint32_t i = off[u] + threadIdx.x; i < off[u + 1]; i += 256) {
      const uint32_t x = adj[i];
      atomicOr((unsigned long long*)&row[x >> 6], 1ull << (x & 63));
    }
  }
}

// Wave-then-block sum of a 64-bit value; one atomic per wave.
__device__ __forceinline__ void add_wave(unsigned long long* dst, uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  // 64-bit sum through two 32-bit reductions with carry handled in 64-bit
  uint64_t s = v;
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
    uint64_t other = ((uint64_t)h2 << 32) | l2;
    s += other;
    lo = (uint32_t)s;
    hi = (uint32_t)(s >> 32);
  }
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

// ---------------------------------------------------------------- level 0
// Level 0 in ascending vertex order: a 0/1 flag per candidate vertex, an
// exclusive scan of the flags, and a scatter.  A deterministic level 0 makes
// a run stopped between trunks correspond to an exact root range.
__global__ __launch_bounds__(BLOCK) void k_level0_flags(const uint32_t* __restrict__ sig, uint32_t tin,
                                                       uint32_t tout, uint32_t rb, uint32_t re, uint32_t rs,
                                                       uint32_t n, uint32_t* __restrict__ flags) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t u = rb + (uint32_t)i * rs;
    flags[i] = u < re && sig[2 * (uint64_t)u] >= tin && sig[2 * (uint64_t)u + 1] >= tout;
  }
}

__global__ __launch_bounds__(BLOCK) void k_level0_scatter(const uint32_t* __restrict__ flags,
                                                         const uint64_t* __restrict__ pos, uint32_t rb, uint32_t rs,
                                                         uint32_t n, uint32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
    if (flags[i]) out[pos[i]] = rb + (uint32_t)i * rs;
}

// Level 0 of at most 64K candidate roots in one launch of one block of
// L0_THREADS: thread t takes a contiguous run of at most 64 candidates and
// keeps a bit per candidate whose signature meets position 0's, the block
// scans the counts in LDS, and each thread writes its run's roots at its
// offset: the rows k_level0_flags, a scan and k_level0_scatter write, in
// the same order.  The total goes to *total in host memory (a system-scope
// store).
constexpr uint32_t L0_THREADS = 1024;
__global__ __launch_bounds__(L0_THREADS) void k_level0_small(const uint32_t* __restrict__ sig, uint32_t tin,
                                                            uint32_t tout, uint32_t rb, uint32_t re, uint32_t rs,
                                                            uint32_t n, uint32_t* __restrict__ out, uint64_t* total,
                                                            unsigned long long* __restrict__ dev_total) {
  __shared__ uint32_t s[L0_THREADS];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (n + L0_THREADS - 1) / L0_THREADS;  // <= 64
  const uint32_t i0 = min(t * per, n), i1 = min(i0 + per, n);
  // eight candidates' signatures in flight at a time
  uint64_t keep = 0;
  for (uint32_t b = i0; b < i1; b += 8) {
    uint2 sg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = b + k, u = rb + i * rs;
      sg[k] = (i < i1 && u < re) ? *(const uint2*)(sig + 2 * (uint64_t)u) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t i = b + k, u = rb + i * rs;
      const bool ok = i < i1 && u < re && sg[k].x >= tin && sg[k].y >= tout;
      keep |= (uint64_t)ok << ((i - i0) & 63);
    }
  }
  const uint32_t c = (uint32_t)__popcll(keep);
  s[t] = c;
  __syncthreads();
  for (uint32_t o = 1; o < L0_THREADS; o <<= 1) {
    const uint32_t v = t >= o ? s[t - o] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  uint32_t pos = s[t] - c;
  for (uint32_t i = i0; i < i1; ++i)
    if ((keep >> (i - i0)) & 1ull) out[pos++] = rb + i * rs;
  if (t == L0_THREADS - 1) {
    __hip_atomic_store(total, (uint64_t)s[t], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (dev_total) *dev_total = s[t];
  }
}

// The run's counters (words 0..13), the last processed root (*last, or
// rows0[lvcnt[0] - 1] when level 0's size stayed on the device), and the
// device-side level counts (16 words, to pub[48..]), written to host memory
// (system-scope stores) for one synchronisation with no copy.
__global__ void k_publish_end(const unsigned long long* __restrict__ ctr, const uint32_t* __restrict__ last,
                              const uint32_t* __restrict__ rows0, const unsigned long long* __restrict__ lvcnt,
                              uint64_t* pub) {
  const uint32_t t = threadIdx.x;
  if (t < 14) {
    __hip_atomic_store(&pub[t], (uint64_t)__hip_atomic_load(&ctr[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  } else if (t == 14 && (last || rows0)) {
    const uint64_t n0 = rows0 ? lvcnt[0] : 1;
    const uint32_t lr = last ? *last : (n0 ? rows0[n0 - 1] : 0u);
    __hip_atomic_store(&pub[14], (uint64_t)lr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  } else if (t >= 32 && t < 32 + kMaxQuery) {
    __hip_atomic_store(&pub[16 + t], (uint64_t)lvcnt[t - 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ------------------------------------------------------------ planning
// Degree-threshold tables of the pair level: |{x in N(u) : deg(x) >= t}| for
// the thresholds of A, B and A & B (null: t <= 1, the plain degree).
struct PairTables {
  const uint32_t *a, *b, *ab;
};

__device__ __forceinline__ uint32_t tbl_or_deg(const GView& g, const uint32_t* tbl, uint32_t u) {
  return tbl ? tbl[u] : g.off_out[u + 1] - g.off_out[u];
}

// The pair level's prefix parts for the prefix r[0..N) of positions
// 0..k-4: |A|, |B| without the candidate of position k-3 (the degree
// tables of u = r[pr_a], w = r[pr_b] minus the prefix vertices inside), and
// for |A & B| the prefix vertices inside both, negated, plus the filtered
// size of N(u) when u = w, else |N(u) & N(w)| from the two-hop index (0
// without it).
template <int N>
__device__ __forceinline__ void pair_bases(const GView& g, const KPlan& P, const uint32_t (&r)[N], const PairTables& pt,
                                           uint32_t& u, uint32_t& w, int64_t& ca, int64_t& cb, int64_t& cab,
                                           uint64_t& bytes) {
  u = 0;
  w = 0;
#pragma unroll
  for (int l = 0; l < N; ++l) {
    if (l == (int)P.pr_a) u = r[l];
    if (l == (int)P.pr_b) w = r[l];
  }
  ca = tbl_or_deg(g, pt.a, u);
  cb = tbl_or_deg(g, pt.b, w);
  cab = 0;
  if (P.pr_a == P.pr_b) cab = (int64_t)tbl_or_deg(g, pt.ab, u);
  else if (g.h2_off && P.pr_tab <= 2) cab = (int64_t)two_hop(g, u, w, bytes);
  else if (g.arc_cn && P.pr_adj && P.pr_tab <= 2) cab = (int64_t)arc_common(g, u, w, bytes);
  bytes += 16 + 4 * N;
#pragma unroll
  for (int l = 0; l < N; ++l) {
    const bool ina = ((P.pr_ma >> l) & 1) || adj_has(g, u, r[l], bytes);
    const bool inb = ((P.pr_mb >> l) & 1) || adj_has(g, w, r[l], bytes);
    const uint32_t d = g.sig[2 * (uint64_t)r[l]];
    ca -= ina && d >= P.pr_ta;
    cb -= inb && d >= P.pr_tb;
    cab -= ina && inb && d >= P.pr_tab;
  }
}

// Per prefix: the constraint whose list is shortest (upper bound on the
// prefix's children).  Ties keep the lowest constraint index.
// With merge (profiling) it adds the merge-model bytes of expanding the
// prefix: 8 per backward neighbour (offset pair) + 4 per element of every
// backward neighbour's list.
// With cbase (fused levels whose closed-form vertex u is an older position)
// it also stores the part of the closed form that only depends on the
// prefix: cnt(u) - |{i < L : r_i in N(u), deg(r_i) >= t}|.
// With pbase (the pair level, L = k-3) it stores the prefix's parts of |A|,
// |B| and |A & B|: the degree-filtered list sizes of u = r[pr_a] and
// w = r[pr_b] minus the prefix vertices inside them, and for |A & B| (as a
// signed word) the prefix vertices inside both, negated, plus -- when u = w --
// the filtered size of N(u), else |N(u) & N(w)| from the two-hop index
// (without it, k_common counts that afterwards, only for prefixes with
// survivors).  When the pair level collapses (pr_collapse) it also counts the
// prefix's completions into counters[0..2] and gives the prefix no work.
// The planner of one prefix row p: the shortest constraint list (minlen,
// minsel), the per-prefix closed-form base (cbase) and the pair level's
// bases (pbase); a collapsing pair level's count is added to acc/surv/sna.
// Returns the list length (0 for a collapsed prefix).
template <int L>
__device__ __forceinline__ uint32_t plan_row(const GView& g, const KPlan& P, const uint32_t* __restrict__ rows,
                                             uint64_t p, uint8_t* __restrict__ minsel, uint32_t* __restrict__ cbase,
                                             const uint32_t* __restrict__ tbl, uint4* __restrict__ pbase,
                                             const PairTables& pt, bool collapse, bool merge, uint64_t& mb,
                                             uint64_t& acc, uint64_t& surv, uint64_t& sna) {
  const uint32_t nj = P.nj[L];
  uint32_t row[L];
#pragma unroll
  for (int l = 0; l < L; ++l) row[l] = rows[p * L + l];
  uint32_t best = 0xFFFFFFFFu, sel = 0;
  uint64_t sum = 0;
  for (uint32_t c = 0; c < nj; ++c) {
    const uint32_t jc = P.jc[L][c];
    const uint32_t pos = jc & 31;
    uint32_t u = 0;
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (l == (int)pos) u = row[l];
    const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
    const uint32_t len = off[u + 1] - off[u];
    sum += len;
    if (len < best) { best = len; sel = c; }
  }
  mb += 8ull * nj + 4ull * sum;
  minsel[p] = (uint8_t)sel;
  if (cbase) {
    uint32_t u = 0;
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (l == (int)P.sc_pos) u = row[l];
    int64_t c = tbl ? (int64_t)tbl[u] : (int64_t)(g.off_out[u + 1] - g.off_out[u]);
    uint64_t dummy = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const bool in = ((P.sc_adjmask >> l) & 1) || adj_has(g, u, row[l], dummy);
      if (in && (P.sc_thr <= 1 || g.sig[2 * (uint64_t)row[l]] >= P.sc_thr)) --c;
    }
    cbase[p] = (uint32_t)c;
  }
  if (pbase) {
    uint32_t u = 0, w = 0;
    int64_t ca = 0, cb = 0, cab = 0;
    uint64_t dummy = 0;
    pair_bases<L>(g, P, row, pt, u, w, ca, cb, cab, dummy);
    pbase[p] = make_uint4((uint32_t)ca, (uint32_t)cb, (uint32_t)(int32_t)cab, 0u);
    if (collapse) {
      // position k-3's survivors are N(u) & N(w) minus the prefix (cab of
      // them: with thresholds <= 2 every common neighbour qualifies), each
      // inside both lists: the prefix's count without expanding it
      const int64_t ns = cab;
      acc += (uint64_t)(ns * ((ca - 1) * (cb - 1) - (cab - 1)));
      surv += (uint64_t)ns;
      sna += (uint64_t)(ns * (ca - 1));
      if (merge) {
        const uint64_t du = g.off_out[u + 1] - g.off_out[u], dw = g.off_out[w + 1] - g.off_out[w];
        mb += (uint64_t)ns * (24 + 4 * du + (uint64_t)(ca - 1) * (24 + 4 * dw));
      }
      best = 0;
    }
  }
  return best;
}

template <int L>
__global__ __launch_bounds__(BLOCK) void k_plan(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                uint32_t* __restrict__ minlen, uint8_t* __restrict__ minsel,
                                                uint32_t* __restrict__ cbase, const uint32_t* __restrict__ tbl,
                                                uint4* __restrict__ pbase, PairTables pt,
                                                unsigned long long* __restrict__ merge,
                                                unsigned long long* __restrict__ counters) {
  const bool collapse = pbase && P.pr_collapse && (g.h2_off || (g.arc_cn && P.pr_adj));
  uint64_t mb = 0, acc = 0, surv = 0, sna = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK)
    minlen[p] = plan_row<L>(g, P, rows, p, minsel, cbase, tbl, pbase, pt, collapse, merge != nullptr, mb, acc, surv,
                            sna);
  if (merge) add_wave(merge, mb);
  if (collapse) {
    add_wave(&counters[0], acc);
    add_wave(&counters[1], surv);
    add_wave(&counters[2], sna);
  }
}


// -------------------------------------------------- exclusive scan (u32->u64)
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const uint32_t lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t lo = __shfl_up((uint32_t)v, o), hi = __shfl_up((uint32_t)(v >> 32), o);
    uint64_t t = ((uint64_t)hi << 32) | lo;
    if (lane >= (uint32_t)o) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread; returns the exclusive
// prefix, *total gets the block sum
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* s_tmp, uint64_t* total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t inc = wave_incl_scan(v);
  if (lane == 63) s_tmp[wave] = inc;
  __syncthreads();
  uint64_t wpre = 0, tot = 0;
  for (int w = 0; w < BLOCK / 64; ++w) {
    if (w < (int)wave) wpre += s_tmp[w];
    tot += s_tmp[w];
  }
  __syncthreads();
  *total = tot;
  return wpre + inc - v;
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                      uint64_t* __restrict__ bsum) {
  __shared__ uint64_t s_tmp[BLOCK / 64];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    uint64_t idx = base + (uint64_t)i * BLOCK + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  uint64_t tot;
  block_excl_scan(s, s_tmp, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of bsum[0..nb) in place; bsum[nb] = total
__global__ __launch_bounds__(BLOCK) void k_scan_partials(uint64_t* __restrict__ bsum, uint64_t nb) {
  __shared__ uint64_t s_tmp[BLOCK / 64];
  uint64_t carry = 0;
  for (uint64_t b0 = 0; b0 < nb; b0 += BLOCK) {
    uint64_t idx = b0 + threadIdx.x;
    uint64_t v = idx < nb ? bsum[idx] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_scan(v, s_tmp, &tot);
    if (idx < nb) bsum[idx] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(BLOCK) void k_scan_down(const uint32_t* __restrict__ in, uint64_t n,
                                                    const uint64_t* __restrict__ bsum, uint64_t nb,
                                                    uint64_t* __restrict__ out) {
  __shared__ uint64_t s_tmp[BLOCK / 64];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  // each thread owns SCAN_IPT consecutive items
  uint32_t v[SCAN_IPT];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_IPT + i;
    v[i] = idx < n ? in[idx] : 0;
    s += v[i];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan(s, s_tmp, &tot) + bsum[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_IPT + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

// k_scan_partials and k_scan_down in one launch for at most SCAN_FUSE_BLOCKS
// tiles: each block sums the tile totals before it itself (a block-wide
// reduction of at most that many words), and the last block writes the
// total.
constexpr uint64_t SCAN_FUSE_BLOCKS = 1024;
__global__ __launch_bounds__(BLOCK) void k_scan_down_fused(const uint32_t* __restrict__ in, uint64_t n,
                                                          const uint64_t* __restrict__ bsum, uint64_t nb,
                                                          uint64_t* __restrict__ out) {
  __shared__ uint64_t s_tmp[BLOCK / 64];
  uint64_t part = 0;
  for (uint64_t i = threadIdx.x; i < blockIdx.x; i += BLOCK) part += bsum[i];
  uint64_t before;
  block_excl_scan(part, s_tmp, &before);
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint32_t v[SCAN_IPT];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_IPT + i;
    v[i] = idx < n ? in[idx] : 0;
    s += v[i];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan(s, s_tmp, &tot) + before;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_IPT + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = before + tot;
}

// Levels of at most SCAN_TILE rows: the whole exclusive scan in one block
// (out[0..n] with out[n] the total), one launch instead of three.
__global__ __launch_bounds__(BLOCK) void k_scan_single(const uint32_t* __restrict__ in, uint64_t n,
                                                      uint64_t* __restrict__ out) {
  __shared__ uint64_t s_tmp[BLOCK / 64];
  uint32_t v[SCAN_IPT];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint64_t idx = (uint64_t)threadIdx.x * SCAN_IPT + i;
    v[i] = idx < n ? in[idx] : 0;
    s += v[i];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan(s, s_tmp, &tot);
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint64_t idx = (uint64_t)threadIdx.x * SCAN_IPT + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) out[n] = tot;
}

// tile_p[t] = prefix owning work item w_begin + t*TILE (t < ntiles),
// tile_p[ntiles] = prefix owning item w_end-1.  Block 0 also clears the
// counters the expansion that follows starts from (its eight tile counters
// and, when given, its output row counter), which saves a fill launch each.
__global__ __launch_bounds__(BLOCK) void k_tile_index(const uint64_t* __restrict__ off, uint64_t n, uint64_t w_begin,
                                                     uint64_t w_end, uint64_t ntiles, uint32_t* __restrict__ tile_p,
                                                     unsigned int* __restrict__ tilectr,
                                                     unsigned int* __restrict__ outctr) {
  if (blockIdx.x == 0) {
    if (threadIdx.x < 8) tilectr[threadIdx.x] = 0u;
    if (outctr && threadIdx.x < 2) outctr[threadIdx.x] = 0u;
  }
  for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t <= ntiles; t += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t w = t < ntiles ? w_begin + t * TILE : w_end - 1;
    uint64_t lo = 0, hi = n + 1;  // upper_bound over off[0..n]
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (off[mid] <= w) lo = mid + 1;
      else hi = mid;
    }
    tile_p[t] = (uint32_t)(lo - 1);
  }
}

// A writing level's rows written and status, to coherent host memory
// (system-scope stores; the row count last, with release order), for a host
// that polls the count word instead of synchronising with the stream.
__device__ __forceinline__ void publish_rows(uint64_t* publish, uint32_t rows, unsigned long long status) {
  __hip_atomic_store(&publish[1], (uint64_t)status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&publish[0], (uint64_t)rows, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The same index without a host readback of the level's work: W = off[n] is
// read on the device.  When it fits (W <= max_items, ceil(W / TILE) + 1 <=
// tile_cap entries, W <= row_cap) the tiles of [0, W) are indexed, dyn =
// {W, ntiles} and W is added to *work; otherwise dyn = {0, 0} -- the
// expansion that reads dyn does nothing -- and bit 0 of *status is set, for
// the host to redo the level with the work it then reads back.  Block 0 also
// clears the expansion's counters, as k_tile_index does.
__global__ __launch_bounds__(BLOCK) void k_tile_index_dyn(const uint64_t* __restrict__ off, uint64_t n,
                                                         uint64_t max_items, uint64_t tile_cap, uint64_t row_cap,
                                                         uint32_t* __restrict__ tile_p,
                                                         unsigned int* __restrict__ tilectr,
                                                         unsigned int* __restrict__ outctr,
                                                         unsigned long long* __restrict__ dyn,
                                                         unsigned long long* __restrict__ status,
                                                         unsigned long long* __restrict__ work, uint64_t* publish) {
  const uint64_t W = off[n];
  const uint64_t nt = (W + TILE - 1) / TILE;
  const bool fits = W <= max_items && nt + 1 <= tile_cap && W <= row_cap;
  if (blockIdx.x == 0) {
    if (threadIdx.x < 8) tilectr[threadIdx.x] = 0u;
    if (outctr && threadIdx.x < 2) outctr[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
      dyn[0] = fits ? W : 0ull;
      dyn[1] = fits ? nt : 0ull;
      if (fits) *work += W;
      else *status |= 1ull;
      // a writing level with no tiles: no expansion block will publish
      if (publish && (!fits || W == 0)) publish_rows(publish, 0u, fits ? 0ull : 1ull);
    }
  }
  if (!fits || W == 0) return;
  for (uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; t <= nt; t += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t w = t < nt ? t * TILE : W - 1;
    uint64_t lo = 0, hi = n + 1;  // upper_bound over off[0..n]
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (off[mid] <= w) lo = mid + 1;
      else hi = mid;
    }
    tile_p[t] = (uint32_t)(lo - 1);
  }
}

// A level of at most SCAN_TILE rows in one block and one launch, in place of
// k_plan + k_scan_single + k_tile_index_dyn: every row planned (thread t
// takes rows t*SCAN_IPT .. +SCAN_IPT-1, the scan's blocked layout), the
// exclusive scan of the list lengths (kept in LDS and written to work_off),
// then the speculative tile index over [0, W) from the LDS copy, with the
// same fit test, dyn / status / work words and counter clears as
// k_tile_index_dyn.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_plan_small(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n_host,
                                                      uint32_t* __restrict__ minlen, uint8_t* __restrict__ minsel,
                                                      uint32_t* __restrict__ cbase, const uint32_t* __restrict__ tbl,
                                                      uint4* __restrict__ pbase, PairTables pt,
                                                      unsigned long long* __restrict__ counters,
                                                      uint64_t* __restrict__ work_off, uint64_t max_items,
                                                      uint64_t tile_cap, uint64_t row_cap, uint32_t* __restrict__ tile_p,
                                                      unsigned int* __restrict__ tilectr,
                                                      unsigned int* __restrict__ outctr,
                                                      unsigned long long* __restrict__ dyn,
                                                      unsigned long long* __restrict__ status,
                                                      unsigned long long* __restrict__ work, uint64_t* publish,
                                                      const unsigned long long* __restrict__ dn,
                                                      unsigned long long* __restrict__ lvc) {
  __shared__ uint64_t s_off[SCAN_TILE + 1];
  __shared__ uint32_t s_len[SCAN_TILE];
  __shared__ uint64_t s_tmp[BLOCK / 64];
  const bool collapse = pbase && P.pr_collapse && (g.h2_off || (g.arc_cn && P.pr_adj));
  // dn non-null: the input level's row count is on the device (the low word
  // of the previous expansion's row counter, read before this launch clears
  // it), recorded in *lvc; more rows than one block plans leave the level
  // undone with the status bit set
  uint64_t n = n_host;
  bool too_big = false;
  if (dn) {
    const uint64_t m = (uint32_t)*dn;
    if (threadIdx.x == 0) *lvc = m;
    too_big = m > (uint64_t)SCAN_TILE;
    n = too_big ? 0 : m;
  }
  uint64_t mb = 0, acc = 0, surv = 0, sna = 0;
  // plan with consecutive threads on consecutive rows (every thread busy
  // for small n), then scan in the blocked layout from LDS
  for (uint64_t p = threadIdx.x; p < n; p += BLOCK) {
    const uint32_t len = plan_row<L>(g, P, rows, p, minsel, cbase, tbl, pbase, pt, collapse, false, mb, acc, surv, sna);
    minlen[p] = len;
    s_len[p] = len;
  }
  __syncthreads();
  uint32_t v[SCAN_IPT];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint64_t p = (uint64_t)threadIdx.x * SCAN_IPT + i;
    v[i] = p < n ? s_len[p] : 0u;
    sum += v[i];
  }
  if (collapse) {
    add_wave(&counters[0], acc);
    add_wave(&counters[1], surv);
    add_wave(&counters[2], sna);
  }
  uint64_t W;
  uint64_t ex = block_excl_scan(sum, s_tmp, &W);
#pragma unroll
  for (int i = 0; i < SCAN_IPT; ++i) {
    const uint64_t p = (uint64_t)threadIdx.x * SCAN_IPT + i;
    if (p < n) {
      work_off[p] = ex;
      s_off[p] = ex;
    }
    ex += v[i];
  }
  if (threadIdx.x == 0) {
    work_off[n] = W;
    s_off[n] = W;
  }
  const uint64_t nt = (W + TILE - 1) / TILE;
  const bool fits = !too_big && W <= max_items && nt + 1 <= tile_cap && W <= row_cap;
  if (threadIdx.x < 8) tilectr[threadIdx.x] = 0u;
  if (outctr && threadIdx.x < 2) outctr[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    dyn[0] = fits ? W : 0ull;
    dyn[1] = fits ? nt : 0ull;
    if (fits) *work += W;
    else *status |= 1ull;
    if (publish && (!fits || W == 0)) publish_rows(publish, 0u, fits ? 0ull : 1ull);
  }
  __syncthreads();
  if (!fits || W == 0) return;
  for (uint64_t t = threadIdx.x; t <= nt; t += BLOCK) {
    const uint64_t w = t < nt ? t * TILE : W - 1;
    uint32_t lo = 0, hi = (uint32_t)n + 1;  // upper_bound over s_off[0..n]
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_off[mid] <= w) lo = mid + 1;
      else hi = mid;
    }
    tile_p[t] = lo - 1;
  }
}

// ------------------------------------------------ closed-form last level
// Number of candidates of the last position given all earlier positions
// r[0..N): |{x in N(u) : deg(x) >= thr}| - |{i : r[i] in N(u), deg(r[i]) >= thr}|
// with u = r[sc_pos] (symmetric graphs only).
// Membership is symmetric (x in N(u) <=> u in N(x)).  When u is the newest
// position -- a different vertex for every work item of a prefix -- the
// search runs in the lists of the older positions instead, which all items
// of the prefix share and which therefore stay in cache.
template <int N>
__device__ __forceinline__ uint64_t closed_form(const GView& g, const KPlan& P, const uint32_t (&r)[N],
                                                const uint32_t* __restrict__ tbl, uint64_t& bytes) {
  uint32_t u = 0;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (i == (int)P.sc_pos) u = r[i];
  const uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
  int64_t c = tbl ? (int64_t)tbl[u] : (int64_t)(hi - lo);
  const bool need_deg = P.sc_thr > 1;
  const bool newest = P.sc_pos == (uint32_t)(N - 1);
  bytes += 8 + (tbl ? 4 : 0);
  // Membership of every position's vertex in N(u) -- or of u in that
  // vertex's list when u is the newest position -- as independent tests
  // issued together: hub rows, then bit words and list bounds, then the
  // lower-bound searches of the other lists in lock-step (one load per
  // search per step), then their final compares.  Positions the query makes
  // adjacent to u are members without a test.
  const uint32_t all = (N >= 32) ? ~0u : ((1u << N) - 1u);
  uint32_t inm = P.sc_adjmask & all;
  const uint32_t need = ~P.sc_adjmask & all;
  int32_t h[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t own = (newest && i != N - 1) ? r[i] : u;
    h[i] = (g.hub_id && ((need >> i) & 1)) ? g.hub_id[own] : -1;
  }
  uint64_t bw[N];
  uint32_t base[N], n[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t own = (newest && i != N - 1) ? r[i] : u;
    const uint32_t val = (newest && i != N - 1) ? u : r[i];
    bw[i] = g.hub_id ? g.bits[h[i] >= 0 ? (uint64_t)h[i] * g.words + (val >> 6) : 0] : 0ull;
    base[i] = g.off_out[own];
    n[i] = g.off_out[own + 1] - base[i];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t val = (newest && i != N - 1) ? u : r[i];
    if (!((need >> i) & 1)) {
      n[i] = 0;
    } else if (h[i] >= 0) {
      if ((bw[i] >> (val & 63)) & 1ull) inm |= 1u << i;
      n[i] = 0;
      bytes += 12;
    } else {
      bytes += (g.hub_id ? 4 : 0) + 8 + bs_bytes(n[i]);
    }
  }
  // lock-step searches: while a list has more than one candidate slot, halve
  // it (base stays <= the value's position < base + n)
  while (true) {
    bool more = false;
#pragma unroll
    for (int i = 0; i < N; ++i) more = more || n[i] > 1;
    if (!more) break;
    uint32_t w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = g.adj_out[n[i] > 1 ? base[i] + (n[i] >> 1) : 0];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const uint32_t val = (newest && i != N - 1) ? u : r[i];
      if (n[i] > 1) {
        const uint32_t half = n[i] >> 1;
        if (w[i] <= val) base[i] += half;
        n[i] -= half;
      }
    }
  }
  {
    uint32_t w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = g.adj_out[n[i] ? base[i] : 0];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const uint32_t val = (newest && i != N - 1) ? u : r[i];
      if (n[i] && w[i] == val) inm |= 1u << i;
    }
  }
  if (need_deg) {
    uint32_t d[N];
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = ((inm >> i) & 1) ? g.sig[2 * (uint64_t)r[i]] : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if ((inm >> i) & 1) {
        bytes += 4;
        if (d[i] >= P.sc_thr) --c;
      }
  } else {
    c -= __popc(inm);
  }
  return (uint64_t)c;
}

template <int L>
__global__ __launch_bounds__(BLOCK) void k_closed_form(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                      const uint32_t* __restrict__ tbl,
                                                      unsigned long long* __restrict__ counters, int stats) {
  uint64_t acc = 0, bytes = 0, mb = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    uint32_t r[L];
#pragma unroll
    for (int l = 0; l < L; ++l) r[l] = rows[p * L + l];
    bytes += 4 * L;
    acc += closed_form<L>(g, P, r, tbl, bytes);
    if (stats) {
      // merge model: the one backward list of the last position
      uint32_t u = 0;
#pragma unroll
      for (int l = 0; l < L; ++l)
        if (l == (int)P.sc_pos) u = r[l];
      mb += 8 + 4ull * (g.off_out[u + 1] - g.off_out[u]);
    }
  }
  add_wave(&counters[0], acc);
  if (stats) {
    add_wave(&counters[3], bytes);
    add_wave(&counters[8], mb);
  }
}

// Degree-threshold table: tbl[u] = |{x in N(u) : deg(x) >= thr}|, one wave per vertex.
__global__ __launch_bounds__(BLOCK) void k_cnt_table(GView g, uint32_t thr, uint32_t* __restrict__ tbl) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / 64);
  for (uint64_t u = (uint64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); u < g.V; u += nw) {
    const uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
    uint32_t c = 0;
    for (uint32_t i = lo + lane; i < hi; i += 64) c += g.sig[2 * (uint64_t)g.adj_out[i]] >= thr;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) tbl[u] = c;
  }
}

// ------------------------------------------------------------ chain level
struct ChainTables {
  const uint64_t* H;           // 2V, k_chain_table
  const uint32_t *ta, *tb;     // degree-threshold tables (null: plain degree)
};

// Completions of the last two positions as a chain for the prefix r[0..N)
// of positions 0..k-3 (N = k-2), u = r[ch_u]:
//   H(u) - sum_{r_i in N(u), deg >= ta} cnt_tb(r_i) - [deg u >= tb] |A|
//        - sum_{i != ch_u, deg(r_i) >= tb} (|N(u) & N(r_i)| - c_i)
// with |A| = cnt_ta(u) - |{i : r_i in N(u), deg(r_i) >= ta}| (returned in
// na) and c_i the prefix vertices inside N(u) & N(r_i) meeting ta.  The
// intersections of positions < N-1 come summed in `shared`; the one with
// r[N-1] is looked up in the two-hop index when h2, else left out (a later
// pass subtracts it).  Adds the demand bytes of its loads to bytes and, with
// stats, to mb the merge-model bytes of expanding the prefix by u's list and
// each of its |A| children by its own list.
template <int N>
__device__ __forceinline__ int64_t chain_closed(const GView& g, const KPlan& P, const uint32_t (&r)[N],
                                                const ChainTables& t, uint32_t shared, bool h2, int64_t& na,
                                                bool stats, uint64_t& mb, uint64_t& bytes) {
  uint64_t& dummy = bytes;
  bytes += 4 * N + 8 + 8 + (t.ta ? 4 : 0);  // degrees of the row, u's offsets, H(u), cnt_ta(u)
  uint32_t u = 0, d[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    if (l == (int)P.ch_u) u = r[l];
    d[l] = g.sig[2 * (uint64_t)r[l]];
  }
  const uint32_t du = g.off_out[u + 1] - g.off_out[u];
  // prefix vertices inside N(u) that meet ta
  uint32_t inu = 0;
#pragma unroll
  for (int l = 0; l < N; ++l) {
    const bool in = l != (int)P.ch_u && (((P.qadj[P.ch_u] >> l) & 1) || adj_has(g, u, r[l], dummy));
    if (in && d[l] >= P.ch_ta) inu |= 1u << l;
  }
  na = (int64_t)(t.ta ? t.ta[u] : du) - __popc(inu);
  int64_t c = (int64_t)t.H[u];
  uint64_t dsum = 0;
#pragma unroll
  for (int l = 0; l < N; ++l)
    if ((inu >> l) & 1) {
      c -= t.tb ? t.tb[r[l]] : g.off_out[r[l] + 1] - g.off_out[r[l]];
      dsum += g.off_out[r[l] + 1] - g.off_out[r[l]];
      bytes += t.tb ? 4 : 8;
    }
  if (P.ch_tb <= 1 || du >= P.ch_tb) c -= na;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i == (int)P.ch_u || d[i] < P.ch_tb) continue;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (j == i || !((inu >> j) & 1)) continue;
      if (((P.qadj[i] >> j) & 1) || adj_has(g, r[i], r[j], dummy)) ++c;
    }
  }
  c -= (int64_t)shared;
  if (h2 && (P.ch_tb <= 1 || d[N - 1] >= P.ch_tb)) c -= two_hop(g, u, r[N - 1], dummy);
  if (stats) mb += 8 + 4ull * du + 24ull * (uint64_t)na + 4ull * (t.H[g.V + u] - dsum);
  return c;
}

// Chain level over rows of L = k-3 words whose position-(k-3) candidates are
// their parent's block of `cand` (built once per parent, constraints and
// threshold checked): one work item per (row, candidate); a candidate equal
// to the row's newest vertex is skipped (injectivity), every other one adds
// the chain count of the row + candidate.  Tiles as in k_expand.
struct SibArgs {
  const uint32_t* rows;
  const uint32_t* cand;
  const uint32_t* shared;      // per row: its own positions' intersections with u
  const uint4* pb;             // per row: the chain closed form's row part (k_sib_base)
  const uint64_t* rstart;      // per row: start of its parent's block in cand
  const uint64_t* work_off;
  const uint32_t* tile_p;
  uint64_t w_begin, w_end, ntiles;
  unsigned long long* counters;
  unsigned int* tilectr;
  uint32_t xcd, stats;
};

// ------------------------------------------------------------ expansion
// Adds to cnt[p] the number of this wave's lanes with `ok` set whose item
// belongs to prefix p.  Lanes hold consecutive work items, so the items of a
// prefix are one run of lanes; the first lane of each run adds the run's
// count (one atomic per prefix and wave instead of one per survivor).
// Called by all lanes of the wave.
__device__ __forceinline__ void add_runs(uint32_t* __restrict__ cnt, uint64_t p, bool ok) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t okm = __ballot(ok);
  if (okm == 0) return;
  const uint32_t pl = (uint32_t)p;
  const uint32_t prev = __shfl_up(pl, 1);
  const bool head = lane == 0 || prev != pl;
  const uint64_t hm = __ballot(head);
  const uint64_t above = lane == 63 ? 0ull : (hm >> (lane + 1)) << (lane + 1);
  const uint64_t upto = above ? (above & (~above + 1)) - 1 : ~0ull;
  const uint64_t mine = okm & upto & (~0ull << lane);
  if (head && mine) atomicAdd(&cnt[p], (uint32_t)__popcll(mine));
}

enum { M_WRITE = 0, M_COUNT = 1, M_FUSED = 2, M_PAIR = 3, M_CHAIN = 4, M_COLLAPSE = 5, M_GWRITE = 6 };

struct ExpandArgs {
  const uint32_t* rows;      // n rows of L words
  const uint64_t* work_off;  // n+1
  const uint8_t* minsel;
  const uint32_t* tile_p;    // ntiles+1
  uint64_t w_begin, w_end, ntiles;
  uint32_t* out_rows;        // M_WRITE: rows of L+1 words
  uint32_t* out_count;
  unsigned long long* counters;
  const uint32_t* tbl;       // M_FUSED closed-form table (null: plain degree)
  const uint32_t* cbase;     // M_FUSED with an older closed-form vertex: per-prefix part
  const uint4* pbase;        // M_PAIR: per-prefix parts of |A|, |B|, |A & B|
  uint32_t* nsurv;           // M_PAIR with u != w: survivors per prefix (zero at launch)
  uint32_t* out_parent;      // M_WRITE: parent prefix of every row written (or null)
  const uint32_t* shared;    // M_CHAIN: per prefix, its older positions' intersections with u
  ChainTables ctab;          // M_CHAIN
  PairTables ptab;           // M_COLLAPSE
  const uint64_t* coff;      // M_GWRITE: start of every prefix's block in cand
  uint32_t* cand;            // M_GWRITE: candidates grouped by prefix
  uint32_t xcd;              // 1: XCD-contiguous tile mapping
  unsigned int* tilectr;     // 8 tile counters, zero at launch
  uint32_t stats;            // 1: accumulate demand bytes into counters[3]
  // non-null: {w_end, ntiles} of this launch, written on the device by
  // k_tile_index_dyn (w_begin is 0), in place of w_end and ntiles
  const unsigned long long* dyn;
  // M_WRITE, non-null: the last block to finish writes the rows written
  // (*out_count) and status[0] to publish[0..1] in host memory; done counts
  // the finished blocks (zero at launch, left zero)
  uint64_t* publish;
  unsigned int* done;
  const unsigned long long* status;
};

// Lock-step membership tests of G values, one per item, each in its own
// sorted list adj[base .. base + n): the halvings of all G searches go out
// together (one load per search per step) instead of one search after
// another.  found[j]: n[j] > 0 and val[j] is in its list.
template <int G>
__device__ __forceinline__ void members_lockstep(const uint32_t* __restrict__ adj, uint32_t (&base)[G],
                                                 uint32_t (&n)[G], const uint32_t (&val)[G], bool (&found)[G]) {
  while (true) {
    bool more = false;
#pragma unroll
    for (int j = 0; j < G; ++j) more = more || n[j] > 1;
    if (!more) break;
    uint32_t w[G];
#pragma unroll
    for (int j = 0; j < G; ++j) w[j] = adj[n[j] > 1 ? base[j] + (n[j] >> 1) : 0];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (n[j] > 1) {
        const uint32_t half = n[j] >> 1;
        if (w[j] <= val[j]) base[j] += half;
        n[j] -= half;
      }
    }
  }
  uint32_t w[G];
#pragma unroll
  for (int j = 0; j < G; ++j) w[j] = adj[n[j] ? base[j] : 0];
#pragma unroll
  for (int j = 0; j < G; ++j) found[j] = n[j] && w[j] == val[j];
}

// The next tile of a block, computed by wave 0 (the same value in every
// lane; ~0 once the block is done).  A launch with no more tiles than
// blocks maps each block to one tile statically, no tile-counter atomic at
// all, keeping each XCD's tiles contiguous: block b runs on XCD b mod 8, the
// c_x = ceil((ntiles - x) / 8) blocks of XCD x take tiles P_x .. P_x + c_x
// - 1 with P_x = c_0 + .. + c_{x-1}.  Otherwise tiles are fetched
// dynamically: block b works first on XCD (b mod 8)'s contiguous share of
// the tiles (neighbouring prefixes share adjacency lists in that XCD's L2)
// and, once that share is drained, takes tiles from the other shares, one
// atomic per 1,024-item tile and no static imbalance between blocks.  Lane
// 0's atomic is nested in wave 0's wave-uniform branch and its result
// broadcast.  `tries` starts at 0 (then: the share offset being drained).
__device__ __forceinline__ uint64_t next_tile(unsigned int* tilectr, uint64_t ntiles, uint32_t nx, uint32_t myx,
                                              uint32_t& tries) {
  if (ntiles <= gridDim.x) {
    if (tries) return ~0ull;
    tries = 1;
    const uint32_t b = blockIdx.x;
    if (b >= ntiles) return ~0ull;
    if (nx == 1) return b;
    const uint32_t x = b & 7u;
    uint64_t p = 0;
    for (uint32_t y = 0; y < x; ++y) p += (ntiles - y + 7) / 8;  // y < x <= b < ntiles
    return p + (b >> 3);
  }
  const uint32_t lane = threadIdx.x & 63;
  while (tries < nx) {
    const uint32_t xx = (myx + tries) % nx;
    const uint64_t lo = ntiles * xx / nx, hi = ntiles * (xx + 1) / nx;
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(&tilectr[xx], 1u);
    const uint64_t i = uni(got);
    if (lo + i < hi) return lo + i;
    // this share is drained: one counter load per lane finds the shares
    // with tiles left (in the same order), so a block's last fetch costs
    // two round trips rather than one per share.  A share once drained
    // stays drained; a stale probe only costs another failed atomic.
    bool left = false;
    if (lane < nx) {
      const uint32_t yy = (myx + lane) % nx;
      const uint64_t c = __hip_atomic_load(&tilectr[yy], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      left = ntiles * yy / nx + c < ntiles * (yy + 1) / nx;
    }
    const uint64_t m = __ballot(left);
    tries = m ? (uint32_t)__builtin_ctzll(m) : nx;
  }
  return ~0ull;
}

template <int L, int MODE>
__global__ __launch_bounds__(BLOCK) void k_expand(GView g, KPlan P, ExpandArgs a) {
  __shared__ uint64_t s_off[LDS_OFF + 1];
  __shared__ uint32_t s_w[BLOCK / 64], s_pre[BLOCK / 64], s_base;
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  // threads per block: BLOCK for the writing expansion; 64, 128 or 256 for
  // the counting ones (a tile is TILE items either way, ipt per lane)
  const uint32_t NT = MODE == M_WRITE ? (uint32_t)BLOCK : (uint32_t)blockDim.x;
  const uint32_t ipt = TILE / NT;
  const uint64_t w_end_ = a.dyn ? a.dyn[0] : a.w_end;
  const uint64_t ntiles_ = a.dyn ? a.dyn[1] : a.ntiles;
  // a grid sized before the tile count was known: blocks beyond it leave
  // at once (before any barrier or tile-counter atomic)
  if (blockIdx.x >= ntiles_) return;
  const uint32_t nj = P.nj[L];
  const uint32_t tin = P.sig_in[L], tout = P.sig_out[L];
  uint64_t acc = 0, valid = 0, pairs = 0, trips = 0, bytes = 0, mb = 0;
  // Tiles (next_tile): static for a launch with no more tiles than blocks,
  // else dynamic and XCD-aware.  The fetch runs in wave 0 under a
  // wave-uniform branch, so the branches at the edges of this barrier loop
  // are all wave-uniform.
  __shared__ uint64_t s_tile;
  const uint32_t nx = a.xcd ? 8u : 1u;
  const uint32_t myx = a.xcd ? (blockIdx.x & 7u) : 0u;
  uint32_t tries = 0;
  while (true) {
    if (uni(tid >> 6) == 0) s_tile = next_tile(a.tilectr, ntiles_, nx, myx, tries);  // same value from every lane
    __syncthreads();
    const uint64_t tile = uni64(s_tile);
    if (tile == ~0ull) break;
    {
    const uint64_t w0 = a.w_begin + tile * TILE;
    const uint64_t w1 = min(w0 + (uint64_t)TILE, w_end_);
    const uint32_t plo = a.tile_p[tile], phi = a.tile_p[tile + 1];
    const uint32_t span = phi - plo + 1;  // offsets plo .. phi+1
    const bool lds = span <= LDS_OFF;
    if (lds)
      for (uint32_t i = tid; i <= span; i += NT) s_off[i] = a.work_off[plo + i];
    __syncthreads();
    {
      // Each thread's items are processed in lock-step stages (search,
      // prefix data, list word, signature, constraints, closed form or
      // output), G items at a time, so the loads of different items are
      // independent and in flight together instead of one dependent chain
      // after another.  Items past w1 are clamped to the tile's last prefix
      // and masked; their list word is never read.
      constexpr int G = (MODE == M_FUSED) ? ((L <= 6) ? FUSED_G_SMALL : (L <= 10 ? FUSED_G_LARGE : 1))
                        : (MODE != M_COUNT && MODE != M_GWRITE && MODE != M_WRITE)
                            ? ((L <= 6) ? 4 : 2)
                            : ((L <= 6) ? 4 : (L <= 10 ? COUNT_G_LARGE : 1));
      for (uint32_t g0 = 0; g0 < ipt; g0 += G) {
        uint64_t w[G];
        uint32_t pi[G];
        bool in[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          w[j] = w0 + (uint64_t)(g0 + j) * NT + tid;
          in[j] = w[j] < w1;
          pi[j] = 0;
        }
        // largest pi with off(pi) <= w: a fixed number of halvings for the
        // whole block (off(0) <= w0 <= w and off(span) >= w1)
        for (uint32_t n = span + 1; n > 1;) {
          const uint32_t half = n >> 1;
#pragma unroll
          for (int j = 0; j < G; ++j) {
            const uint64_t o = lds ? s_off[pi[j] + half] : a.work_off[plo + pi[j] + half];
            pi[j] = (o <= w[j]) ? pi[j] + half : pi[j];
          }
          n -= half;
        }
        uint32_t row[G][L], sel[G], jc[G], idx[G], lb[G], v[G], cb[G];
        uint4 pb[G];
        bool ok[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const uint32_t pj = min(pi[j], span - 1);
          const uint64_t p = (uint64_t)plo + pj;
          const uint64_t pbase = lds ? s_off[pj] : a.work_off[p];
          idx[j] = (uint32_t)(w[j] - pbase);
#pragma unroll
          for (int l = 0; l < L; ++l) row[j][l] = a.rows[p * L + l];
          sel[j] = a.minsel[p];
          cb[j] = (MODE == M_FUSED && a.cbase) ? a.cbase[p] : 0u;
          pb[j] = MODE == M_PAIR ? a.pbase[p] : make_uint4(0u, 0u, 0u, 0u);
          if (a.stats) {
            bytes += 12;
            if (in[j] && idx[j] == 0) bytes += 4 * L + 9 + 8 * nj;
          }
        }
        // hub rows of the first constraints' vertices: they depend on the
        // row only, so their loads go out with the list base's instead of
        // after the candidate's
        constexpr int NJP = 3;
        int32_t hpre[NJP][G];
#pragma unroll
        for (int cc = 0; cc < NJP; ++cc) {
          const uint32_t jcc = cc < (int)nj ? P.jc[L][cc] : 0u;
#pragma unroll
          for (int j = 0; j < G; ++j) {
            uint32_t uu = 0;
#pragma unroll
            for (int l = 0; l < L; ++l)
              if (l == (int)(jcc & 31)) uu = row[j][l];
            hpre[cc][j] = (g.hub_id && cc < (int)nj) ? g.hub_id[uu] : -1;
          }
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          jc[j] = P.jc[L][sel[j]];
          uint32_t u = 0;
#pragma unroll
          for (int l = 0; l < L; ++l)
            if (l == (int)(jc[j] & 31)) u = row[j][l];
          lb[j] = ((jc[j] >> 7) ? g.off_in : g.off_out)[u];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) v[j] = in[j] ? ((jc[j] >> 7) ? g.adj_in : g.adj_out)[lb[j] + idx[j]] : 0u;
        uint2 sg[G];
#pragma unroll
        for (int j = 0; j < G; ++j) sg[j] = in[j] ? ((jc[j] >> 7) ? g.nsig_in : g.nsig_out)[lb[j] + idx[j]] : make_uint2(0u, 0u);
#pragma unroll
        for (int j = 0; j < G; ++j) {
          ok[j] = in[j] && sg[j].x >= tin && sg[j].y >= tout;
#pragma unroll
          for (int l = 0; l < L; ++l) ok[j] = ok[j] && v[j] != row[j][l];
        }
        for (uint32_t c = 0; c < nj; ++c) {
          const uint32_t jc2 = P.jc[L][c];
          uint32_t u2[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            u2[j] = 0;
#pragma unroll
            for (int l = 0; l < L; ++l)
              if (l == (int)(jc2 & 31)) u2[j] = row[j][l];
          }
          // bit tests for hub lists; the other items' lower-bound searches
          // in lock-step (symmetric graphs with hubs: every constraint is an
          // out-list)
          const bool hubs = g.hub_id != nullptr;
          const uint32_t* off2 = (!hubs && (jc2 >> 7)) ? g.off_in : g.off_out;
          const uint32_t* adj2 = (!hubs && (jc2 >> 7)) ? g.adj_in : g.adj_out;
          bool srch[G];
          if (hubs) {
            int32_t h[G];
            uint64_t bw[G];
#pragma unroll
            for (int j = 0; j < G; ++j) {
              int32_t hh = 0;
#pragma unroll
              for (int cc = 0; cc < NJP; ++cc)
                if ((int)c == cc) hh = hpre[cc][j];
              h[j] = c < (uint32_t)NJP ? hh : g.hub_id[u2[j]];
            }
#pragma unroll
            for (int j = 0; j < G; ++j)
              bw[j] = g.bits[h[j] >= 0 ? (uint64_t)h[j] * g.words + (v[j] >> 6) : 0];
#pragma unroll
            for (int j = 0; j < G; ++j) {
              srch[j] = false;
              if (!ok[j] || c == sel[j]) continue;
              if (h[j] >= 0) {
                ok[j] = (bw[j] >> (v[j] & 63)) & 1ull;
                if (a.stats) bytes += 12;
              } else {
                srch[j] = true;
              }
            }
          } else {
#pragma unroll
            for (int j = 0; j < G; ++j) srch[j] = ok[j] && c != sel[j];
          }
          uint32_t sb[G], sn[G];
          bool fnd[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            sb[j] = srch[j] ? off2[u2[j]] : 0u;
            sn[j] = srch[j] ? off2[u2[j] + 1] - sb[j] : 0u;
            if (a.stats && srch[j]) bytes += (hubs ? 12 : 8) + bs_bytes(sn[j]);
          }
          members_lockstep<G>(adj2, sb, sn, v, fnd);
#pragma unroll
          for (int j = 0; j < G; ++j)
            if (srch[j]) ok[j] = fnd[j];
        }
        if (MODE == M_WRITE) {
          // survivors of the G items: one block-wide reservation on
          // out_count, rows in ballot order within each wave
          uint64_t m[G];
          uint32_t cnt = 0;
#pragma unroll
          for (int j = 0; j < G; ++j) {
            m[j] = __ballot(ok[j]);
            cnt += (uint32_t)__popcll(m[j]);
          }
          if ((tid & 63) == 0) s_w[wave] = cnt;
          __syncthreads();
          if (tid == 0) {
            uint32_t t = 0;
            for (int ww = 0; ww < BLOCK / 64; ++ww) {
              s_pre[ww] = t;
              t += s_w[ww];
            }
            s_base = t ? atomicAdd(a.out_count, t) : 0;
          }
          __syncthreads();
          uint64_t o = (uint64_t)s_base + s_pre[wave];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (ok[j]) {
              const uint64_t at = o + lane_rank(m[j]);
              uint32_t* dst = a.out_rows + at * (L + 1);
#pragma unroll
              for (int l = 0; l < L; ++l) dst[l] = row[j][l];
              dst[L] = v[j];
              if (a.out_parent) a.out_parent[at] = plo + min(pi[j], span - 1);
              if (a.stats) bytes += 4 * (L + 1);
            }
            o += (uint64_t)__popcll(m[j]);
          }
        } else if (MODE == M_COUNT) {
#pragma unroll
          for (int j = 0; j < G; ++j) acc += ok[j];
        } else if (MODE == M_GWRITE) {
          // survivors into their prefix's block (order within a block free)
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            const uint64_t p = (uint64_t)plo + min(pi[j], span - 1);
            a.cand[a.coff[p] + atomicAdd(&a.nsurv[p], 1u)] = v[j];
            if (a.stats) bytes += 16;  // block start, slot counter, candidate
          }
        } else if (MODE == M_COLLAPSE) {
          // a survivor v completes the prefix of positions 0..k-4 of a
          // collapsing pair level: its |A & B| candidates of position k-3
          // each add (|A|-1)(|B|-1) - (|A & B|-1)
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            ++valid;
            uint32_t r[L + 1];
#pragma unroll
            for (int l = 0; l < L; ++l) r[l] = row[j][l];
            r[L] = v[j];
            uint32_t u, w;
            int64_t ca, cb, cab;
            pair_bases<L + 1>(g, P, r, a.ptab, u, w, ca, cb, cab, bytes);
            acc += (uint64_t)(cab * ((ca - 1) * (cb - 1) - (cab - 1)));
            pairs += (uint64_t)cab;
            trips += (uint64_t)(cab * (ca - 1));
            if (a.stats) {
              // merge model: the survivor's node and its expansion by the
              // lists of u and w; each of its cab children's node and
              // expansion by u's list; each of their (ca-1) children's node
              // and expansion by w's list
              const uint64_t du = g.off_out[u + 1] - g.off_out[u], dw = g.off_out[w + 1] - g.off_out[w];
              mb += 16 + 16 + 4 * (du + dw) + (uint64_t)cab * (24 + 4 * du + (uint64_t)(ca - 1) * (24 + 4 * dw));
            }
          }
        } else if (MODE == M_CHAIN) {
          // a survivor v completes the prefix of positions 0..k-3, whose
          // chain of the last two positions is counted in closed form
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            ++valid;
            uint32_t r[L + 1];
#pragma unroll
            for (int l = 0; l < L; ++l) r[l] = row[j][l];
            r[L] = v[j];
            const uint32_t sh = a.shared[plo + min(pi[j], span - 1)];
            int64_t na = 0;
            acc += (uint64_t)chain_closed<L + 1>(g, P, r, a.ctab, sh, true, na, a.stats != 0, mb, bytes);
            pairs += (uint64_t)na;
            if (a.stats) mb += 16;  // the survivor's node, written and read back
          }
        } else if (MODE == M_PAIR) {
          // a survivor v completes the prefix of positions 0..k-3; its
          // |A| * |B| - |A & B| completions take the prefix's bases minus v's
          // own membership (certain when v's position is adjacent to pr_a /
          // pr_b in the query, a bit test or a search otherwise)
          const bool ca = (P.pr_ma >> L) & 1, cbb = (P.pr_mb >> L) & 1;
          uint32_t ua[G], ub[G];
          int32_t ha[G], hb[G];
          uint64_t bwa[G], bwb[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            ua[j] = 0;
            ub[j] = 0;
#pragma unroll
            for (int l = 0; l < L; ++l) {
              if (l == (int)P.pr_a) ua[j] = row[j][l];
              if (l == (int)P.pr_b) ub[j] = row[j][l];
            }
            ha[j] = (!ca && g.hub_id) ? g.hub_id[ua[j]] : -1;
            hb[j] = (!cbb && g.hub_id) ? g.hub_id[ub[j]] : -1;
          }
#pragma unroll
          for (int j = 0; j < G; ++j) {
            bwa[j] = ha[j] >= 0 ? g.bits[(uint64_t)ha[j] * g.words + (v[j] >> 6)] : 0;
            bwb[j] = hb[j] >= 0 ? g.bits[(uint64_t)hb[j] * g.words + (v[j] >> 6)] : 0;
          }
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            ++valid;
            bool ina = ca, inb = cbb;
            if (!ina) {
              if (ha[j] >= 0) {
                ina = (bwa[j] >> (v[j] & 63)) & 1ull;
              } else {
                const uint32_t lo2 = g.off_out[ua[j]], hi2 = g.off_out[ua[j] + 1];
                ina = member_safe(g.adj_out, lo2, hi2, v[j]);
                if (a.stats) bytes += 8 + bs_bytes(hi2 - lo2);
              }
            }
            if (!inb) {
              if (hb[j] >= 0) {
                inb = (bwb[j] >> (v[j] & 63)) & 1ull;
              } else {
                const uint32_t lo2 = g.off_out[ub[j]], hi2 = g.off_out[ub[j] + 1];
                inb = member_safe(g.adj_out, lo2, hi2, v[j]);
                if (a.stats) bytes += 8 + bs_bytes(hi2 - lo2);
              }
            }
            const uint32_t dv = sg[j].x;
            const uint64_t na = pb[j].x - (uint32_t)(ina && dv >= P.pr_ta);
            const uint64_t nb = pb[j].y - (uint32_t)(inb && dv >= P.pr_tb);
            // |A & B|: without the two-hop index and u != w, N(u) & N(w)
            // itself is subtracted per prefix afterwards by k_common;
            // modulo-2^64 sums are exact
            const int64_t nab = (int64_t)(int32_t)pb[j].z - (int64_t)(ina && inb && dv >= P.pr_tab);
            acc += na * nb - (uint64_t)nab;
            pairs += na;
            if (a.stats) {
              bytes += 16 + ((ha[j] >= 0) ? 12 : 0) + ((hb[j] >= 0) ? 12 : 0);
              // merge model: the survivor's node and its expansion by the
              // list of pr_a; each of its na children's node and expansion
              // by the list of pr_b
              mb += 16 + 8 + 4ull * (g.off_out[ua[j] + 1] - g.off_out[ua[j]]) +
                    na * (16 + 8 + 4ull * (g.off_out[ub[j] + 1] - g.off_out[ub[j]]));
            }
          }
          if (a.nsurv) {
#pragma unroll
            for (int j = 0; j < G; ++j) add_runs(a.nsurv, plo + min(pi[j], span - 1), ok[j]);
          }
        } else if (a.cbase) {
          // closed form = per-prefix base - [v in N(u), deg(v) >= t]; v is
          // certainly in N(u) when it was drawn from u's own list
          uint32_t us[G];
          int32_t h[G];
          uint64_t bw[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            us[j] = 0;
#pragma unroll
            for (int l = 0; l < L; ++l)
              if (l == (int)P.sc_pos) us[j] = row[j][l];
            h[j] = g.hub_id ? g.hub_id[us[j]] : -1;
          }
#pragma unroll
          for (int j = 0; j < G; ++j) bw[j] = h[j] >= 0 ? g.bits[(uint64_t)h[j] * g.words + (v[j] >> 6)] : 0;
          // the other survivors' searches of v in N(u), in lock-step
          uint32_t sb[G], sn[G];
          bool fnd[G];
#pragma unroll
          for (int j = 0; j < G; ++j) {
            const bool srch = ok[j] && (jc[j] & 31) != P.sc_pos && h[j] < 0;
            sb[j] = srch ? g.off_out[us[j]] : 0u;
            sn[j] = srch ? g.off_out[us[j] + 1] - sb[j] : 0u;
            if (a.stats && srch) bytes += 8 + bs_bytes(sn[j]);
          }
          members_lockstep<G>(g.adj_out, sb, sn, v, fnd);
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            ++valid;
            bool inn = (jc[j] & 31) == P.sc_pos;
            if (!inn) inn = h[j] >= 0 ? (bool)((bw[j] >> (v[j] & 63)) & 1ull) : fnd[j];
            const bool cnt_v = inn && (P.sc_thr <= 1 || sg[j].x >= P.sc_thr);
            acc += (uint64_t)cb[j] - (cnt_v ? 1 : 0);
            if (a.stats) {
              bytes += 8;
              // merge model: the survivor's trie node (written and read back)
              // and the last position's one backward list
              mb += 16 + 8 + 4ull * (g.off_out[us[j] + 1] - g.off_out[us[j]]);
            }
          }
        } else {
#pragma unroll
          for (int j = 0; j < G; ++j) {
            if (!ok[j]) continue;
            ++valid;
            uint32_t r[L + 1];
#pragma unroll
            for (int l = 0; l < L; ++l) r[l] = row[j][l];
            r[L] = v[j];
            acc += closed_form<L + 1>(g, P, r, a.tbl, bytes);
            if (a.stats) {
              uint32_t us = v[j];
#pragma unroll
              for (int l = 0; l < L; ++l)
                if (l == (int)P.sc_pos) us = row[j][l];
              mb += 16 + 8 + 4ull * (g.off_out[us + 1] - g.off_out[us]);
            }
          }
        }
      }
    }
    __syncthreads();
    }
  }
  if constexpr (MODE == M_WRITE) {
    if (a.publish && tid == 0) {
      // every block that took part (the grid, or one block per tile when
      // there are fewer tiles) counts itself once its reservations on
      // out_count are done; the last one publishes the total
      const unsigned int parts = (unsigned int)min<uint64_t>(ntiles_, gridDim.x);
      const unsigned int prev = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (prev + 1 == parts) {
        const uint32_t m = __hip_atomic_load(a.out_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long st = __hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&a.publish[1], (uint64_t)st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&a.publish[0], (uint64_t)m, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
  }
  if (MODE != M_WRITE) {
    add_wave(&a.counters[0], acc);
    if (MODE == M_FUSED || MODE == M_PAIR || MODE == M_CHAIN || MODE == M_COLLAPSE) add_wave(&a.counters[1], valid);
    if (MODE == M_PAIR || MODE == M_CHAIN || MODE == M_COLLAPSE) add_wave(&a.counters[2], pairs);
    if (MODE == M_COLLAPSE) add_wave(&a.counters[9], trips);
  }
  if (a.stats) {
    add_wave(&a.counters[3], bytes);
    if (MODE == M_FUSED || MODE == M_PAIR || MODE == M_CHAIN || MODE == M_COLLAPSE) add_wave(&a.counters[8], mb);
  }
}

// Per-row part of the chain closed form for the rows of level k-4 (L = k-3
// words) whose position-(k-3) candidates are their parent's block: every
// term of chain_closed that does not involve the candidate v, computed once
// per row instead of once per (row, candidate).  With p = L the candidate's
// position, u = r[ch_u], inu_old / Dold the older positions inside N(u)
// meeting ta / meeting tb, the row's base is
//   H(u) - sum_{inu_old} cnt_tb(r_l) - [deg u >= tb] na_old
//        + sum_{i in Dold} sum_{j in inu_old, j != i} [r_i ~ r_j] - shared
// and a candidate v adds, with inv = [v in N(u), deg v >= ta]:
//   inv (-cnt_tb(v) + [deg u >= tb] + sum_{i in Dold} [r_i ~ v])
//   + [deg v >= tb] sum_{j in inu_old} [v ~ r_j] - [deg v >= tb] |N(u) & N(v)|
// (chain_sib_item); na = na_old - inv.  Stored per row: base (int64), na_old,
// and the masks (inu_old in bits 0..14, Dold in 15..29, [deg u >= tb] in 30)
// in pb, the start of the parent's block in rstart, its length in minlen.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_sib_base(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                   const uint32_t* __restrict__ parent,
                                                   const uint32_t* __restrict__ par_len,
                                                   const uint64_t* __restrict__ par_off,
                                                   const uint32_t* __restrict__ shared, ChainTables t,
                                                   uint32_t* __restrict__ minlen, uint4* __restrict__ pb,
                                                   uint64_t* __restrict__ rstart) {
  uint64_t dummy = 0;
  for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < n; q += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t par = parent[q];
    minlen[q] = par_len[par];
    rstart[q] = par_off[par];
    uint32_t r[L], d[L], u = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      r[l] = rows[q * L + l];
      if (l == (int)P.ch_u) u = r[l];
    }
#pragma unroll
    for (int l = 0; l < L; ++l) d[l] = g.sig[2 * (uint64_t)r[l]];
    const uint32_t du = g.off_out[u + 1] - g.off_out[u];
    uint32_t inu = 0, dold = 0;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (l == (int)P.ch_u) continue;
      const bool in = ((P.qadj[P.ch_u] >> l) & 1) || adj_has(g, u, r[l], dummy);
      if (in && d[l] >= P.ch_ta) inu |= 1u << l;
      if (d[l] >= P.ch_tb) dold |= 1u << l;
    }
    const int64_t na_old = (int64_t)(t.ta ? t.ta[u] : du) - __popc(inu);
    int64_t c = (int64_t)t.H[u];
#pragma unroll
    for (int l = 0; l < L; ++l)
      if ((inu >> l) & 1) c -= t.tb ? t.tb[r[l]] : g.off_out[r[l] + 1] - g.off_out[r[l]];
    const bool dtb = P.ch_tb <= 1 || du >= P.ch_tb;
    if (dtb) c -= na_old;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      if (!((dold >> i) & 1)) continue;
#pragma unroll
      for (int j = 0; j < L; ++j) {
        if (j == i || !((inu >> j) & 1)) continue;
        if (((P.qadj[i] >> j) & 1) || adj_has(g, r[i], r[j], dummy)) ++c;
      }
    }
    c -= (int64_t)shared[q];
    pb[q] = make_uint4((uint32_t)(uint64_t)c, (uint32_t)((uint64_t)c >> 32), (uint32_t)na_old,
                       inu | (dold << 15) | ((dtb ? 1u : 0u) << 30));
  }
}

// The candidate's part of the chain closed form on top of the row's base
// (k_sib_base); na receives na_old - inv.
template <int L>
__device__ __forceinline__ int64_t chain_sib_item(const GView& g, const KPlan& P, const uint32_t (&r)[L], uint32_t v,
                                                  uint4 pbq, const ChainTables& t, int64_t& na, uint64_t& bytes) {
  const uint32_t inu = pbq.w & 0x7FFFu, dold = (pbq.w >> 15) & 0x7FFFu;
  const bool dtb = (pbq.w >> 30) & 1u;
  int64_t c = (int64_t)(((uint64_t)pbq.y << 32) | pbq.x);
  uint32_t u = 0;
#pragma unroll
  for (int l = 0; l < L; ++l)
    if (l == (int)P.ch_u) u = r[l];
  const uint32_t dv = g.sig[2 * (uint64_t)v];
  bytes += 4;
  const bool inv = dv >= P.ch_ta && (((P.qadj[P.ch_u] >> L) & 1) || adj_has(g, u, v, bytes));
  na = (int64_t)pbq.z - (inv ? 1 : 0);
  if (inv) {
    c -= t.tb ? t.tb[v] : g.off_out[v + 1] - g.off_out[v];
    bytes += t.tb ? 4 : 8;
    if (dtb) ++c;
#pragma unroll
    for (int i = 0; i < L; ++i)
      if (((dold >> i) & 1) && (((P.qadj[i] >> L) & 1) || adj_has(g, r[i], v, bytes))) ++c;
  }
  if (P.ch_tb <= 1 || dv >= P.ch_tb) {
#pragma unroll
    for (int j = 0; j < L; ++j)
      if (((inu >> j) & 1) && (((P.qadj[L] >> j) & 1) || adj_has(g, v, r[j], bytes))) ++c;
    c -= two_hop(g, u, v, bytes);
  }
  return c;
}

template <int L>
__global__ __launch_bounds__(BLOCK) void k_chain_sib(GView g, KPlan P, SibArgs a, ChainTables t) {
  __shared__ uint64_t s_off[LDS_OFF + 1];
  __shared__ uint64_t s_tile;
  const uint32_t tid = threadIdx.x;
  uint64_t acc = 0, valid = 0, pairs = 0, bytes = 0, mb = 0;
  const uint32_t nx = a.xcd ? 8u : 1u;
  const uint32_t myx = a.xcd ? (blockIdx.x & 7u) : 0u;
  uint32_t tries = 0;
  while (true) {
    if (uni(tid >> 6) == 0) s_tile = next_tile(a.tilectr, a.ntiles, nx, myx, tries);
    __syncthreads();
    const uint64_t tile = uni64(s_tile);
    if (tile == ~0ull) break;
    const uint64_t w0 = a.w_begin + tile * TILE;
    const uint64_t w1 = min(w0 + (uint64_t)TILE, a.w_end);
    const uint32_t plo = a.tile_p[tile], phi = a.tile_p[tile + 1];
    const uint32_t span = phi - plo + 1;
    const bool lds = span <= LDS_OFF;
    if (lds)
      for (uint32_t i = tid; i <= span; i += BLOCK) s_off[i] = a.work_off[plo + i];
    __syncthreads();
    if (a.stats) {
      // profiling: the full closed form per item (it also counts the merge
      // model's bytes)
      for (int it = 0; it < IPT; ++it) {
        const uint64_t w = w0 + (uint64_t)it * BLOCK + tid;
        if (w >= w1) continue;
        uint32_t lo = 0, hi = span + 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          const uint64_t o = lds ? s_off[mid] : a.work_off[plo + mid];
          if (o <= w) lo = mid + 1;
          else hi = mid;
        }
        const uint32_t pi = lo - 1;
        const uint64_t q = (uint64_t)plo + pi;
        const uint64_t qbase = lds ? s_off[pi] : a.work_off[q];
        const uint32_t v = a.cand[a.rstart[q] + (w - qbase)];
        uint32_t r[L + 1];
#pragma unroll
        for (int l = 0; l < L; ++l) r[l] = a.rows[q * L + l];
        r[L] = v;
        bytes += 4 * L + 16;
        if (v == r[L - 1]) continue;
        ++valid;
        int64_t na = 0;
        acc += (uint64_t)chain_closed<L + 1>(g, P, r, t, a.shared[q], true, na, true, mb, bytes);
        pairs += (uint64_t)na;
        mb += 16;  // the survivor's node, written and read back
      }
    } else {
      // the IPT items of a thread in lock-step stages (prefix search, row
      // data and base, candidate word, candidate part of the closed form)
      uint64_t w[IPT];
      uint32_t pi[IPT];
      bool in[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        w[j] = w0 + (uint64_t)j * BLOCK + tid;
        in[j] = w[j] < w1;
        pi[j] = 0;
      }
      for (uint32_t nn = span + 1; nn > 1;) {
        const uint32_t half = nn >> 1;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          const uint64_t o = lds ? s_off[pi[j] + half] : a.work_off[plo + pi[j] + half];
          pi[j] = (o <= w[j]) ? pi[j] + half : pi[j];
        }
        nn -= half;
      }
      uint64_t q[IPT], cidx[IPT];
      uint4 pbq[IPT];
      uint32_t r[IPT][L], v[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t pj = min(pi[j], span - 1);
        q[j] = (uint64_t)plo + pj;
        const uint64_t qbase = lds ? s_off[pj] : a.work_off[q[j]];
        cidx[j] = a.rstart[q[j]] + (w[j] - qbase);
        pbq[j] = a.pb[q[j]];
#pragma unroll
        for (int l = 0; l < L; ++l) r[j][l] = a.rows[q[j] * L + l];
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) v[j] = in[j] ? a.cand[cidx[j]] : 0u;
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        if (!in[j] || v[j] == r[j][L - 1]) continue;
        ++valid;
        int64_t na = 0;
        acc += (uint64_t)chain_sib_item<L>(g, P, r[j], v[j], pbq[j], t, na, bytes);
        pairs += (uint64_t)na;
      }
    }
    __syncthreads();
  }
  add_wave(&a.counters[0], acc);
  add_wave(&a.counters[1], valid);
  add_wave(&a.counters[2], pairs);
  if (a.stats) {
    add_wave(&a.counters[3], bytes);
    add_wave(&a.counters[8], mb);
  }
}

// ------------------------------------- intersections |N(u) & N(w)|
// Levels that count the last two positions together need, per row, the
// size of N(u) & N(w) for two of the row's vertices u = r[pa], w = r[pb]
// (degree-filtered).  minlen/minsel give each row's work: the shorter of
// the two lists (minsel 0: u's, 1: w's), or 0 for rows that need none.
struct CommonArgs {
  uint32_t pa, pb;  // row positions of u and w
  uint32_t thr;     // degree threshold of a common neighbour
  uint32_t mode;    // 0: count -= weight[p] per hit; 1: count -= 1 per hit; 2: weight[p] += hits
  // null, or per prefix {first list word of the walked vertex, hub row of the
  // tested one (~0u: none), its list bounds}: one 16-byte load per item in
  // place of the row, selector, offset and hub-row loads
  const uint4* desc;
};

// Pair level: the work of the rows (prefixes) that have survivors.  A
// prefix whose u or w has a hub row takes survivors x |N(u) & N(w)| off the
// count here and needs no walk.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_pair_prep(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                    const uint32_t* __restrict__ nsurv, uint32_t* __restrict__ minlen,
                                                    uint8_t* __restrict__ minsel, uint4* __restrict__ desc,
                                                    unsigned long long* __restrict__ counters) {
  uint64_t acc = 0;
  const bool hub = g.hr_id && P.pr_tab <= 2;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    uint32_t len = 0, sel = 0;
    const uint32_t ns = nsurv[p];
    const uint32_t u = ns ? rows[p * L + P.pr_a] : 0u, w = ns ? rows[p * L + P.pr_b] : 0u;
    const uint32_t hc = (ns && hub) ? hub_common(g, u, w) : ~0u;
    if (hc != ~0u) {
      acc -= (uint64_t)ns * hc;
    } else if (ns) {
      const uint32_t ul = g.off_out[u], uh = g.off_out[u + 1], wl = g.off_out[w], wh = g.off_out[w + 1];
      const uint32_t du = uh - ul, dw = wh - wl;
      len = min(du, dw);
      sel = du <= dw ? 0u : 1u;
      // walk the shorter list, test membership in the other
      const uint32_t oth = sel ? u : w;
      const int32_t h = g.hub_id ? g.hub_id[oth] : -1;
      desc[p] = sel ? make_uint4(wl, h >= 0 ? (uint32_t)h : ~0u, ul, uh)
                    : make_uint4(ul, h >= 0 ? (uint32_t)h : ~0u, wl, wh);
    }
    minlen[p] = len;
    minsel[p] = (uint8_t)sel;
  }
  add_wave(&counters[0], acc);
}

// Chain level, shared part: the work of N(u) & N(r_i) for one older
// position i of every parent prefix (rows of L words, u = r[ch_u]); zero
// when deg(r_i) is below the last position's threshold (such r_i take no
// completions away), and when u or r_i has a hub row (the count is added to
// the prefix's weight here).
template <int L>
__global__ __launch_bounds__(BLOCK) void k_chain_prep(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                     uint32_t i, uint32_t* __restrict__ minlen,
                                                     uint8_t* __restrict__ minsel, uint32_t* __restrict__ weight) {
  const bool hub = g.hr_id && P.ch_ta <= 2;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t u = rows[p * L + P.ch_u], w = rows[p * L + i];
    const uint32_t du = g.off_out[u + 1] - g.off_out[u], dw = g.off_out[w + 1] - g.off_out[w];
    bool need = P.ch_tb <= 1 || g.sig[2 * (uint64_t)w] >= P.ch_tb;
    if (need && hub) {
      const uint32_t hc = hub_common(g, u, w);
      if (hc != ~0u) {
        weight[p] += hc;
        need = false;
      }
    }
    minlen[p] = need ? min(du, dw) : 0u;
    minsel[p] = (uint8_t)(du <= dw ? 0u : 1u);
  }
}

// One work item per element x of a row's shorter list: x is a common
// neighbour when deg(x) meets the threshold and x lies in the other list.
// Tiles are fetched as in k_expand (dynamic, XCD-aware, wave-uniform loop
// edges).
template <int L>
__global__ __launch_bounds__(BLOCK) void k_common(GView g, KPlan P, ExpandArgs a, CommonArgs c) {
  __shared__ uint64_t s_off[LDS_OFF + 1];
  __shared__ uint64_t s_tile;
  const uint32_t tid = threadIdx.x;
  uint64_t acc = 0, bytes = 0;
  const uint64_t w_end_ = a.dyn ? a.dyn[0] : a.w_end;
  const uint64_t ntiles_ = a.dyn ? a.dyn[1] : a.ntiles;
  // a grid sized before the tile count was known: blocks beyond it leave
  // at once (before any barrier or tile-counter atomic)
  if (blockIdx.x >= ntiles_) return;
  const uint32_t nx = a.xcd ? 8u : 1u;
  const uint32_t myx = a.xcd ? (blockIdx.x & 7u) : 0u;
  uint32_t tries = 0;
  while (true) {
    if (uni(tid >> 6) == 0) s_tile = next_tile(a.tilectr, ntiles_, nx, myx, tries);
    __syncthreads();
    const uint64_t tile = uni64(s_tile);
    if (tile == ~0ull) break;
    const uint64_t w0 = a.w_begin + tile * TILE;
    const uint64_t w1 = min(w0 + (uint64_t)TILE, w_end_);
    const uint32_t plo = a.tile_p[tile], phi = a.tile_p[tile + 1];
    const uint32_t span = phi - plo + 1;
    const bool lds = span <= LDS_OFF;
    if (lds)
      for (uint32_t i = tid; i <= span; i += BLOCK) s_off[i] = a.work_off[plo + i];
    __syncthreads();
    // the IPT items of a thread in lock-step stages (prefix search, row,
    // list word, membership) so that their loads are in flight together
    uint64_t w[IPT];
    uint32_t pi[IPT];
    bool in[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      w[j] = w0 + (uint64_t)j * BLOCK + tid;
      in[j] = w[j] < w1;
      pi[j] = 0;
    }
    // largest pi with off(pi) <= w: a fixed number of halvings for the whole
    // block (off(0) <= w0 <= w and off(span) >= w1)
    for (uint32_t n = span + 1; n > 1;) {
      const uint32_t half = n >> 1;
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint64_t o = lds ? s_off[pi[j] + half] : a.work_off[plo + pi[j] + half];
        pi[j] = (o <= w[j]) ? pi[j] + half : pi[j];
      }
      n -= half;
    }
    uint64_t p[IPT];
    uint32_t own[IPT], oth[IPT], li[IPT], x[IPT];
    bool hit[IPT];
    if (c.desc) {
      // per-prefix descriptor: walked list's base, tested vertex's hub row
      // or list bounds
      uint4 d[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t pj = min(pi[j], span - 1);
        p[j] = (uint64_t)plo + pj;
        d[j] = c.desc[p[j]];
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t pj = min(pi[j], span - 1);
        const uint64_t pbase = lds ? s_off[pj] : a.work_off[p[j]];
        li[j] = d[j].x + (uint32_t)(w[j] - pbase);
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) x[j] = in[j] ? g.adj_out[li[j]] : 0u;
      uint64_t bw[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        bw[j] = (in[j] && d[j].y != ~0u) ? g.bits[(uint64_t)d[j].y * g.words + (x[j] >> 6)] : 0ull;
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        if (!in[j]) {
          hit[j] = false;
        } else if (d[j].y != ~0u) {
          hit[j] = (bw[j] >> (x[j] & 63)) & 1ull;
          if (a.stats) bytes += 8;
        } else {
          hit[j] = member_safe(g.adj_out, d[j].z, d[j].w, x[j]);
          if (a.stats) bytes += bs_bytes(d[j].w - d[j].z);
        }
        if (a.stats && in[j]) bytes += 16;  // the descriptor
      }
    } else {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t pj = min(pi[j], span - 1);
        p[j] = (uint64_t)plo + pj;
        const uint64_t pbase = lds ? s_off[pj] : a.work_off[p[j]];
        const uint32_t u = a.rows[p[j] * L + c.pa], wv = a.rows[p[j] * L + c.pb];
        const bool sel = a.minsel[p[j]] != 0;
        own[j] = sel ? wv : u;
        oth[j] = sel ? u : wv;
        li[j] = g.off_out[own[j]] + (uint32_t)(w[j] - pbase);
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) x[j] = in[j] ? g.adj_out[li[j]] : 0u;
      if (g.hub_id) {
        int32_t h[IPT];
        uint64_t bw[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) h[j] = g.hub_id[oth[j]];
#pragma unroll
        for (int j = 0; j < IPT; ++j) bw[j] = (in[j] && h[j] >= 0) ? g.bits[(uint64_t)h[j] * g.words + (x[j] >> 6)] : 0;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          if (!in[j]) {
            hit[j] = false;
          } else if (h[j] >= 0) {
            hit[j] = (bw[j] >> (x[j] & 63)) & 1ull;
            if (a.stats) bytes += 12;
          } else {
            const uint32_t lo2 = g.off_out[oth[j]], hi2 = g.off_out[oth[j] + 1];
            hit[j] = member_safe(g.adj_out, lo2, hi2, x[j]);
            if (a.stats) bytes += 12 + bs_bytes(hi2 - lo2);
          }
        }
      } else {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          hit[j] = false;
          if (!in[j]) continue;
          const uint32_t lo2 = g.off_out[oth[j]], hi2 = g.off_out[oth[j] + 1];
          hit[j] = member_safe(g.adj_out, lo2, hi2, x[j]);
          if (a.stats) bytes += 8 + bs_bytes(hi2 - lo2);
        }
      }
    }
    // a common neighbour of two distinct vertices has degree >= 2, so only
    // higher thresholds need the degree
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (a.stats && in[j]) bytes += 4;
      if (c.thr > 2 && hit[j]) {
        hit[j] = g.nsig_out[li[j]].x >= c.thr;
        if (a.stats) bytes += 8;
      }
    }
    if (c.mode == 2) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) add_runs(a.nsurv, p[j], hit[j]);  // all lanes; items past w1 never hit
    } else {
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if (hit[j]) acc -= c.mode == 0 ? (uint64_t)a.nsurv[p[j]] : 1ull;
    }
    __syncthreads();
  }
  if (c.mode != 2) add_wave(&a.counters[0], acc);
  if (a.stats) add_wave(&a.counters[3], bytes);
}

// Chain level, shared part from the two-hop index: per parent prefix (rows
// of L words), the sum of |N(u) & N(r_i)| over its positions i != ch_u
// whose degree takes completions away.
// Chain level without the two-hop index, shared part for the older
// positions i in `mask` (adjacent to u's position in the query, so r_i is in
// N(u)): sum of |N(u) & N(r_i)| from the per-arc table, stored per parent
// row before the k_common passes of the other positions add theirs.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_chain_shared_arc(GView g, KPlan P, const uint32_t* __restrict__ rows,
                                                           uint64_t n, uint32_t mask, uint32_t* __restrict__ shared) {
  uint64_t dummy = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t u = rows[p * L + P.ch_u];
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      if (!((mask >> i) & 1)) continue;
      const uint32_t w = rows[p * L + i];
      if (P.ch_tb <= 1 || g.sig[2 * (uint64_t)w] >= P.ch_tb) t += arc_common(g, u, w, dummy);
    }
    shared[p] = t;
  }
}

template <int L>
__global__ __launch_bounds__(BLOCK) void k_chain_shared_h2(GView g, KPlan P, const uint32_t* __restrict__ rows,
                                                          uint64_t n, uint32_t* __restrict__ shared) {
  uint64_t dummy = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t u = rows[p * L + P.ch_u];
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      if (i == (int)P.ch_u) continue;
      const uint32_t w = rows[p * L + i];
      if (P.ch_tb <= 1 || g.sig[2 * (uint64_t)w] >= P.ch_tb) t += two_hop(g, u, w, dummy);
    }
    shared[p] = t;
  }
}

// Chain level with shared candidates, rows of L = k-3 words (level k-4):
// every row's work is the filled part of its parent's block of
// position-(k-3) candidates (cnt = the parents' filled lengths).
// Chain tables for thresholds (ta, tb): H[u] = sum over x in N(u) with
// deg(x) >= ta of |{y in N(x) : deg(y) >= tb}|, H[V + u] = sum of those
// deg(x).  One wave per vertex.
__global__ __launch_bounds__(BLOCK) void k_chain_table(GView g, uint32_t ta, const uint32_t* __restrict__ tb_tbl,
                                                      uint64_t* __restrict__ H) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / 64);
  for (uint64_t u = (uint64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6); u < g.V; u += nw) {
    const uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
    uint64_t h = 0, hd = 0;
    for (uint32_t i = lo + lane; i < hi; i += 64) {
      const uint32_t x = g.adj_out[i];
      const uint32_t dx = g.off_out[x + 1] - g.off_out[x];
      if (dx >= ta) {
        h += tb_tbl ? tb_tbl[x] : dx;
        hd += dx;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      h += ((uint64_t)__shfl_xor((uint32_t)(h >> 32), o) << 32) | __shfl_xor((uint32_t)h, o);
      hd += ((uint64_t)__shfl_xor((uint32_t)(hd >> 32), o) << 32) | __shfl_xor((uint32_t)hd, o);
    }
    if (lane == 0) {
      H[u] = h;
      H[g.V + u] = hd;
    }
  }
}

// Chain level over stored rows of L = k-2 words (positions 0..k-3):
// chain_closed per row; the intersection with the newest vertex v = r[L-1]
// from the two-hop index, or left to k_common (minlen/minsel set here).
// Level k-2 size += |A|.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_chain_row(GView g, KPlan P, const uint32_t* __restrict__ rows, uint64_t n,
                                                    const uint32_t* __restrict__ parent,
                                                    const uint32_t* __restrict__ shared, ChainTables t,
                                                    uint32_t* __restrict__ minlen, uint8_t* __restrict__ minsel,
                                                    unsigned long long* __restrict__ counters, int stats) {
  uint64_t acc = 0, na_sum = 0, mb = 0, bytes = 0;
  const bool h2 = g.h2_off && P.ch_ta <= 2;
  // without the index, |N(u) & N(r_{L-1})| comes from the per-arc table when
  // position L-1 is adjacent to u's in the query, else from a k_common pass
  const bool arc = !h2 && g.arc_cn && ((P.qadj[P.ch_u] >> (L - 1)) & 1) && P.ch_ta <= 2;
  for (uint64_t p = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; p < n; p += (uint64_t)gridDim.x * BLOCK) {
    uint32_t r[L];
#pragma unroll
    for (int l = 0; l < L; ++l) r[l] = rows[p * L + l];
    int64_t na = 0;
    int64_t c = chain_closed<L>(g, P, r, t, shared[parent[p]], h2, na, stats != 0, mb, bytes);
    bytes += 4 * L + 8;  // the row and its parent's shared sum
    na_sum += (uint64_t)na;
    uint32_t u = 0;
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (l == (int)P.ch_u) u = r[l];
    const uint32_t v = r[L - 1];
    const uint32_t du = g.off_out[u + 1] - g.off_out[u], dv = g.off_out[v + 1] - g.off_out[v];
    const bool need = P.ch_tb <= 1 || g.sig[2 * (uint64_t)v] >= P.ch_tb;
    bool known = h2 || arc;
    if (arc && need) c -= arc_common(g, u, v, bytes);
    if (!known && need && g.hr_id && P.ch_ta <= 2) {
      const uint32_t hc = hub_common(g, u, v);
      if (hc != ~0u) {
        c -= hc;
        known = true;
      }
    }
    acc += (uint64_t)c;
    minlen[p] = (!known && need) ? min(du, dv) : 0u;
    minsel[p] = (uint8_t)(du <= dv ? 0u : 1u);
  }
  add_wave(&counters[0], acc);
  add_wave(&counters[1], na_sum);
  if (stats) {
    add_wave(&counters[8], mb);
    add_wave(&counters[3], bytes);
  }
}

// ---------------------------------------------------------- dispatch
#define CUTS_DISPATCH_L(L, CALL) \
  switch (L) {                   \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 3: CALL(3); break;      \
    case 4: CALL(4); break;      \
    case 5: CALL(5); break;      \
    case 6: CALL(6); break;      \
    case 7: CALL(7); break;      \
    case 8: CALL(8); break;      \
    case 9: CALL(9); break;      \
    case 10: CALL(10); break;    \
    case 11: CALL(11); break;    \
    case 12: CALL(12); break;    \
    case 13: CALL(13); break;    \
    case 14: CALL(14); break;    \
    case 15: CALL(15); break;    \
    default: throw Error(-3, "query position out of range"); \
  }

// the pair level is position k-3 <= 13
#define CUTS_DISPATCH_PAIR(L, CALL) \
  switch (L) {                      \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 3: CALL(3); break;         \
    case 4: CALL(4); break;         \
    case 5: CALL(5); break;         \
    case 6: CALL(6); break;         \
    case 7: CALL(7); break;         \
    case 8: CALL(8); break;         \
    case 9: CALL(9); break;         \
    case 10: CALL(10); break;       \
    case 11: CALL(11); break;       \
    case 12: CALL(12); break;       \
    case 13: CALL(13); break;       \
    default: throw Error(-3, "query position out of range"); \
  }

uint32_t grid_for(uint64_t work, uint64_t per_block, uint32_t cap) {
  uint64_t b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  return (uint32_t)std::min<uint64_t>(b, cap);
}

template <typename T>
void dev_alloc(T*& p, uint64_t n) {
  if (p) HIPCHECK(hipFree(p));
  p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(-5, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
}

// Device buffer of n elements: hipMalloc'd (the old one freed), or carved
// from the arena when there is one (nothing freed; sizes are rounded up to a
// power of two so that regrowth wastes at most half of what was carved).
template <typename T>
void dev_alloc(ScratchArena* a, T*& p, uint64_t n) {
  if (!a) {
    dev_alloc(p, n);
    return;
  }
  uint64_t m = 1;
  while (m < std::max<uint64_t>(n, 1)) m <<= 1;
  p = (T*)a->carve(m * sizeof(T));
}

// Makes sure level buffer `lv` can hold `rows` rows of `words` words plus
// the planning arrays for them.
void ensure_rows(ScratchArena* ar, LevelBuf& b, uint64_t rows, uint32_t words) {
  if (b.cap >= rows && b.rows) return;
  uint64_t cap = std::max<uint64_t>(rows, 1024);
  if (b.cap) cap = std::max<uint64_t>(cap, b.cap + b.cap / 2);
  dev_alloc(ar, b.rows, cap * words);
  dev_alloc(ar, b.minlen, cap);
  dev_alloc(ar, b.minsel, cap);
  dev_alloc(ar, b.work_off, cap + 1);
  dev_alloc(ar, b.cbase, cap);
  b.cap = cap;
}
void ensure_rows_exact(ScratchArena* ar, LevelBuf& b, uint64_t rows, uint32_t words) {
  // like ensure_rows but never grows beyond `rows` (budget-bound levels)
  if (b.cap >= rows && b.rows) return;
  dev_alloc(ar, b.rows, std::max<uint64_t>(rows, 1) * words);
  dev_alloc(ar, b.minlen, std::max<uint64_t>(rows, 1));
  dev_alloc(ar, b.minsel, std::max<uint64_t>(rows, 1));
  dev_alloc(ar, b.work_off, std::max<uint64_t>(rows, 1) + 1);
  dev_alloc(ar, b.cbase, std::max<uint64_t>(rows, 1));
  b.cap = rows;
}
// The per-row arrays of the pair and chain levels, allocated only for the
// levels that use them (each grows to the level's row capacity).
template <typename T>
void ensure_aux(ScratchArena* ar, T*& p, uint64_t& cap, uint64_t rows) {
  if (cap >= rows && p) return;
  dev_alloc(ar, p, std::max<uint64_t>(rows, 1));
  cap = std::max<uint64_t>(rows, 1);
}

void ensure_bsum(ScratchArena* ar, LevelBuf& b, uint64_t nb) {
  if (b.bsum_cap >= nb + 1 && b.bsum) return;
  dev_alloc(ar, b.bsum, nb + 1);
  b.bsum_cap = nb + 1;
}
void ensure_tiles(ScratchArena* ar, LevelBuf& b, uint64_t nt) {
  if (b.tile_cap >= nt + 1 && b.tile_p) return;
  dev_alloc(ar, b.tile_p, nt + 1);
  b.tile_cap = nt + 1;
}

struct Matcher {
  Context& c;
  DeviceGraph& dg;
  const Plan& plan;
  const MatchOpts& o;
  MatchResult& r;
  KPlan kp;
  GView g;
  uint32_t k;
  bool closed_last = false, fuse = false, write = false, pair = false, chain = false, chain_fused = false;
  bool collapse_fused = false, sib = false;
  bool sib_live = false;      // candidates of the current level-(k-4) input are built
  uint64_t cand_limit = 0;    // most candidates one build may hold
  const uint32_t* tbl = nullptr;
  PairTables ptbl{};
  ChainTables ctab{};
  uint64_t budget = 0;
  uint64_t level_cap[kMaxQuery] = {0};
  uint32_t grid_cap = 0;
  // blocks of a tiled kernel of class cls (CUTS_K_*): the class's own
  // blocks per CU when one is set, else the run's
  uint32_t gcap(int cls) const {
    const uint32_t b = o.class_bpc[cls] ? o.class_bpc[cls] : (o.blocks_per_cu ? o.blocks_per_cu : 6u);
    return (uint32_t)c.cus * b;
  }

  // streamed embedding output: chunks of the last level copied so far (the
  // output buffer of chunk i is i & 1), and whether copies are streamed
  uint64_t emb_chunks = 0;
  bool stream_emb = false;
  // where the run's buffers live: the context's level buffers and counters
  // (hipMalloc'd on demand), or a scratch arena and caller-owned level
  // buffers (the no-allocation entry point)
  LevelBuf* lv;
  unsigned long long* ctr;
  ScratchArena* arena = nullptr;

  Matcher(Context& c_, DeviceGraph& dg_, const Plan& p_, const MatchOpts& o_, MatchResult& r_)
      : c(c_), dg(dg_), plan(p_), o(o_), r(r_), lv(c_.lv), ctr(c_.counters) {}

  bool prof = false;
  bool has_deadline = false, aborted = false;
  std::chrono::steady_clock::time_point deadline;
  // the pair level's |N(u) & N(w)| base comes from an index (two-hop, or the
  // per-arc table when u and w are adjacent) rather than a k_common pass
  bool pair_cab_known() const {
    return kp.pr_tab <= 2 && (g.h2_off || (g.arc_cn && kp.pr_adj));
  }
  bool past_deadline() {
    if (has_deadline && !aborted && std::chrono::steady_clock::now() >= deadline) aborted = true;
    return aborted;
  }
  // deterministic work cap: the candidates examined so far
  bool add_work(uint64_t w) {
    r.work += w;
    if (o.max_work && r.work > o.max_work) {
      aborted = true;
      r.work_capped = 1;
    }
    return aborted;
  }
  void pbegin(bool dev_bytes) {
    if (!prof) return;
    if (dev_bytes) {
      HIPCHECK(hipMemsetAsync(ctr + 3, 0, 8, c.stream));
      HIPCHECK(hipMemsetAsync(ctr + 8, 0, 8, c.stream));
    }
    HIPCHECK(hipEventRecord(c.kev[0], c.stream));
  }
  void pend(int cls, double host_bytes, bool dev_bytes) {
    if (!prof) return;
    HIPCHECK(hipEventRecord(c.kev[1], c.stream));
    HIPCHECK(hipEventSynchronize(c.kev[1]));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c.kev[0], c.kev[1]));
    r.kernel_ms[cls] += ms;
    r.kernel_launches[cls]++;
    r.kernel_bytes[cls] += host_bytes + (dev_bytes ? (double)read_u64(ctr + 3) : 0.0);
    if (dev_bytes) r.merge_bytes[cls] += (double)read_u64(ctr + 8);
  }

  uint64_t read_u64(const void* dev) {
    HIPCHECK(hipMemcpyAsync(c.host_pinned, dev, 8, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    return c.host_pinned[0];
  }
  uint32_t read_u32(const void* dev) {
    HIPCHECK(hipMemcpyAsync(c.host_pinned, dev, 4, hipMemcpyDeviceToHost, c.stream));
    HIPCHECK(hipStreamSynchronize(c.stream));
    return (uint32_t)(c.host_pinned[0] & 0xFFFFFFFFu);
  }

  // exclusive scan of b.minlen[0..n) into b.work_off[0..n] (total at n)
  void scan_launch(LevelBuf& b, uint64_t n) {
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
      pbegin(false);
      hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(BLOCK), 0, c.stream, b.minlen, n, b.work_off);
      pend(2, 16.0 * n, false);
      return;
    }
    ensure_bsum(arena, b, nb);
    pbegin(false);
    hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(BLOCK), 0, c.stream, b.minlen, n, b.bsum);
    if (nb <= SCAN_FUSE_BLOCKS) {
      hipLaunchKernelGGL(k_scan_down_fused, dim3((uint32_t)nb), dim3(BLOCK), 0, c.stream, b.minlen, n, b.bsum, nb,
                         b.work_off);
    } else {
      hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(BLOCK), 0, c.stream, b.bsum, nb);
      hipLaunchKernelGGL(k_scan_down, dim3((uint32_t)nb), dim3(BLOCK), 0, c.stream, b.minlen, n, b.bsum, nb,
                         b.work_off);
    }
    pend(2, 16.0 * n, false);
  }
  // the same, returning the total (a readback)
  uint64_t scan(LevelBuf& b, uint64_t n) {
    scan_launch(b, n);
    return read_u64(b.work_off + n);
  }

  // Speculative levels: the level's work is not read back before it is
  // expanded.  k_tile_index_dyn checks on the device that the work fits the
  // tile index and the output rows already allocated, and the expansion
  // takes its range from the device; a writing level reads back its row
  // count together with the status word (one readback instead of two), a
  // counting level none.  A writing level that did not fit is redone with
  // the readback at once; a counting level that did not fit leaves the
  // status bit for the end of the run, which then repeats the whole query
  // with readbacks.  Off for profiling, work caps, deadlines, the
  // no-allocation entry point and CUTS_FLAG_SYNC_LEVELS.
  bool no_spec = false;
  bool spec_overflow = false;
  // level sizes of the last complete run of this query, or null; bit l:
  // level l's size was handed on on the device (read from lvcnt at the end)
  const uint64_t* hint = nullptr;
  uint32_t chained_mask = 0;
  // level L can take its input size from the device: a speculative level
  // planned by one block (k_plan_small) with no host step that needs the
  // size (chain rows and shared sums, shared candidates, the last level's
  // closed form, written embeddings, the pair level's intersection pass)
  bool chainable(uint32_t L) const {
    if (L >= k || !speculate() || o.trunks > 1) return false;
    if (chain && !chain_fused && L == k - 2) return false;
    if (chain && L == k - 3) return false;
    if (sib && L == k - 4) return false;
    if (L == k - 1 && closed_last) return false;
    if (L == k - 1 && write) return false;
    if (pair && L == k - 3 && !(collapse_fused && L == k - 4) && kp.pr_a != kp.pr_b && !pair_cab_known()) return false;
    return true;
  }
  bool speculate() const {
    return !no_spec && !prof && !o.max_work && !has_deadline && !arena && !(o.flags & 8192u);
  }
  // Word w of the coherent host words once the stream has drained: true
  // when a kernel published it (it holds ~0 until then).  Polling the word
  // instead, so that the next level's launches queue while the expansion's
  // last waves retire, measured the same per step (249.6 vs 248.3 ms).
  bool wait_published(int w) {
    HIPCHECK(hipStreamSynchronize(c.stream));
    return ((const volatile uint64_t*)c.host_res)[w] != ~0ull;
  }
  // tile index over [0, W) with W read on the device; returns the launch
  // shape bound (the most tiles the index may hold)
  // status: ctr[11] for a writing level (read back with its row count),
  // ctr[12] for the counting levels (read at the end of the run)
  // (a writing level: at most MAX_CHUNK items; a counting one MAX_SPEC_ITEMS)
  uint64_t tile_index_dyn(LevelBuf& in, uint64_t n, uint64_t row_cap, unsigned int* outctr, uint64_t* publish = nullptr) {
    ensure_tiles(arena, in, std::max<uint64_t>(in.tile_cap, 4097) - 1);
    const uint64_t tcap = in.tile_cap;
    const uint32_t tg = (uint32_t)std::min<uint64_t>((tcap + BLOCK - 1) / BLOCK, 1024);
    hipLaunchKernelGGL(k_tile_index_dyn, dim3(tg), dim3(BLOCK), 0, c.stream, in.work_off, n,
                       outctr ? MAX_CHUNK : MAX_SPEC_ITEMS, tcap, row_cap,
                       in.tile_p, (unsigned int*)(ctr + 4), outctr, ctr + 14, outctr ? ctr + 11 : ctr + 12, ctr + 13,
                       publish);
    HIPCHECK(hipGetLastError());
    return tcap - 1;
  }

  // rows: level L-1 table (n rows of L words) living in lv[L-1].rows (or an
  // offset into it); produces position L.
  // dn non-null: the input level's size is on the device (its expansion's
  // row counter) and n is only its bound; the caller has checked
  // chainable(L)
  void level(uint32_t L, const uint32_t* rows, uint64_t n, const unsigned long long* dn = nullptr) {
    if (n == 0) return;
    LevelBuf& in = lv[L - 1];
    if (chain && !chain_fused && L == k - 2) {
      chain_rows(L, rows, n);
      return;
    }
    if (chain && L == k - 3) {
      chain_shared(L, rows, n, in);
      if (aborted) return;
      if (sib && sib_live) {
        sib_level(L, rows, n);
        return;
      }
    }
    if (sib && L == k - 4) {
      build_candidates(L, rows, n, in);
      if (aborted) return;
    }
    if (L == k - 1 && closed_last) {
      if (add_work(n)) return;
      const uint32_t grid = grid_for(n, BLOCK, grid_cap);
      pbegin(true);
#define CF(LL) hipLaunchKernelGGL((k_closed_form<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, tbl, ctr, (int)prof)
      CUTS_DISPATCH_L(L, CF)
#undef CF
      HIPCHECK(hipGetLastError());
      pend(7, 0, true);
      return;
    }
    // plan + scan over the input rows; arrays live with the input level
    const bool use_cbase = fuse && L == k - 2 && kp.sc_pos < L && !(o.flags & 64u);
    int mode = M_WRITE;
    if (collapse_fused && L == k - 4) mode = M_COLLAPSE;
    else if (pair && L == k - 3) mode = M_PAIR;
    else if (chain_fused && L == k - 3) mode = M_CHAIN;
    else if (L == k - 1 && !write) mode = M_COUNT;
    else if (L == k - 2 && fuse) mode = M_FUSED;
    const int mclass = mode == M_WRITE ? 4 : mode == M_COUNT ? 5 : mode == M_FUSED ? 6 : (mode == M_PAIR || mode == M_COLLAPSE) ? 8 : 10;
    if (mode == M_PAIR) ensure_aux(arena, in.pbase, in.pbase_cap, n);
    const bool spec = speculate() && !(mode == M_WRITE && L == k - 1);
    // a speculative level of at most SCAN_TILE rows plans, scans and indexes
    // its tiles in one launch (k_plan_small, inside level_spec)
    const bool small = spec && n <= (uint64_t)SCAN_TILE;
    if (!small) {
      const uint32_t grid = grid_for(n, BLOCK, grid_cap);
      unsigned long long* merge = prof ? ctr + 8 : nullptr;
      if (merge) HIPCHECK(hipMemsetAsync(merge, 0, 8, c.stream));
      pbegin(false);
#define PL(LL) hipLaunchKernelGGL((k_plan<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.minlen, in.minsel, use_cbase ? in.cbase : nullptr, tbl, mode == M_PAIR ? in.pbase : nullptr, ptbl, merge, ctr)
      CUTS_DISPATCH_L(L, PL)
#undef PL
      HIPCHECK(hipGetLastError());
      pend(1, (double)n * (4.0 * L + 8.0 * kp.nj[L] + 5.0), false);
      // the prefixes' list reads belong to the expansion that performs them
      if (merge) r.merge_bytes[mclass] += (double)read_u64(merge);
    }
    if (spec && level_spec(L, rows, n, in, mode, use_cbase, small, dn)) return;
    // (a small level that did not fit has been planned: the scan below reads it)
    const uint64_t W = scan(in, n);
    if (W == 0) return;
    if (add_work(W)) return;
    // a counting level sizes the tile index for all of W, so that its next
    // speculative run fits in one launch
    if (mode != M_WRITE && W <= MAX_SPEC_ITEMS) ensure_tiles(arena, in, (W + TILE - 1) / TILE);
    uint64_t chunk = MAX_CHUNK;
    LevelBuf& out = lv[L];
    if (mode == M_WRITE) {
      const uint64_t cap = level_cap[L];
      chunk = std::min<uint64_t>(chunk, std::max<uint64_t>(cap, 1));
      // the written last level goes out in chunks of at most 2^22 candidates,
      // so that the copy of one chunk overlaps the expansion of the next
      if (L == k - 1 && o.emb_rows) chunk = std::min<uint64_t>(chunk, 1ull << 22);
      // a buffer about to be replaced may still be the source of a copy
      if (L == k - 1 && emb_chunks && (out.cap < std::min(chunk, W) || !out.rows2))
        HIPCHECK(hipStreamSynchronize(c.copy_stream));
      ensure_rows_exact(arena, out, std::min<uint64_t>(chunk, W), L + 1);
      chunk = std::min<uint64_t>(chunk, out.cap);
      if (L == k - 1 && stream_emb) {
        ensure_aux(arena, out.rows2, out.rows2_cap, (uint64_t)out.cap * (L + 1));
      }
    }
    const bool pair_common = mode == M_PAIR && kp.pr_a != kp.pr_b && !pair_cab_known();
    if (pair_common) {
      ensure_aux(arena, in.nsurv, in.nsurv_cap, n);
      HIPCHECK(hipMemsetAsync(in.nsurv, 0, n * sizeof(uint32_t), c.stream));
    }
    for (uint64_t w0 = 0; w0 < W; w0 += chunk) {
      if (past_deadline()) return;
      const uint64_t w1 = std::min(W, w0 + chunk);
      const uint64_t nt = (w1 - w0 + TILE - 1) / TILE;
      ensure_tiles(arena, in, nt);
      pbegin(false);
      hipLaunchKernelGGL(k_tile_index, dim3(grid_for(nt + 1, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, in.work_off,
                         n, w0, w1, nt, in.tile_p, (unsigned int*)(ctr + 4), mode == M_WRITE ? (unsigned int*)(ctr + 10) : nullptr);
      pend(3, (double)(nt + 1) * (8.0 * (64 - __builtin_clzll(n + 2)) + 4.0), false);
      ExpandArgs a{};
      a.rows = rows;
      a.work_off = in.work_off;
      a.minsel = in.minsel;
      a.tile_p = in.tile_p;
      a.w_begin = w0;
      a.w_end = w1;
      a.ntiles = nt;
      a.counters = ctr;
      a.tbl = tbl;
      a.cbase = use_cbase ? in.cbase : nullptr;
      a.pbase = mode == M_PAIR ? in.pbase : nullptr;
      a.nsurv = pair_common ? in.nsurv : nullptr;
      a.out_parent = nullptr;
      if ((chain && !chain_fused && mode == M_WRITE && L == k - 3) || (sib && mode == M_WRITE && L == k - 4)) {
        ensure_aux(arena, out.parent, out.parent_cap, out.cap);
        a.out_parent = out.parent;
      }
      a.shared = in.nsurv;
      a.ctab = ctab;
      a.ptab = ptbl;
      a.xcd = (o.flags & 8u) ? 0 : 1;
      a.tilectr = (unsigned int*)(ctr + 4);  // 8 x uint32 in counters[4..7], cleared by k_tile_index
      a.stats = prof ? 1 : 0;
      uint32_t grid = grid_for(nt, 1, gcap(mclass));
      // counting expansions: nthr threads per block, cgrow times the blocks
      // (as many lanes), at most one block per tile
      const uint32_t nthr = o.expand_threads ? o.expand_threads : (uint32_t)BLOCK;
      const uint32_t cgrow = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(BLOCK / nthr, nt / std::max<uint64_t>(grid, 1)));
      r.chunks++;
      if (mode == M_WRITE) {
        const bool streamed = L == k - 1 && stream_emb;
        const uint32_t ob = (uint32_t)(emb_chunks & 1);
        a.out_rows = (streamed && ob) ? out.rows2 : out.rows;
        a.out_count = (uint32_t*)(ctr + 10);
        // the buffer's previous copy to the host must be done before it is refilled
        if (streamed) HIPCHECK(hipStreamWaitEvent(c.stream, c.copy_ev[ob], 0));
        pbegin(true);
#define EW(LL) hipLaunchKernelGGL((k_expand<LL, M_WRITE>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_L(L, EW)
#undef EW
        HIPCHECK(hipGetLastError());
        pend(4, 0, true);
        const uint64_t m = read_u32(ctr + 10);
        r.level[L] += m;
        if (prof && L < k - 1) r.merge_bytes[4] += 16.0 * (double)m;  // trie nodes written and read back
        if (L == k - 1) {
          // embeddings: copy rows to the host sink
          r.count += m;
          const uint64_t room = o.emb_cap > r.emb_written ? o.emb_cap - r.emb_written : 0;
          const uint64_t take = std::min(room, m);
          if (take && o.emb_rows && streamed) {
            // on the copy stream, while the next chunk expands (the count
            // read above already waited for this chunk's rows)
            HIPCHECK(hipMemcpyAsync(o.emb_rows + r.emb_written * k, a.out_rows, take * k * 4, hipMemcpyDeviceToHost,
                                    c.copy_stream));
            HIPCHECK(hipEventRecord(c.copy_ev[ob], c.copy_stream));
            ++emb_chunks;
            r.emb_written += take;
          } else if (take && o.emb_rows) {
            HIPCHECK(hipMemcpyAsync(o.emb_rows + r.emb_written * k, out.rows, take * k * 4, hipMemcpyDeviceToHost,
                                    c.stream));
            HIPCHECK(hipStreamSynchronize(c.stream));
            r.emb_written += take;
          }
        } else {
          level(L + 1, out.rows, m);
          if (aborted) return;
        }
      } else if (mode == M_COUNT) {
        pbegin(true);
#define EC(LL) hipLaunchKernelGGL((k_expand<LL, M_COUNT>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_L(L, EC)
#undef EC
        HIPCHECK(hipGetLastError());
        pend(5, 0, true);
      } else if (mode == M_COLLAPSE) {
        pbegin(true);
#define EK(LL) hipLaunchKernelGGL((k_expand<LL, M_COLLAPSE>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_PAIR(L, EK)
#undef EK
        HIPCHECK(hipGetLastError());
        pend(8, 0, true);
      } else if (mode == M_CHAIN) {
        pbegin(true);
#define EC2(LL) hipLaunchKernelGGL((k_expand<LL, M_CHAIN>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_PAIR(L, EC2)
#undef EC2
        HIPCHECK(hipGetLastError());
        pend(10, 0, true);
      } else if (mode == M_FUSED) {
        pbegin(true);
#define EF(LL) hipLaunchKernelGGL((k_expand<LL, M_FUSED>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_L(L, EF)
#undef EF
        HIPCHECK(hipGetLastError());
        pend(6, 0, true);
      } else {
        pbegin(true);
#define EP(LL) hipLaunchKernelGGL((k_expand<LL, M_PAIR>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
        CUTS_DISPATCH_PAIR(L, EP)
#undef EP
        HIPCHECK(hipGetLastError());
        pend(8, 0, true);
      }
    }
    if (pair_common && !aborted) pair_intersections(L, rows, n, in);
  }

  // One level without the readback of its work (see speculate()).  Returns
  // false when a writing level did not fit: the caller then runs it with the
  // readback (its scan is repeated; nothing was written or counted).
  bool level_spec(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in, int mode, bool use_cbase, bool small,
                  const unsigned long long* dn = nullptr) {
    if (!small) scan_launch(in, n);
    LevelBuf& out = lv[L];
    uint64_t row_cap = ~0ull;
    const bool want_parent = (chain && !chain_fused && mode == M_WRITE && L == k - 3) ||
                             (sib && mode == M_WRITE && L == k - 4);
    if (mode == M_WRITE) {
      // the rows already allocated for this level (at least 64K, within
      // the level's share of the budget)
      ensure_rows_exact(arena, out, std::min<uint64_t>(std::max<uint64_t>(out.cap, 1ull << 16), level_cap[L]), L + 1);
      row_cap = std::min<uint64_t>(out.cap, level_cap[L]);
      if (want_parent) ensure_aux(arena, out.parent, out.parent_cap, out.cap);
    }
    const bool pair_common = mode == M_PAIR && kp.pr_a != kp.pr_b && !pair_cab_known();
    if (pair_common) {
      ensure_aux(arena, in.nsurv, in.nsurv_cap, n);
      HIPCHECK(hipMemsetAsync(in.nsurv, 0, n * sizeof(uint32_t), c.stream));
    }
    // a writing level's row count and status arrive in host memory (a slot
    // per level), from the tile index (no tiles) or the expansion's last
    // block
    uint64_t* pub_dev = c.host_res_dev + 32 + 2 * L;
    volatile uint64_t* pub = c.host_res + 32 + 2 * L;
    if (mode == M_WRITE) pub[0] = ~0ull;
    // a level whose input size is on the device cannot be redone here: a
    // misfit goes to the end-of-run status word and repeats the run
    unsigned long long* st_word = (mode == M_WRITE && !dn) ? ctr + 11 : ctr + 12;
    uint64_t tmax;
    if (small) {
      // plan + scan + tile index in one launch
      ensure_tiles(arena, in, std::max<uint64_t>(in.tile_cap, 4097) - 1);
      tmax = in.tile_cap - 1;
      unsigned int* oc = mode == M_WRITE ? (unsigned int*)(ctr + 10) : nullptr;
#define PSM(LL) hipLaunchKernelGGL((k_plan_small<LL>), dim3(1), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.minlen, in.minsel, use_cbase ? in.cbase : nullptr, tbl, mode == M_PAIR ? in.pbase : nullptr, ptbl, ctr, in.work_off, oc ? MAX_CHUNK : MAX_SPEC_ITEMS, in.tile_cap, row_cap, in.tile_p, (unsigned int*)(ctr + 4), oc, ctr + 14, st_word, ctr + 13, oc ? pub_dev : nullptr, dn, dn ? c.lvcnt + (L - 1) : nullptr)
      CUTS_DISPATCH_L(L, PSM)
#undef PSM
      HIPCHECK(hipGetLastError());
    } else {
      tmax = tile_index_dyn(in, n, row_cap, mode == M_WRITE ? (unsigned int*)(ctr + 10) : nullptr,
                            mode == M_WRITE ? pub_dev : nullptr);
    }
    ExpandArgs a{};
    a.rows = rows;
    a.work_off = in.work_off;
    a.minsel = in.minsel;
    a.tile_p = in.tile_p;
    a.dyn = ctr + 14;
    a.counters = ctr;
    a.tbl = tbl;
    a.cbase = use_cbase ? in.cbase : nullptr;
    a.pbase = mode == M_PAIR ? in.pbase : nullptr;
    a.nsurv = pair_common ? in.nsurv : nullptr;
    a.out_parent = want_parent ? out.parent : nullptr;
    a.shared = in.nsurv;
    a.ctab = ctab;
    a.ptab = ptbl;
    a.xcd = (o.flags & 8u) ? 0 : 1;
    a.tilectr = (unsigned int*)(ctr + 4);
    a.stats = 0;
    const int mclass = mode == M_WRITE ? 4 : mode == M_COUNT ? 5 : mode == M_FUSED ? 6 : (mode == M_PAIR || mode == M_COLLAPSE) ? 8 : 10;
    const uint32_t grid = grid_for(tmax, 1, gcap(mclass));
    const uint32_t nthr = o.expand_threads ? o.expand_threads : (uint32_t)BLOCK;
    const uint32_t cgrow = (uint32_t)(BLOCK / nthr);
    r.chunks++;
    if (mode == M_WRITE) {
      a.out_rows = out.rows;
      a.out_count = (uint32_t*)(ctr + 10);
      // rows written (ctr[10]) and this level's status word (ctr[11]),
      // written to host memory by the expansion's last block; with no
      // block taking part (no work, or a level that did not fit) they are
      // read back
      a.publish = pub_dev;
      a.done = (unsigned int*)(ctr + 8);
      a.status = st_word;
#define EW(LL) hipLaunchKernelGGL((k_expand<LL, M_WRITE>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, a)
      CUTS_DISPATCH_L(L, EW)
#undef EW
      HIPCHECK(hipGetLastError());
      // the last complete run's size of this level fits one planning block:
      // the next level takes it from the device, with no synchronisation
      if (hint && hint[L] > 0 && hint[L] <= (uint64_t)SCAN_TILE && chainable(L + 1)) {
        chained_mask |= 1u << L;
        level(L + 1, out.rows, SCAN_TILE, ctr + 10);
        return true;
      }
      HIPCHECK(hipStreamSynchronize(c.stream));
      uint64_t got[2] = {pub[0], pub[1]};
      if (got[0] == ~0ull) {
        HIPCHECK(hipMemcpyAsync(c.host_pinned, ctr + 10, 16, hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        got[0] = c.host_pinned[0];
        got[1] = c.host_pinned[1];
        if (dn) got[1] = 0;  // a misfit of this level is in the end-of-run word
      }
      const uint64_t m = got[0] & 0xFFFFFFFFu;
      // a level whose input size was on the device and that did not fit:
      // the end-of-run word repeats the run
      if (dn && (got[1] & 1ull)) return true;
      if (got[1] & 1ull) {
        HIPCHECK(hipMemsetAsync(ctr + 11, 0, 8, c.stream));
        r.chunks--;
        return false;
      }
      r.level[L] += m;
      if (m) level(L + 1, out.rows, m);
      return true;
    }
#define EXS(LL, MD) hipLaunchKernelGGL((k_expand<LL, MD>), dim3(grid * cgrow), dim3(nthr), 0, c.stream, g, kp, a)
#define ESC(LL) EXS(LL, M_COUNT)
#define ESF(LL) EXS(LL, M_FUSED)
#define ESK(LL) EXS(LL, M_COLLAPSE)
#define ESH(LL) EXS(LL, M_CHAIN)
#define ESP(LL) EXS(LL, M_PAIR)
    if (mode == M_COUNT) {
      CUTS_DISPATCH_L(L, ESC)
    } else if (mode == M_FUSED) {
      CUTS_DISPATCH_L(L, ESF)
    } else if (mode == M_COLLAPSE) {
      CUTS_DISPATCH_PAIR(L, ESK)
    } else if (mode == M_CHAIN) {
      CUTS_DISPATCH_PAIR(L, ESH)
    } else {
      CUTS_DISPATCH_PAIR(L, ESP)
    }
#undef ESP
#undef ESH
#undef ESK
#undef ESF
#undef ESC
#undef EXS
    HIPCHECK(hipGetLastError());
    if (pair_common && !aborted) pair_intersections(L, rows, n, in, true);
    return true;
  }

  // Second pass of the pair level with u != w: subtracts, for every prefix
  // with survivors, survivors x |N(u) & N(w)| (degree-filtered).  The
  // planning arrays of the input level are reused once the expansion is
  // done.
  void pair_intersections(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in, bool spec = false) {
    const uint32_t grid = grid_for(n, BLOCK, grid_cap);
    pbegin(false);
#define PP(LL) hipLaunchKernelGGL((k_pair_prep<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.nsurv, in.minlen, in.minsel, in.pbase, ctr)
    CUTS_DISPATCH_PAIR(L, PP)
#undef PP
    HIPCHECK(hipGetLastError());
    pend(1, (double)n * (4.0 * 2 + 16.0 + 5.0), false);
    // the pair level's bases (in.pbase) were read by the expansion that has
    // run; the same per-prefix slots now take the walk's descriptors
    common_pass(L, rows, n, in, CommonArgs{kp.pr_a, kp.pr_b, kp.pr_tab, 0u, in.pbase}, in.nsurv, spec);
  }

  // Counts common neighbours over the work planned in in.minlen / in.minsel
  // (k_common; mode and weights as in ca), in chunks of the work space.
  void common_pass(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in, CommonArgs ca, uint32_t* weight,
                   bool spec = false) {
    if (spec) {
      // the pair level's second pass without the readback of its work
      scan_launch(in, n);
      const uint64_t tmax = tile_index_dyn(in, n, ~0ull, nullptr);
      ExpandArgs a{};
      a.rows = rows;
      a.work_off = in.work_off;
      a.minsel = in.minsel;
      a.tile_p = in.tile_p;
      a.dyn = ctr + 14;
      a.counters = ctr;
      a.nsurv = weight;
      a.xcd = (o.flags & 8u) ? 0 : 1;
      a.tilectr = (unsigned int*)(ctr + 4);
      a.stats = 0;
      r.chunks++;
#define PCS(LL) hipLaunchKernelGGL((k_common<LL>), dim3(grid_for(tmax, 1, gcap(9))), dim3(BLOCK), 0, c.stream, g, kp, a, ca)
      CUTS_DISPATCH_L(L, PCS)
#undef PCS
      HIPCHECK(hipGetLastError());
      return;
    }
    const uint64_t W = scan(in, n);
    if (W == 0 || add_work(W)) return;
    if (W <= MAX_SPEC_ITEMS) ensure_tiles(arena, in, (W + TILE - 1) / TILE);
    for (uint64_t w0 = 0; w0 < W; w0 += MAX_CHUNK) {
      if (past_deadline()) return;
      const uint64_t w1 = std::min(W, w0 + MAX_CHUNK);
      const uint64_t nt = (w1 - w0 + TILE - 1) / TILE;
      ensure_tiles(arena, in, nt);
      pbegin(false);
      hipLaunchKernelGGL(k_tile_index, dim3(grid_for(nt + 1, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, in.work_off,
                         n, w0, w1, nt, in.tile_p, (unsigned int*)(ctr + 4), nullptr);
      pend(3, (double)(nt + 1) * (8.0 * (64 - __builtin_clzll(n + 2)) + 4.0), false);
      ExpandArgs a{};
      a.rows = rows;
      a.work_off = in.work_off;
      a.minsel = in.minsel;
      a.tile_p = in.tile_p;
      a.w_begin = w0;
      a.w_end = w1;
      a.ntiles = nt;
      a.counters = ctr;
      a.nsurv = weight;
      a.xcd = (o.flags & 8u) ? 0 : 1;
      a.tilectr = (unsigned int*)(ctr + 4);
      a.stats = prof ? 1 : 0;
      r.chunks++;
      pbegin(true);
#define PC(LL) hipLaunchKernelGGL((k_common<LL>), dim3(grid_for(nt, 1, gcap(9))), dim3(BLOCK), 0, c.stream, g, kp, a, ca)
      CUTS_DISPATCH_L(L, PC)
#undef PC
      HIPCHECK(hipGetLastError());
      pend(9, 0, true);
    }
  }

  // Chain level, before the rows of level k-3 are written from the rows of
  // level k-4 (L = k-3 words): per parent row, the sum over its older
  // positions i of |N(u) & N(r_i)| (those whose degree takes completions
  // away), shared by all its children; into in.nsurv.
  void chain_shared(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in) {
    ensure_aux(arena, in.nsurv, in.nsurv_cap, n);
    if (g.h2_off && kp.ch_ta <= 2) {
      const uint32_t grid = grid_for(n, BLOCK, grid_cap);
      pbegin(false);
#define CS(LL) hipLaunchKernelGGL((k_chain_shared_h2<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.nsurv)
      CUTS_DISPATCH_PAIR(L, CS)
#undef CS
      HIPCHECK(hipGetLastError());
      pend(1, (double)n * (4.0 * L + 4.0), false);
      return;
    }
    // older positions adjacent to u's in the query: per-arc table lookups
    uint32_t arc_mask = 0;
    if (g.arc_cn && kp.ch_ta <= 2)
      for (uint32_t i = 0; i < L; ++i)
        if (i != kp.ch_u && ((kp.qadj[kp.ch_u] >> i) & 1)) arc_mask |= 1u << i;
    if (arc_mask) {
      const uint32_t grid = grid_for(n, BLOCK, grid_cap);
      pbegin(false);
#define CA(LL) hipLaunchKernelGGL((k_chain_shared_arc<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, arc_mask, in.nsurv)
      CUTS_DISPATCH_PAIR(L, CA)
#undef CA
      HIPCHECK(hipGetLastError());
      pend(1, (double)n * (4.0 * L + 4.0), false);
    } else {
      HIPCHECK(hipMemsetAsync(in.nsurv, 0, n * sizeof(uint32_t), c.stream));
    }
    for (uint32_t i = 0; i < L && !aborted; ++i) {
      if (i == kp.ch_u || ((arc_mask >> i) & 1)) continue;
      const uint32_t grid = grid_for(n, BLOCK, grid_cap);
      pbegin(false);
#define CP(LL) hipLaunchKernelGGL((k_chain_prep<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, i, in.minlen, in.minsel, in.nsurv)
      CUTS_DISPATCH_PAIR(L, CP)
#undef CP
      HIPCHECK(hipGetLastError());
      pend(1, (double)n * (4.0 * 2 + 16.0 + 5.0), false);
      common_pass(L, rows, n, in, CommonArgs{kp.ch_u, i, kp.ch_ta, 2u, nullptr}, in.nsurv);
    }
  }

  // Expansion chunks of the rows (L words) under plan pk whose survivors go
  // into their row's block of in.cand (M_GWRITE), over the work planned in
  // in.minlen / in.minsel.
  void write_candidates(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in, const KPlan& pk) {
    const uint64_t W = scan(in, n);
    if (W == 0 || add_work(W)) return;
    for (uint64_t w0 = 0; w0 < W; w0 += MAX_CHUNK) {
      if (past_deadline()) return;
      const uint64_t w1 = std::min(W, w0 + MAX_CHUNK);
      const uint64_t nt = (w1 - w0 + TILE - 1) / TILE;
      ensure_tiles(arena, in, nt);
      hipLaunchKernelGGL(k_tile_index, dim3(grid_for(nt + 1, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, in.work_off,
                         n, w0, w1, nt, in.tile_p, (unsigned int*)(ctr + 4), nullptr);
      ExpandArgs a{};
      a.rows = rows;
      a.work_off = in.work_off;
      a.minsel = in.minsel;
      a.tile_p = in.tile_p;
      a.w_begin = w0;
      a.w_end = w1;
      a.ntiles = nt;
      a.counters = ctr;
      a.nsurv = in.nsurv;
      a.coff = in.coff;
      a.cand = in.cand;
      a.xcd = (o.flags & 8u) ? 0 : 1;
      a.tilectr = (unsigned int*)(ctr + 4);
      a.stats = prof ? 1 : 0;
      r.chunks++;
      const uint32_t grid = grid_for(nt, 1, gcap(4));
      pbegin(true);
#define GW(LL) hipLaunchKernelGGL((k_expand<LL, M_GWRITE>), dim3(grid), dim3(BLOCK), 0, c.stream, g, pk, a)
      CUTS_DISPATCH_PAIR(L, GW)
#undef GW
      HIPCHECK(hipGetLastError());
      pend(4, 0, true);  // an expansion that writes (candidates)
    }
  }

  // Chain level with shared candidates, before the rows of level k-4 are
  // written from the rows of level k-5 (L = k-4 words): the candidates of
  // position k-3 for every row of level k-5 (its constraints and threshold
  // checked, injective against the row), grouped by row into in.cand with
  // block starts in in.coff.
  void build_candidates(uint32_t L, const uint32_t* rows, uint64_t n, LevelBuf& in) {
    KPlan vk = kp;  // position k-3 evaluated at position k-4
    vk.nj[L] = kp.nj[L + 1];
    for (uint32_t i = 0; i < 2 * kMaxQuery; ++i) vk.jc[L][i] = kp.jc[L + 1][i];
    vk.sig_in[L] = kp.sig_in[L + 1];
    vk.sig_out[L] = kp.sig_out[L + 1];
    const uint32_t grid = grid_for(n, BLOCK, grid_cap);
    auto plan_vk = [&]() {
      pbegin(false);
#define PV(LL) hipLaunchKernelGGL((k_plan<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, vk, rows, n, in.minlen, in.minsel, nullptr, nullptr, nullptr, ptbl, nullptr, ctr)
      CUTS_DISPATCH_PAIR(L, PV)
#undef PV
      HIPCHECK(hipGetLastError());
      pend(1, (double)n * (4.0 * L + 8.0 * vk.nj[L] + 5.0), false);
    };
    sib_live = false;
    ensure_aux(arena, in.nsurv, in.nsurv_cap, n);
    ensure_aux(arena, in.coff, in.coff_cap, n + 1);
    // a row's shortest constraint list bounds its candidates: blocks sized by
    // the bounds (the scan of the planner's lengths), filled in one pass, with
    // the filled length of every block in in.nsurv
    plan_vk();
    const uint64_t total = scan(in, n);
    // beyond the limit, the children of these rows find their candidates
    // themselves (the fused chain expansion)
    if (total > cand_limit) return;
    HIPCHECK(hipMemcpyAsync(in.coff, in.work_off, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToDevice, c.stream));
    ensure_aux(arena, in.cand, in.cand_cap, total);
    HIPCHECK(hipMemsetAsync(in.nsurv, 0, n * sizeof(uint32_t), c.stream));
    write_candidates(L, rows, n, in, vk);
    sib_live = !aborted;
  }

  // Chain level with shared candidates over the rows of level k-4 (L = k-3
  // words): every row paired with its parent's candidates (k_chain_sib).
  void sib_level(uint32_t L, const uint32_t* rows, uint64_t n) {
    LevelBuf& in = lv[L - 1];
    LevelBuf& par = lv[L - 2];
    const uint32_t grid = grid_for(n, BLOCK, grid_cap);
    if (prof) {
      // merge model: the expansion of these rows (the planner's figure)
      unsigned long long* merge = ctr + 8;
      HIPCHECK(hipMemsetAsync(merge, 0, 8, c.stream));
#define PS(LL) hipLaunchKernelGGL((k_plan<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.minlen, in.minsel, nullptr, nullptr, nullptr, ptbl, merge, ctr)
      CUTS_DISPATCH_PAIR(L, PS)
#undef PS
      HIPCHECK(hipGetLastError());
      r.merge_bytes[10] += (double)read_u64(merge);
    }
    // per row: its parent's block (start and length) and the candidate-free
    // part of the chain closed form
    ensure_aux(arena, in.pbase, in.pbase_cap, n);
    ensure_aux(arena, in.coff, in.coff_cap, n + 1);
    pbegin(false);
#define SBB(LL) hipLaunchKernelGGL((k_sib_base<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.parent, par.nsurv, par.coff, in.nsurv, ctab, in.minlen, in.pbase, in.coff)
    CUTS_DISPATCH_PAIR(L, SBB)
#undef SBB
    HIPCHECK(hipGetLastError());
    pend(1, (double)n * (4.0 * L + 60.0), false);
    const uint64_t W = scan(in, n);
    if (W == 0 || add_work(W)) return;
    for (uint64_t w0 = 0; w0 < W; w0 += MAX_CHUNK) {
      if (past_deadline()) return;
      const uint64_t w1 = std::min(W, w0 + MAX_CHUNK);
      const uint64_t nt = (w1 - w0 + TILE - 1) / TILE;
      ensure_tiles(arena, in, nt);
      hipLaunchKernelGGL(k_tile_index, dim3(grid_for(nt + 1, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, in.work_off,
                         n, w0, w1, nt, in.tile_p, (unsigned int*)(ctr + 4), nullptr);
      SibArgs a{};
      a.rows = rows;
      a.pb = in.pbase;
      a.rstart = in.coff;
      a.cand = par.cand;
      a.shared = in.nsurv;
      a.work_off = in.work_off;
      a.tile_p = in.tile_p;
      a.w_begin = w0;
      a.w_end = w1;
      a.ntiles = nt;
      a.counters = ctr;
      a.tilectr = (unsigned int*)(ctr + 4);
      a.xcd = (o.flags & 8u) ? 0 : 1;
      a.stats = prof ? 1 : 0;
      r.chunks++;
      pbegin(true);
#define SB(LL) hipLaunchKernelGGL((k_chain_sib<LL>), dim3(grid_for(nt, 1, gcap(10))), dim3(BLOCK), 0, c.stream, g, kp, a, ctab)
      CUTS_DISPATCH_PAIR(L, SB)
#undef SB
      HIPCHECK(hipGetLastError());
      pend(10, 0, true);
    }
  }

  // Chain level over the rows of level k-3 (L = k-2 words): closed form per
  // row (k_chain_row), then the intersection with the newest vertex.
  void chain_rows(uint32_t L, const uint32_t* rows, uint64_t n) {
    LevelBuf& in = lv[L - 1];
    const uint32_t grid = grid_for(n, BLOCK, grid_cap);
    pbegin(true);
#define CR(LL) hipLaunchKernelGGL((k_chain_row<LL>), dim3(grid), dim3(BLOCK), 0, c.stream, g, kp, rows, n, in.parent, lv[L - 2].nsurv, ctab, in.minlen, in.minsel, ctr, (int)prof)
    CUTS_DISPATCH_L(L, CR)
#undef CR
    HIPCHECK(hipGetLastError());
    pend(10, 0, true);
    const bool arc = g.arc_cn && ((kp.qadj[kp.ch_u] >> (L - 1)) & 1) && kp.ch_ta <= 2;
    if (!(g.h2_off && kp.ch_ta <= 2) && !arc)
      common_pass(L, rows, n, in, CommonArgs{kp.ch_u, L - 1, kp.ch_ta, 1u, nullptr}, nullptr);
  }

  void run() {
    k = plan.k;
    if (k == 0 || k > kMaxQuery) throw Error(-3, "query must have between 1 and 16 vertices");
    kp = make_kplan(plan, dg.symmetric, dg.loops == 0);
    g = view_of(dg);
    if (o.flags & 32u) g.hub_id = nullptr;  // CUTS_FLAG_NO_BITMAP
    if (o.flags & 512u) g.h2_off = nullptr;  // CUTS_FLAG_NO_TWO_HOP
    if (o.flags & 512u) g.h2_did = nullptr;
    if (o.flags & 512u) g.hr_id = nullptr;
    if (o.flags & 2048u) g.arc_cn = nullptr;  // CUTS_FLAG_NO_ARC_COMMON
    write = (o.flags & 4u) != 0;
    // CUTS_FLAG_SYNC_WRITE: each chunk's rows copied before the next chunk runs
    stream_emb = write && o.emb_rows && !(o.flags & 4096u);
    prof = (o.flags & 16u) != 0;
    closed_last = !write && !(o.flags & 1u) && kp.sc_ok && k >= 2;
    fuse = closed_last && !(o.flags & 2u) && k >= 3;
    pair = fuse && kp.pr_ok && !(o.flags & 128u);
    chain = fuse && !pair && kp.ch_ok && !(o.flags & 256u);
    // 6 blocks of 4 waves per CU: measured best on the power-law workloads
    // (the tiles are taken dynamically, so the grid only sets residency)
    if (o.expand_threads != 0 && o.expand_threads != 64 && o.expand_threads != 128 && o.expand_threads != 256)
      throw Error(-1, "expand_threads must be 64, 128 or 256 (or 0)");
    if (o.blocks_per_cu > 64) throw Error(-1, "blocks_per_cu must be at most 64");
    for (int i = 0; i < 16; ++i)
      if (o.class_bpc[i] > 64) throw Error(-1, "class_blocks_per_cu entries must be at most 64");
    grid_cap = (uint32_t)c.cus * (o.blocks_per_cu ? o.blocks_per_cu : 6);
    budget = o.budget ? o.budget : c.default_budget;
    // split the budget evenly over the materialised levels 1..k-2 (or k-1
    // when writing embeddings); each row of level l costs (l+1)*4 bytes plus
    // 13 bytes of planning arrays.
    uint32_t nlev = 0;
    for (uint32_t l = 1; l < k; ++l)
      if (l < k - 1 || write) ++nlev;
    for (uint32_t l = 1; l < k; ++l) {
      const uint64_t per = nlev ? budget / nlev : budget;
      // rows, planning arrays (minlen, minsel, scan entry, closed-form base)
      // and, for the levels that feed the pair / chain level, their bases,
      // survivor sums and parent indices
      level_cap[l] = std::max<uint64_t>(per / ((l + 1) * 4 + 17 + 24), 1024);
    }
    // shared chain candidates: up to one level's share of the budget
    cand_limit = std::max<uint64_t>((nlev ? budget / nlev : budget) / 4, 1u << 10);
    if (closed_last && kp.sc_thr > 1) tbl = cnt_table(c, dg, arena, kp.sc_thr);
    // with the two-hop index the chain is counted inside the expansion of
    // position k-3 (no rows of level k-3 are stored)
    chain_fused = chain && g.h2_off && kp.ch_ta <= 2;
    // a collapsing pair level is counted inside the expansion of position
    // k-4 (no rows of level k-4 are stored); with a deadline, in the planner
    // of level k-3 instead (the rolled-back counters are those of 0..2)
    collapse_fused = pair && kp.pr_collapse && pair_cab_known() && k >= 5 && !(o.time_limit_ms > 0);
    // the fused chain with position k-3's candidates shared by siblings
    sib = chain_fused && kp.ch_sib && !(o.flags & 1024u);
    if (chain) {
      ctab.ta = kp.ch_ta > 1 ? cnt_table(c, dg, arena, kp.ch_ta) : nullptr;
      ctab.tb = kp.ch_tb > 1 ? cnt_table(c, dg, arena, kp.ch_tb) : nullptr;
      ctab.H = chain_table(c, dg, arena, kp.ch_ta, kp.ch_tb);
    }
    if (pair) {
      ptbl.a = kp.pr_ta > 1 ? cnt_table(c, dg, arena, kp.pr_ta) : nullptr;
      ptbl.b = kp.pr_tb > 1 ? cnt_table(c, dg, arena, kp.pr_tb) : nullptr;
      ptbl.ab = kp.pr_tab > 1 ? cnt_table(c, dg, arena, kp.pr_tab) : nullptr;
    }
    // level sizes of this query's last complete run (a hint only: a level
    // that outgrows it repeats the run)
    std::string hint_key = std::to_string((uintptr_t)&dg) + "/" + std::to_string((uintptr_t)&plan);
    for (uint32_t v : plan.order) hint_key += "," + std::to_string(v);
    for (uint32_t v : plan.par) hint_key += ";" + std::to_string(v);
    hint_key += "/" + std::to_string(o.root_begin) + "/" + std::to_string(o.root_end) + "/" +
                std::to_string(o.root_stride) + "/" + std::to_string(o.flags) + "/" + std::to_string(o.trunks);
    if (!arena) {
      const auto it = c.level_hints.find(hint_key);
      hint = it == c.level_hints.end() ? nullptr : it->second.data();
    }
    // level 0
    const uint32_t rb = o.root_stride ? o.root_begin : 0;
    const uint32_t re = o.root_stride ? std::min(o.root_end, dg.V) : dg.V;
    const uint32_t rs = o.root_stride ? o.root_stride : 1;
    const uint32_t n0c = re > rb ? (re - rb + rs - 1) / rs : 0;
    LevelBuf& l0 = lv[0];
    ensure_rows(arena, l0, std::max<uint32_t>(n0c, 1), 1);
    // counts, tile counters, merge bytes, level counts, rows written,
    // speculation status words and device-side work: one clear
    HIPCHECK(hipMemsetAsync(ctr, 0, 14 * 8, c.stream));
    HIPCHECK(hipEventRecord(c.ev0, c.stream));
    const auto t_start = std::chrono::steady_clock::now();
    uint64_t n0 = 0;
    bool last_root_read = false;
    // level 0 of at most one planning block's roots in the last complete
    // run: level 1 takes its size from the device
    const bool chain0 = hint && hint[0] > 0 && hint[0] <= (uint64_t)SCAN_TILE && k > 1 && !(o.time_limit_ms > 0) &&
                        n0c && n0c <= (1u << 16) && chainable(1);
    if (n0c && n0c <= (1u << 16)) {
      pbegin(false);
      c.host_res[2] = ~0ull;
      hipLaunchKernelGGL(k_level0_small, dim3(1), dim3(L0_THREADS), 0, c.stream, dg.sig, kp.sig_in[0], kp.sig_out[0], rb,
                         re, rs, n0c, l0.rows, c.host_res_dev + 2, chain0 ? ctr + 10 : nullptr);
      HIPCHECK(hipGetLastError());
      pend(0, 16.0 * n0c, false);
      if (chain0) {
        chained_mask |= 1u;
        n0 = SCAN_TILE;  // a bound; the size is read at the end
      } else {
        if (!wait_published(2)) throw Error(-4, "level 0: no size published");
        n0 = ((const volatile uint64_t*)c.host_res)[2];
      }
    } else if (n0c) {
      pbegin(false);
      hipLaunchKernelGGL(k_level0_flags, dim3(grid_for(n0c, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, dg.sig,
                         kp.sig_in[0], kp.sig_out[0], rb, re, rs, n0c, l0.minlen);
      HIPCHECK(hipGetLastError());
      pend(0, 8.0 * n0c + 4.0 * n0c, false);
      n0 = scan(l0, n0c);
      pbegin(false);
      hipLaunchKernelGGL(k_level0_scatter, dim3(grid_for(n0c, BLOCK, 65535)), dim3(BLOCK), 0, c.stream, l0.minlen,
                         l0.work_off, rb, rs, n0c, l0.rows);
      HIPCHECK(hipGetLastError());
      pend(0, 12.0 * n0c + 4.0 * n0, false);
    }
    r.level[0] = n0;
    r.complete = 1;
    if (prof) r.merge_bytes[0] = 8.0 * n0c + (k > 1 ? 16.0 * (double)n0 : 0.0);  // signature scan, level-0 nodes
    if (k == 1) {
      r.count = n0;
      r.roots_done = n0;
    } else if (n0) {
      uint32_t trunks = o.trunks ? o.trunks : 1;
      if (trunks > n0) trunks = (uint32_t)n0;
      const uint64_t ts = (n0 + trunks - 1) / trunks;
      if (o.time_limit_ms > 0 && !write) {
        has_deadline = true;
        deadline = t_start + std::chrono::microseconds((int64_t)(o.time_limit_ms * 1000.0));
      }
      for (uint64_t s = 0; s < n0; s += ts) {
        // snapshot so that a trunk interrupted by the deadline can be undone
        uint64_t snap[3] = {0, 0, 0};
        uint64_t snap_level[kMaxQuery];
        std::memcpy(snap_level, r.level, sizeof(snap_level));
        if (has_deadline && !write) {
          if (past_deadline()) break;
          HIPCHECK(hipMemcpyAsync(c.host_pinned, ctr, 24, hipMemcpyDeviceToHost, c.stream));
          HIPCHECK(hipStreamSynchronize(c.stream));
          snap[0] = c.host_pinned[0];
          snap[1] = c.host_pinned[1];
          snap[2] = c.host_pinned[2];
        }
        level(1, l0.rows + s, std::min(ts, n0 - s), (chained_mask & 1u) ? ctr + 10 : nullptr);
        if (aborted) {
          c.host_pinned[0] = snap[0];
          c.host_pinned[1] = snap[1];
          c.host_pinned[2] = snap[2];
          HIPCHECK(hipMemcpyAsync(ctr, c.host_pinned, 24, hipMemcpyHostToDevice, c.stream));
          HIPCHECK(hipStreamSynchronize(c.stream));
          std::memcpy(r.level, snap_level, sizeof(snap_level));
          break;
        }
        r.roots_done = std::min(n0, s + ts);
      }
      if (aborted || r.roots_done < n0) r.complete = 0;
      // counts, the work of speculative levels (summed on the device) and
      // their status, and the last processed root: one readback
      const volatile uint64_t* hp = c.host_res + 16;
      c.host_res[16] = ~0ull;
      const bool dev0 = (chained_mask & 1u) != 0;
      hipLaunchKernelGGL(k_publish_end, dim3(1), dim3(64), 0, c.stream, ctr,
                         (r.roots_done && !dev0) ? l0.rows + r.roots_done - 1 : nullptr,
                         dev0 ? (const uint32_t*)l0.rows : nullptr, (const unsigned long long*)c.lvcnt,
                         c.host_res_dev + 16);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(c.stream));
      if (r.roots_done) {
        r.last_root = (uint32_t)(hp[14] & 0xFFFFFFFFu);
        last_root_read = true;
      }
      spec_overflow = (hp[12] & 1ull) != 0;
      r.work += hp[13];
      // sizes of the levels whose successor read them on the device (level
      // 0's: every root processed in one trunk)
      for (uint32_t l = 0; l < k; ++l)
        if ((chained_mask >> l) & 1) r.level[l] = hp[48 + l];
      if (dev0) {
        r.roots_done = r.level[0];
        last_root_read = r.roots_done > 0;
        if (last_root_read) r.last_root = (uint32_t)(hp[14] & 0xFFFFFFFFu);
      }
      if (!write) {
        r.count = hp[0];
        r.level[k - 1] = r.count;
        if (collapse_fused) {
          r.level[k - 4] = hp[1];
          r.level[k - 3] = hp[2];
          r.level[k - 2] = hp[9];
        } else if (pair || chain_fused) {
          r.level[k - 3] = hp[1];
          r.level[k - 2] = hp[2];
        } else if (fuse) {
          r.level[k - 2] = hp[1];
        }
      }
      if (r.work_capped) r.count = 0;
    }
    if (r.roots_done && !last_root_read) {
      // the last processed root, for callers that restate the sample
      HIPCHECK(hipMemcpyAsync(c.host_pinned + 2, l0.rows + r.roots_done - 1, 4, hipMemcpyDeviceToHost, c.stream));
      HIPCHECK(hipStreamSynchronize(c.stream));
      r.last_root = (uint32_t)(c.host_pinned[2] & 0xFFFFFFFFu);
    }
    if (emb_chunks) {
      // the timed window ends once every embedding row is in the caller's buffer
      HIPCHECK(hipStreamWaitEvent(c.stream, c.copy_ev[0], 0));
      HIPCHECK(hipStreamWaitEvent(c.stream, c.copy_ev[1], 0));
    }
    HIPCHECK(hipEventRecord(c.ev1, c.stream));
    HIPCHECK(hipEventSynchronize(c.ev1));
    if (emb_chunks) HIPCHECK(hipStreamSynchronize(c.copy_stream));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    r.device_ms = ms;
    if (!arena && r.complete && !spec_overflow && !aborted && !r.work_capped) {
      // at most 2^16 queries remembered (about 10 MB); past that the
      // table starts over
      if (c.level_hints.size() >= (1u << 16) && !c.level_hints.count(hint_key)) c.level_hints.clear();
      std::array<uint64_t, kMaxQuery> sizes{};
      for (uint32_t l = 0; l < k; ++l) sizes[l] = r.level[l];
      c.level_hints[hint_key] = sizes;
    }
  }
};

// Two-hop index on the GPU (symmetric graphs): for every vertex u, the
// vertices v != u that share a neighbour with u, ascending, with
// |N(u) & N(v)|.  Each block owns a V-word counter row, zero between
// vertices.  For its vertex u the block counts every wedge u-x-v with an
// atomic (a wave per x, its lanes over N(x)), then reads the row in vertex
// order -- each wave a contiguous quarter, so no barrier inside the sweeps:
// pass 1 counts the distinct v (u's entries), pass 2 writes them at u's
// offset (and into u's dense uint16 row when u has one) -- and clears it.
// Counter reads are agent-scope atomic loads, so they see the atomics'
// results at L2 rather than a stale L1 line.
__global__ __launch_bounds__(BLOCK) void k_two_hop(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                  uint32_t V, uint32_t* __restrict__ rows, int pass,
                                                  unsigned long long* __restrict__ entries,
                                                  const uint64_t* __restrict__ h2_off, uint32_t* __restrict__ hv,
                                                  uint32_t* __restrict__ hc, const int32_t* __restrict__ did,
                                                  uint16_t* __restrict__ dense) {
  __shared__ uint32_t s_cnt[BLOCK / 64];
  uint32_t* cnt = rows + (uint64_t)blockIdx.x * V;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
  constexpr uint32_t NW = BLOCK / 64;
  const uint32_t q0 = (uint32_t)((uint64_t)V * wave / NW), q1 = (uint32_t)((uint64_t)V * (wave + 1) / NW);
  for (uint32_t u = blockIdx.x; u < V; u += gridDim.x) {
    const uint32_t a0 = off[u], a1 = off[u + 1];
    for (uint32_t i = a0 + wave; i < a1; i += NW) {
      const uint32_t x = adj[i];
      const uint32_t b1 = off[x + 1];
      for (uint32_t j = off[x] + lane; j < b1; j += 64) {
        const uint32_t v = adj[j];
        if (v != u) atomicAdd(&cnt[v], 1u);
      }
    }
    __syncthreads();
    // sweep A: nonzero entries of this wave's quarter
    uint32_t mine = 0;
    for (uint32_t v0 = q0; v0 < q1; v0 += 64) {
      const uint32_t v = v0 + lane;
      const uint32_t c = v < q1 ? __hip_atomic_load(&cnt[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      mine += (uint32_t)__popcll(__ballot(c != 0));
      if (pass == 1 && c) __hip_atomic_store(&cnt[v], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_cnt[wave] = mine;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) {
      const uint32_t t = s_cnt[w];
      if (w < wave) before += t;
      total += t;
    }
    if (pass == 1) {
      if (wave == 0) entries[u] = total;  // the same value from every lane of wave 0
    } else {
      // sweep B: write the entries in vertex order, clear the row
      const int32_t dr = did ? did[u] : -1;
      uint64_t o = h2_off[u] + before;
      for (uint32_t v0 = q0; v0 < q1; v0 += 64) {
        const uint32_t v = v0 + lane;
        const uint32_t c = v < q1 ? __hip_atomic_load(&cnt[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
        const uint64_t m = __ballot(c != 0);
        if (c) {
          const uint64_t at = o + lane_rank(m);
          hv[at] = v;
          hc[at] = c;
          if (dr >= 0) dense[(uint64_t)dr * V + v] = (uint16_t)c;
          __hip_atomic_store(&cnt[v], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        o += (uint64_t)__popcll(m);
      }
    }
    __syncthreads();
  }
}

// Hub rows (graphs without the two-hop index): row r holds |N(u) & N(v)|
// for u = hubs[r] and every v != u as uint16 (a count stays below the
// largest degree, < 2^16).  One block per job (row, first of up to HR_JOB
// consecutive entries x of u's list): each wave takes an x, its lanes add
// one to the count of every v in N(x).  Two counts share a 32-bit word, so
// an add is a 32-bit atomic of 1 or 1 << 16 into its half, which never
// carries (its count ends below 2^16).
constexpr uint32_t HR_JOB = 64;
__global__ __launch_bounds__(BLOCK) void k_hub_rows(const uint32_t* __restrict__ off, const uint32_t* __restrict__ adj,
                                                   uint32_t V, const uint32_t* __restrict__ hubs,
                                                   const uint2* __restrict__ jobs, uint32_t* __restrict__ words) {
  const uint2 jb = jobs[blockIdx.x];
  const uint32_t u = hubs[jb.x];
  const uint32_t a1 = min(jb.y + HR_JOB, off[u + 1]);
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)jb.x * V;
  for (uint32_t i = jb.y + wave; i < a1; i += BLOCK / 64) {
    const uint32_t x = adj[i];
    const uint32_t b1 = off[x + 1];
    for (uint32_t j = off[x] + lane; j < b1; j += 64) {
      const uint32_t v = adj[j];
      if (v != u) {
        const uint64_t gi = base + v;
        atomicAdd(&words[gi >> 1], 1u << ((uint32_t)(gi & 1) * 16));
      }
    }
  }
}

// Plain streaming copy of n 16-byte words: the calibration kernel for HBM
// byte counters (it moves exactly 16n bytes each way).
__global__ __launch_bounds__(BLOCK) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
    dst[i] = src[i];
}

}  // namespace

void* ScratchArena::carve(uint64_t bytes) {
  const uint64_t at = (used + 255) & ~255ull;
  if (at + bytes > size)
    throw Error(-5, "no-allocation scratch of " + std::to_string(size) + " bytes exhausted (" +
                        std::to_string(at + bytes) + " needed)");
  used = at + bytes;
  return base + at;
}

// Degree-threshold table: cached on the device graph, or, with an arena,
// built in the scratch for this call only (and not cached).
const uint32_t* cnt_table(Context& c, DeviceGraph& dg, ScratchArena* arena, uint32_t thr) {
  // concurrent matchers (cuts_match_many) share the device graph
  std::lock_guard<std::recursive_mutex> lk(dg.tables_mu);
  auto it = dg.cnt_tables.find(thr);
  if (it != dg.cnt_tables.end()) return it->second;
  uint32_t* t = nullptr;
  dev_alloc(arena, t, dg.V + 1);
  const GView g = view_of(dg);
  hipLaunchKernelGGL(k_cnt_table, dim3(grid_for(dg.V, BLOCK / 64, 65535)), dim3(BLOCK), 0, c.stream, g, thr, t);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (!arena) dg.cnt_tables.emplace(thr, t);
  return t;
}
const uint32_t* cnt_table(Context& c, DeviceGraph& dg, uint32_t thr) { return cnt_table(c, dg, nullptr, thr); }

const uint64_t* chain_table(Context& c, DeviceGraph& dg, ScratchArena* arena, uint32_t ta, uint32_t tb) {
  std::lock_guard<std::recursive_mutex> lk(dg.tables_mu);
  const auto key = std::make_pair(ta, tb);
  auto it = dg.chain_tables.find(key);
  if (it != dg.chain_tables.end()) return it->second;
  uint64_t* t = nullptr;
  dev_alloc(arena, t, 2 * (uint64_t)dg.V + 1);
  const GView g = view_of(dg);
  const uint32_t* tbt = tb > 1 ? cnt_table(c, dg, arena, tb) : nullptr;
  hipLaunchKernelGGL(k_chain_table, dim3(grid_for(dg.V, BLOCK / 64, 65535)), dim3(BLOCK), 0, c.stream, g, ta, tbt, t);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (!arena) dg.chain_tables.emplace(key, t);
  return t;
}
const uint64_t* chain_table(Context& c, DeviceGraph& dg, uint32_t ta, uint32_t tb) {
  return chain_table(c, dg, nullptr, ta, tb);
}

double copy_probe(Context& c, uint64_t bytes, uint32_t iters) {
  HIPCHECK(hipSetDevice(c.device));
  const uint64_t n = std::max<uint64_t>(bytes / 16, 1);
  uint4 *a = nullptr, *b = nullptr;
  dev_alloc(a, n);
  dev_alloc(b, n);
  HIPCHECK(hipMemsetAsync(a, 1, n * 16, c.stream));
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n + BLOCK - 1) / BLOCK, (uint64_t)c.cus * 16);
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(BLOCK), 0, c.stream, a, b, n);  // warm
  HIPCHECK(hipEventRecord(c.kev[0], c.stream));
  for (uint32_t i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(BLOCK), 0, c.stream, a, b, n);
  HIPCHECK(hipEventRecord(c.kev[1], c.stream));
  HIPCHECK(hipEventSynchronize(c.kev[1]));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, c.kev[0], c.kev[1]));
  (void)hipFree(a);
  (void)hipFree(b);
  return (double)iters * 2.0 * (double)n * 16.0 / (ms / 1e3) / 1e9;
}

KPlan make_kplan(const Plan& p, bool symmetric, bool loop_free) {
  KPlan kp;
  std::memset(&kp, 0, sizeof(kp));
  kp.k = p.k;
  for (uint32_t i = 0; i < p.k; ++i) {
    kp.sig_in[i] = p.sig_in[i];
    kp.sig_out[i] = p.sig_out[i];
    uint32_t n = 0;
    if (symmetric) {
      const uint32_t m = p.pmask[i] | p.cmask[i];
      for (uint32_t j = 0; j < i; ++j)
        if (m >> j & 1) kp.jc[i][n++] = (uint8_t)j;
    } else {
      for (uint32_t j = 0; j < i; ++j)
        if (p.pmask[i] >> j & 1) kp.jc[i][n++] = (uint8_t)j;
      for (uint32_t j = 0; j < i; ++j)
        if (p.cmask[i] >> j & 1) kp.jc[i][n++] = (uint8_t)(j | 0x80);
    }
    kp.nj[i] = n;
  }
  const uint32_t last = p.k - 1;
  if (symmetric && p.k >= 2 && kp.nj[last] == 1) {
    kp.sc_ok = 1;
    kp.sc_pos = kp.jc[last][0] & 31;
    kp.sc_thr = std::max(p.sig_in[last], p.sig_out[last]);
    uint32_t adj = 0;
    for (uint32_t i = 0; i < last; ++i) {
      if (i == kp.sc_pos) continue;
      const uint32_t a = std::max(i, kp.sc_pos), b = std::min(i, kp.sc_pos);
      if (((p.pmask[a] | p.cmask[a]) >> b) & 1) adj |= 1u << i;
    }
    kp.sc_adjmask = adj;
  }
  // the last two positions together: a = k-2 and b = k-1 not adjacent, one
  // backward neighbour each, both older than k-3 (so |A & B| depends on the
  // rows of level k-4 only)
  if (kp.sc_ok && loop_free && p.k >= 4 && kp.nj[last - 1] == 1) {
    const uint32_t pa = kp.jc[last - 1][0] & 31, pb = kp.sc_pos;
    if (pb != last - 1 && pa + 3 < p.k && pb + 3 < p.k) {
      kp.pr_ok = 1;
      kp.pr_a = pa;
      kp.pr_b = pb;
      kp.pr_ta = std::max(p.sig_in[last - 1], p.sig_out[last - 1]);
      kp.pr_tb = kp.sc_thr;
      kp.pr_tab = std::max(kp.pr_ta, kp.pr_tb);
      auto adjacent = [&](uint32_t i, uint32_t j) {
        const uint32_t a = std::max(i, j), b = std::min(i, j);
        return i != j && (((p.pmask[a] | p.cmask[a]) >> b) & 1);
      };
      for (uint32_t i = 0; i + 2 < p.k; ++i) {
        if (adjacent(i, pa)) kp.pr_ma |= 1u << i;
        if (adjacent(i, pb)) kp.pr_mb |= 1u << i;
      }
      kp.pr_adj = adjacent(pa, pb);
      const uint32_t c = last - 2;
      if (pa != pb && kp.nj[c] == 2 && std::max(p.sig_in[c], p.sig_out[c]) <= 2 && kp.pr_tab <= 2) {
        const uint32_t j0 = kp.jc[c][0] & 31, j1 = kp.jc[c][1] & 31;
        kp.pr_collapse = (j0 == pa && j1 == pb) || (j0 == pb && j1 == pa);
      }
    }
  }
  for (uint32_t i = 0; i < p.k; ++i)
    for (uint32_t j = 0; j < p.k; ++j)
      if (i != j && ((p.pmask[std::max(i, j)] | p.cmask[std::max(i, j)]) >> std::min(i, j) & 1)) kp.qadj[i] |= 1u << j;
  // the last two positions as a chain: k-1 hangs on k-2 only, k-2 on one
  // position older than k-3 only
  if (kp.sc_ok && loop_free && p.k >= 4 && kp.sc_pos == last - 1 && kp.nj[last - 1] == 1) {
    const uint32_t u = kp.jc[last - 1][0] & 31;
    if (u + 3 < p.k) {
      kp.ch_ok = 1;
      kp.ch_u = u;
      kp.ch_ta = std::max(p.sig_in[last - 1], p.sig_out[last - 1]);
      kp.ch_tb = kp.sc_thr;
      // position k-3's constraints all older than k-4 (and at least one)
      const uint32_t c = last - 2;
      bool sib = p.k >= 5 && kp.nj[c] >= 1;
      for (uint32_t i = 0; i < kp.nj[c]; ++i) sib = sib && (kp.jc[c][i] & 31) + 1 < c;
      kp.ch_sib = sib;
    }
  }
  return kp;
}

DeviceGraph::~DeviceGraph() {
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(device);
  for (uint32_t* p : {off_out, adj_out, sig}) (void)hipFree(p);
  if (nsig_out) (void)hipFree(nsig_out);
  if (!symmetric) {
    (void)hipFree(off_in);
    (void)hipFree(adj_in);
    if (nsig_in) (void)hipFree(nsig_in);
  }
  for (auto& kv : cnt_tables) (void)hipFree(kv.second);
  for (auto& kv : chain_tables) (void)hipFree(kv.second);
  if (hub_id) (void)hipFree(hub_id);
  if (bits) (void)hipFree(bits);
  if (arc_cn) (void)hipFree(arc_cn);
  for (void* p : {(void*)h2_off, (void*)h2_v, (void*)h2_c, (void*)h2_did, (void*)h2_dense, (void*)hr_id, (void*)hr})
    if (p) (void)hipFree(p);
  (void)hipSetDevice(prev);
}

Context::~Context() {
  (void)hipSetDevice(device);
  for (auto& b : lv) {
    for (void* p : {(void*)b.rows, (void*)b.rows2, (void*)b.minlen, (void*)b.minsel, (void*)b.work_off, (void*)b.bsum,
                    (void*)b.tile_p, (void*)b.cbase, (void*)b.pbase,
                    (void*)b.nsurv, (void*)b.parent, (void*)b.coff, (void*)b.cand})
      if (p) (void)hipFree(p);
  }
  if (counters) (void)hipFree(counters);
  if (host_pinned) (void)hipHostFree(host_pinned);
  if (host_res) (void)hipHostFree(host_res);
  if (lvcnt) (void)hipFree(lvcnt);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  for (hipEvent_t e : kev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : copy_ev)
    if (e) (void)hipEventDestroy(e);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (stream) (void)hipStreamDestroy(stream);
}

void context_init(Context& c, int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(-6, "no HIP device available");
  if (device < 0 || device >= n) throw Error(-1, "device index out of range");
  c.device = device;
  HIPCHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  c.cus = prop.multiProcessorCount;
  HIPCHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&c.ev0));
  HIPCHECK(hipEventCreate(&c.ev1));
  HIPCHECK(hipEventCreate(&c.kev[0]));
  HIPCHECK(hipEventCreate(&c.kev[1]));
  HIPCHECK(hipStreamCreateWithFlags(&c.copy_stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&c.copy_ev[0], hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&c.copy_ev[1], hipEventDisableTiming));
  HIPCHECK(hipMalloc(&c.counters, 16 * sizeof(unsigned long long)));
  HIPCHECK(hipHostMalloc(&c.host_pinned, 256, hipHostMallocDefault));
  HIPCHECK(hipHostMalloc(&c.host_res, 1024, hipHostMallocCoherent | hipHostMallocMapped));
  HIPCHECK(hipMalloc(&c.lvcnt, kMaxQuery * sizeof(unsigned long long)));
  HIPCHECK(hipHostGetDevicePointer((void**)&c.host_res_dev, c.host_res, 0));
  size_t fr = 0, tot = 0;
  HIPCHECK(hipMemGetInfo(&fr, &tot));
  // trie tables may take 40% of the free HBM by default
  c.default_budget = (uint64_t)(fr * 0.4);
  c.scratch = scratch_address(device);
}

void upload_graph(Context& c, const Graph& g, DeviceGraph& dg) {
  HIPCHECK(hipSetDevice(c.device));
  dg.device = c.device;
  dg.V = g.V;
  dg.A = g.arcs();
  dg.symmetric = g.symmetric;
  uint32_t md = 0;
  uint64_t loops = 0;
  for (uint32_t v = 0; v < g.V; ++v) {
    md = std::max(md, g.out_off[v + 1] - g.out_off[v]);
    loops += std::binary_search(g.out_adj.begin() + g.out_off[v], g.out_adj.begin() + g.out_off[v + 1], v);
  }
  dg.max_deg = md;
  dg.loops = loops;
  auto up = [&](uint32_t*& d, const std::vector<uint32_t>& h) {
    dev_alloc(d, h.size());
    HIPCHECK(hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  };
  up(dg.off_out, g.out_off);
  up(dg.adj_out, g.out_adj);
  up(dg.sig, g.sig);
  if (g.symmetric) {
    dg.off_in = dg.off_out;
    dg.adj_in = dg.adj_out;
  } else {
    up(dg.off_in, g.in_off);
    up(dg.adj_in, g.in_adj);
  }
  auto up_nsig = [&](uint2*& d, const std::vector<uint32_t>& adj) {
    std::vector<uint2> h(adj.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)adj.size(); ++i) h[i] = make_uint2(g.sig[2 * (uint64_t)adj[i]], g.sig[2 * (uint64_t)adj[i] + 1]);
    dev_alloc(d, std::max<size_t>(h.size(), 1));
    if (!h.empty()) HIPCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(uint2), hipMemcpyHostToDevice));
  };
  up_nsig(dg.nsig_out, g.out_adj);
  if (g.symmetric) dg.nsig_in = dg.nsig_out;
  else up_nsig(dg.nsig_in, g.in_adj);
  // Hub bitmaps: the lowest power-of-two degree threshold (>= 32) whose
  // bitmaps fit the budget (1/32 of free HBM, at most 4 GiB).
  if (g.symmetric && g.V >= 64) {
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    uint64_t cap = std::min<uint64_t>(fr / 32, 4ull << 30);
    // CUTS_HUB_BITMAP_MB replaces the cap (at most half of free HBM)
    if (const char* e = std::getenv("CUTS_HUB_BITMAP_MB"))
      cap = std::min<uint64_t>((uint64_t)std::strtoull(e, nullptr, 10) << 20, fr / 2);
    const uint32_t words = (g.V + 63) / 64;
    std::vector<uint32_t> deg(g.V);
    for (uint32_t v = 0; v < g.V; ++v) deg[v] = g.out_off[v + 1] - g.out_off[v];
    uint32_t t = 32;
    // CUTS_HUB_MIN_DEG raises the starting threshold (fewer, larger hubs:
    // a bitmap set small enough to stay in L2)
    if (const char* e = std::getenv("CUTS_HUB_MIN_DEG")) t = std::max<uint32_t>(t, (uint32_t)std::strtoul(e, nullptr, 10));
    for (; t <= (1u << 30); t *= 2) {
      uint64_t nh = 0;
      for (uint32_t d : deg) nh += d >= t;
      if (nh * words * 8 <= cap) break;
    }
    std::vector<uint32_t> hubs;
    std::vector<int32_t> hid(g.V, -1);
    for (uint32_t v = 0; v < g.V; ++v)
      if (deg[v] >= t) {
        hid[v] = (int32_t)hubs.size();
        hubs.push_back(v);
      }
    if (!hubs.empty()) {
      dg.words = words;
      dg.hubs = (uint32_t)hubs.size();
      dg.hub_min_deg = t;
      dev_alloc(dg.hub_id, g.V);
      HIPCHECK(hipMemcpy(dg.hub_id, hid.data(), g.V * 4, hipMemcpyHostToDevice));
      dev_alloc(dg.bits, (uint64_t)hubs.size() * words);
      HIPCHECK(hipMemset(dg.bits, 0, (uint64_t)hubs.size() * words * 8));
      uint32_t* dh = nullptr;
      dev_alloc(dh, hubs.size());
      HIPCHECK(hipMemcpy(dh, hubs.data(), hubs.size() * 4, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_hub_bits, dim3((uint32_t)std::min<size_t>(hubs.size(), 65535)), dim3(256), 0, c.stream,
                         dg.off_out, dg.adj_out, dh, (uint32_t)hubs.size(), words, dg.bits);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(c.stream));
      (void)hipFree(dh);
    }
  }
  // Per-arc common-neighbour counts (symmetric loop-free graphs), built on
  // the device with the hub bitmaps above: 4 bytes per arc while they fit
  // 1/16 of free HBM.
  if (g.symmetric && dg.loops == 0 && dg.A > 0) {
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    if (dg.A * 4 <= fr / 16) {
      const auto t0 = std::chrono::steady_clock::now();
      dev_alloc(dg.arc_cn, dg.A);
      GView gv = view_of(dg);
      gv.arc_cn = nullptr;
      const uint32_t grid = (uint32_t)std::min<uint64_t>((dg.A + 255) / 256, (uint64_t)c.cus * 32);
      hipLaunchKernelGGL(k_arc_common, dim3(grid), dim3(256), 0, c.stream, gv, dg.A, dg.arc_cn);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(c.stream));
      dg.arc_cn_build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  }
  // Two-hop index: built on the host when the wedges (ordered neighbour
  // pairs of every vertex, an upper bound on the entries) fit 1/16 of free
  // HBM at 8 bytes each, at most 2^28 of them, on graphs of at most 2^25
  // vertices.
  if (g.symmetric && g.V >= 2) {
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t wedges = 0;
    for (uint32_t v = 0; v < g.V; ++v) {
      const uint64_t d = g.out_off[v + 1] - g.out_off[v];
      wedges += d * (d ? d - 1 : 0);
    }
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    // the host build keeps two V-word arrays per thread: bounded V.
    // CUTS_H2_MAX_WEDGES lowers the wedge cap (0: no index, hub rows instead)
    uint64_t wcap = std::min<uint64_t>(fr / 16 / 8, 1ull << 28);
    bool no_index = false;
    if (const char* e = std::getenv("CUTS_H2_MAX_WEDGES")) {
      const uint64_t w = std::strtoull(e, nullptr, 10);
      no_index = w == 0;
      wcap = std::min<uint64_t>(wcap, w);
    }
    if (!no_index && wedges <= wcap && g.V <= (1u << 25)) {
      const uint32_t V = g.V;
      // dense rows for the longest lists (at least V/64 entries, or
      // V/CUTS_H2_DENSE_DIV; 1 turns them off), while their bytes stay within
      // 1/32 of free HBM; counts fit uint16 below degree 2^16.  A lookup in a
      // row is one load instead of a lower-bound search; below V/64 entries
      // the search is short and cached and rows gain nothing more
      auto dense_rows = [&](const std::vector<uint64_t>& off) {
        std::vector<uint32_t> cand;
        if (dg.max_deg >= 65536) return cand;
        uint64_t div = 64;
        if (const char* e = std::getenv("CUTS_H2_DENSE_DIV")) div = std::max<uint64_t>(1, std::strtoull(e, nullptr, 10));
        for (uint32_t u = 0; u < V; ++u)
          if ((off[u + 1] - off[u]) * div >= (uint64_t)V) cand.push_back(u);
        std::sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) {
          return off[a + 1] - off[a] > off[b + 1] - off[b] || (off[a + 1] - off[a] == off[b + 1] - off[b] && a < b);
        });
        const uint64_t rows = std::min<uint64_t>(cand.size(), fr / 32 / std::max<uint64_t>(2ull * V, 1));
        cand.resize(rows);
        return cand;
      };
      std::vector<uint64_t> off((size_t)V + 1, 0);
      if (V <= (1u << 17)) {
        // on the GPU (k_two_hop): pass 1 counts every vertex's entries, the
        // host sums them into offsets and picks the dense rows, pass 2 writes
        // the lists and the rows; one V-word counter row per block
        dg.h2_on_gpu = true;
        const uint32_t G = std::min<uint32_t>(V, 128);
        uint32_t* cnt_rows = nullptr;
        unsigned long long* ent = nullptr;
        dev_alloc(cnt_rows, (uint64_t)G * V);
        dev_alloc(ent, V);
        HIPCHECK(hipMemsetAsync(cnt_rows, 0, (uint64_t)G * V * 4, c.stream));
        hipLaunchKernelGGL(k_two_hop, dim3(G), dim3(BLOCK), 0, c.stream, dg.off_out, dg.adj_out, V, cnt_rows, 1, ent,
                           (const uint64_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, (const int32_t*)nullptr,
                           (uint16_t*)nullptr);
        HIPCHECK(hipGetLastError());
        std::vector<unsigned long long> T(V);
        HIPCHECK(hipMemcpyAsync(T.data(), ent, (uint64_t)V * 8, hipMemcpyDeviceToHost, c.stream));
        HIPCHECK(hipStreamSynchronize(c.stream));
        for (uint32_t u = 0; u < V; ++u) off[u + 1] = off[u] + T[u];
        dg.h2_entries = off[V];
        const std::vector<uint32_t> cand = dense_rows(off);
        dev_alloc(dg.h2_off, V + 1);
        HIPCHECK(hipMemcpy(dg.h2_off, off.data(), (V + 1) * 8, hipMemcpyHostToDevice));
        dev_alloc(dg.h2_v, std::max<uint64_t>(off[V], 1));
        dev_alloc(dg.h2_c, std::max<uint64_t>(off[V], 1));
        if (!cand.empty()) {
          std::vector<int32_t> did(V, -1);
          for (uint64_t r = 0; r < cand.size(); ++r) did[cand[r]] = (int32_t)r;
          dev_alloc(dg.h2_did, V);
          HIPCHECK(hipMemcpy(dg.h2_did, did.data(), (uint64_t)V * 4, hipMemcpyHostToDevice));
          dev_alloc(dg.h2_dense, cand.size() * V);
          HIPCHECK(hipMemset(dg.h2_dense, 0, cand.size() * 2ull * V));
          dg.h2_dense_rows = (uint32_t)cand.size();
        }
        hipLaunchKernelGGL(k_two_hop, dim3(G), dim3(BLOCK), 0, c.stream, dg.off_out, dg.adj_out, V, cnt_rows, 2, ent,
                           (const uint64_t*)dg.h2_off, dg.h2_v, dg.h2_c, (const int32_t*)dg.h2_did, dg.h2_dense);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c.stream));
        (void)hipFree(cnt_rows);
        (void)hipFree(ent);
      } else {
        std::vector<uint32_t> hv, hc;
        build_two_hop(g, off, hv, hc);
        dg.h2_entries = off[V];
        const std::vector<uint32_t> cand = dense_rows(off);
        if (!cand.empty()) {
          const uint64_t rows = cand.size();
          std::vector<int32_t> did(V, -1);
          std::vector<uint16_t> dense(rows * V, 0);
#pragma omp parallel for schedule(dynamic, 4)
          for (int64_t r = 0; r < (int64_t)rows; ++r) {
            const uint32_t u = cand[r];
            uint16_t* row = dense.data() + (uint64_t)r * V;
            for (uint64_t i = off[u]; i < off[u + 1]; ++i) row[hv[i]] = (uint16_t)hc[i];
          }
          for (uint64_t r = 0; r < rows; ++r) did[cand[r]] = (int32_t)r;
          dev_alloc(dg.h2_did, V);
          HIPCHECK(hipMemcpy(dg.h2_did, did.data(), (uint64_t)V * 4, hipMemcpyHostToDevice));
          dev_alloc(dg.h2_dense, rows * V);
          HIPCHECK(hipMemcpy(dg.h2_dense, dense.data(), rows * 2ull * V, hipMemcpyHostToDevice));
          dg.h2_dense_rows = (uint32_t)rows;
        }
        dev_alloc(dg.h2_off, V + 1);
        HIPCHECK(hipMemcpy(dg.h2_off, off.data(), (V + 1) * 8, hipMemcpyHostToDevice));
        dev_alloc(dg.h2_v, hv.size());
        HIPCHECK(hipMemcpy(dg.h2_v, hv.data(), hv.size() * 4, hipMemcpyHostToDevice));
        dev_alloc(dg.h2_c, hc.size());
        HIPCHECK(hipMemcpy(dg.h2_c, hc.data(), hc.size() * 4, hipMemcpyHostToDevice));
      }
      dg.h2_build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else if (dg.max_deg < 65536 && dg.loops == 0) {
      // hub rows in place of the index: the vertices of degree >= 256
      // (CUTS_H2_HUB_MIN_DEG), longest lists first, at most CUTS_H2_HUB_ROWS
      // rows (default 16384; 0 turns them off) within 1/8 of free HBM, built
      // on the device
      const uint32_t V = g.V;
      uint64_t cap_rows = 16384, min_deg = 256;
      if (const char* e = std::getenv("CUTS_H2_HUB_ROWS")) cap_rows = std::strtoull(e, nullptr, 10);
      if (const char* e = std::getenv("CUTS_H2_HUB_MIN_DEG")) min_deg = std::max<uint64_t>(1, std::strtoull(e, nullptr, 10));
      cap_rows = std::min<uint64_t>(cap_rows, fr / 8 / (2ull * V));
      std::vector<uint32_t> hubs;
      for (uint32_t u = 0; u < V; ++u)
        if (g.out_off[u + 1] - g.out_off[u] >= min_deg) hubs.push_back(u);
      auto deg = [&](uint32_t u) { return g.out_off[u + 1] - g.out_off[u]; };
      std::sort(hubs.begin(), hubs.end(),
                [&](uint32_t a, uint32_t b) { return deg(a) > deg(b) || (deg(a) == deg(b) && a < b); });
      hubs.resize(std::min<uint64_t>(hubs.size(), cap_rows));
      if (!hubs.empty()) {
        const uint64_t H = hubs.size();
        std::vector<uint2> jobs;
        std::vector<int32_t> id(V, -1);
        for (uint64_t r = 0; r < H; ++r) {
          const uint32_t u = hubs[r];
          id[u] = (int32_t)r;
          for (uint32_t i = g.out_off[u]; i < g.out_off[u + 1]; i += HR_JOB) jobs.push_back(make_uint2((uint32_t)r, i));
        }
        uint32_t* d_hubs = nullptr;
        uint2* d_jobs = nullptr;
        dev_alloc(d_hubs, H);
        dev_alloc(d_jobs, jobs.size());
        dev_alloc(dg.hr_id, V);
        dev_alloc(dg.hr, (H * V + 1) & ~1ull);
        HIPCHECK(hipMemcpy(d_hubs, hubs.data(), H * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_jobs, jobs.data(), jobs.size() * sizeof(uint2), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(dg.hr_id, id.data(), (uint64_t)V * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemsetAsync(dg.hr, 0, ((H * V + 1) & ~1ull) * 2, c.stream));
        hipLaunchKernelGGL(k_hub_rows, dim3((uint32_t)jobs.size()), dim3(BLOCK), 0, c.stream, dg.off_out, dg.adj_out, V,
                           d_hubs, d_jobs, (uint32_t*)dg.hr);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c.stream));
        (void)hipFree(d_hubs);
        (void)hipFree(d_jobs);
        dg.hr_rows = (uint32_t)H;
        dg.hr_build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
    }
  }
}

void run_match(Context& c, DeviceGraph& dg, const Plan& p, const MatchOpts& o, MatchResult& r) {
  HIPCHECK(hipSetDevice(c.device));
  r = MatchResult();
  Matcher m(c, dg, p, o, r);
  m.run();
  if (m.spec_overflow) {
    // a counting level outgrew what the speculative pass had allocated:
    // the query again, every level with its readback
    r = MatchResult();
    Matcher m2(c, dg, p, o, r);
    m2.no_spec = true;
    m2.run();
  }
}

uint64_t noalloc_scratch_bytes() { return kScratchBytes; }

void run_match_noalloc(Context& c, DeviceGraph& dg, const Plan& p, const MatchOpts& o, MatchResult& r) {
  HIPCHECK(hipSetDevice(c.device));
  if (!c.scratch) throw Error(-4, "the context has no scratch address");
  // level 0 holds one row per candidate root (at most V) with its planning
  // arrays, rounded up to powers of two: about 32 bytes per vertex
  const uint64_t l0 = 32ull * ((uint64_t)dg.V + 1);
  if (l0 > kScratchBytes / 4)
    throw Error(-1, "graph of " + std::to_string(dg.V) + " vertices exceeds the no-allocation scratch (" +
                        std::to_string(kScratchBytes) + " bytes: at most " + std::to_string(kScratchBytes / 128 - 1) +
                        " vertices)");
  ScratchArena a;
  a.base = c.scratch;
  a.size = kScratchBytes;
  auto* ctr = (unsigned long long*)a.carve(16 * sizeof(unsigned long long));
  HIPCHECK(hipMemsetAsync(ctr, 0, 16 * sizeof(unsigned long long), c.stream));
  LevelBuf lv[kMaxQuery];
  // trie rows take at most an eighth of the scratch; the scan, tile,
  // pair/chain arrays and regrowth take the rest
  MatchOpts o2 = o;
  o2.budget = o.budget ? std::min<uint64_t>(o.budget, kScratchBytes / 8) : kScratchBytes / 8;
  r = MatchResult();
  Matcher m(c, dg, p, o2, r);
  m.lv = lv;
  m.ctr = ctr;
  m.arena = &a;
  m.run();
}

}  // namespace cuts
