// Batched matching: many queries against one resident data graph.  No
// reference counterpart (cuTS runs one query per process and syncs the host
// after every level, src/host_funcs.cu:141-157; its MPI driver distributes
// level-0 trunks, main_mpi.cu:257-376).
//
// Level-synchronous device work queue of (query, level, row chunk) items.
// All items of level L exist before the launch that processes level L:
//   * launch L runs a persistent grid; a block takes the next item of level
//     L with one atomicAdd and exits when none is left.  No block ever waits
//     for another, so every wave ends, whatever the grid size or residency;
//   * an item's rows are expanded with the same tests as the single-query
//     kernels (signature, injectivity, constraint membership, closed-form
//     last level, fusion of the last two levels); surviving rows are staged
//     in LDS and flushed to the arena half of level L+1 as new items.  A
//     flush happens when the stage holds RMAX rows or when the staged rows'
//     candidate work for level L+1 (their shortest constraint lists) reaches
//     `item_work`, so hub rows do not pile up in one item;
//   * the kernel boundary between launches publishes level L+1 to every
//     XCD (the XCDs' L2 caches are not kept coherent with each other inside
//     a launch, so no item is consumed in the launch that wrote it);
//   * all queries advance together: the launch sequence (init + one launch
//     per level up to the largest query size) is captured once in a hipGraph
//     and replayed.
// Loops that contain block barriers keep every branch around their edges
// wave-uniform: the single-lane work (ticket, arena allocation, item push)
// is done by all 64 lanes of wave 0 under a wave-uniform branch, with atomic
// operands masked to lane 0 and results broadcast by readfirstlane.  A
// `threadIdx.x == 0` region at a loop's back edge lets the compiler split
// the loop into lane-0 and other-lane latches (nested loops after
// structurisation); the other lanes then pass the next barrier without lane
// 0's update and spin on a stale item.
// Per query: exact uint64 count, per-level table sizes and the wall-clock
// tick at which its last item finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "batch.hpp"

namespace cuts {
namespace {

constexpr int BB = 256;      // threads per block (4 wave64)
constexpr int NW = BB / 64;  // waves per block
constexpr int RMAX = 256;    // rows per item (one per thread in the scan)
// LDS words per row (input and staged): a template parameter of at least
// kmax-1, odd so that rows of different lanes fall in different banks.

struct GV {
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
  uint32_t V;
  const int32_t* hub_id;  // hub adjacency bitmaps (symmetric graphs) or null
  const uint2 *nsig_out, *nsig_in;  // signature of every list entry
  const uint64_t* bits;
  uint32_t words;
  const uint64_t* h2_off;  // two-hop index or null
  const uint32_t *h2_v, *h2_c;
  const int32_t* h2_did;   // dense rows of the index or null
  const uint16_t* h2_dense;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
// readfirstlane returns int: both halves are taken as uint32 before they
// are widened, or a low word >= 2^31 would be sign-extended into the high one
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t wave_id() { return uni(threadIdx.x >> 6); }
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint32_t lane_rank64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ bool bmember(const uint32_t* __restrict__ adj, uint32_t lo, uint32_t hi, uint32_t x) {
  const uint32_t end = hi;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (adj[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < end && adj[lo] == x;
}

// x in out-list(u) of a symmetric graph (bit test for hubs)
__device__ __forceinline__ bool has_arc(const GV& g, uint32_t u, uint32_t x) {
  if (g.hub_id) {
    const int32_t h = g.hub_id[u];
    if (h >= 0) return (g.bits[(uint64_t)h * g.words + (x >> 6)] >> (x & 63)) & 1ull;
  }
  return bmember(g.adj_out, g.off_out[u], g.off_out[u + 1], x);
}

// |N(u) & N(w)| for u != w from the two-hop index (lower-bound search of w
// in u's list; common neighbours of distinct vertices have degree >= 2)
__device__ __forceinline__ uint32_t two_hop(const GV& g, uint32_t u, uint32_t w) {
  if (g.h2_did) {
    const int32_t dr = g.h2_did[u];
    if (dr >= 0) return g.h2_dense[(uint64_t)dr * g.V + w];
  }
  uint64_t lo = g.h2_off[u], hi = g.h2_off[u + 1];
  const uint64_t end = hi;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (g.h2_v[mid] < w) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && g.h2_v[lo] == w) ? g.h2_c[lo] : 0u;
}

// Chain of the last two positions (k-1 on k-2 alone, k-2 on ch_u alone)
// for the prefix row[0..L) + v of positions 0..k-3, u = row[ch_u]:
//   H(u) - sum_{r_i in N(u), deg >= ta} cnt_tb(r_i) - [deg u >= tb] |A|
//        - sum_{i != ch_u, deg(r_i) >= tb} (|N(u) & N(r_i)| - c_i)
// with |A| = cnt_ta(u) - |{i : r_i in N(u), deg(r_i) >= ta}| (into na) and
// c_i the prefix vertices inside N(u) & N(r_i) meeting ta.  The
// intersections of the row's own positions come summed in `shared`; the one
// with v is looked up in the two-hop index.
__device__ int64_t chain_closed_rt(const GV& g, const KPlan& P, const uint32_t* row, uint32_t L, uint32_t v,
                                   const uint64_t* __restrict__ H, const uint32_t* __restrict__ ta_tbl,
                                   const uint32_t* __restrict__ tb_tbl, uint32_t shared, int64_t& na) {
  const uint32_t N = L + 1;
  const uint32_t cu = P.ch_u, ta = P.ch_ta, tb = P.ch_tb;
  const uint32_t u = row[cu];
  const uint32_t du = g.off_out[u + 1] - g.off_out[u];
  uint32_t inu = 0;
  for (uint32_t l = 0; l < N; ++l) {
    const uint32_t x = l == L ? v : row[l];
    const bool in = l != cu && (((P.qadj[cu] >> l) & 1) || has_arc(g, u, x));
    if (in && g.sig[2 * (uint64_t)x] >= ta) inu |= 1u << l;
  }
  na = (int64_t)(ta_tbl ? ta_tbl[u] : du) - __popc(inu);
  int64_t c = (int64_t)H[u];
  for (uint32_t l = 0; l < N; ++l)
    if ((inu >> l) & 1) {
      const uint32_t x = l == L ? v : row[l];
      c -= tb_tbl ? tb_tbl[x] : g.off_out[x + 1] - g.off_out[x];
    }
  if (tb <= 1 || du >= tb) c -= na;
  for (uint32_t i = 0; i < N; ++i) {
    const uint32_t xi = i == L ? v : row[i];
    if (i == cu || g.sig[2 * (uint64_t)xi] < tb) continue;
    for (uint32_t j = 0; j < N; ++j) {
      if (j == i || !((inu >> j) & 1)) continue;
      if (((P.qadj[i] >> j) & 1) || has_arc(g, xi, j == L ? v : row[j])) ++c;
    }
  }
  c -= (int64_t)shared;
  if (tb <= 1 || g.sig[2 * (uint64_t)v] >= tb) c -= two_hop(g, u, v);
  return c;
}

// Closed-form count of the last position given positions r[0..n) (LDS row)
// plus, when has_v, the candidate v at position n.
__device__ uint64_t closed_form_rt(const GV& g, const KPlan& P, const uint32_t* r, uint32_t n, bool has_v, uint32_t v,
                                   const uint32_t* __restrict__ tbl) {
  const uint32_t npos = n + (has_v ? 1 : 0);
  const uint32_t u = (has_v && P.sc_pos == n) ? v : r[P.sc_pos];
  const bool newest = has_v && P.sc_pos == n;
  const uint32_t lo = g.off_out[u], hi = g.off_out[u + 1];
  int64_t c = tbl ? (int64_t)tbl[u] : (int64_t)(hi - lo);
  const bool need_deg = P.sc_thr > 1;
  for (uint32_t i = 0; i < npos; ++i) {
    const uint32_t x = (i == n) ? v : r[i];
    bool in = (P.sc_adjmask >> i) & 1;
    if (!in) in = (newest && i != n) ? has_arc(g, x, u) : has_arc(g, u, x);
    if (in && (!need_deg || g.sig[2 * (uint64_t)x] >= P.sc_thr)) --c;
  }
  return (uint64_t)c;
}

// Candidate work of the row row[0..L) + v at its next position L+1: the
// length of its shortest constraint list (what the consumer will scan), or
// the closed form's membership tests when L+1 is a closed-form last position.
__device__ __forceinline__ uint64_t next_work(const GV& g, const KPlan& P, const uint32_t* row, uint32_t L, uint32_t v,
                                              bool closed) {
  const uint32_t p = L + 1;
  if (p == P.k - 1 && closed) return p;
  uint32_t best = 0xFFFFFFFFu;
  for (uint32_t c = 0; c < P.nj[p]; ++c) {
    const uint32_t jc = P.jc[p][c];
    const uint32_t pos = jc & 31;
    const uint32_t u = pos == L ? v : row[pos];
    const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
    best = min(best, off[u + 1] - off[u]);
  }
  return best;
}

template <int WROW>
struct Shared {
  Item item;
  uint32_t cmd;
  uint32_t nstage;                // rows waiting in `stage`
  uint64_t wstage;                // their candidate work at the next level
  uint64_t base;
  uint32_t wcnt[NW];
  uint64_t wwork[NW];
  uint64_t off[RMAX + 1];         // exclusive scan of the rows' shortest list lengths
  uint8_t minsel[RMAX];
  uint32_t cbase[RMAX];           // fused level: prefix part of the closed form
  uint32_t rows[RMAX * WROW];     // input rows of the item
  uint32_t stage[RMAX * WROW];    // output rows waiting for a flush
};

__device__ uint64_t wave_sum(uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  uint64_t s = v;
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
    s += ((uint64_t)h2 << 32) | l2;
    lo = (uint32_t)s;
    hi = (uint32_t)(s >> 32);
  }
  return s;
}

__device__ uint64_t wave_incl_scan(uint64_t x) {
  const uint32_t lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t lo = __shfl_up((uint32_t)x, o), hi = __shfl_up((uint32_t)(x >> 32), o);
    if (lane >= (uint32_t)o) x += ((uint64_t)hi << 32) | lo;
  }
  return x;
}

// One atomic per wave (operand masked to lane 0; the wave total is uniform).
__device__ void add_block(unsigned long long* dst, uint64_t v) {
  const uint64_t s = uni64(wave_sum(v));
  if (s) atomicAdd(dst, lane_id() == 0 ? (unsigned long long)s : 0ull);
}

// Called by all 64 lanes of one wave.
__device__ void push_item(const BatchArgs& a, uint32_t level, const Item& it) {
  const bool l0 = lane_id() == 0;
  const unsigned long long slot = uni64(atomicAdd(&a.st->tail[level].v, l0 ? 1ull : 0ull));
  if (slot >= a.qcap) {
    atomicOr(&a.st->abort.v, l0 ? 1ull : 0ull);
    return;
  }
  atomicAdd(&a.qpend[it.qL >> 8], l0 ? 1ull : 0ull);
  a.items[(uint64_t)(level & 1) * a.qcap + slot] = it;  // same 16 bytes from every lane
}

// Writes the staged rows (length `words`, i.e. level `words`) to that
// level's arena half and publishes them as one item.  Whole block.
template <int WROW>
__device__ void flush_stage(const BatchArgs& a, Shared<WROW>& s, uint32_t q, uint32_t words) {
  __syncthreads();
  const uint32_t n = uni(s.nstage);
  if (n == 0) return;
  const bool l0 = lane_id() == 0;
  if (wave_id() == 0) {
    const unsigned long long need = (unsigned long long)n * words;
    const unsigned long long o = uni64(atomicAdd(&a.st->arena[words].v, l0 ? need : 0ull));
    unsigned long long base = o;
    if (o + need > a.half_words) {
      atomicOr(&a.st->abort.v, l0 ? 2ull : 0ull);
      base = ~0ull;
    }
    s.base = base;
  }
  __syncthreads();
  const uint64_t base = uni64(s.base);
  if (base != ~0ull) {
    uint32_t* dst = a.arena + (uint64_t)(words & 1) * a.half_words + base;
    for (uint32_t i = threadIdx.x; i < n * words; i += BB) dst[i] = s.stage[(i / words) * WROW + i % words];
    if (wave_id() == 0) {
      Item it;
      it.qL = (q << 8) | words;
      it.n = n;
      it.off = base;
      push_item(a, words, it);
    }
  }
  s.nstage = 0;  // every thread read nstage before the barrier above
  s.wstage = 0;
  __syncthreads();
}

// Block-wide append of the survivors row[0..L) + v (one per thread with ok)
// to the stage, flushing first when they would overfill it.
template <int WROW>
__device__ void stage_append(const BatchArgs& a, Shared<WROW>& s, uint32_t q, bool ok, const uint32_t* row, uint32_t L,
                             uint32_t v, uint64_t work) {
  const uint32_t wave = wave_id();
  const uint64_t m = __ballot(ok);
  const uint64_t wsum = uni64(wave_sum(ok ? work : 0));
  s.wcnt[wave] = (uint32_t)__popcll(m);  // same value from every lane of the wave
  s.wwork[wave] = wsum;
  __syncthreads();
  uint32_t total = 0, pre = 0;
  uint64_t wtot = 0;
  for (uint32_t w = 0; w < (uint32_t)NW; ++w) {
    if (w == wave) pre = total;
    total += uni(s.wcnt[w]);
    wtot += uni64(s.wwork[w]);
  }
  uint32_t nst = uni(s.nstage);
  uint64_t wst = uni64(s.wstage);
  if (total) {
    if (nst + total > RMAX || (nst > 0 && wst + wtot > a.item_work)) {
      flush_stage(a, s, q, L + 1);
      nst = 0;
      wst = 0;
    }
    if (ok) {
      uint32_t* dst = &s.stage[(nst + pre + lane_rank64(m)) * WROW];
      for (uint32_t l = 0; l < L; ++l) dst[l] = row[l];
      dst[L] = v;
    }
    __syncthreads();
    s.nstage = nst + total;  // same value from every thread
    s.wstage = wst + wtot;
  }
  __syncthreads();
}

// Exclusive block scan of one length per row (thread tid < n holds row
// tid's) into s.off[0..n]; returns the total.  Whole block.
template <int WROW>
__device__ uint64_t block_scan(Shared<WROW>& s, uint32_t n, uint64_t len) {
  const uint32_t tid = threadIdx.x;
  const uint64_t incl = wave_incl_scan(len);
  if ((tid & 63) == 63) s.wwork[tid >> 6] = incl;
  __syncthreads();
  uint64_t before = 0, W = 0;
  for (uint32_t w = 0; w < (uint32_t)NW; ++w) {
    const uint64_t x = uni64(s.wwork[w]);
    if (w < wave_id()) before += x;
    W += x;
  }
  if (tid < n) s.off[tid] = before + incl - len;
  s.off[n] = W;  // same value from every thread; no thread writes index n above
  __syncthreads();
  return W;
}

template <int WROW>
__device__ void process(const BatchArgs& a, Shared<WROW>& s, uint32_t L) {
  const GV g{a.off_out, a.adj_out, a.off_in, a.adj_in, a.sig, a.V, a.hub_id, a.nsig_out, a.nsig_in, a.bits, a.words,
             a.h2_off, a.h2_v, a.h2_c, a.h2_did, a.h2_dense};
  Item it;
  it.qL = uni(s.item.qL);
  it.n = uni(s.item.n);
  it.off = uni64(s.item.off);
  const uint32_t q = it.qL >> 8;
  const KPlan& P = a.plans[q];
  const uint32_t k = uni(P.k);
  const uint32_t tid = threadIdx.x;
  const uint32_t* tbl = a.tbls[3 * q];
  const bool closed = (a.flags & 1u) == 0 && uni(P.sc_ok);
  const bool fuse = closed && (a.flags & 2u) == 0 && k >= 3;
  const bool pair = fuse && uni(P.pr_ok) && (a.flags & 8u) == 0;
  // the chain level needs the two-hop index (threshold of A at most 2)
  const bool chain = fuse && !pair && uni(P.ch_ok) && (a.flags & 16u) == 0 && g.h2_off != nullptr && uni(P.ch_ta) <= 2;
  uint64_t cnt = 0, surv = 0, sna = 0;

  if (L == 0) {
    // level-0 generation over roots off .. off+n-1 of the query's root sequence
    const uint32_t rfirst = uni(a.roots[3 * q]), rstride = uni(a.roots[3 * q + 1]);
    for (uint32_t i0 = 0; i0 < it.n; i0 += BB) {
      const uint32_t i = i0 + tid;
      const bool in = i < it.n;
      const uint32_t u = rfirst + (uint32_t)(it.off + (in ? i : 0)) * rstride;  // it.n >= 1 whenever this loop runs
      const bool ok = in && (g.sig[2 * (uint64_t)u] >= P.sig_in[0]) & (g.sig[2 * (uint64_t)u + 1] >= P.sig_out[0]);
      surv += ok;
      if (k > 1) stage_append(a, s, q, ok, nullptr, 0, u, ok ? next_work(g, P, nullptr, 0, u, closed) : 0);
    }
    if (k > 1) flush_stage(a, s, q, 1);
    add_block(&a.levels[q * kMaxQuery + 0], surv);
    if (k == 1) add_block(&a.counts[q], surv);
    return;
  }

  const uint32_t n = it.n;
  const uint32_t* src = a.arena + (uint64_t)(L & 1) * a.half_words + it.off;
  for (uint32_t i = tid; i < n * L; i += BB) s.rows[(i / L) * WROW + i % L] = src[i];
  __syncthreads();

  if (L == k - 1 && closed) {
    if (tid < n) cnt = closed_form_rt(g, P, &s.rows[tid * WROW], L, false, 0, tbl);
    add_block(&a.counts[q], cnt);
    add_block(&a.levels[q * kMaxQuery + L], cnt);
    return;
  }

  // shortest constraint list per row (one row per thread), then a block scan
  const uint32_t nj = uni(P.nj[L]);
  const int mode = (L == k - 1) ? 1
                   : (pair && L == k - 3)  ? 3
                   : (chain && L == k - 3) ? 4
                   : (L == k - 2 && fuse)  ? 2
                                           : 0;
  // fused level with an older closed-form vertex u: the closed form splits
  // into a per-prefix base, cnt(u) - |{i < L : r_i in N(u), deg(r_i) >= t}|,
  // and one term for the candidate itself
  const uint32_t sc_pos = uni(P.sc_pos), sc_thr = uni(P.sc_thr);
  const bool use_cbase = mode == 2 && sc_pos < L && (a.flags & 4u) == 0;
  // pair level (L = k-3): a survivor v completes the prefix and has
  // |A| * |B| - |A & B| completions (A = N(row[pr_a]), B = N(row[pr_b]) minus
  // the prefix, degree-filtered).  Per row: the prefix parts of |A|, |B| and
  // of |A & B| (signed: for pr_a != pr_b the size of N(u) & N(w) is
  // subtracted afterwards, per survivor, by a second pass over the rows), and
  // the row's survivor count.  The stage is empty on this level (nothing is
  // flushed), so these live there.
  const uint32_t pa = uni(P.pr_a), pb = uni(P.pr_b);
  // |N(u) & N(w)| from the two-hop index where it serves the threshold
  const bool h2 = mode == 3 && pa != pb && g.h2_off != nullptr && uni(P.pr_tab) <= 2;
  uint32_t* const pxa = s.stage;
  uint32_t* const pxb = s.stage + RMAX;
  uint32_t* const pxz = s.stage + 2 * RMAX;
  uint32_t* const pns = s.stage + 3 * RMAX;
  uint64_t len = 0;
  bool collapsed = false;
  if (tid < n) {
    if (mode == 3) {
      const uint32_t* row = &s.rows[tid * WROW];
      const uint32_t u = row[pa], w = row[pb];
      const uint32_t* ta_tbl = a.tbls[3 * q + 1];
      const uint32_t* tab_tbl = a.tbls[3 * q + 2];
      int64_t ca = ta_tbl ? (int64_t)ta_tbl[u] : (int64_t)(g.off_out[u + 1] - g.off_out[u]);
      int64_t cb = tbl ? (int64_t)tbl[w] : (int64_t)(g.off_out[w + 1] - g.off_out[w]);
      int64_t cab = pa != pb ? (h2 ? (int64_t)two_hop(g, u, w) : 0)
                             : (tab_tbl ? (int64_t)tab_tbl[u] : (int64_t)(g.off_out[u + 1] - g.off_out[u]));
      for (uint32_t l = 0; l < L; ++l) {
        const bool ina = ((P.pr_ma >> l) & 1) || has_arc(g, u, row[l]);
        const bool inb = ((P.pr_mb >> l) & 1) || has_arc(g, w, row[l]);
        const uint32_t d = g.sig[2 * (uint64_t)row[l]];
        ca -= ina && d >= P.pr_ta;
        cb -= inb && d >= P.pr_tb;
        cab -= ina && inb && d >= P.pr_tab;
      }
      pxa[tid] = (uint32_t)ca;
      pxb[tid] = (uint32_t)cb;
      pxz[tid] = (uint32_t)(int32_t)cab;
      pns[tid] = 0;
      if (h2 && P.pr_collapse) {
        // position k-3's survivors are N(u) & N(w) minus the prefix, each
        // adding the same number: no candidate is expanded
        const int64_t ns = cab;
        cnt += (uint64_t)(ns * ((ca - 1) * (cb - 1) - (cab - 1)));
        surv += (uint64_t)ns;
        sna += (uint64_t)(ns * (ca - 1));
        collapsed = true;
      }
    }
    if (mode == 4) {
      // chain level: the intersections of N(u) with the row's own positions
      const uint32_t* row = &s.rows[tid * WROW];
      const uint32_t cu = P.ch_u, u = row[cu];
      uint32_t t = 0;
      for (uint32_t i = 0; i < L; ++i) {
        if (i == cu) continue;
        if (P.ch_tb <= 1 || g.sig[2 * (uint64_t)row[i]] >= P.ch_tb) t += two_hop(g, u, row[i]);
      }
      pxa[tid] = t;
    }
    if (use_cbase) {
      const uint32_t* row = &s.rows[tid * WROW];
      const uint32_t u = row[sc_pos];
      int64_t c = tbl ? (int64_t)tbl[u] : (int64_t)(g.off_out[u + 1] - g.off_out[u]);
      for (uint32_t l = 0; l < L; ++l) {
        const bool in = ((P.sc_adjmask >> l) & 1) || has_arc(g, u, row[l]);
        if (in && (sc_thr <= 1 || g.sig[2 * (uint64_t)row[l]] >= sc_thr)) --c;
      }
      s.cbase[tid] = (uint32_t)c;
    }
    uint32_t best = 0xFFFFFFFFu, sel = 0;
    for (uint32_t c = 0; c < nj; ++c) {
      const uint32_t jc = P.jc[L][c];
      const uint32_t u = s.rows[tid * WROW + (jc & 31)];
      const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
      const uint32_t l = off[u + 1] - off[u];
      if (l < best) {
        best = l;
        sel = c;
      }
    }
    s.minsel[tid] = (uint8_t)sel;
    len = collapsed ? 0 : best;
  }
  const uint64_t W = block_scan(s, n, len);

  const uint32_t tin = P.sig_in[L], tout = P.sig_out[L];
  for (uint64_t w0 = 0; w0 < W; w0 += BB) {
    const uint64_t w = w0 + tid;
    bool ok = false;
    uint32_t v = 0, r = 0;
    uint64_t fc = 0;  // fused: closed-form count of the survivor
    if (w < W) {
      uint32_t lo = 0, hi = n + 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.off[mid] <= w) lo = mid + 1;
        else hi = mid;
      }
      r = lo - 1;
      const uint32_t idx = (uint32_t)(w - s.off[r]);
      const uint32_t* row = &s.rows[r * WROW];
      const uint32_t sel = s.minsel[r];
      const uint32_t jc = P.jc[L][sel];
      const uint32_t u = row[jc & 31];
      const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
      const uint32_t* adj = (jc >> 7) ? g.adj_in : g.adj_out;
      const uint32_t li = off[u] + idx;
      v = adj[li];
      const uint2 sgv = ((jc >> 7) ? g.nsig_in : g.nsig_out)[li];
      ok = sgv.x >= tin && sgv.y >= tout;
      for (uint32_t l = 0; l < L && ok; ++l) ok = v != row[l];
      for (uint32_t c = 0; c < nj && ok; ++c) {
        if (c == sel) continue;
        const uint32_t jc2 = P.jc[L][c];
        const uint32_t u2 = row[jc2 & 31];
        if (g.hub_id) {
          ok = has_arc(g, u2, v);
        } else {
          const uint32_t* off2 = (jc2 >> 7) ? g.off_in : g.off_out;
          const uint32_t* adj2 = (jc2 >> 7) ? g.adj_in : g.adj_out;
          ok = bmember(adj2, off2[u2], off2[u2 + 1], v);
        }
      }
      if (mode == 3 && ok) {
        const bool ina = ((P.pr_ma >> L) & 1) || has_arc(g, row[pa], v);
        const bool inb = ((P.pr_mb >> L) & 1) || has_arc(g, row[pb], v);
        const uint64_t na = pxa[r] - (uint32_t)(ina && sgv.x >= P.pr_ta);
        const uint64_t nb = pxb[r] - (uint32_t)(inb && sgv.x >= P.pr_tb);
        const int64_t nab = (int64_t)(int32_t)pxz[r] - (int64_t)(ina && inb && sgv.x >= P.pr_tab);
        fc = na * nb - (uint64_t)nab;  // modulo-2^64 sums are exact
        sna += na;
        if (pa != pb && !h2) atomicAdd(&pns[r], 1u);
      }
      if (mode == 4 && ok) {
        int64_t na = 0;
        fc = (uint64_t)chain_closed_rt(g, P, row, L, v, a.chH[q], a.tbls[3 * q + 1], tbl, pxa[r], na);
        sna += (uint64_t)na;
      }
      if (mode == 2 && ok) {
        if (use_cbase) {
          // v is certainly in N(u) when it was drawn from u's own list
          const uint32_t us = row[sc_pos];
          const bool in = (jc & 31) == sc_pos || has_arc(g, us, v);
          fc = (uint64_t)s.cbase[r] - ((in && (sc_thr <= 1 || g.sig[2 * (uint64_t)v] >= sc_thr)) ? 1 : 0);
        } else {
          fc = closed_form_rt(g, P, row, L, true, v, tbl);
        }
      }
    }
    const uint32_t* row = &s.rows[r * WROW];
    if (mode == 1) {
      cnt += ok;
    } else if (mode >= 2) {
      if (ok) {
        ++surv;
        cnt += fc;
      }
    } else {
      surv += ok;
      stage_append(a, s, q, ok, row, L, v, ok ? next_work(g, P, row, L, v, closed) : 0);
    }
  }
  if (mode == 0) flush_stage(a, s, q, L + 1);
  if (mode == 3 && pa != pb && !h2) {
    // second pass: every x in N(u) & N(w) (degree-filtered) takes one
    // completion from each survivor of its row; per row with survivors the
    // shorter of the two lists is walked and x looked up in the other
    __syncthreads();
    uint64_t len2 = 0;
    if (tid < n && pns[tid]) {
      const uint32_t u = s.rows[tid * WROW + pa], w = s.rows[tid * WROW + pb];
      const uint32_t du = g.off_out[u + 1] - g.off_out[u], dw = g.off_out[w + 1] - g.off_out[w];
      len2 = min(du, dw);
      s.minsel[tid] = du <= dw ? 0 : 1;
    }
    const uint64_t W2 = block_scan(s, n, len2);
    for (uint64_t w0 = 0; w0 < W2; w0 += BB) {
      const uint64_t w = w0 + tid;
      if (w < W2) {
        uint32_t lo = 0, hi = n + 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s.off[mid] <= w) lo = mid + 1;
          else hi = mid;
        }
        const uint32_t r = lo - 1;
        const uint32_t idx = (uint32_t)(w - s.off[r]);
        const uint32_t u = s.rows[r * WROW + pa], wv = s.rows[r * WROW + pb];
        const bool sel = s.minsel[r] != 0;
        const uint32_t li = g.off_out[sel ? wv : u] + idx;
        const uint32_t x = g.adj_out[li];
        if ((P.pr_tab <= 1 || g.nsig_out[li].x >= P.pr_tab) && has_arc(g, sel ? u : wv, x)) cnt -= pns[r];
      }
    }
  }
  if (mode == 3 || mode == 4) {
    add_block(&a.counts[q], cnt);
    add_block(&a.levels[q * kMaxQuery + L], surv);
    add_block(&a.levels[q * kMaxQuery + L + 1], sna);
    add_block(&a.levels[q * kMaxQuery + L + 2], cnt);
  } else if (mode == 1) {
    add_block(&a.counts[q], cnt);
    add_block(&a.levels[q * kMaxQuery + L], cnt);
  } else if (mode == 2) {
    add_block(&a.counts[q], cnt);
    add_block(&a.levels[q * kMaxQuery + L], surv);
    add_block(&a.levels[q * kMaxQuery + L + 1], cnt);
  } else {
    add_block(&a.levels[q * kMaxQuery + L], surv);
  }
}

// Processes every item of level L: published by the earlier launches, or by
// the earlier rounds of the persistent kernel (read after its barrier).
template <int WROW>
__device__ __forceinline__ void level_items(const BatchArgs& a, Shared<WROW>& s, uint32_t L) {
  const bool l0 = lane_id() == 0;
  const unsigned long long n_items = uni64(std::min<unsigned long long>(
      __hip_atomic_load(&a.st->tail[L].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), a.qcap));
  s.nstage = 0;  // same value from every thread
  s.wstage = 0;
  uint32_t prev_q = 0xFFFFFFFFu;  // query of the item finished in the last round
  while (true) {
    if (wave_id() == 0) {
      // retire the previous item, then take the next ticket
      if (prev_q != 0xFFFFFFFFu) {
        const unsigned long long left = uni64(atomicSub(&a.qpend[prev_q], l0 ? 1ull : 0ull));
        if (left == 1ull) a.done[prev_q] = wall_clock64();
      }
      uint32_t cmd = 1;
      if (uni64(__hip_atomic_load(&a.st->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
        const unsigned long long t = uni64(atomicAdd(&a.st->head[L].v, l0 ? 1ull : 0ull));
        if (t < n_items) {
          const Item* ip = &a.items[(uint64_t)(L & 1) * a.qcap + t];
          Item it;
          it.qL = uni(ip->qL);
          it.n = uni(ip->n);
          it.off = uni64(ip->off);
          // an item outside the shapes the kernels write means corrupt
          // memory: stop instead of reading out of bounds
          const bool sane = (it.qL >> 8) < a.nq && (it.qL & 0xFF) == L &&
                            (L == 0 ? (it.off + it.n <= a.roots[3 * (it.qL >> 8) + 2])
                                    : (it.n >= 1 && it.n <= (uint32_t)RMAX &&
                                       it.off + (uint64_t)it.n * L <= a.half_words));
          if (sane) {
            s.item = it;
            cmd = 0;
          } else {
            atomicOr(&a.st->abort.v, l0 ? 4ull : 0ull);
          }
        }
      }
      s.cmd = cmd;
    }
    __syncthreads();
    if (uni(s.cmd)) break;
    prev_q = uni(s.item.qL) >> 8;
    process(a, s, L);
    __syncthreads();
  }
}

// Device-wide barrier between two levels of the persistent kernel (round r
// = the level just finished).  Wave 0 of every block releases the block's
// writes (agent-scope fence: the XCDs' L2 caches are written back), counts
// itself in, and waits until all gridDim.x blocks of the round are in, then
// acquires (the L2 lines other XCDs wrote are invalidated).  The wait is
// bounded: an abort raised anywhere, or a wait longer than the run's limit
// (which raises abort 8), ends it, and the barrier returns false in every
// block, so every block leaves.  The arrival counter only grows (round r
// waits for (r+1) x gridDim.x); k_batch_init clears it.  Every branch
// around the wait loop is wave-uniform.
template <int WROW>
__device__ bool level_barrier(const BatchArgs& a, Shared<WROW>& s, uint32_t round) {
  __syncthreads();
  if (wave_id() == 0) {
    const bool l0 = lane_id() == 0;
    __threadfence();
    atomicAdd(&a.st->bar.v, l0 ? 1ull : 0ull);
    const unsigned long long target = (unsigned long long)(round + 1) * gridDim.x;
    const unsigned long long t0 = uni64(wall_clock64());
    uint32_t stop = 0;
    while (true) {
      if (uni64(__hip_atomic_load(&a.st->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0) {
        stop = 1;
        break;
      }
      if (uni64(__hip_atomic_load(&a.st->bar.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= target) break;
      if (uni64(wall_clock64()) - t0 > a.bar_ticks) {
        atomicOr(&a.st->abort.v, l0 ? 8ull : 0ull);
        stop = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    __threadfence();
    s.cmd = stop;
  }
  __syncthreads();
  return uni(s.cmd) == 0;
}

}  // namespace

// Resets the queue on the device (so the sequence replays inside a hipGraph)
// and publishes the level-0 items (up to vchunk roots of one query each).
__global__ __launch_bounds__(BB) void k_batch_init(BatchArgs a) {
  const uint64_t tid = (uint64_t)blockIdx.x * BB + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * BB;
  for (uint64_t i = tid; i < (uint64_t)a.nq * kMaxQuery; i += nthreads) a.levels[i] = 0;
  for (uint64_t q = tid; q < a.nq; q += nthreads) {
    a.counts[q] = 0;
    a.qpend[q] = a.qinit[q];
    a.done[q] = 0;
  }
  for (uint64_t i = tid; i < a.n_init; i += nthreads) a.items[i] = a.init_items[i];
  if (tid == 0) {
    for (uint32_t l = 0; l <= kMaxQuery; ++l) {
      a.st->head[l].v = 0;
      a.st->tail[l].v = 0;
      a.st->arena[l].v = 0;
    }
    a.st->tail[0].v = a.n_init;
    a.st->abort.v = 0;
    a.st->bar.v = 0;
    a.st->t_start.v = wall_clock64();
  }
}

// Processes every item of level L (published by the earlier launches).
// At least 6 waves per SIMD: 80 VGPRs; with row widths 7 and 11 the LDS also
// fits 6 blocks of 4 waves per CU.
template <int WROW>
__global__ __launch_bounds__(BB, 6) void k_batch_level(BatchArgs a, uint32_t L) {
  __shared__ Shared<WROW> s;
  level_items<WROW>(a, s, L);
}

// The persistent kernel: one cooperative launch (every block resident at
// once) runs levels 0 .. kmax-1, pulling the (query, level, chunk) items of
// each level from the device queue, with a device-wide barrier between
// levels in place of the launch boundary.
template <int WROW>
__global__ __launch_bounds__(BB, 6) void k_batch_all(BatchArgs a, uint32_t kmax) {
  __shared__ Shared<WROW> s;
  for (uint32_t L = 0; L < kmax; ++L) {
    if (L > 0 && !level_barrier<WROW>(a, s, L - 1)) break;
    level_items<WROW>(a, s, L);
  }
}

}  // namespace cuts

// ------------------------------------------------------------------ host
namespace cuts {

namespace {
template <typename T>
T* balloc(Batch& b, int slot, uint64_t n) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<uint64_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(-5, "batch: device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  b.allocs[slot] = p;
  return (T*)p;
}
}  // namespace

Batch::~Batch() {
  if (ctx) (void)hipSetDevice(ctx->device);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (void* p : allocs)
    if (p) (void)hipFree(p);
}

void batch_create(Context& c, DeviceGraph& dg, const std::vector<const Plan*>& plans, const BatchOpts& o, Batch& b) {
  if (o.ticket_queue)
    throw Error(-1, "the asynchronous ticket-queue batch kernel is disabled: it stalled on MI355X (no progress, "
                    "no fault) in its first test and again after a fix, with the cause not established");
  HIPCHECK(hipSetDevice(c.device));
  if (plans.empty()) throw Error(-1, "batch: no queries");
  if (plans.size() >= (1u << 24)) throw Error(-1, "batch: too many queries");
  b.ctx = &c;
  b.dg = &dg;
  b.nq = (uint32_t)plans.size();
  b.vchunk = o.vchunk ? o.vchunk : 4096;
  b.use_graph = o.use_graph != 0;
  b.persistent = o.persistent != 0;
  uint32_t bpc = o.blocks_per_cu ? std::min<uint32_t>(o.blocks_per_cu, 8) : 6;
  b.grid = (uint32_t)c.cus * bpc;
  const uint64_t qcap = o.queue_cap ? o.queue_cap : (1ull << 22);
  // position-0 roots per query (first, stride, count) and the level-0 items
  std::vector<uint32_t> roots(3 * (size_t)b.nq);
  std::vector<unsigned long long> qinit(b.nq);
  std::vector<Item> init;
  for (uint32_t q = 0; q < b.nq; ++q) {
    uint32_t first = 0, end = dg.V, stride = 1;
    if (o.roots) {
      first = o.roots[3 * q];
      end = std::min(o.roots[3 * q + 1], dg.V);
      stride = o.roots[3 * q + 2];
      if (stride == 0) throw Error(-1, "batch: root stride must be positive");
    }
    const uint32_t cnt = first < end ? (uint32_t)(((uint64_t)end - first + stride - 1) / stride) : 0;
    roots[3 * q] = first;
    roots[3 * q + 1] = stride;
    roots[3 * q + 2] = cnt;
    for (uint64_t off = 0; off < cnt; off += b.vchunk) {
      Item it;
      it.qL = q << 8;
      it.n = (uint32_t)std::min<uint64_t>(b.vchunk, cnt - off);
      it.off = off;
      init.push_back(it);
      ++qinit[q];
    }
  }
  if (init.size() > qcap) throw Error(-1, "batch: queue too small for the level-0 items");
  const uint64_t half_words = (o.arena_bytes ? o.arena_bytes : (8ull << 30)) / 8;
  if (half_words < 16) throw Error(-1, "batch: arena too small");
  std::vector<KPlan> kp(b.nq);
  std::vector<const uint32_t*> tb(3 * (size_t)b.nq, nullptr);
  std::vector<const uint64_t*> chh(b.nq, nullptr);
  b.kmax = 0;
  for (uint32_t q = 0; q < b.nq; ++q) {
    if (plans[q]->k == 0 || plans[q]->k > kMaxQuery) throw Error(-3, "batch: query size out of range");
    kp[q] = make_kplan(*plans[q], dg.symmetric, dg.loops == 0);
    b.k.push_back(plans[q]->k);
    b.kmax = std::max(b.kmax, plans[q]->k);
    if (kp[q].sc_ok && kp[q].sc_thr > 1 && !(o.flags & 1u)) tb[3 * q] = cnt_table(c, dg, kp[q].sc_thr);
    if (kp[q].pr_ok && !(o.flags & (1u | 2u | 128u))) {
      if (kp[q].pr_ta > 1) tb[3 * q + 1] = cnt_table(c, dg, kp[q].pr_ta);
      if (kp[q].pr_tab > 1) tb[3 * q + 2] = cnt_table(c, dg, kp[q].pr_tab);
    }
    // chain level (never together with the pair level): A's table in slot 1
    if (!kp[q].pr_ok && kp[q].ch_ok && dg.h2_off && kp[q].ch_ta <= 2 && !(o.flags & (1u | 2u | 256u | 512u))) {
      if (kp[q].ch_ta > 1) tb[3 * q + 1] = cnt_table(c, dg, kp[q].ch_ta);
      chh[q] = chain_table(c, dg, kp[q].ch_ta, kp[q].ch_tb);
    }
  }
  BatchArgs& a = b.args;
  a.off_out = dg.off_out;
  a.adj_out = dg.adj_out;
  a.off_in = dg.off_in;
  a.adj_in = dg.adj_in;
  a.sig = dg.sig;
  a.V = dg.V;
  a.hub_id = (o.flags & 32u) ? nullptr : dg.hub_id;
  a.nsig_out = dg.nsig_out;
  a.nsig_in = dg.nsig_in;
  a.bits = dg.bits;
  a.words = dg.words;
  a.h2_off = (o.flags & 512u) ? nullptr : dg.h2_off;
  a.h2_v = dg.h2_v;
  a.h2_c = dg.h2_c;
  a.h2_did = (o.flags & 512u) ? nullptr : dg.h2_did;
  a.h2_dense = dg.h2_dense;
  KPlan* dplans = balloc<KPlan>(b, 0, b.nq);
  HIPCHECK(hipMemcpy(dplans, kp.data(), b.nq * sizeof(KPlan), hipMemcpyHostToDevice));
  a.plans = dplans;
  const uint32_t** dtb = balloc<const uint32_t*>(b, 1, tb.size());
  HIPCHECK(hipMemcpy(dtb, tb.data(), tb.size() * sizeof(void*), hipMemcpyHostToDevice));
  a.tbls = dtb;
  const uint64_t** dchh = balloc<const uint64_t*>(b, 12, b.nq);
  HIPCHECK(hipMemcpy(dchh, chh.data(), b.nq * sizeof(void*), hipMemcpyHostToDevice));
  a.chH = dchh;
  a.st = balloc<BState>(b, 2, 1);
  HIPCHECK(hipMemset(a.st, 0, sizeof(BState)));
  a.items = balloc<Item>(b, 3, 2 * qcap);
  a.qcap = qcap;
  a.arena = balloc<uint32_t>(b, 5, 2 * half_words);
  a.half_words = half_words;
  a.counts = balloc<unsigned long long>(b, 6, b.nq);
  a.levels = balloc<unsigned long long>(b, 7, (uint64_t)b.nq * kMaxQuery);
  a.qpend = balloc<unsigned long long>(b, 8, b.nq);
  a.done = balloc<unsigned long long>(b, 9, b.nq);
  uint32_t* droots = balloc<uint32_t>(b, 4, roots.size());
  HIPCHECK(hipMemcpy(droots, roots.data(), roots.size() * 4, hipMemcpyHostToDevice));
  a.roots = droots;
  unsigned long long* dqinit = balloc<unsigned long long>(b, 10, b.nq);
  HIPCHECK(hipMemcpy(dqinit, qinit.data(), b.nq * 8, hipMemcpyHostToDevice));
  a.qinit = dqinit;
  Item* dinit = balloc<Item>(b, 11, init.size());
  if (!init.empty()) HIPCHECK(hipMemcpy(dinit, init.data(), init.size() * sizeof(Item), hipMemcpyHostToDevice));
  a.init_items = dinit;
  a.n_init = init.size();
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device) == hipSuccess && khz > 0)
    b.tick_khz = khz;
  // 0xFFFFFFFF: no wait at all (the first barrier aborts the run; tests of the abort path)
  a.bar_ticks = o.barrier_timeout_ms == 0xFFFFFFFFu
                    ? 0ull
                    : (unsigned long long)((o.barrier_timeout_ms ? o.barrier_timeout_ms : 60000u) * b.tick_khz);
  if (b.persistent) {
    // every block of a cooperative launch must be resident at once
    int coop = 0, nb = 0;
    HIPCHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c.device));
    if (!coop) throw Error(-1, "batch: the device does not support cooperative launches (persistent kernel)");
    const void* f = b.kmax <= 8    ? (const void*)k_batch_all<7>
                    : b.kmax <= 12 ? (const void*)k_batch_all<11>
                                   : (const void*)k_batch_all<15>;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, BB, 0));
    if (nb <= 0) throw Error(-1, "batch: the persistent kernel does not fit a CU");
    bpc = std::min<uint32_t>(bpc, (uint32_t)nb);
    b.grid = (uint32_t)c.cus * bpc;
  }
  a.item_work = o.item_work ? o.item_work : 16384;
  a.flags = (o.flags & 3u) | ((o.flags & 64u) ? 4u : 0u) | ((o.flags & 128u) ? 8u : 0u) | ((o.flags & 256u) ? 16u : 0u);
  a.nq = b.nq;
}

// Enqueues init + the level launches (or init + the persistent kernel);
// returns the first launch error.
static hipError_t batch_launch(Batch& b, hipStream_t s) {
  const uint64_t nitems = std::max<uint64_t>(b.args.n_init, b.nq);
  const uint32_t igrid = (uint32_t)std::min<uint64_t>((nitems + BB - 1) / BB, 4096);
  hipLaunchKernelGGL(k_batch_init, dim3(std::max<uint32_t>(igrid, 1)), dim3(BB), 0, s, b.args);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (b.persistent) {
    hipLaunchAttribute attr[1];
    std::memset(attr, 0, sizeof(attr));
    attr[0].id = hipLaunchAttributeCooperative;
    attr[0].val.cooperative = 1;
    hipLaunchConfig_t cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.gridDim = dim3(b.grid);
    cfg.blockDim = dim3(BB);
    cfg.stream = s;
    cfg.attrs = attr;
    cfg.numAttrs = 1;
    uint32_t kmax = b.kmax;
    void* args[] = {(void*)&b.args, (void*)&kmax};
    const void* f = b.kmax <= 8    ? (const void*)k_batch_all<7>
                    : b.kmax <= 12 ? (const void*)k_batch_all<11>
                                   : (const void*)k_batch_all<15>;
    return hipLaunchKernelExC(&cfg, f, args);
  }
  for (uint32_t L = 0; L < b.kmax; ++L) {
    if (b.kmax <= 8) hipLaunchKernelGGL(k_batch_level<7>, dim3(b.grid), dim3(BB), 0, s, b.args, L);
    else if (b.kmax <= 12) hipLaunchKernelGGL(k_batch_level<11>, dim3(b.grid), dim3(BB), 0, s, b.args, L);
    else hipLaunchKernelGGL(k_batch_level<15>, dim3(b.grid), dim3(BB), 0, s, b.args, L);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void batch_run(Batch& b, std::vector<BatchQueryResult>& out, double& device_ms, uint64_t& arena_words) {
  Context& c = *b.ctx;
  HIPCHECK(hipSetDevice(c.device));
  if (b.use_graph && !b.exec) {
    HIPCHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = batch_launch(b, c.stream);
    hipGraph_t gr = nullptr;
    const hipError_t ce = hipStreamEndCapture(c.stream, &gr);  // always ends the capture
    if (le != hipSuccess || ce != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      (void)hipGetLastError();
      throw Error(-4, std::string("batch: capturing the launch sequence failed: ") +
                          hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b.graph = gr;
    HIPCHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  }
  HIPCHECK(hipEventRecord(c.ev0, c.stream));
  if (b.use_graph) HIPCHECK(hipGraphLaunch(b.exec, c.stream));
  else HIPCHECK(batch_launch(b, c.stream));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c.ev1, c.stream));
  HIPCHECK(hipEventSynchronize(c.ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  device_ms = ms;
  BState st;
  HIPCHECK(hipMemcpy(&st, b.args.st, sizeof(BState), hipMemcpyDeviceToHost));
  if (st.abort.v) {
    const char* why = (st.abort.v & 4)   ? "corrupt work item"
                      : (st.abort.v & 8) ? "a block waited at the level barrier longer than the limit"
                      : (st.abort.v & 1) ? "work queue full"
                                         : "row arena full";
    if (st.abort.v & 8) throw Error(-4, std::string("batch aborted: ") + why);
    throw Error(-5, std::string("batch aborted: ") + why + "; run fewer queries per batch or raise the limits");
  }
  std::vector<unsigned long long> cnt(b.nq), lev((size_t)b.nq * kMaxQuery), done(b.nq);
  HIPCHECK(hipMemcpy(cnt.data(), b.args.counts, b.nq * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(lev.data(), b.args.levels, lev.size() * 8, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(done.data(), b.args.done, b.nq * 8, hipMemcpyDeviceToHost));
  out.assign(b.nq, BatchQueryResult());
  for (uint32_t q = 0; q < b.nq; ++q) {
    out[q].count = cnt[q];
    for (uint32_t l = 0; l < kMaxQuery; ++l) out[q].level[l] = lev[(size_t)q * kMaxQuery + l];
    out[q].latency_ms = done[q] > st.t_start.v ? (double)(done[q] - st.t_start.v) / b.tick_khz : 0.0;
  }
  arena_words = 0;
  for (uint32_t l = 0; l <= kMaxQuery; ++l) arena_words = std::max<uint64_t>(arena_words, st.arena[l].v);
}

}  // namespace cuts
