// Batched matching: many queries against one resident data graph in a single
// persistent launch.  No reference counterpart (cuTS runs one query per
// process and syncs the host after every level, src/host_funcs.cu:141-157;
// its MPI driver distributes level-0 trunks, main_mpi.cu:257-376).
//
// Work items (query, level, row chunk) live in a device ticket queue:
//   * a block takes the next ticket with one atomicAdd and waits until that
//     slot is published (ready[slot] == epoch) -- tickets are never reused,
//     so there is no ABA race and no CAS loop;
//   * an item's rows are expanded with the same tests as the single-query
//     kernels (signature, injectivity, constraint membership, closed-form
//     last level); surviving rows are staged in LDS, flushed to a bump
//     arena with an exact allocation, and published as new items;
//   * `pending` counts published-but-unfinished items; a block whose ticket
//     lies beyond the final tail exits once pending reaches 0, and every
//     wait is also bounded by a wall-clock limit, so the grid always drains;
//   * the epoch is bumped on the device by the init kernel, so the whole
//     launch sequence can be captured once in a hipGraph and replayed
//     without resetting the ready flags.
// Per query: exact uint64 count, per-level table sizes and the wall-clock
// tick at which its last item finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "batch.hpp"

namespace cuts {
namespace {

constexpr int BB = 256;      // threads per block (4 wave64)
constexpr int RMAX = 256;    // rows per item
constexpr int WROW = 16;     // LDS words per staged row (k <= 16)

struct GV {
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
  uint32_t V;
  const int32_t* hub_id;  // hub adjacency bitmaps (symmetric graphs) or null
  const uint64_t* bits;
  uint32_t words;
};

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

template <typename T>
__device__ __forceinline__ T aload(T* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
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

struct Shared {
  Item item;
  int cmd;
  uint32_t rows[RMAX * WROW];    // input rows of the item
  uint32_t minlen[RMAX];
  uint8_t minsel[RMAX];
  uint32_t off[RMAX + 1];        // exclusive scan of minlen
  uint32_t stage[RMAX * WROW];   // output rows waiting for a flush
  uint32_t nstage;
  uint32_t wcnt[BB / 64], wpre[BB / 64];
  uint64_t base;
};

// Data handed from one block to another inside the persistent kernel (items,
// arena rows) goes through agent-scope atomic loads and stores: the XCDs of
// MI355X have separate L2 caches that ordinary loads do not keep coherent,
// so a plain load could return a line cached on the reader's XCD before the
// writer on another XCD filled it.
__device__ __forceinline__ void store_item(Item* dst, const Item& it) {
  unsigned long long* d = reinterpret_cast<unsigned long long*>(dst);
  const unsigned long long w0 = ((unsigned long long)it.n << 32) | it.qL;
  __hip_atomic_store(d, w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(d + 1, (unsigned long long)it.off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ Item load_item(const Item* src) {
  unsigned long long* p = const_cast<unsigned long long*>(reinterpret_cast<const unsigned long long*>(src));
  const unsigned long long w0 = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long w1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  Item it;
  it.qL = (uint32_t)w0;
  it.n = (uint32_t)(w0 >> 32);
  it.off = w1;
  return it;
}
__device__ __forceinline__ void store_word(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t load_word(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ bool push_item(const BatchArgs& a, uint32_t epoch, Item it) {
  const uint32_t q = it.qL >> 8;
  atomicAdd(&a.qpend[q], 1ull);
  atomicAdd(&a.st->pending, 1ull);
  const unsigned long long slot = atomicAdd(&a.st->tail, 1ull);
  if (slot >= a.qcap) {
    atomicExch(&a.st->abort, 1u);
    atomicSub(&a.st->pending, 1ull);
    atomicSub(&a.qpend[q], 1ull);
    return false;
  }
  store_item(&a.items[slot], it);
  __threadfence();
  __hip_atomic_store(&a.ready[slot], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Writes the staged rows (row length `words`) to the arena and publishes
// them as one item of level `L`.  Called by the whole block.
__device__ void flush_stage(const BatchArgs& a, Shared& s, uint32_t epoch, uint32_t q, uint32_t words, uint32_t L) {
  __syncthreads();
  const uint32_t n = s.nstage;
  if (n == 0) return;
  if (threadIdx.x == 0) {
    const unsigned long long need = (unsigned long long)n * words;
    const unsigned long long o = atomicAdd(&a.st->arena_head, need);
    if (o + need > a.arena_cap) {
      atomicExch(&a.st->abort, 2u);
      s.base = ~0ull;
    } else {
      s.base = o;
    }
  }
  __syncthreads();
  const uint64_t base = s.base;
  if (base != ~0ull)
    for (uint32_t i = threadIdx.x; i < n * words; i += BB)
      store_word(&a.arena[base + i], s.stage[(i / words) * WROW + i % words]);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (base != ~0ull) {
      __threadfence();
      Item it;
      it.qL = (q << 8) | L;
      it.n = n;
      it.off = base;
      push_item(a, epoch, it);
    }
    s.nstage = 0;
  }
  __syncthreads();
}

// Block-wide append of per-thread survivors (row + optional v) to the stage,
// flushing first when they would not fit.
__device__ void stage_append(const BatchArgs& a, Shared& s, uint32_t epoch, uint32_t q, bool ok,
                             const uint32_t* row, uint32_t L, bool has_v, uint32_t v) {
  const uint32_t tid = threadIdx.x, wave = tid >> 6;
  const uint64_t m = __ballot(ok);
  if ((tid & 63) == 0) s.wcnt[wave] = __popcll(m);
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
    for (int w = 0; w < BB / 64; ++w) { s.wpre[w] = t; t += s.wcnt[w]; }
    s.wcnt[0] = t;  // reused as the round total after the prefix is taken
  }
  __syncthreads();
  const uint32_t total = s.wcnt[0];
  const uint32_t words = L + (has_v ? 1 : 0);
  if (s.nstage + total > RMAX) flush_stage(a, s, epoch, q, words, words);
  if (ok) {
    const uint32_t slot = s.nstage + s.wpre[wave] + lane_rank64(m);
    uint32_t* dst = &s.stage[slot * WROW];
    for (uint32_t l = 0; l < L; ++l) dst[l] = row[l];
    if (has_v) dst[L] = v;
  }
  __syncthreads();
  if (tid == 0) s.nstage += total;
  __syncthreads();
}

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

__device__ void add_block(unsigned long long* dst, uint64_t v) {
  const uint64_t s = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

__device__ void process(const BatchArgs& a, Shared& s, uint32_t epoch) {
  const GV g{a.off_out, a.adj_out, a.off_in, a.adj_in, a.sig, a.V, a.hub_id, a.bits, a.words};
  const Item it = s.item;
  const uint32_t q = it.qL >> 8, L = it.qL & 0xFF;
  const KPlan& P = a.plans[q];
  const uint32_t k = P.k;
  const uint32_t tid = threadIdx.x;
  const uint32_t* tbl = a.tbls[q];
  const bool closed = (a.flags & 1u) == 0 && P.sc_ok;
  const bool fuse = closed && (a.flags & 2u) == 0 && k >= 3;
  uint64_t cnt = 0, surv = 0;

  if (L == 0) {
    // level-0 generation over vertices [off, off + n)
    for (uint32_t i0 = 0; i0 < it.n; i0 += BB) {
      const uint32_t i = i0 + tid;
      const uint32_t u = (uint32_t)it.off + i;
      const bool ok = i < it.n && g.sig[2 * (uint64_t)u] >= P.sig_in[0] && g.sig[2 * (uint64_t)u + 1] >= P.sig_out[0];
      surv += ok;
      if (k == 1) continue;
      stage_append(a, s, epoch, q, ok, &u, 1, false, 0);
    }
    if (k > 1) flush_stage(a, s, epoch, q, 1, 1);
    add_block(&a.levels[q * kMaxQuery + 0], surv);
    if (k == 1) add_block(&a.counts[q], surv);
    return;
  }

  const uint32_t n = it.n;
  const uint32_t* src = a.arena + it.off;
  for (uint32_t i = tid; i < n * L; i += BB) s.rows[(i / L) * WROW + i % L] = load_word(&src[i]);
  __syncthreads();

  if (L == k - 1 && closed) {
    for (uint32_t r = tid; r < n; r += BB) cnt += closed_form_rt(g, P, &s.rows[r * WROW], L, false, 0, tbl);
    add_block(&a.counts[q], cnt);
    add_block(&a.levels[q * kMaxQuery + L], cnt);
    return;
  }

  // shortest constraint list per row
  const uint32_t nj = P.nj[L];
  for (uint32_t r = tid; r < n; r += BB) {
    uint32_t best = 0xFFFFFFFFu, sel = 0;
    for (uint32_t c = 0; c < nj; ++c) {
      const uint32_t jc = P.jc[L][c];
      const uint32_t u = s.rows[r * WROW + (jc & 31)];
      const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
      const uint32_t len = off[u + 1] - off[u];
      if (len < best) { best = len; sel = c; }
    }
    s.minlen[r] = best;
    s.minsel[r] = (uint8_t)sel;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
    for (uint32_t r = 0; r < n; ++r) { s.off[r] = t; t += s.minlen[r]; }
    s.off[n] = t;
  }
  __syncthreads();
  const uint32_t W = s.off[n];
  const int mode = (L == k - 1) ? 1 : ((L == k - 2 && fuse) ? 2 : 0);
  const uint32_t tin = P.sig_in[L], tout = P.sig_out[L];
  for (uint32_t w0 = 0; w0 < W; w0 += BB) {
    const uint32_t w = w0 + tid;
    bool ok = false;
    uint32_t v = 0, r = 0;
    if (w < W) {
      uint32_t lo = 0, hi = n + 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.off[mid] <= w) lo = mid + 1;
        else hi = mid;
      }
      r = lo - 1;
      const uint32_t idx = w - s.off[r];
      const uint32_t* row = &s.rows[r * WROW];
      const uint32_t sel = s.minsel[r];
      const uint32_t jc = P.jc[L][sel];
      const uint32_t u = row[jc & 31];
      const uint32_t* off = (jc >> 7) ? g.off_in : g.off_out;
      const uint32_t* adj = (jc >> 7) ? g.adj_in : g.adj_out;
      v = adj[off[u] + idx];
      ok = g.sig[2 * (uint64_t)v] >= tin && g.sig[2 * (uint64_t)v + 1] >= tout;
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
    }
    if (mode == 1) {
      cnt += ok;
    } else if (mode == 2) {
      if (ok) {
        ++surv;
        cnt += closed_form_rt(g, P, &s.rows[r * WROW], L, true, v, tbl);
      }
    } else {
      surv += ok;
      stage_append(a, s, epoch, q, ok, &s.rows[r * WROW], L, true, v);
    }
  }
  if (mode == 0) flush_stage(a, s, epoch, q, L + 1, L + 1);
  if (mode == 1) {
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

}  // namespace

// Resets the queue (device side, so it can be replayed inside a hipGraph):
// bumps the epoch, clears counters and publishes the level-0 items.
__global__ __launch_bounds__(BB) void k_batch_init(BatchArgs a, uint32_t nq, uint32_t vchunk, uint32_t nch) {
  const uint64_t tid = (uint64_t)blockIdx.x * BB + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * BB;
  const uint32_t epoch = a.st->epoch + 1;  // every thread reads the old value first
  __syncthreads();
  for (uint64_t i = tid; i < (uint64_t)nq * kMaxQuery; i += nthreads) a.levels[i] = 0;
  for (uint64_t q = tid; q < nq; q += nthreads) {
    a.counts[q] = 0;
    a.qpend[q] = nch;
    a.done[q] = 0;
  }
  for (uint64_t i = tid; i < (uint64_t)nq * nch; i += nthreads) {
    const uint32_t q = (uint32_t)(i / nch), c = (uint32_t)(i % nch);
    Item it;
    it.qL = q << 8;
    it.off = (uint64_t)c * vchunk;
    it.n = (uint32_t)std::min<uint64_t>(vchunk, a.V - it.off);
    store_item(&a.items[i], it);
    __threadfence();
    __hip_atomic_store(&a.ready[i], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (tid == 0) {
    a.st->head = 0;
    a.st->tail = (unsigned long long)nq * nch;
    a.st->pending = (unsigned long long)nq * nch;
    a.st->arena_head = 0;
    a.st->abort = 0;
    a.st->t_start = wall_clock64();
  }
}

// Publishes the new epoch after k_batch_init has finished (same stream).
__global__ void k_batch_epoch(BatchArgs a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) a.st->epoch = a.st->epoch + 1;
}

__global__ __launch_bounds__(BB) void k_batch(BatchArgs a) {
  __shared__ Shared s;
  const uint32_t epoch = aload(&a.st->epoch);
  if (threadIdx.x == 0) s.nstage = 0;
  while (true) {
    if (threadIdx.x == 0) {
      int cmd = 1;
      const unsigned long long t = atomicAdd(&a.st->head, 1ull);
      if (t < a.qcap) {
        const uint64_t t0 = wall_clock64();
        for (uint32_t spin = 0;; ++spin) {
          if (aload(&a.ready[t]) == epoch) {
            __threadfence();
            const Item it = load_item(&a.items[t]);
            // an item outside the shapes this kernel writes means the queue
            // is corrupt: stop everything instead of looping over garbage
            const uint32_t q = it.qL >> 8, L = it.qL & 0xFF;
            const bool sane = q < a.nq && L < kMaxQuery &&
                              (L == 0 ? (it.off < a.V && it.n <= a.V - it.off)
                                      : (it.n >= 1 && it.n <= (uint32_t)RMAX &&
                                         it.off + (uint64_t)it.n * L <= a.arena_cap));
            if (!sane) {
              atomicExch(&a.st->abort, 4u);
              break;
            }
            s.item = it;
            cmd = 0;
            break;
          }
          // the shared counters are consulted on every 8th round only
          if ((spin & 7) == 7) {
            if (aload(&a.st->abort)) break;
            if (aload(&a.st->pending) == 0) {
              __threadfence();
              if (aload(&a.st->tail) <= t) break;
            }
            if (wall_clock64() - t0 > a.wait_ticks) {
              atomicExch(&a.st->abort, 3u);
              break;
            }
          }
          __builtin_amdgcn_s_sleep(32);
        }
      } else {
        atomicExch(&a.st->abort, 1u);
      }
      s.cmd = cmd;
    }
    __syncthreads();
    if (s.cmd) break;
    process(a, s, epoch);
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t q = s.item.qL >> 8;
      __threadfence();
      if (atomicSub(&a.qpend[q], 1ull) == 1ull) a.done[q] = wall_clock64();
      atomicSub(&a.st->pending, 1ull);
    }
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
  HIPCHECK(hipSetDevice(c.device));
  if (plans.empty()) throw Error(-1, "batch: no queries");
  if (plans.size() >= (1u << 24)) throw Error(-1, "batch: too many queries");
  b.ctx = &c;
  b.dg = &dg;
  b.nq = (uint32_t)plans.size();
  b.vchunk = o.vchunk ? o.vchunk : 4096;
  b.nch = (dg.V + b.vchunk - 1) / b.vchunk;
  if (b.nch == 0) b.nch = 1;
  b.use_graph = o.use_graph != 0;
  const uint32_t bpc = o.blocks_per_cu ? std::min<uint32_t>(o.blocks_per_cu, 4) : 1;
  b.grid = (uint32_t)c.cus * bpc;
  const uint64_t qcap = o.queue_cap ? o.queue_cap : (1ull << 22);
  if ((uint64_t)b.nq * b.nch > qcap) throw Error(-1, "batch: queue too small for the level-0 items");
  const uint64_t arena_words = (o.arena_bytes ? o.arena_bytes : (8ull << 30)) / 4;
  std::vector<KPlan> kp(b.nq);
  std::vector<const uint32_t*> tb(b.nq, nullptr);
  for (uint32_t q = 0; q < b.nq; ++q) {
    if (plans[q]->k == 0 || plans[q]->k > kMaxQuery) throw Error(-3, "batch: query size out of range");
    kp[q] = make_kplan(*plans[q], dg.symmetric);
    b.k.push_back(plans[q]->k);
    if (kp[q].sc_ok && kp[q].sc_thr > 1 && !(o.flags & 1u)) tb[q] = cnt_table(c, dg, kp[q].sc_thr);
  }
  BatchArgs& a = b.args;
  a.off_out = dg.off_out;
  a.adj_out = dg.adj_out;
  a.off_in = dg.off_in;
  a.adj_in = dg.adj_in;
  a.sig = dg.sig;
  a.V = dg.V;
  a.hub_id = (o.flags & 32u) ? nullptr : dg.hub_id;
  a.bits = dg.bits;
  a.words = dg.words;
  KPlan* dplans = balloc<KPlan>(b, 0, b.nq);
  HIPCHECK(hipMemcpy(dplans, kp.data(), b.nq * sizeof(KPlan), hipMemcpyHostToDevice));
  a.plans = dplans;
  const uint32_t** dtb = balloc<const uint32_t*>(b, 1, b.nq);
  HIPCHECK(hipMemcpy(dtb, tb.data(), b.nq * sizeof(void*), hipMemcpyHostToDevice));
  a.tbls = dtb;
  a.st = balloc<BState>(b, 2, 1);
  HIPCHECK(hipMemset(a.st, 0, sizeof(BState)));
  a.items = balloc<Item>(b, 3, qcap);
  a.ready = balloc<uint32_t>(b, 4, qcap);
  HIPCHECK(hipMemset(a.ready, 0, qcap * 4));
  a.qcap = qcap;
  a.arena = balloc<uint32_t>(b, 5, arena_words);
  a.arena_cap = arena_words;
  a.counts = balloc<unsigned long long>(b, 6, b.nq);
  a.levels = balloc<unsigned long long>(b, 7, (uint64_t)b.nq * kMaxQuery);
  a.qpend = balloc<unsigned long long>(b, 8, b.nq);
  a.done = balloc<unsigned long long>(b, 9, b.nq);
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c.device) == hipSuccess && khz > 0)
    b.tick_khz = khz;
  const double lim = o.wait_limit_s > 0 ? o.wait_limit_s : 300.0;
  a.wait_ticks = (uint64_t)(lim * 1000.0 * b.tick_khz);
  a.flags = o.flags & 3u;
  a.nq = b.nq;
}

static void batch_launch(Batch& b, hipStream_t s) {
  const uint64_t nitems = (uint64_t)b.nq * b.nch;
  const uint32_t igrid = (uint32_t)std::min<uint64_t>((nitems + BB - 1) / BB, 4096);
  hipLaunchKernelGGL(k_batch_init, dim3(std::max<uint32_t>(igrid, 1)), dim3(BB), 0, s, b.args, b.nq, b.vchunk,
                     b.nch);
  hipLaunchKernelGGL(k_batch_epoch, dim3(1), dim3(64), 0, s, b.args);
  hipLaunchKernelGGL(k_batch, dim3(b.grid), dim3(BB), 0, s, b.args);
}

void batch_run(Batch& b, std::vector<BatchQueryResult>& out, double& device_ms, uint64_t& arena_words) {
  Context& c = *b.ctx;
  HIPCHECK(hipSetDevice(c.device));
  if (b.use_graph && !b.exec) {
    HIPCHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    batch_launch(b, c.stream);
    HIPCHECK(hipStreamEndCapture(c.stream, &b.graph));
    HIPCHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
  }
  HIPCHECK(hipEventRecord(c.ev0, c.stream));
  if (b.use_graph) HIPCHECK(hipGraphLaunch(b.exec, c.stream));
  else batch_launch(b, c.stream);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c.ev1, c.stream));
  HIPCHECK(hipEventSynchronize(c.ev1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  device_ms = ms;
  BState st;
  HIPCHECK(hipMemcpy(&st, b.args.st, sizeof(BState), hipMemcpyDeviceToHost));
  if (st.abort) {
    const char* why = st.abort == 1   ? "work queue full"
                      : st.abort == 2 ? "row arena full"
                      : st.abort == 3 ? "wait limit reached"
                                      : "corrupt work item";
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
    out[q].latency_ms = done[q] > st.t_start ? (double)(done[q] - st.t_start) / b.tick_khz : 0.0;
  }
  arena_words = st.arena_head;
}

}  // namespace cuts
