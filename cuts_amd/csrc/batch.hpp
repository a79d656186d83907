// Batched matcher: device structures shared by batch.hip and the C ABI.
// Not part of the public ABI.
#pragma once
#include "device.hpp"

namespace cuts {

struct Item {      // 16 bytes
  uint32_t qL;     // query << 8 | level (row length in words; 0 = level-0 generation)
  uint32_t n;      // rows (or roots for level 0)
  uint64_t off;    // arena word offset of the rows (for level 0: index of the first root)
};

// One counter per 256-byte line: blocks of a launch hit head[L] while
// others bump tail[L+1] and arena[L+1].
struct alignas(256) PadCtr {
  unsigned long long v;
};

struct BState {
  PadCtr head[kMaxQuery + 1];   // items of level L taken
  PadCtr tail[kMaxQuery + 1];   // items of level L published
  PadCtr arena[kMaxQuery + 1];  // arena words of level L allocated (in half L & 1)
  PadCtr abort;                 // 1 item queue full, 2 arena full, 4 corrupt item, 8 barrier wait too long
  PadCtr t_start;
  PadCtr bar;                   // persistent kernel: barrier arrivals (grows by the grid size per level)
};

struct BatchArgs {
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
  uint32_t V;
  const int32_t* hub_id;
  const uint2 *nsig_out, *nsig_in;
  const uint64_t* bits;
  uint32_t words;
  const uint64_t* h2_off;       // two-hop index or null
  const uint32_t *h2_v, *h2_c;
  const int32_t* h2_did;        // dense rows of the index or null
  const uint16_t* h2_dense;
  const KPlan* plans;
  const uint32_t* const* tbls;  // per query: 3 degree-threshold tables or null -- the closed form's
                                // (= B's of the pair level), A's and A & B's (A's of the chain level)
  const uint64_t* const* chH;   // per query: chain-level vertex tables or null
  BState* st;
  Item* items;                  // 2 halves of qcap items; level L uses half L & 1
  uint64_t qcap;
  uint32_t* arena;              // 2 halves of half_words; level L rows in half L & 1
  uint64_t half_words;
  unsigned long long *counts, *levels, *qpend, *done;
  const uint32_t* roots;        // per query: first image, stride, count of position-0 roots
  const unsigned long long* qinit;  // per query: level-0 items
  const Item* init_items;       // level-0 items, copied into the queue by every replay
  uint64_t n_init;
  uint64_t item_work;           // candidate work at which a staged item is flushed
  uint32_t flags;               // 1: no closed form, 2: no fusion, 4: no per-prefix base, 8: no pair level,
                                // 16: no chain level
  uint32_t nq;                  // queries in the batch
  unsigned long long bar_ticks; // persistent kernel: longest barrier wait (wall-clock ticks) before it aborts
};

struct BatchOpts {
  uint32_t flags = 0;
  uint32_t blocks_per_cu = 0;
  uint32_t vchunk = 0;
  uint32_t use_graph = 1;
  uint64_t queue_cap = 0;
  uint64_t arena_bytes = 0;
  uint32_t ticket_queue = 0;
  uint32_t item_work = 0;
  const uint32_t* roots = nullptr;  // null or 3 words per query: first, end, stride
  uint32_t persistent = 0;          // 1: one cooperative launch for every level
  uint32_t barrier_timeout_ms = 0;  // persistent: longest barrier wait; 0 = 60000
};

struct BatchQueryResult {
  uint64_t count = 0;
  uint64_t level[kMaxQuery] = {0};
  double latency_ms = 0;
};

struct Batch {
  Context* ctx = nullptr;
  DeviceGraph* dg = nullptr;
  uint32_t nq = 0, vchunk = 0, grid = 0, kmax = 0;
  BatchArgs args{};
  std::vector<uint32_t> k;   // query sizes
  void* allocs[13] = {nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  bool persistent = false;
  double tick_khz = 100000.0;
  ~Batch();
};

void batch_create(Context& c, DeviceGraph& dg, const std::vector<const Plan*>& plans, const BatchOpts& o, Batch& b);
// Runs the whole batch once; fills per-query results, the device time of the
// launch sequence and the peak arena words of one level.
void batch_run(Batch& b, std::vector<BatchQueryResult>& out, double& device_ms, uint64_t& arena_words);

// Degree-threshold count table of the closed form (built once per threshold).
const uint32_t* cnt_table(Context& c, DeviceGraph& dg, uint32_t thr);
// Chain-level vertex tables for thresholds (ta, tb), 2V uint64 (built once).
const uint64_t* chain_table(Context& c, DeviceGraph& dg, uint32_t ta, uint32_t tb);
// The same, built in `arena` for one call (not cached) when the graph has
// not cached them and arena is non-null.
const uint32_t* cnt_table(Context& c, DeviceGraph& dg, ScratchArena* arena, uint32_t thr);
const uint64_t* chain_table(Context& c, DeviceGraph& dg, ScratchArena* arena, uint32_t ta, uint32_t tb);

}  // namespace cuts
