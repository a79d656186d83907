// Batched persistent matcher: device structures shared by batch.hip and the
// C ABI.  Not part of the public ABI.
#pragma once
#include "device.hpp"

namespace cuts {

struct Item {      // 16 bytes
  uint32_t qL;     // query << 8 | level (row length in words; 0 = level-0 generation)
  uint32_t n;      // rows (or vertices for level 0)
  uint64_t off;    // arena word offset of the rows (first vertex for level 0)
};

// Queue state.  Every field has a 256-byte line of its own: waiting blocks
// poll some of them while working blocks update others, and on one shared
// line those accesses would serialise in a single L2 channel.
struct BState {
  unsigned long long head;
  char pad0[248];
  unsigned long long tail;
  char pad1[248];
  unsigned long long pending;
  char pad2[248];
  unsigned long long arena_head;
  char pad3[248];
  uint32_t abort;  // 1 queue full, 2 arena full, 3 wait limit, 4 corrupt item
  char pad4[252];
  uint32_t epoch;
  char pad5[252];
  unsigned long long t_start;
};

struct BatchArgs {
  const uint32_t *off_out, *adj_out, *off_in, *adj_in, *sig;
  uint32_t V;
  const int32_t* hub_id;
  const uint64_t* bits;
  uint32_t words;
  const KPlan* plans;
  const uint32_t* const* tbls;  // per query: closed-form table or null
  BState* st;
  Item* items;
  uint32_t* ready;
  uint64_t qcap;
  uint32_t* arena;
  uint64_t arena_cap;            // words
  unsigned long long *counts, *levels, *qpend, *done;
  uint64_t wait_ticks;
  uint32_t flags;                // 1: no closed form, 2: no fusion
  uint32_t nq;                   // queries in the batch
};

struct BatchOpts {
  uint32_t flags = 0;
  uint32_t blocks_per_cu = 0;
  uint32_t vchunk = 0;
  uint32_t use_graph = 1;
  uint64_t queue_cap = 0;
  uint64_t arena_bytes = 0;
  double wait_limit_s = 0;
};

struct BatchQueryResult {
  uint64_t count = 0;
  uint64_t level[kMaxQuery] = {0};
  double latency_ms = 0;
};

struct Batch {
  Context* ctx = nullptr;
  DeviceGraph* dg = nullptr;
  uint32_t nq = 0, nch = 0, vchunk = 0, grid = 0;
  BatchArgs args{};
  std::vector<uint32_t> k;   // query sizes
  void* allocs[12] = {nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  double tick_khz = 100000.0;
  ~Batch();
};

void batch_create(Context& c, DeviceGraph& dg, const std::vector<const Plan*>& plans, const BatchOpts& o, Batch& b);
// Runs the whole batch once; fills per-query results, the device time of the
// launch sequence and the arena words used.
void batch_run(Batch& b, std::vector<BatchQueryResult>& out, double& device_ms, uint64_t& arena_words);

// Degree-threshold count table of the closed form (built once per threshold).
const uint32_t* cnt_table(Context& c, DeviceGraph& dg, uint32_t thr);

}  // namespace cuts
