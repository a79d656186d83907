// extern "C" boundary (include/cuts.h).  Every entry point converts
// exceptions into status codes; nothing here calls exit().
#include "../../include/cuts.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "mpnn.hpp"

struct cuts_graph {
  cuts::Graph g;
};
struct cuts_plan {
  cuts::Plan p;
};
struct cuts_context {
  cuts::Context c;
  // cuts_match_many: one context (stream, trie arena, counters) per worker
  // thread, created on first use and kept
  std::vector<std::unique_ptr<cuts::Context>> workers;
  std::mutex workers_mu;
};
struct cuts_device_graph {
  cuts::DeviceGraph dg;
};
struct cuts_batch {
  cuts::Batch b;
};
struct cuts_mpnn {
  cuts::Mpnn m;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <typename F>
int guard(F&& f) {
  try {
    return f();
  } catch (const cuts::Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CUTS_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CUTS_E_ARG, e.what());
  }
}

int check(const std::string& err, int code) {
  if (!err.empty()) throw cuts::Error(code, err);
  return CUTS_OK;
}

// src/common.cpp:19-26: text after the last '/', or "" when there is none.
std::string base_name(const std::string& s) {
  size_t i = s.rfind('/');
  return i == std::string::npos ? std::string() : s.substr(i + 1);
}

// Default ostream formatting of a float, as `cout << time_milli_sec`.
std::string ostream_float(float f) {
  std::ostringstream os;
  os << f;
  return os.str();
}

void write_matches(const std::string& path, const cuts::Plan& p, const uint32_t* rows, uint64_t n) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw cuts::Error(CUTS_E_IO, "cannot open " + path);
  std::string buf;
  buf.reserve(1 << 22);
  char tmp[32];
  for (uint64_t r = 0; r < n; ++r) {
    for (uint32_t i = 0; i < p.k; ++i) {
      int len = snprintf(tmp, sizeof tmp, "(%u,%u)", p.order[i], rows[r * p.k + i]);
      buf.append(tmp, len);
    }
    buf += '\n';
    if (buf.size() > (1u << 22)) { out << buf; buf.clear(); }
  }
  out << buf;
  if (!out.good()) throw cuts::Error(CUTS_E_IO, "write failed for " + path);
}

void plan_for(const cuts::Graph& q, const uint32_t* order, cuts::Plan& p) {
  if (q.V == 0) throw cuts::Error(CUTS_E_QUERY, "empty query graph");
  if (q.V > cuts::kMaxQuery) throw cuts::Error(CUTS_E_QUERY, "query has more than 16 vertices");
  if (!cuts::weakly_connected(q)) throw cuts::Error(CUTS_E_QUERY, "query graph is not weakly connected");
  std::vector<uint32_t> ord;
  if (order) ord.assign(order, order + q.V);
  else check(cuts::heuristic_order(q, ord), CUTS_E_QUERY);
  check(cuts::make_plan(q, ord, p), CUTS_E_QUERY);
}

cuts::MatchOpts to_opts(const cuts_match_options* o) {
  cuts::MatchOpts m;
  if (!o) return m;
  m.trunks = o->trunks;
  m.flags = o->flags;
  m.budget = o->budget_bytes;
  m.root_begin = o->root_begin;
  m.root_end = o->root_end;
  m.root_stride = o->root_stride;
  m.blocks_per_cu = o->blocks_per_cu;
  m.emb_rows = o->emb_rows;
  m.emb_cap = o->emb_cap;
  m.time_limit_ms = o->time_limit_ms;
  m.max_work = o->max_work;
  m.expand_threads = o->expand_threads;
  for (int i = 0; i < 16; ++i) m.class_bpc[i] = o->class_blocks_per_cu[i];
  return m;
}

std::string json_escape(const std::string& in) {
  std::string o;
  for (char ch : in) {
    if (ch == '"' || ch == '\\') { o += '\\'; o += ch; }
    else if ((unsigned char)ch < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", ch); o += b; }
    else o += ch;
  }
  return o;
}

void log_tables(const std::string& path, const std::string& label, const cuts::Plan& p, uint64_t count,
                const uint64_t* level, uint64_t work, double ms, int complete) {
  std::ofstream out(path, std::ios::app);
  if (!out) throw cuts::Error(CUTS_E_IO, "cannot append to " + path);
  std::ostringstream os;
  os << "{\"label\": \"" << json_escape(label) << "\", \"k\": " << p.k << ", \"order\": [";
  for (uint32_t i = 0; i < p.k; ++i) os << (i ? ", " : "") << p.order[i];
  os << "], \"level_sizes\": [";
  for (uint32_t i = 0; i < p.k; ++i) os << (i ? ", " : "") << level[i];
  char msb[64];
  snprintf(msb, sizeof msb, "%.6f", ms);
  os << "], \"count\": " << count << ", \"work\": " << work << ", \"device_ms\": " << msb
     << ", \"complete\": " << complete << "}\n";
  out << os.str();
  if (!out.good()) throw cuts::Error(CUTS_E_IO, "write failed for " + path);
}

// search / search_dfs_bfs_strategy (src/host_funcs.cu:39-171).
int search_impl(const char* query_file, const char* data_file, int write_to_disk, bool trunked, uint32_t trunks,
                uint64_t* count) {
  std::cout << "start loading graph file from disk to memory..." << std::endl;
  cuts::Graph q, d;
  check(cuts::load_graph(query_file, q), CUTS_E_IO);
  check(cuts::load_graph(data_file, d), CUTS_E_IO);
  std::cout << "graph loading complete..." << std::endl;
  cuts::Plan p;
  plan_for(q, nullptr, p);
  const std::string qb = base_name(query_file), db = base_name(data_file);
  auto ctx = std::make_unique<cuts::Context>();
  cuts::context_init(*ctx, 0);
  std::cout << "start copying graph to gpu..." << std::endl;
  auto dg = std::make_unique<cuts::DeviceGraph>();
  cuts::upload_graph(*ctx, d, *dg);
  std::cout << (trunked ? "finish copying graph to gpu..." : "end copying graph to gpu...") << std::endl;
  cuts::MatchOpts o;
  o.trunks = trunked ? trunks : 0;
  cuts::MatchResult r;
  cuts::run_match(*ctx, *dg, p, o, r);
  std::cout << db << "," << qb << "," << ostream_float((float)r.device_ms) << "ms," << r.count << std::endl;
  if (const char* lp = std::getenv("CUTS_TABLE_LOG"))
    log_tables(lp, db + "," + qb, p, r.count, r.level, r.work, r.device_ms, (int)r.complete);
  if (write_to_disk) {
    std::cout << "start writting matching results to disk,ans.txt" << std::endl;
    const uint64_t words = r.count * (uint64_t)p.k;
    if (words > (1ull << 34)) throw cuts::Error(CUTS_E_NOMEM, "too many embeddings to write");
    std::vector<uint32_t> rows(std::max<uint64_t>(words, 1));
    cuts::MatchOpts ow = o;
    ow.flags |= CUTS_FLAG_WRITE;
    ow.emb_rows = rows.data();
    ow.emb_cap = r.count;
    cuts::MatchResult rw;
    cuts::run_match(*ctx, *dg, p, ow, rw);
    write_matches("ans.txt", p, rows.data(), rw.emb_written);
    std::cout << "finish writting matching results to disk,ans.txt" << std::endl;
  }
  if (count) *count = r.count;
  return CUTS_OK;
}
}  // namespace

extern "C" {

const char* cuts_last_error(void) { return g_err.c_str(); }
const char* cuts_version(void) { return "cuts_amd 0.1 (gfx950)"; }

int cuts_graph_load(const char* path, cuts_graph** out) {
  return guard([&]() -> int {
    if (!path || !out) return fail(CUTS_E_ARG, "null argument");
    auto g = std::make_unique<cuts_graph>();
    check(cuts::load_graph(path, g->g), CUTS_E_IO);
    *out = g.release();
    return CUTS_OK;
  });
}

int cuts_graph_from_arcs(uint32_t V, const uint32_t* src, const uint32_t* dst, uint64_t m, int symmetrize,
                         cuts_graph** out) {
  return guard([&]() -> int {
    if (!out || (m && (!src || !dst))) return fail(CUTS_E_ARG, "null argument");
    std::vector<uint64_t> arcs(m);
    for (uint64_t i = 0; i < m; ++i) arcs[i] = ((uint64_t)src[i] << 32) | dst[i];
    auto g = std::make_unique<cuts_graph>();
    check(cuts::build_graph(V, std::move(arcs), symmetrize != 0, g->g), CUTS_E_ARG);
    *out = g.release();
    return CUTS_OK;
  });
}

int cuts_graph_save(const cuts_graph* g, const char* path) {
  return guard([&]() -> int {
    if (!g || !path) return fail(CUTS_E_ARG, "null argument");
    return check(cuts::save_graph(path, g->g), CUTS_E_IO);
  });
}

int cuts_graph_two_hop(const cuts_graph* g, uint64_t* off, uint32_t* v, uint32_t* c, uint64_t cap,
                       uint64_t* entries) {
  if (!g || !entries) return fail(CUTS_E_ARG, "null argument");
  if (!g->g.symmetric) return fail(CUTS_E_ARG, "two-hop index: the graph is not symmetric");
  try {
    std::vector<uint64_t> o;
    std::vector<uint32_t> hv, hc;
    cuts::build_two_hop(g->g, o, hv, hc);
    *entries = o.back();
    if (cap == 0 && !off) return CUTS_OK;
    if (!off || (o.back() && (!v || !c))) return fail(CUTS_E_ARG, "null output buffer");
    if (o.back() > cap) return fail(CUTS_E_ARG, "two-hop index: " + std::to_string(o.back()) + " entries exceed cap");
    std::copy(o.begin(), o.end(), off);
    std::copy(hv.begin(), hv.begin() + o.back(), v);
    std::copy(hc.begin(), hc.begin() + o.back(), c);
    return CUTS_OK;
  } catch (const std::exception& e) {
    return fail(CUTS_E_ARG, e.what());
  }
}

int cuts_graph_view_of(const cuts_graph* g, cuts_graph_view* v) {
  if (!g || !v) return fail(CUTS_E_ARG, "null argument");
  v->V = g->g.V;
  v->arcs = g->g.arcs();
  v->symmetric = g->g.symmetric;
  v->out_off = g->g.out_off.data();
  v->out_adj = g->g.out_adj.data();
  v->in_off = g->g.in_off.data();
  v->in_adj = g->g.in_adj.data();
  v->sig = g->g.sig.data();
  return CUTS_OK;
}

void cuts_graph_free(cuts_graph* g) { delete g; }

int cuts_gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint32_t max_deg, uint64_t seed, cuts_graph** out) {
  return guard([&]() -> int {
    if (!out) return fail(CUTS_E_ARG, "null argument");
    std::vector<uint64_t> e;
    check(cuts::gen_chung_lu(n, m, gamma, max_deg, seed, e), CUTS_E_ARG);
    auto g = std::make_unique<cuts_graph>();
    check(cuts::build_graph(n, std::move(e), true, g->g), CUTS_E_ARG);
    *out = g.release();
    return CUTS_OK;
  });
}

int cuts_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                  cuts_graph** out) {
  return guard([&]() -> int {
    if (!out) return fail(CUTS_E_ARG, "null argument");
    std::vector<uint64_t> e;
    check(cuts::gen_rmat(scale, edge_factor, a, b, c, seed, e), CUTS_E_ARG);
    auto g = std::make_unique<cuts_graph>();
    check(cuts::build_graph(1u << scale, std::move(e), true, g->g), CUTS_E_ARG);
    *out = g.release();
    return CUTS_OK;
  });
}

int cuts_find_pattern(const cuts_graph* d, uint32_t k, const uint32_t* edges, uint32_t ne, uint64_t seed,
                      uint32_t* witness) {
  return guard([&]() -> int {
    if (!d || !edges || !witness) return fail(CUTS_E_ARG, "null argument");
    std::vector<uint32_t> pe(edges, edges + 2 * (size_t)ne), w;
    check(cuts::find_pattern(d->g, k, pe, seed, 100000, w), CUTS_E_QUERY);
    std::copy(w.begin(), w.end(), witness);
    return CUTS_OK;
  });
}

int cuts_walk_query(const cuts_graph* d, uint32_t k, uint32_t n_edges, uint64_t seed, uint32_t* edges,
                    uint32_t* n_out, uint32_t* witness) {
  return guard([&]() -> int {
    if (!d || !edges || !n_out || !witness) return fail(CUTS_E_ARG, "null argument");
    std::vector<uint32_t> qe, w;
    check(cuts::walk_query(d->g, k, n_edges, seed, qe, w), CUTS_E_QUERY);
    std::copy(qe.begin(), qe.end(), edges);
    *n_out = (uint32_t)(qe.size() / 2);
    std::copy(w.begin(), w.end(), witness);
    return CUTS_OK;
  });
}

int cuts_plan_create(const cuts_graph* query, const uint32_t* order, cuts_plan** out) {
  return guard([&]() -> int {
    if (!query || !out) return fail(CUTS_E_ARG, "null argument");
    auto p = std::make_unique<cuts_plan>();
    plan_for(query->g, order, p->p);
    *out = p.release();
    return CUTS_OK;
  });
}

int cuts_plan_info(const cuts_plan* p, uint32_t* k, uint32_t* order, uint32_t* par_off, uint32_t* par,
                   uint32_t* chi_off, uint32_t* chi) {
  if (!p) return fail(CUTS_E_ARG, "null argument");
  const cuts::Plan& P = p->p;
  if (k) *k = P.k;
  if (order) std::copy(P.order.begin(), P.order.end(), order);
  if (par_off) std::copy(P.par_off.begin(), P.par_off.end(), par_off);
  if (par) std::copy(P.par.begin(), P.par.end(), par);
  if (chi_off) std::copy(P.chi_off.begin(), P.chi_off.end(), chi_off);
  if (chi) std::copy(P.chi.begin(), P.chi.end(), chi);
  return CUTS_OK;
}

void cuts_plan_free(cuts_plan* p) { delete p; }

int cuts_context_create(int device, cuts_context** out) {
  return guard([&]() -> int {
    if (!out) return fail(CUTS_E_ARG, "null argument");
    auto c = std::make_unique<cuts_context>();
    cuts::context_init(c->c, device);
    *out = c.release();
    return CUTS_OK;
  });
}

void cuts_context_destroy(cuts_context* ctx) { delete ctx; }

int cuts_device_graph_upload(cuts_context* ctx, const cuts_graph* g, cuts_device_graph** out) {
  return guard([&]() -> int {
    if (!ctx || !g || !out) return fail(CUTS_E_ARG, "null argument");
    auto d = std::make_unique<cuts_device_graph>();
    cuts::upload_graph(ctx->c, g->g, d->dg);
    *out = d.release();
    return CUTS_OK;
  });
}

void cuts_device_graph_free(cuts_device_graph* dg) { delete dg; }

int cuts_device_graph_query(const cuts_device_graph* dg, cuts_device_graph_info* info) {
  return guard([&]() -> int {
    if (!dg || !info) return fail(CUTS_E_ARG, "null argument");
    const cuts::DeviceGraph& d = dg->dg;
    std::memset(info, 0, sizeof(*info));
    info->V = d.V;
    info->arcs = d.A;
    info->symmetric = d.symmetric ? 1 : 0;
    info->self_loops = d.loops;
    info->hubs = d.hubs;
    info->hub_min_deg = d.hub_min_deg;
    info->two_hop_entries = d.h2_off ? d.h2_entries : 0;
    info->two_hop_dense_rows = d.h2_dense_rows;
    info->arc_common = d.arc_cn ? 1 : 0;
    info->two_hop_build_s = d.h2_build_s;
    info->arc_common_build_s = d.arc_cn_build_s;
    info->two_hop_on_device = d.h2_on_gpu ? 1 : 0;
    info->hub_rows = d.hr_rows;
    info->hub_rows_build_s = d.hr_build_s;
    return CUTS_OK;
  });
}

int cuts_device_graph_arc_common(const cuts_device_graph* dg, uint32_t* out, uint64_t n) {
  return guard([&]() -> int {
    if (!dg || !out) return fail(CUTS_E_ARG, "null argument");
    const cuts::DeviceGraph& d = dg->dg;
    if (!d.arc_cn) return fail(CUTS_E_ARG, "the device graph has no per-arc common-neighbour table");
    if (n != d.A) return fail(CUTS_E_ARG, "n must equal the number of arcs");
    cuts::hip_check(hipSetDevice(d.device), "hipSetDevice");
    cuts::hip_check(hipMemcpy(out, d.arc_cn, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return CUTS_OK;
  });
}

int cuts_device_graph_two_hop(const cuts_device_graph* dg, uint64_t* off, uint32_t* v, uint32_t* c, uint64_t cap,
                              int32_t* dense_id, uint16_t* dense, uint64_t dense_cap) {
  return guard([&]() -> int {
    if (!dg || !off) return fail(CUTS_E_ARG, "null argument");
    const cuts::DeviceGraph& d = dg->dg;
    if (!d.h2_off) return fail(CUTS_E_ARG, "the device graph has no two-hop index");
    const uint64_t n = d.h2_entries;
    if (n > cap) return fail(CUTS_E_ARG, "two-hop index: " + std::to_string(n) + " entries exceed cap");
    if (n && (!v || !c)) return fail(CUTS_E_ARG, "null output buffer");
    const uint64_t dn = (uint64_t)d.h2_dense_rows * d.V;
    if (dense_id && d.h2_dense_rows && (!dense || dense_cap < dn))
      return fail(CUTS_E_ARG, "two-hop index: the dense rows need " + std::to_string(dn) + " entries");
    cuts::hip_check(hipSetDevice(d.device), "hipSetDevice");
    cuts::hip_check(hipMemcpy(off, d.h2_off, ((uint64_t)d.V + 1) * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (n) {
      cuts::hip_check(hipMemcpy(v, d.h2_v, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
      cuts::hip_check(hipMemcpy(c, d.h2_c, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    if (dense_id) {
      if (d.h2_dense_rows) {
        cuts::hip_check(hipMemcpy(dense_id, d.h2_did, (uint64_t)d.V * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        cuts::hip_check(hipMemcpy(dense, d.h2_dense, dn * 2, hipMemcpyDeviceToHost), "hipMemcpy");
      } else {
        std::fill(dense_id, dense_id + d.V, -1);
      }
    }
    return CUTS_OK;
  });
}

int cuts_device_graph_hub_rows(const cuts_device_graph* dg, int32_t* row_id, uint16_t* rows, uint64_t cap) {
  return guard([&]() -> int {
    if (!dg || !row_id || !rows) return fail(CUTS_E_ARG, "null argument");
    const cuts::DeviceGraph& d = dg->dg;
    if (!d.hr_rows) return fail(CUTS_E_ARG, "the device graph has no hub rows");
    const uint64_t n = (uint64_t)d.hr_rows * d.V;
    if (cap < n) return fail(CUTS_E_ARG, "hub rows: " + std::to_string(n) + " entries exceed cap");
    cuts::hip_check(hipSetDevice(d.device), "hipSetDevice");
    cuts::hip_check(hipMemcpy(row_id, d.hr_id, (uint64_t)d.V * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    cuts::hip_check(hipMemcpy(rows, d.hr, n * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    return CUTS_OK;
  });
}

void cuts_match_options_default(cuts_match_options* o) {
  if (o) std::memset(o, 0, sizeof(*o));
}

static void fill_result(const cuts::MatchResult& r, cuts_match_result* res) {
  std::memset(res, 0, sizeof(*res));
  res->count = r.count;
  for (uint32_t i = 0; i < cuts::kMaxQuery; ++i) res->level_sizes[i] = r.level[i];
  res->device_ms = r.device_ms;
  res->emb_written = r.emb_written;
  res->chunks = r.chunks;
  res->complete = r.complete;
  res->roots_done = r.roots_done;
  res->last_root = r.last_root;
  res->work_capped = r.work_capped;
  res->work = r.work;
  for (int i = 0; i < 11; ++i) {
    res->kernel_ms[i] = r.kernel_ms[i];
    res->kernel_launches[i] = r.kernel_launches[i];
    res->kernel_bytes[i] = r.kernel_bytes[i];
    res->merge_bytes[i] = r.merge_bytes[i];
  }
}

int cuts_match(cuts_context* ctx, const cuts_device_graph* dg, const cuts_plan* plan, const cuts_match_options* opt,
               cuts_match_result* res) {
  return guard([&]() -> int {
    if (!ctx || !dg || !plan || !res) return fail(CUTS_E_ARG, "null argument");
    if (opt && (opt->flags & CUTS_FLAG_WRITE) && opt->emb_cap && !opt->emb_rows)
      return fail(CUTS_E_ARG, "CUTS_FLAG_WRITE needs emb_rows");
    cuts::MatchResult r;
    cuts::run_match(ctx->c, const_cast<cuts::DeviceGraph&>(dg->dg), plan->p, to_opts(opt), r);
    fill_result(r, res);
    return CUTS_OK;
  });
}

void cuts_batch_options_default(cuts_batch_options* o) {
  if (o) std::memset(o, 0, sizeof(*o));
}

int cuts_batch_create(cuts_context* ctx, const cuts_device_graph* dg, const cuts_plan* const* plans, uint32_t n,
                      const cuts_batch_options* opt, cuts_batch** out) {
  return guard([&]() -> int {
    if (!ctx || !dg || !plans || !out || n == 0) return fail(CUTS_E_ARG, "invalid argument");
    std::vector<const cuts::Plan*> ps(n);
    for (uint32_t i = 0; i < n; ++i) {
      if (!plans[i]) return fail(CUTS_E_ARG, "null plan");
      ps[i] = &plans[i]->p;
    }
    cuts::BatchOpts o;
    if (opt) {
      o.flags = opt->flags;
      o.blocks_per_cu = opt->blocks_per_cu;
      o.vchunk = opt->vchunk;
      o.use_graph = opt->no_graph ? 0 : 1;
      o.queue_cap = opt->queue_cap;
      o.arena_bytes = opt->arena_bytes;
      o.ticket_queue = opt->ticket_queue;
      o.item_work = opt->item_work;
      o.roots = opt->roots;
      o.persistent = opt->persistent;
      o.barrier_timeout_ms = opt->barrier_timeout_ms;
    }
    auto b = std::make_unique<cuts_batch>();
    cuts::batch_create(ctx->c, const_cast<cuts::DeviceGraph&>(dg->dg), ps, o, b->b);
    *out = b.release();
    return CUTS_OK;
  });
}

int cuts_batch_run(cuts_batch* b, cuts_batch_query_result* results, double* device_ms, uint64_t* arena_bytes) {
  return guard([&]() -> int {
    if (!b || !results) return fail(CUTS_E_ARG, "null argument");
    std::vector<cuts::BatchQueryResult> r;
    double ms = 0;
    uint64_t words = 0;
    cuts::batch_run(b->b, r, ms, words);
    for (size_t i = 0; i < r.size(); ++i) {
      results[i].count = r[i].count;
      for (uint32_t l = 0; l < 16; ++l) results[i].level_sizes[l] = r[i].level[l];
      results[i].latency_ms = r[i].latency_ms;
    }
    if (device_ms) *device_ms = ms;
    if (arena_bytes) *arena_bytes = words * 4;
    return CUTS_OK;
  });
}

void cuts_batch_free(cuts_batch* b) { delete b; }

uint32_t cuts_mpnn_num_params(void) { return (uint32_t)cuts::kMpnnParams; }

int cuts_mpnn_create(cuts_context* ctx, const float* params, cuts_mpnn** out) {
  return guard([&]() -> int {
    if (!ctx || !params || !out) return fail(CUTS_E_ARG, "null argument");
    auto m = std::make_unique<cuts_mpnn>();
    cuts::mpnn_init(ctx->c, m->m, params);
    *out = m.release();
    return CUTS_OK;
  });
}

int cuts_mpnn_get_params(cuts_mpnn* m, float* params) {
  return guard([&]() -> int {
    if (!m || !params) return fail(CUTS_E_ARG, "null argument");
    cuts::mpnn_get(m->m, params);
    return CUTS_OK;
  });
}

int cuts_mpnn_set_params(cuts_mpnn* m, const float* params) {
  return guard([&]() -> int {
    if (!m || !params) return fail(CUTS_E_ARG, "null argument");
    cuts::mpnn_set(m->m, params);
    return CUTS_OK;
  });
}

int cuts_mpnn_forward(cuts_mpnn* m, uint32_t n, const float* x, const uint16_t* adj, const uint32_t* meta, float* y) {
  return guard([&]() -> int {
    if (!m || (n && (!x || !adj || !meta || !y))) return fail(CUTS_E_ARG, "null argument");
    cuts::mpnn_forward(m->m, n, x, adj, meta, y);
    return CUTS_OK;
  });
}

int cuts_mpnn_gradient(cuts_mpnn* m, uint32_t n, const float* x, const uint16_t* adj, const uint32_t* meta,
                       const float* target, float* grad, double* loss) {
  return guard([&]() -> int {
    if (!m || !n || !x || !adj || !meta || !target) return fail(CUTS_E_ARG, "invalid argument");
    const double l = cuts::mpnn_gradient(m->m, n, x, adj, meta, target, grad);
    if (loss) *loss = l;
    return CUTS_OK;
  });
}

int cuts_mpnn_train_step(cuts_mpnn* m, uint32_t n, const float* x, const uint16_t* adj, const uint32_t* meta,
                         const float* target, float lr, double* loss) {
  return guard([&]() -> int {
    if (!m || !n || !x || !adj || !meta || !target) return fail(CUTS_E_ARG, "invalid argument");
    const double l = cuts::mpnn_train_step(m->m, n, x, adj, meta, target, lr);
    if (loss) *loss = l;
    return CUTS_OK;
  });
}

void cuts_mpnn_free(cuts_mpnn* m) { delete m; }

int cuts_match_many(cuts_context* ctx, const cuts_device_graph* dg, const cuts_plan* const* plans, uint32_t n,
                    const cuts_match_options* opts, uint32_t n_opts, uint32_t workers, cuts_match_result* res,
                    double* wall_ms) {
  return guard([&]() -> int {
    if (!ctx || !dg || (n && (!plans || !res))) return fail(CUTS_E_ARG, "null argument");
    if (n_opts != 0 && n_opts != 1 && n_opts != n) return fail(CUTS_E_ARG, "n_opts must be 0, 1 or n");
    if (n_opts && !opts) return fail(CUTS_E_ARG, "null options");
    for (uint32_t i = 0; i < n; ++i)
      if (!plans[i]) return fail(CUTS_E_ARG, "null plan");
    for (uint32_t i = 0; i < n_opts; ++i)
      if ((opts[i].flags & CUTS_FLAG_WRITE) && opts[i].emb_cap && !opts[i].emb_rows)
        return fail(CUTS_E_ARG, "CUTS_FLAG_WRITE needs emb_rows");
    if (workers == 0) workers = 4;
    workers = std::min<uint32_t>(workers, 16);
    {
      std::lock_guard<std::mutex> lk(ctx->workers_mu);
      while (ctx->workers.size() < workers) {
        auto w = std::make_unique<cuts::Context>();
        cuts::context_init(*w, ctx->c.device);
        ctx->workers.push_back(std::move(w));
      }
      // the trie tables of the workers share the main context's budget
      for (auto& w : ctx->workers) w->default_budget = ctx->c.default_budget / workers;
    }
    std::atomic<uint32_t> next{0};
    std::mutex err_mu;
    int err_code = CUTS_OK;
    std::string err_msg;
    auto body = [&](cuts::Context* wc) {
      for (;;) {
        const uint32_t i = next.fetch_add(1);
        if (i >= n) break;
        try {
          cuts::MatchResult r;
          const cuts_match_options* o = n_opts == 0 ? nullptr : (n_opts == 1 ? opts : opts + i);
          cuts::run_match(*wc, const_cast<cuts::DeviceGraph&>(dg->dg), plans[i]->p, to_opts(o), r);
          fill_result(r, res + i);
        } catch (const cuts::Error& e) {
          std::lock_guard<std::mutex> lk(err_mu);
          if (err_code == CUTS_OK) { err_code = e.code; err_msg = e.what(); }
          next = n;  // stop handing out queries
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(err_mu);
          if (err_code == CUTS_OK) { err_code = CUTS_E_ARG; err_msg = e.what(); }
          next = n;
        }
      }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < workers; ++t) th.emplace_back(body, ctx->workers[t].get());
    for (auto& t : th) t.join();
    if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (err_code != CUTS_OK) return fail(err_code, err_msg);
    return CUTS_OK;
  });
}

int cuts_match_noalloc(cuts_context* ctx, const cuts_device_graph* dg, const cuts_plan* plan,
                       const cuts_match_options* opt, cuts_match_result* res) {
  return guard([&]() -> int {
    if (!ctx || !dg || !plan || !res) return fail(CUTS_E_ARG, "null argument");
    if (opt && (opt->flags & CUTS_FLAG_WRITE) && opt->emb_cap && !opt->emb_rows)
      return fail(CUTS_E_ARG, "CUTS_FLAG_WRITE needs emb_rows");
    cuts::MatchResult r;
    cuts::run_match_noalloc(ctx->c, const_cast<cuts::DeviceGraph&>(dg->dg), plan->p, to_opts(opt), r);
    fill_result(r, res);
    return CUTS_OK;
  });
}

uint64_t cuts_noalloc_scratch_bytes(void) { return cuts::noalloc_scratch_bytes(); }

int cuts_host_alloc(uint64_t bytes, void** out) {
  return guard([&]() -> int {
    if (!out) return fail(CUTS_E_ARG, "null argument");
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      *out = nullptr;
      return fail(CUTS_E_NOMEM, "pinned host allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return CUTS_OK;
  });
}

void cuts_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int cuts_log_tables(const char* path, const char* label, const cuts_plan* plan, const cuts_match_result* r) {
  return guard([&]() -> int {
    if (!path || !plan || !r) return fail(CUTS_E_ARG, "null argument");
    log_tables(path, label ? label : "", plan->p, r->count, r->level_sizes, r->work, r->device_ms, (int)r->complete);
    return CUTS_OK;
  });
}

int cuts_copy_probe(cuts_context* ctx, uint64_t bytes, uint32_t iters, double* gbps) {
  return guard([&]() -> int {
    if (!ctx || !gbps || !iters) return fail(CUTS_E_ARG, "invalid argument");
    *gbps = cuts::copy_probe(ctx->c, bytes, iters);
    return CUTS_OK;
  });
}

int cuts_search(const char* query_file, const char* data_file, int write_to_disk, uint64_t* count) {
  return guard([&]() -> int {
    if (!query_file || !data_file) return fail(CUTS_E_ARG, "null argument");
    return search_impl(query_file, data_file, write_to_disk, false, 0, count);
  });
}

int cuts_search_dfs_bfs_strategy(const char* query_file, const char* data_file, int write_to_disk, uint32_t trunks,
                                 uint64_t* count) {
  return guard([&]() -> int {
    if (!query_file || !data_file) return fail(CUTS_E_ARG, "null argument");
    return search_impl(query_file, data_file, write_to_disk, true, trunks, count);
  });
}

int cuts_write_matches(const char* path, const cuts_plan* plan, const uint32_t* rows, uint64_t n) {
  return guard([&]() -> int {
    if (!path || !plan || (n && !rows)) return fail(CUTS_E_ARG, "null argument");
    write_matches(path, plan->p, rows, n);
    return CUTS_OK;
  });
}

}  // extern "C"
