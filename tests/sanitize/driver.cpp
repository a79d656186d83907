// Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer
// build (tests/test_sanitize.py): runs the edge-list loader on well-formed
// and malformed files, the CSR build, the generators, the walk-query
// extractor, the two-hop index builder, the score order and plan builder,
// and the CPU oracle and trie baseline, checking a few invariants so that a
// sanitizer report or a failed check fails the test.  No GPU code.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../cuts_amd/csrc/host.hpp"
extern "C" {
#include "../../oracle/oracle.h"
}

#define CHECK(c)                                                       \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(2);                                                    \
    }                                                                  \
  } while (0)

static oracle_graph view(const cuts::Graph& g) {
  oracle_graph o;
  o.V = g.V;
  o.off_out = g.out_off.data();
  o.adj_out = g.out_adj.data();
  o.off_in = g.in_off.data();
  o.adj_in = g.in_adj.data();
  o.sig = g.sig.data();
  return o;
}

int main(int argc, char** argv) {
  // 1. loader: every file named on the command line; the expected outcome
  //    ("ok" or "err") follows each path
  for (int i = 1; i + 1 < argc; i += 2) {
    cuts::Graph g;
    const std::string err = cuts::load_graph(argv[i], g);
    const bool want_ok = std::string(argv[i + 1]) == "ok";
    std::printf("load %s: %s\n", argv[i], err.empty() ? "ok" : err.c_str());
    CHECK(err.empty() == want_ok);
    if (err.empty()) CHECK(g.out_off.size() == g.V + 1 && g.sig.size() == 2ull * g.V);
  }
  // 2. generators, CSR, two-hop index, walk queries, order, oracle, trie
  for (uint64_t seed = 0; seed < 3; ++seed) {
    std::vector<uint64_t> e;
    CHECK(cuts::gen_chung_lu(200, 700, 2.2, 40, seed, e).empty());
    cuts::Graph g;
    CHECK(cuts::build_graph(200, std::move(e), true, g).empty());
    CHECK(g.symmetric);
    std::vector<uint64_t> off;
    std::vector<uint32_t> hv, hc;
    cuts::build_two_hop(g, off, hv, hc);
    CHECK(off.size() == g.V + 1 && off[g.V] == hv.size() && hv.size() == hc.size());
    std::vector<uint64_t> r;
    CHECK(cuts::gen_rmat(10, 8, 0.57, 0.19, 0.19, seed, r).empty());
    cuts::Graph rg;
    CHECK(cuts::build_graph(1u << 10, std::move(r), true, rg).empty());
    const oracle_graph og = view(g);
    for (uint32_t k = 3; k <= 5; ++k) {
      std::vector<uint32_t> qe, wit;
      CHECK(cuts::walk_query(g, k, k + 1, seed * 31 + k, qe, wit).empty());
      std::vector<uint64_t> arcs;
      for (size_t j = 0; j + 1 < qe.size(); j += 2) arcs.push_back(((uint64_t)qe[j] << 32) | qe[j + 1]);
      cuts::Graph q;
      CHECK(cuts::build_graph(k, std::move(arcs), true, q).empty());
      std::vector<uint32_t> order;
      CHECK(cuts::heuristic_order(q, order).empty());
      cuts::Plan p;
      CHECK(cuts::make_plan(q, order, p).empty());
      std::vector<uint8_t> qa(k * k, 0);
      for (uint32_t u = 0; u < k; ++u)
        for (uint32_t j = q.out_off[u]; j < q.out_off[u + 1]; ++j) qa[u * k + q.out_adj[j]] = 1;
      const uint64_t c = oracle_count(&og, k, qa.data(), 2);
      trie_cpu_result tr;
      CHECK(trie_cpu_run(&og, k, qa.data(), order.data(), 1, 1 << 12, 2, &tr) == 0);
      CHECK(tr.count == c);
      CHECK(c >= 1);  // the walk's own vertices are an embedding
      std::printf("seed %llu k %u: count %llu\n", (unsigned long long)seed, k, (unsigned long long)c);
    }
  }
  std::printf("sanitize driver ok\n");
  return 0;
}
