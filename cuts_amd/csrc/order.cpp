// Matching order and backward-neighbour constraints of a query.
//
// heuristic_order() computes the greedy order of src/graph.cpp:62-93 with the
// three scores of src/score.cpp:18-80 and the lexicographic strict compare of
// src/score.cpp:81-99:
//   position 0: the first vertex of maximal out-degree;
//   then repeatedly the first unplaced vertex (in id order) whose score
//   (s1, s2, s3) is lexicographically strictly larger than every earlier
//   candidate's, starting from (0,0,0), where for the placed set S
//     s1 = |{u in S : v in out(u)}| + |out(v) & S|
//     s2 = |(out(S) \ S) & out(v)|
//     s3 = |(V \ (S u out(S))) & out(v)|.
// The reference reuses a stale vertex when every score is (0,0,0), which only
// happens for a query that is not weakly connected; such queries are
// rejected here instead.
//
// make_plan() builds P[i] = {j<i : q[j]->q[i]} and C[i] = {j<i : q[i]->q[j]}
// as in src/graph.cpp:94-133, plus per-position signatures and bit masks.
#include "host.hpp"

#include <algorithm>

namespace cuts {

bool weakly_connected(const Graph& q) {
  if (q.V == 0) return false;
  std::vector<char> seen(q.V, 0);
  std::vector<uint32_t> st{0};
  seen[0] = 1;
  uint32_t n = 1;
  while (!st.empty()) {
    uint32_t u = st.back();
    st.pop_back();
    auto visit = [&](uint32_t w) {
      if (!seen[w]) { seen[w] = 1; ++n; st.push_back(w); }
    };
    for (uint32_t i = q.out_off[u]; i < q.out_off[u + 1]; ++i) visit(q.out_adj[i]);
    for (uint32_t i = q.in_off[u]; i < q.in_off[u + 1]; ++i) visit(q.in_adj[i]);
  }
  return n == q.V;
}

std::string heuristic_order(const Graph& q, std::vector<uint32_t>& order) {
  const uint32_t V = q.V;
  order.clear();
  if (V == 0) return "empty query graph";
  uint32_t best = 0, idx = 0;
  for (uint32_t v = 0; v < V; ++v) {
    uint32_t od = q.out_off[v + 1] - q.out_off[v];
    if (best < od) { best = od; idx = v; }
  }
  order.push_back(idx);
  std::vector<char> inS(V, 0), inOut(V, 0);
  auto add = [&](uint32_t u) {
    inS[u] = 1;
    for (uint32_t i = q.out_off[u]; i < q.out_off[u + 1]; ++i) inOut[q.out_adj[i]] = 1;
  };
  add(idx);
  while (order.size() < V) {
    uint32_t m1 = 0, m2 = 0, m3 = 0;
    bool found = false;
    for (uint32_t v = 0; v < V; ++v) {
      if (inS[v]) continue;
      uint32_t s1 = 0, s2 = 0, s3 = 0;
      for (uint32_t u : order) {  // u in S with arc u->v
        const uint32_t* b = q.out_adj.data() + q.out_off[u];
        const uint32_t* e = q.out_adj.data() + q.out_off[u + 1];
        if (std::binary_search(b, e, v)) ++s1;
      }
      for (uint32_t i = q.out_off[v]; i < q.out_off[v + 1]; ++i) {
        uint32_t w = q.out_adj[i];
        if (inS[w]) ++s1;
        else if (inOut[w]) ++s2;
        else ++s3;
      }
      bool better = s1 > m1 || (s1 == m1 && (s2 > m2 || (s2 == m2 && s3 > m3)));
      if (better) { m1 = s1; m2 = s2; m3 = s3; idx = v; found = true; }
    }
    if (!found) return "query graph is not weakly connected";
    order.push_back(idx);
    add(idx);
  }
  return "";
}

std::string make_plan(const Graph& q, const std::vector<uint32_t>& order, Plan& p) {
  const uint32_t k = q.V;
  if (order.size() != k) return "order length differs from the query size";
  std::vector<int> pos(k, -1);
  for (uint32_t i = 0; i < k; ++i) {
    if (order[i] >= k || pos[order[i]] >= 0) return "order is not a permutation";
    pos[order[i]] = (int)i;
  }
  p = Plan();
  p.k = k;
  p.order = order;
  p.par_off.assign(k + 1, 0);
  p.chi_off.assign(k + 1, 0);
  p.pmask.assign(k, 0);
  p.cmask.assign(k, 0);
  p.sig_in.resize(k);
  p.sig_out.resize(k);
  for (uint32_t i = 0; i < k; ++i) {
    const uint32_t v = order[i];
    p.sig_in[i] = q.sig[2 * v];
    p.sig_out[i] = q.sig[2 * v + 1];
    std::vector<uint32_t> P, C;
    for (uint32_t j = 0; j < i; ++j) {
      const uint32_t t = order[j];
      // arc v->t makes j a "child" position, arc t->v a "parent" position
      if (std::binary_search(q.out_adj.begin() + q.out_off[v], q.out_adj.begin() + q.out_off[v + 1], t))
        C.push_back(j);
      if (std::binary_search(q.in_adj.begin() + q.in_off[v], q.in_adj.begin() + q.in_off[v + 1], t))
        P.push_back(j);
    }
    if (i > 0 && P.empty() && C.empty())
      return "order position " + std::to_string(i) + " has no earlier neighbour";
    for (uint32_t j : P) { p.par.push_back(j); if (j < 32) p.pmask[i] |= 1u << j; }
    for (uint32_t j : C) { p.chi.push_back(j); if (j < 32) p.cmask[i] |= 1u << j; }
    p.par_off[i + 1] = (uint32_t)p.par.size();
    p.chi_off[i + 1] = (uint32_t)p.chi.size();
  }
  return "";
}

}  // namespace cuts
