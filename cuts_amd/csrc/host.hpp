// Host-side data model of the trie matcher: graphs in CSR form, query plans
// (matching order + backward-neighbour constraints) and the synthetic
// generators.  Everything here is plain C++17 + OpenMP; device code lives in
// the .hip files and talks to this through the structs below.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cuts {

// Largest query the device kernels accept (positions are held in 16-bit masks).
constexpr uint32_t kMaxQuery = 16;

// A directed graph in the layout the reference's Graph class builds
// (src/graph.cpp:18-61): sorted, duplicate-free out-lists and in-lists in CSR
// form plus a per-vertex signature [in-degree, out-degree].
struct Graph {
  uint32_t V = 0;
  std::vector<uint32_t> out_off, out_adj;  // CSR of out-neighbours
  std::vector<uint32_t> in_off, in_adj;    // CSR of in-neighbours
  std::vector<uint32_t> sig;               // 2*V: sig[2v]=in, sig[2v+1]=out
  bool symmetric = false;                  // out-lists == in-lists
  uint64_t arcs() const { return out_adj.size(); }
};

// Builds the CSR/signature form from a list of arcs (duplicates collapse,
// self-loops are kept, as the reference's std::set-based reader does).
// If `symmetrize`, every arc s->t also adds t->s.  Returns "" or an error.
std::string build_graph(uint32_t V, std::vector<uint64_t>&& arcs, bool symmetrize, Graph& g);

// Edge-list reader with the semantics of src/util.cpp:19-52.
std::string load_graph(const std::string& path, Graph& g);
// Writes a graph in the same edge-list format (header V, then "s\tt").
std::string save_graph(const std::string& path, const Graph& g);

// Two-hop index of a symmetric graph: for every vertex u, the vertices
// v != u that share a neighbour with u, ascending (v over off[u] ..
// off[u+1]), and |N(u) & N(v)| for each (c).  OpenMP over u.
void build_two_hop(const Graph& g, std::vector<uint64_t>& off, std::vector<uint32_t>& v, std::vector<uint32_t>& c);

// A query plan: the matching order and, per order position i, the earlier
// positions j whose image constrains position i's candidates
// (src/graph.cpp:94-133).
struct Plan {
  uint32_t k = 0;
  std::vector<uint32_t> order;            // order[i] = query vertex at position i
  std::vector<uint32_t> par_off, par;     // P[i] = {j<i : arc q[j]->q[i]}  (CSR)
  std::vector<uint32_t> chi_off, chi;     // C[i] = {j<i : arc q[i]->q[j]}  (CSR)
  std::vector<uint32_t> sig_in, sig_out;  // degrees of q[order[i]]
  std::vector<uint32_t> pmask, cmask;     // bit j set <=> j in P[i] / C[i]
};

// Greedy score order of src/graph.cpp:62-93 + src/score.cpp:18-99.
std::string heuristic_order(const Graph& q, std::vector<uint32_t>& order);
// Builds P/C and signatures for a given order (any permutation whose every
// prefix is weakly connected).
std::string make_plan(const Graph& q, const std::vector<uint32_t>& order, Plan& p);
bool weakly_connected(const Graph& q);

// ---- synthetic inputs -----------------------------------------------------
struct Rng {  // xoshiro256** seeded through splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed);
  uint64_t next();
  double uniform();                       // [0,1)
  uint64_t below(uint64_t n);             // [0,n)
};
// Chung-Lu: weights (i+1)^(-1/(gamma-1)); endpoints drawn in proportion to
// the weights until exactly m distinct non-loop undirected edges exist.  If
// max_deg > 0, weights are capped so that no expected degree exceeds it and
// any edge that would push a vertex above max_deg is rejected.
std::string gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint32_t max_deg,
                         uint64_t seed, std::vector<uint64_t>& edges);
// R-MAT with (a,b,c,1-a-b-c); 2^scale vertices, edge_factor*2^scale sampled
// edges; self-loops dropped.  Returned edges are u<v keys (dedup happens in
// build_graph).
std::string gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                     uint64_t seed, std::vector<uint64_t>& edges);
// Finds one occurrence of the undirected pattern (k vertices, edge list)
// in the symmetric data graph by a seeded randomized walk-and-extend search.
// witness[i] = data vertex of pattern vertex i.
std::string find_pattern(const Graph& d, uint32_t k, const std::vector<uint32_t>& pe,
                         uint64_t seed, uint64_t max_tries, std::vector<uint32_t>& witness);
// Random-walk query: walks from a random vertex until k distinct vertices
// are seen; the walk's first-visit edges form a spanning tree and further
// edges of the induced data subgraph are added (seeded shuffle) until
// n_edges (or all induced edges).  Query vertex ids follow first visit.
std::string walk_query(const Graph& d, uint32_t k, uint32_t n_edges, uint64_t seed,
                       std::vector<uint32_t>& qedges, std::vector<uint32_t>& witness);

}  // namespace cuts
