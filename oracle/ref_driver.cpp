// TEST INFRASTRUCTURE ONLY.  Links the reference's own host C++
// (src/common.cpp, src/graph.cpp, src/score.cpp, src/util.cpp, compiled
// from where they lie by oracle/Makefile) and exposes the pieces the tests
// compare against: the matching order and P/C constraints computed by
// Graph::sort_search_order, the CSR/signature arrays built by
// Graph::Graph, and the ans.txt writer write_match_to_disk.
#include <cstring>

#include "graph.h"

extern "C" {

// Query plan of Graph(0, path).  Arrays must hold V (order), V+1 (offsets)
// and V*V (parents/children) entries.
int ref_query_plan(const char *path, unsigned *V, unsigned *order, unsigned *par_off, unsigned *par,
                   unsigned *chi_off, unsigned *chi) {
  Graph q(0, path);
  *V = q.V;
  std::memcpy(order, q.order_sequence, q.V * sizeof(unsigned));
  std::memcpy(par_off, q.parents_offset, (q.V + 1) * sizeof(unsigned));
  std::memcpy(chi_off, q.children_offset, (q.V + 1) * sizeof(unsigned));
  std::memcpy(par, q.parents, q.parents_offset[q.V] * sizeof(unsigned));
  std::memcpy(chi, q.children, q.children_offset[q.V] * sizeof(unsigned));
  return 0;
}

// Sizes of Graph(1, path): V and E (arcs).
int ref_graph_size(const char *path, unsigned *V, unsigned *E) {
  Graph d(1, path);
  *V = d.V;
  *E = d.E;
  return 0;
}

// CSR and signatures of Graph(1, path) into caller arrays sized from
// ref_graph_size.
int ref_graph_csr(const char *path, unsigned *off, unsigned *adj, unsigned *roff, unsigned *radj, unsigned *sig) {
  Graph d(1, path);
  std::memcpy(off, d.neighbors_offset, (d.V + 1) * sizeof(unsigned));
  std::memcpy(roff, d.r_neighbors_offset, (d.V + 1) * sizeof(unsigned));
  std::memcpy(adj, d.neighbors, d.neighbors_offset[d.V] * sizeof(unsigned));
  std::memcpy(radj, d.r_neighbors, d.r_neighbors_offset[d.V] * sizeof(unsigned));
  std::memcpy(sig, d.signatures, 2 * d.V * sizeof(unsigned));
  return 0;
}

// write_match_to_disk(count, row_ptrs, V, order, results): writes ans.txt
// in the current directory.
void ref_write_matches(unsigned long long count, unsigned long long *row_ptrs, unsigned V, unsigned *order,
                       unsigned *results) {
  write_match_to_disk(count, row_ptrs, V, order, results);
}
}
