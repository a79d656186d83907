// Edge-list I/O and CSR construction.
//
// Reader semantics follow src/util.cpp:19-52 (first line is V; every other
// line holding a TAB is an arc "s<TAB>t", each side parsed the way std::stoi
// parses it; lines without a TAB are skipped; duplicate arcs collapse) and
// the CSR/signature layout follows src/graph.cpp:18-61.  The reference
// builds std::set adjacency; here the file is parsed in parallel chunks and
// the arcs are radix-sorted and deduplicated, which yields the same sorted
// duplicate-free lists.  Unlike the reference, ids >= V, negative ids and
// unparsable numbers are reported as errors instead of being undefined.
#include "host.hpp"

#include <omp.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace cuts {
namespace {

// Parses one integer the way std::stoi does on [p, e): optional leading
// whitespace, optional sign, at least one digit; trailing text is ignored.
// Returns false where std::stoi would throw.
bool stoi_like(const char* p, const char* e, long long& out) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f'))
    ++p;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) { neg = (*p == '-'); ++p; }
  if (p >= e || *p < '0' || *p > '9') return false;
  long long v = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    v = v * 10 + (*p - '0');
    if (v > (long long)INT_MAX + 1) return false;  // std::stoi: out_of_range
    ++p;
  }
  v = neg ? -v : v;
  if (v > INT_MAX || v < INT_MIN) return false;
  out = v;
  return true;
}

// LSD radix sort of 64-bit keys over the given bit ranges (low range first).
void radix_sort(std::vector<uint64_t>& a, const std::vector<std::pair<int, int>>& ranges) {
  const size_t n = a.size();
  if (n < 2) return;
  std::vector<uint64_t> tmp(n);
  const int B = 11, NB = 1 << B;
  int nt = omp_get_max_threads();
  if (n < (1u << 16)) nt = 1;
  std::vector<size_t> hist((size_t)nt * NB);
  for (auto [lo, width] : ranges) {
    for (int sh = lo; sh < lo + width; sh += B) {
      const int w = std::min(B, lo + width - sh);
      const uint64_t mask = (1ull << w) - 1;
      std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
      {
        int t = omp_get_thread_num();
        size_t b = n * t / nt, e = n * (t + 1) / nt;
        size_t* h = &hist[(size_t)t * NB];
        for (size_t i = b; i < e; ++i) h[(a[i] >> sh) & mask]++;
#pragma omp barrier
#pragma omp single
        {
          size_t run = 0;
          for (int d = 0; d < NB; ++d)
            for (int u = 0; u < nt; ++u) {
              size_t c = hist[(size_t)u * NB + d];
              hist[(size_t)u * NB + d] = run;
              run += c;
            }
        }
        for (size_t i = b; i < e; ++i) tmp[h[(a[i] >> sh) & mask]++] = a[i];
      }
      a.swap(tmp);
    }
  }
}

int bits_for(uint32_t V) {
  int b = 1;
  while (b < 32 && (1ull << b) < V) ++b;
  return b;
}

void csr_from_sorted(uint32_t V, const std::vector<uint64_t>& keys, bool swap,
                     std::vector<uint32_t>& off, std::vector<uint32_t>& adj) {
  off.assign((size_t)V + 1, 0);
  adj.resize(keys.size());
  for (size_t i = 0; i < keys.size(); ++i) {
    uint32_t hi = (uint32_t)(keys[i] >> 32), lo = (uint32_t)keys[i];
    uint32_t s = swap ? lo : hi, t = swap ? hi : lo;
    off[s + 1]++;
    adj[i] = t;
  }
  for (uint32_t v = 0; v < V; ++v) off[v + 1] += off[v];
}

}  // namespace

std::string build_graph(uint32_t V, std::vector<uint64_t>&& arcs, bool symmetrize, Graph& g) {
  for (uint64_t a : arcs)
    if ((a >> 32) >= V || (uint32_t)a >= V) return "vertex id out of range";
  if (symmetrize) {
    size_t n = arcs.size();
    arcs.resize(2 * n);
#pragma omp parallel for
    for (size_t i = 0; i < n; ++i) arcs[n + i] = (arcs[i] << 32) | (arcs[i] >> 32);
  }
  const int bw = bits_for(V);
  // out-lists: sort by (s, t)
  radix_sort(arcs, {{0, bw}, {32, bw}});
  arcs.erase(std::unique(arcs.begin(), arcs.end()), arcs.end());
  g.V = V;
  csr_from_sorted(V, arcs, false, g.out_off, g.out_adj);
  // in-lists: sort the swapped keys by (t, s)
#pragma omp parallel for
  for (size_t i = 0; i < arcs.size(); ++i) arcs[i] = (arcs[i] << 32) | (arcs[i] >> 32);
  radix_sort(arcs, {{0, bw}, {32, bw}});
  csr_from_sorted(V, arcs, false, g.in_off, g.in_adj);
  g.sig.resize(2 * (size_t)V);
#pragma omp parallel for
  for (long long v = 0; v < (long long)V; ++v) {
    g.sig[2 * v] = g.in_off[v + 1] - g.in_off[v];
    g.sig[2 * v + 1] = g.out_off[v + 1] - g.out_off[v];
  }
  g.symmetric = (g.out_off == g.in_off) && (g.out_adj == g.in_adj);
  return "";
}

std::string load_graph(const std::string& path, Graph& g) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return "load graph file failed";
  std::vector<char> buf;
  {
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) { fclose(f); return "load graph file failed"; }
    buf.resize((size_t)sz);
    if (sz && fread(buf.data(), 1, (size_t)sz, f) != (size_t)sz) { fclose(f); return "load graph file failed"; }
    fclose(f);
  }
  const char* p = buf.data();
  const char* end = p + buf.size();
  // (memchr needs a non-null pointer even for an empty file)
  const char* nl = buf.empty() ? nullptr : (const char*)memchr(p, '\n', buf.size());
  const char* l0e = nl ? nl : end;
  long long Vll;
  if (!stoi_like(p, l0e, Vll) || Vll < 0) return "invalid vertex count on the first line of " + path;
  const uint32_t V = (uint32_t)Vll;
  const char* body = nl ? nl + 1 : end;
  const size_t blen = end - body;
  int nt = omp_get_max_threads();
  if (blen < (1u << 20)) nt = 1;
  std::vector<std::vector<uint64_t>> parts(nt);
  std::vector<std::string> errs(nt);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    // chunk [b, e) aligned to line starts
    const char* b = body + blen * t / nt;
    const char* e = body + blen * (t + 1) / nt;
    if (t > 0) {
      const char* q = b;
      if (q[-1] != '\n') {
        const char* x = (const char*)memchr(q, '\n', end - q);
        b = x ? x + 1 : end;
      }
    }
    if (t + 1 < nt) {
      if (e[-1] != '\n') {
        const char* x = (const char*)memchr(e, '\n', end - e);
        e = x ? x + 1 : end;
      }
    } else {
      e = end;
    }
    auto& out = parts[t];
    const char* q = b;
    while (q < e) {
      const char* le = (const char*)memchr(q, '\n', e - q);
      if (!le) le = e;
      const char* tab = (const char*)memchr(q, '\t', le - q);
      if (tab) {
        long long s, d;
        if (!stoi_like(q, tab, s) || !stoi_like(tab + 1, le, d)) {
          if (errs[t].empty()) errs[t] = "invalid arc line '" + std::string(q, le) + "'";
        } else if (s < 0 || d < 0 || s >= (long long)V || d >= (long long)V) {
          if (errs[t].empty()) errs[t] = "vertex id out of range in line '" + std::string(q, le) + "'";
        } else {
          out.push_back(((uint64_t)s << 32) | (uint64_t)d);
        }
      }
      q = le + 1;
    }
  }
  for (auto& e : errs)
    if (!e.empty()) return e;
  size_t total = 0;
  for (auto& v : parts) total += v.size();
  std::vector<uint64_t> arcs;
  arcs.reserve(total);
  for (auto& v : parts) {
    arcs.insert(arcs.end(), v.begin(), v.end());
    std::vector<uint64_t>().swap(v);
  }
  return build_graph(V, std::move(arcs), false, g);
}

std::string save_graph(const std::string& path, const Graph& g) {
  std::ofstream out(path, std::ios::binary);
  if (!out) return "cannot open " + path + " for writing";
  std::string s = std::to_string(g.V) + "\n";
  out << s;
  std::string line;
  line.reserve(1 << 20);
  for (uint32_t v = 0; v < g.V; ++v) {
    for (uint32_t i = g.out_off[v]; i < g.out_off[v + 1]; ++i) {
      line += std::to_string(v);
      line += '\t';
      line += std::to_string(g.out_adj[i]);
      line += '\n';
    }
    if (line.size() > (1u << 20)) { out << line; line.clear(); }
  }
  out << line;
  return out.good() ? "" : "write failed for " + path;
}

}  // namespace cuts

namespace cuts {

void build_two_hop(const Graph& g, std::vector<uint64_t>& off, std::vector<uint32_t>& hv, std::vector<uint32_t>& hc) {
  const uint32_t V = g.V;
  off.assign((size_t)V + 1, 0);
  // every thread keeps one V-word array: at most 4 GiB of them in total
  const int nt = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)omp_get_max_threads(),
                                                               (4ull << 30) / (4ull * std::max<uint32_t>(V, 1))));
  // pass 1: distinct two-hop vertices per u (a stamp per vertex, u + 1)
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> mark(V, 0);
#pragma omp for schedule(dynamic, 64)
    for (int64_t u = 0; u < (int64_t)V; ++u) {
      uint64_t cnt = 0;
      for (uint32_t i = g.out_off[u]; i < g.out_off[u + 1]; ++i) {
        const uint32_t x = g.out_adj[i];
        for (uint32_t j = g.out_off[x]; j < g.out_off[x + 1]; ++j) {
          const uint32_t v = g.out_adj[j];
          if (v != (uint32_t)u && mark[v] != (uint32_t)u + 1) {
            mark[v] = (uint32_t)u + 1;
            ++cnt;
          }
        }
      }
      off[u + 1] = cnt;
    }
  }
  for (uint32_t u = 0; u < V; ++u) off[u + 1] += off[u];
  hv.assign(std::max<uint64_t>(off[V], 1), 0);
  hc.assign(std::max<uint64_t>(off[V], 1), 0);
  // pass 2: the vertices, ascending, with their common-neighbour counts
#pragma omp parallel num_threads(nt)
  {
    std::vector<uint32_t> cnt(V, 0), touched;
#pragma omp for schedule(dynamic, 64)
    for (int64_t u = 0; u < (int64_t)V; ++u) {
      touched.clear();
      for (uint32_t i = g.out_off[u]; i < g.out_off[u + 1]; ++i) {
        const uint32_t x = g.out_adj[i];
        for (uint32_t j = g.out_off[x]; j < g.out_off[x + 1]; ++j) {
          const uint32_t v = g.out_adj[j];
          if (v != (uint32_t)u && cnt[v]++ == 0) touched.push_back(v);
        }
      }
      std::sort(touched.begin(), touched.end());
      uint64_t o = off[u];
      for (uint32_t v : touched) {
        hv[o] = v;
        hc[o] = cnt[v];
        cnt[v] = 0;
        ++o;
      }
    }
  }
}

}  // namespace cuts
