// Deterministic synthetic inputs: Chung-Lu and R-MAT data graphs, and
// queries located in a data graph by seeded random walks.  None of this has
// a counterpart in the reference (which reads SNAP files); the generators
// exist so that every workload can be rebuilt bit-for-bit from a seed.
#include "host.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <unordered_set>

namespace cuts {

static inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (auto& w : s) w = splitmix64(x);
}
uint64_t Rng::next() {
  const uint64_t r = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return r;
}
double Rng::uniform() { return (next() >> 11) * 0x1.0p-53; }
uint64_t Rng::below(uint64_t n) {  // Lemire-free simple rejection for exactness
  if (n <= 1) return 0;
  uint64_t lim = UINT64_MAX - UINT64_MAX % n;
  uint64_t r;
  do r = next(); while (r >= lim);
  return r % n;
}

std::string gen_chung_lu(uint32_t n, uint64_t m, double gamma, uint32_t max_deg, uint64_t seed,
                         std::vector<uint64_t>& edges) {
  if (n < 2 || gamma <= 1.0) return "invalid Chung-Lu parameters";
  if (m > (uint64_t)n * (n - 1) / 2) return "too many edges requested";
  std::vector<double> w(n);
  const double ex = -1.0 / (gamma - 1.0);
  for (uint32_t i = 0; i < n; ++i) w[i] = std::pow((double)(i + 1), ex);
  if (max_deg > 0) {
    // cap so that the expected degree 2m*w_i/sum(w) stays <= max_deg
    for (int it = 0; it < 100; ++it) {
      double sum = 0;
      for (double x : w) sum += x;
      const double cap = max_deg * sum / (2.0 * (double)m);
      bool changed = false;
      for (double& x : w)
        if (x > cap * (1 + 1e-12)) { x = cap; changed = true; }
      if (!changed) break;
    }
  }
  std::vector<double> cdf(n);
  double run = 0;
  for (uint32_t i = 0; i < n; ++i) { run += w[i]; cdf[i] = run; }
  Rng rng(seed);
  auto draw = [&]() -> uint32_t {
    double u = rng.uniform() * run;
    uint32_t i = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    return i < n ? i : n - 1;
  };
  std::unordered_set<uint64_t> seen;
  seen.reserve(m * 2);
  std::vector<uint32_t> deg(n, 0);
  edges.clear();
  edges.reserve(m);
  uint64_t attempts = 0;
  while (edges.size() < m) {
    if (++attempts > 1000 * m + 1000000) return "Chung-Lu sampling did not converge";
    uint32_t a = draw(), b = draw();
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    if (max_deg && (deg[a] >= max_deg || deg[b] >= max_deg)) continue;
    uint64_t key = ((uint64_t)a << 32) | b;
    if (!seen.insert(key).second) continue;
    deg[a]++;
    deg[b]++;
    edges.push_back(key);
  }
  return "";
}

std::string gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                     std::vector<uint64_t>& edges) {
  if (scale < 1 || scale > 31) return "invalid R-MAT scale";
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1) return "invalid R-MAT probabilities";
  const uint64_t M = (uint64_t)edge_factor << scale;
  const uint64_t BLK = 1u << 16;
  const uint64_t nblk = (M + BLK - 1) / BLK;
  std::vector<uint64_t> raw(M);
  std::vector<char> keep(M);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (long long blk = 0; blk < (long long)nblk; ++blk) {
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(blk + 1));
    Rng rng(splitmix64(x));
    uint64_t e0 = blk * BLK, e1 = std::min(M, e0 + BLK);
    for (uint64_t e = e0; e < e1; ++e) {
      uint32_t u = 0, v = 0;
      for (uint32_t bit = 0; bit < scale; ++bit) {
        double r = rng.uniform();
        uint32_t bu = 0, bv = 0;
        if (r < a) {
        } else if (r < ab) {
          bv = 1;
        } else if (r < abc) {
          bu = 1;
        } else {
          bu = 1;
          bv = 1;
        }
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      raw[e] = ((uint64_t)u << 32) | v;
      keep[e] = (u != v);
    }
  }
  edges.clear();
  edges.reserve(M);
  for (uint64_t e = 0; e < M; ++e)
    if (keep[e]) edges.push_back(raw[e]);
  return "";
}

namespace {
bool has_arc(const Graph& d, uint32_t u, uint32_t v) {
  const uint32_t* b = d.out_adj.data() + d.out_off[u];
  const uint32_t* e = d.out_adj.data() + d.out_off[u + 1];
  return std::binary_search(b, e, v);
}
}  // namespace

std::string find_pattern(const Graph& d, uint32_t k, const std::vector<uint32_t>& pe, uint64_t seed,
                         uint64_t max_tries, std::vector<uint32_t>& witness) {
  if (k == 0 || pe.size() % 2) return "invalid pattern";
  std::vector<std::vector<uint32_t>> adj(k);
  for (size_t i = 0; i < pe.size(); i += 2) {
    if (pe[i] >= k || pe[i + 1] >= k || pe[i] == pe[i + 1]) return "invalid pattern edge";
    adj[pe[i]].push_back(pe[i + 1]);
    adj[pe[i + 1]].push_back(pe[i]);
  }
  // BFS order of the pattern from vertex 0
  std::vector<uint32_t> ord{0};
  std::vector<char> in(k, 0);
  in[0] = 1;
  for (size_t h = 0; h < ord.size(); ++h)
    for (uint32_t w : adj[ord[h]])
      if (!in[w]) { in[w] = 1; ord.push_back(w); }
  if (ord.size() != k) return "pattern is not connected";
  std::vector<uint32_t> nz;
  for (uint32_t v = 0; v < d.V; ++v)
    if (d.out_off[v + 1] > d.out_off[v]) nz.push_back(v);
  if (nz.empty()) return "data graph has no edges";
  Rng rng(seed);
  std::vector<int64_t> img(k);
  for (uint64_t t = 0; t < max_tries; ++t) {
    std::fill(img.begin(), img.end(), -1);
    img[ord[0]] = nz[rng.below(nz.size())];
    bool ok = true;
    for (uint32_t h = 1; h < k && ok; ++h) {
      const uint32_t pv = ord[h];
      // earlier-matched pattern neighbours; walk from the first one
      std::vector<uint32_t> back;
      for (uint32_t w : adj[pv])
        if (img[w] >= 0) back.push_back((uint32_t)img[w]);
      const uint32_t from = back[0];
      std::vector<uint32_t> cand;
      for (uint32_t i = d.out_off[from]; i < d.out_off[from + 1]; ++i) {
        uint32_t x = d.out_adj[i];
        bool good = true;
        for (uint32_t y : back)
          if (y == x || !has_arc(d, y, x)) { good = false; break; }
        for (uint32_t j = 0; j < k && good; ++j)
          if (img[j] == (int64_t)x) good = false;
        if (good) cand.push_back(x);
      }
      if (cand.empty()) ok = false;
      else img[pv] = cand[rng.below(cand.size())];
    }
    if (ok) {
      witness.assign(k, 0);
      for (uint32_t i = 0; i < k; ++i) witness[i] = (uint32_t)img[i];
      return "";
    }
  }
  return "pattern not found within the try budget";
}

std::string walk_query(const Graph& d, uint32_t k, uint32_t n_edges, uint64_t seed,
                       std::vector<uint32_t>& qedges, std::vector<uint32_t>& witness) {
  if (k < 2) return "walk query needs k >= 2";
  std::vector<uint32_t> nz;
  for (uint32_t v = 0; v < d.V; ++v)
    if (d.out_off[v + 1] > d.out_off[v]) nz.push_back(v);
  if (nz.empty()) return "data graph has no edges";
  Rng rng(seed);
  for (int attempt = 0; attempt < 1000; ++attempt) {
    std::vector<uint32_t> S{nz[rng.below(nz.size())]};
    std::vector<uint32_t> tree;  // pairs of query ids
    uint32_t cur = 0;            // query id of the walk position
    for (uint32_t step = 0; step < 64 * k && S.size() < k; ++step) {
      const uint32_t u = S[cur];
      const uint32_t deg = d.out_off[u + 1] - d.out_off[u];
      if (!deg) break;
      const uint32_t w = d.out_adj[d.out_off[u] + rng.below(deg)];
      auto it = std::find(S.begin(), S.end(), w);
      if (it == S.end()) {
        tree.push_back(cur);
        tree.push_back((uint32_t)S.size());
        S.push_back(w);
        cur = (uint32_t)S.size() - 1;
      } else {
        cur = (uint32_t)(it - S.begin());
      }
    }
    if (S.size() < k) continue;
    std::vector<uint32_t> extra;
    for (uint32_t i = 0; i < k; ++i)
      for (uint32_t j = i + 1; j < k; ++j) {
        bool is_tree = false;
        for (size_t t = 0; t < tree.size(); t += 2)
          if ((tree[t] == i && tree[t + 1] == j) || (tree[t] == j && tree[t + 1] == i)) is_tree = true;
        if (!is_tree && has_arc(d, S[i], S[j])) { extra.push_back(i); extra.push_back(j); }
      }
    // seeded Fisher-Yates over the extra edge pairs
    const size_t ne = extra.size() / 2;
    for (size_t i = ne; i > 1; --i) {
      size_t j = rng.below(i);
      std::swap(extra[2 * (i - 1)], extra[2 * j]);
      std::swap(extra[2 * (i - 1) + 1], extra[2 * j + 1]);
    }
    qedges = tree;
    const size_t want = n_edges > k - 1 ? n_edges - (k - 1) : 0;
    for (size_t i = 0; i < std::min(want, ne); ++i) {
      qedges.push_back(extra[2 * i]);
      qedges.push_back(extra[2 * i + 1]);
    }
    witness = S;
    return "";
  }
  return "random walk did not reach k distinct vertices";
}

}  // namespace cuts
