"""The synthetic workloads the matcher is measured on.

Data graphs are rebuilt deterministically from seeds by the native
generators; query shapes are fixed, and each is located once in its data
graph by a seeded random-walk search (the witness proves count >= 1).

Query vertex numbering follows the walk: shape vertices are numbered in the
order the walk reaches them.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from typing import Dict, List, Tuple

import numpy as np

from . import FLAG_NO_CHAIN, FLAG_NO_PAIR, Graph

Edge = Tuple[int, int]


@dataclass(frozen=True)
class Shape:
    name: str
    k: int
    edges: Tuple[Edge, ...]


# triangle 0-1-2 plus a 2-edge tail 2-3-4 (5 vertices, 5 edges)
TRI_TAIL2 = Shape("triangle+2-edge-tail", 5, ((0, 1), (1, 2), (2, 0), (2, 3), (3, 4)))
# 4-cycle 0-1-2-3 with pendants 4 (on 0) and 5 (on 2)
C4_2PEND = Shape("4-cycle+2-pendants", 6, ((0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (2, 5)))
# K4 minus edge 0-3 (diamond) with a 2-edge tail 3-4-5
K4E_TAIL2 = Shape("K4-e+2-tail", 6, ((0, 1), (0, 2), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5)))
# 5-cycle 0-1-2-3-4 with chord 0-2, pendants 5 (on 3) and 6 (on 4)
C5CHORD_2TAIL = Shape("5-cycle+chord+2-tails", 7,
                      ((0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2), (3, 5), (4, 6)))


def shape_arcs(s: Shape) -> List[Edge]:
    out = []
    for a, b in s.edges:
        out += [(a, b), (b, a)]
    return out


def query_graph(s: Shape) -> Graph:
    return Graph.from_edges(s.k, list(s.edges))


@lru_cache(maxsize=None)
def config1_graph() -> Graph:
    """Chung-Lu, n=4,096, m=32,768, exponent 2.3, seed 0."""
    return Graph.chung_lu(4096, 32768, 2.3, 0, 0)


@lru_cache(maxsize=None)
def config2_graph() -> Graph:
    """Chung-Lu, n=36,692, m=183,831, exponent 2.1, degree capped at 1,400, seed 0."""
    return Graph.chung_lu(36692, 183831, 2.1, 1400, 0)


@lru_cache(maxsize=None)
def config3_graph() -> Graph:
    """R-MAT scale 20, edge factor 16, (a,b,c)=(0.57,0.19,0.19), seed 0,
    deduplicated, loops removed, symmetrised."""
    return Graph.rmat(20, 16, 0.57, 0.19, 0.19, 0)


@lru_cache(maxsize=None)
def rmat18_graph(seed: int) -> Graph:
    return Graph.rmat(18, 16, 0.57, 0.19, 0.19, seed)


CONFIGS = {
    1: ("chung-lu-4k", config1_graph, [TRI_TAIL2]),
    2: ("chung-lu-36.7k", config2_graph, [C4_2PEND, K4E_TAIL2]),
    3: ("rmat-20", config3_graph, [C5CHORD_2TAIL]),
}


def witness(cfg: int, shape: Shape, seed: int = 0) -> np.ndarray:
    g = CONFIGS[cfg][1]()
    return g.find_pattern(shape.k, list(shape.edges), seed)


def csr_fingerprint(g: Graph) -> str:
    """FNV-1a-style 64-bit fold over (a deterministic sample of at most 64Ki)
    8-byte words of out_off || out_adj, plus the byte length."""
    v = g.view()
    h = 0xCBF29CE484222325
    data = np.concatenate([v["out_off"], v["out_adj"]]).astype("<u4").tobytes()
    # vectorised FNV is awkward; fold 8-byte words instead of bytes
    words = np.frombuffer(data + b"\0" * (-len(data) % 8), dtype="<u8")
    for w in words[:: max(1, len(words) // 65536)]:  # bounded sample for large graphs
        h ^= int(w)
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}:{len(data)}"


# ---------------------------------------------------------------- batches
def batch_candidates(g: Graph, seed: int = 0):
    """Endless seeded stream of random-walk queries: k uniform in [4, 12],
    edge/vertex ratio uniform in [1.0, 2.5] (edges clamped to
    [k-1, k(k-1)/2]; a walk's induced subgraph may hold fewer)."""
    rng = np.random.default_rng(seed)
    i = 0
    while True:
        k = int(rng.integers(4, 13))
        ratio = float(rng.uniform(1.0, 2.5))
        ne = max(k - 1, min(int(round(ratio * k)), k * (k - 1) // 2))
        edges, _ = g.walk_query(k, ne, seed * 1_000_003 + i)
        yield i, k, ne, [tuple(map(int, e)) for e in edges.tolist()]
        i += 1


# Admission rule for fixed query sets.  Walk queries of 5+ vertices on the
# config-2 graph have counts far beyond what any run finishes (12-vertex ones
# beyond 2^64), so each query is matched on a sample of the data vertices
# that order position 0 may take: the residue class v = s-1 (mod s), with s
# a power of two.  These classes are nested (s' = 2s keeps a subset), so work
# and table sizes fall monotonically with s, and the rule picks the smallest
# s whose score-order run finishes within MAX_WORK candidates and
# materialises at most MAX_ROWS trie rows (levels 1..k-3; the last two are
# fused), with a non-zero count.  Counts are exact for that sample.  Sizes
# stay uniform: each k in [KMIN, KMAX] gets an equal quota, filled by the
# first candidates of that size.  Everything is deterministic (work and table
# sizes do not depend on timing).
MAX_WORK = 1 << 28
MAX_ROWS = 1 << 21
MAX_STRIDE_LOG = 15
KMIN, KMAX = 4, 12


def materialised_rows(level_sizes) -> int:
    k = len(level_sizes)
    return int(sum(level_sizes[1:max(1, k - 2)]))


def root_sample(V: int, stride: int) -> Tuple[int, int, int]:
    """(first, end, stride) of the residue class v = stride-1 (mod stride)."""
    return (stride - 1, V, stride)


def quotas(n: int, kmin: int = KMIN, kmax: int = KMAX) -> Dict[int, int]:
    """n split over the sizes kmin..kmax as evenly as possible (the
    remainder goes to the smallest sizes)."""
    m = kmax - kmin + 1
    return {k: n // m + (1 if k - kmin < n % m else 0) for k in range(kmin, kmax + 1)}


# execution mode of the admission rule (see admit)
ADMIT_FLAGS = FLAG_NO_PAIR | FLAG_NO_CHAIN


def admit(ctx, dg, g: Graph, n: int, seed: int = 0, max_work: int = MAX_WORK, max_rows: int = MAX_ROWS,
          kmin: int = KMIN, kmax: int = KMAX, max_tries_per_k: int = 3000, log=None):
    """Per size k, the first quotas(n)[k] candidates of size k from
    batch_candidates(g, seed) that pass the rule at some root sample, in
    stream order, each with its densest passing sample.  A size whose quota
    is not met within max_tries_per_k candidates keeps what it found.
    Returns (queries, tried per size)."""
    from . import Plan
    want = quotas(n, kmin, kmax)
    got = {k: 0 for k in want}
    tried = {k: 0 for k in want}
    out = []

    def run(p, j):
        # work (candidates examined) depends on the execution mode; the rule
        # is applied with the last two levels expanded level by level, so the
        # admitted set does not move when faster modes are added
        r = ctx.match(dg, p, roots=root_sample(g.V, 1 << j), max_work=max_work, flags=ADMIT_FLAGS)
        return (not r.work_capped and materialised_rows(r.level_sizes) <= max_rows), r

    for i, k, ne, edges in batch_candidates(g, seed):
        if all(got[x] >= want[x] or tried[x] >= max_tries_per_k for x in want):
            break
        if k not in want or got[k] >= want[k] or tried[k] >= max_tries_per_k:
            continue
        tried[k] += 1
        if log and sum(tried.values()) % 200 == 0:
            log(f"tried {sum(tried.values())} candidates, admitted {len(out)}/{n}")
        p = Plan(Graph.from_edges(k, edges))
        ok, r = run(p, MAX_STRIDE_LOG)
        if not ok:
            continue
        lo, hi, best = 0, MAX_STRIDE_LOG, r
        while lo < hi:  # smallest passing stride exponent
            mid = (lo + hi) // 2
            ok, r = run(p, mid)
            if ok:
                hi, best = mid, r
            else:
                lo = mid + 1
        if best.count == 0:
            continue
        got[k] += 1
        out.append(dict(index=i, k=k, edges=edges, roots=list(root_sample(g.V, 1 << hi)), count=best.count,
                        level_sizes=best.level_sizes, work=best.work, order=p.order))
        if log:
            log(f"admitted {len(out)}/{n} (candidate {i}, k={k}: {got[k]}/{want[k]} after {tried[k]} tries, "
                f"stride {1 << hi}, count={best.count})")
    return out, tried
