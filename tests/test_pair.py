"""The last two query positions counted together.

Pair level: when positions k-2 and k-1 are not adjacent in the query and
each has one backward neighbour older than position k-3 (pendants,
typically), a prefix of positions 0..k-3 has |A| * |B| - |A & B|
completions, A and B being the degree-filtered neighbour lists of those
backward neighbours minus the prefix; the expansion of position k-3 adds
that number per survivor and levels k-2 and k-1 are never expanded.

Chain level: when k-1 hangs on k-2 alone and k-2 on one position u older
than k-3 alone (a 2-edge tail), a prefix of positions 0..k-3 has
sum_{x in A} |N(x) - prefix - x| completions, computed from per-vertex
tables and one intersection N(u) & N(r_i) per other prefix vertex (those of
older positions shared by the rows of one parent prefix).

The intersections come from the two-hop index built at upload (small
graphs) or from flattened list walks (FLAG_NO_TWO_HOP, and graphs too large
for the index).  Counts must equal the oracle's and the same run without
these levels; level sizes and merge-model bytes must equal the CPU trie
walk's, which materialises every level.
"""
import random

import numpy as np
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

SHAPES = {
    "star3": [(0, 1), (0, 2), (0, 3)],
    "star4": [(0, 1), (0, 2), (0, 3), (0, 4)],
    "triangle+2pend-same": [(0, 1), (1, 2), (2, 0), (0, 3), (0, 4)],
    "triangle+2pend-two": [(0, 1), (1, 2), (2, 0), (0, 3), (1, 4)],
    "4cycle+2pend": list(W.C4_2PEND.edges),
    "diamond+2pend": [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3), (0, 4), (3, 5)],
    "k4+2pend": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (1, 4), (2, 5)],
    "5cycle+3pend": [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 5), (2, 6), (3, 7)],
    # chain level: a 2-edge tail on a vertex older than position k-3
    "triangle+tail2": list(W.TRI_TAIL2.edges),
    "k4e+tail2": list(W.K4E_TAIL2.edges),
    "path5": [(0, 1), (1, 2), (2, 3), (3, 4)],
    "4cycle+tail2": [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (4, 5)],
    "star3+tail2": [(0, 1), (0, 2), (0, 3), (3, 4), (4, 5)],
    "k4+tail2": [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5)],
    # neither: the second pendant hangs on position k-3, or the tail is longer
    "5cycle+chord+2tails": list(W.C5CHORD_2TAIL.edges),
    "triangle+tail3": [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)],
}


def backward(p):
    """Backward positions of every order position (undirected query)."""
    po, pa = list(p.parents_offset), list(p.parents)
    co, ch = list(p.children_offset), list(p.children)
    return [set(pa[po[i]:po[i + 1]]) | set(ch[co[i]:co[i + 1]]) for i in range(p.k)]


def pair_eligible(p):
    k = p.k
    if k < 4:
        return False
    b = backward(p)
    if len(b[k - 1]) != 1 or len(b[k - 2]) != 1:
        return False
    (pa,), (pb,) = b[k - 2], b[k - 1]
    return pb != k - 2 and pa <= k - 4 and pb <= k - 4


def collapse_eligible(p):
    """Pair level whose position k-3 has exactly the pair's two list owners
    as constraints, all thresholds <= 2: counted per prefix of level k-4
    from the two-hop index, with no expansion at all."""
    if not pair_eligible(p):
        return False
    k = p.k
    b = backward(p)
    deg = [len(b[i]) + sum(1 for j in range(i + 1, k) if i in b[j]) for i in range(k)]
    owners = b[k - 2] | b[k - 1]
    return len(owners) == 2 and b[k - 3] == owners and max(deg[k - 3], deg[k - 2], deg[k - 1]) <= 2


def chain_eligible(p):
    k = p.k
    if k < 4:
        return False
    b = backward(p)
    if b[k - 1] != {k - 2} or len(b[k - 2]) != 1:
        return False
    (u,) = b[k - 2]
    return u <= k - 4


def graph_for(seed):
    rng = random.Random(seed)
    n = rng.randint(50, 140)
    return cuts_amd.Graph.chung_lu(n, rng.randint(2 * n, 4 * n), rng.uniform(2.2, 2.8), 0, seed)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SHAPES))
@pytest.mark.parametrize("seed", range(3))
def test_pair_level_vs_oracle_and_cpu_trie(ctx, name, seed):
    edges = SHAPES[name]
    k = 1 + max(max(e) for e in edges)
    g = graph_for(7000 + 13 * seed + k)
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(k, oracle.undirected(edges))
    want = oracle.count(og, qa)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    cpu = oracle.trie_cpu(og, qa, p.order, nthreads=4)
    assert cpu.count == want
    dg = ctx.upload(g)
    prof = cuts_amd.FLAG_PROFILE
    for flags in (0, cuts_amd.FLAG_NO_BITMAP, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_TWO_HOP | cuts_amd.FLAG_NO_BITMAP,
                  cuts_amd.FLAG_NO_PAIR, cuts_amd.FLAG_NO_CHAIN, cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN,
                  cuts_amd.FLAG_NO_SHARED, cuts_amd.FLAG_NO_SHARED | cuts_amd.FLAG_NO_BITMAP,
                  cuts_amd.FLAG_NO_TWO_HOP | cuts_amd.FLAG_NO_ARC_COMMON):
        r = ctx.match(dg, p, flags=flags | prof)
        assert r.count == want, (name, flags)
        assert r.level_sizes == [cpu.level[i] for i in range(k)], (name, flags)
        assert sum(r.merge_bytes.values()) == cpu.bytes, (name, flags, r.merge_bytes)
        used = r.kernels["pair"][1] + r.kernels["chain"][1] > 0
        expect = (pair_eligible(p) and not flags & cuts_amd.FLAG_NO_PAIR) or \
            (chain_eligible(p) and not flags & cuts_amd.FLAG_NO_CHAIN)
        assert not used or expect, (name, flags, p.order)
        collapsed = collapse_eligible(p) and not flags & (cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_TWO_HOP)
        assert used or not expect or collapsed or r.level_sizes[k - 3] == 0, (name, flags, p.order)
    for m in (dict(trunks=5), dict(budget=2048), dict(flags=cuts_amd.FLAG_NO_XCD),
              dict(budget=2048, flags=cuts_amd.FLAG_NO_TWO_HOP)):
        assert ctx.match(dg, p, **m).count == want, (name, m)


def test_shapes_cover_every_case():
    """The shape table holds pair shapes with one and with two distinct list
    owners (some of the latter collapsing position k-3), chain shapes with
    k = 5 and k >= 6, and shapes that are neither (plans only, no GPU work)."""
    kinds = set()
    for name, edges in SHAPES.items():
        k = 1 + max(max(e) for e in edges)
        p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
        assert not (pair_eligible(p) and chain_eligible(p))
        if pair_eligible(p):
            b = backward(p)
            kinds.add("pair-same" if b[k - 1] == b[k - 2] else "pair-two")
            if collapse_eligible(p):
                kinds.add("pair-collapse")
        elif chain_eligible(p):
            kinds.add("chain-5" if k == 5 else "chain-6+")
        else:
            kinds.add("neither")
    assert kinds == {"pair-same", "pair-two", "pair-collapse", "chain-5", "chain-6+", "neither"}


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(12))
def test_pair_level_random_queries(ctx, seed):
    """Random-walk queries on random graphs: eligible or not, every mode
    counts the oracle's number."""
    rng = random.Random(9100 + seed)
    g = graph_for(9100 + seed)
    k = rng.randint(4, 7)
    e, _ = g.walk_query(k, rng.randint(k - 1, k + 2), seed + 5)
    edges = e.tolist()
    og = oracle.OGraph.from_view(g.view())
    want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    dg = ctx.upload(g)
    for flags in (0, cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN, cuts_amd.FLAG_NO_BITMAP, cuts_amd.FLAG_NO_TWO_HOP,
                  cuts_amd.FLAG_NO_SHARED):
        assert ctx.match(dg, p, flags=flags).count == want, (edges, flags)


@pytest.mark.gpu
@pytest.mark.parametrize("div", ["1", "8", "1000000"])
def test_two_hop_dense_rows_and_searches_agree(ctx, monkeypatch, div):
    """Two-hop lookups by lower-bound search only (CUTS_H2_DENSE_DIV=1: no
    vertex's list reaches V entries), by dense rows for the longest lists, and
    by dense rows for every vertex: the same counts as the oracle."""
    monkeypatch.setenv("CUTS_H2_DENSE_DIV", div)
    for seed, names in ((9301, ("4cycle+2pend", "k4e+tail2", "4cycle+tail2")),
                        (9302, ("triangle+2pend-two", "k4+tail2", "path5"))):
        g = graph_for(seed)
        og = oracle.OGraph.from_view(g.view())
        dg = ctx.upload(g)
        for name in names:
            edges = SHAPES[name]
            k = 1 + max(max(e) for e in edges)
            want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
            p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
            assert ctx.match(dg, p).count == want, (div, seed, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["4cycle+2pend", "triangle+2pend-two", "k4e+tail2", "star4"])
def test_last_two_levels_under_a_deadline(ctx, name):
    """With a time limit the run stops between chunks and rolls back a cut
    trunk; the collapsed pair level is then counted by the planner over
    stored rows.  Whatever it stopped at, the count equals the count of the
    first roots_done level-0 vertices, and a generous limit gives the full,
    oracle-equal count."""
    edges = SHAPES[name]
    k = 1 + max(max(e) for e in edges)
    g = graph_for(7700 + k)
    og = oracle.OGraph.from_view(g.view())
    want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    full = ctx.match(dg, p, time_limit_ms=60_000)
    assert full.complete and full.count == want
    n0 = full.level_sizes[0]
    part = ctx.match(dg, p, trunks=n0, time_limit_ms=0.05)
    if part.roots_done:
        again = ctx.match(dg, p, roots=(0, part.last_root + 1, 1))
        assert part.count == again.count and again.level_sizes[0] == part.roots_done


def _chain_formula(N, deg, P, cu, ta, tb):
    """The chain closed form as the kernels evaluate it (chain_closed in
    cuts_amd/csrc/match.hip): completions of positions k-2 (on P[cu]) and
    k-1 (on k-2) for the prefix P, from per-vertex tables and intersections.
    Valid on loop-free symmetric graphs with ta <= 2."""
    u = P[cu]
    cnt = lambda x, t: sum(1 for y in N[x] if deg[y] >= t)
    H = sum(cnt(x, tb) for x in N[u] if deg[x] >= ta)
    inu = [l for l in range(len(P)) if l != cu and P[l] in N[u] and deg[P[l]] >= ta]
    na = cnt(u, ta) - len(inu)
    c = H - sum(cnt(P[l], tb) for l in inu)
    if deg[u] >= tb:
        c -= na
    for i in range(len(P)):
        if i == cu or deg[P[i]] < tb:
            continue
        c -= len(N[u] & N[P[i]])
        c += sum(1 for j in inu if j != i and P[j] in N[P[i]])
    return c


def _chain_brute(N, deg, P, cu, ta, tb):
    u, used = P[cu], set(P)
    return sum(1 for x in N[u] if deg[x] >= ta and x not in used
               for y in N[x] if deg[y] >= tb and y not in used and y != x)


@pytest.mark.parametrize("seed", range(6))
def test_chain_formula_vs_brute_force_loop_free(seed):
    """CPU pin of the chain closed form on random loop-free symmetric graphs,
    random injective prefixes (with and without edges among them) and every
    threshold pair the planner admits (ta <= 2)."""
    rng = random.Random(4200 + seed)
    n = rng.randint(12, 40)
    N = [set() for _ in range(n)]
    for _ in range(rng.randint(n, 4 * n)):
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b:
            N[a].add(b)
            N[b].add(a)
    deg = [len(s) for s in N]
    for _ in range(40):
        L = rng.randint(2, 6)
        P = rng.sample(range(n), L)
        cu = rng.randrange(L)
        for ta in (1, 2):
            for tb in (1, 2, 3, 5):
                assert _chain_formula(N, deg, P, cu, ta, tb) == _chain_brute(N, deg, P, cu, ta, tb)


def _with_loops(g, loops):
    v = g.view()
    src = np.repeat(np.arange(v["V"], dtype=np.uint32), np.diff(v["out_off"]).astype(np.int64))
    dst = v["out_adj"]
    lo = np.asarray(loops, dtype=np.uint32)
    return cuts_amd.Graph.from_arcs(v["V"], np.concatenate([src, lo]), np.concatenate([dst, lo]), symmetrize=True)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_symmetric_graph_with_self_loops(ctx, seed):
    """Symmetric data graphs with self-loops (the loader keeps them; degrees
    count them): the pair, chain and collapsed closed forms assume u is not
    in N(u), so the planner turns them off and expands those positions; the
    last-level closed form counts u in N(u) itself and stays on.  Every mode
    and the batch kernel give the oracle's count, and no pair or chain kernel
    runs."""
    g0 = graph_for(9500 + seed)
    V = g0.V
    deg = np.diff(g0.view()["out_off"])
    rng = random.Random(9500 + seed)
    hubs = np.argsort(-deg)[:3].tolist()
    g = _with_loops(g0, hubs + rng.sample(range(V), V // 4))
    v = g.view()
    assert v["symmetric"]
    og = oracle.OGraph.from_view(v)
    dg = ctx.upload(g)
    plans, wants = [], []
    for name in ("triangle+tail2", "k4e+tail2", "4cycle+2pend", "triangle+2pend-two", "star3", "path5",
                 "4cycle+tail2", "5cycle+chord+2tails"):
        edges = SHAPES[name]
        k = 1 + max(max(e) for e in edges)
        want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
        p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
        for flags in (0, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_SHARED, cuts_amd.FLAG_NO_BITMAP,
                      cuts_amd.FLAG_NO_FUSE, cuts_amd.FLAG_NO_SHORTCUT):
            r = ctx.match(dg, p, flags=flags | cuts_amd.FLAG_PROFILE)
            assert r.count == want, (name, flags)
            assert r.kernels["pair"][1] + r.kernels["chain"][1] + r.kernels["common"][1] == 0, (name, flags)
        plans.append(p)
        wants.append(want)
    assert cuts_amd.Batch(ctx, dg, plans).run().counts == wants
    # the same graph without its loops still takes the closed forms
    dg0 = ctx.upload(g0)
    r = ctx.match(dg0, plans[1], flags=cuts_amd.FLAG_PROFILE)
    assert r.kernels["chain"][1] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_arc_common_table_vs_matrix_product(ctx, seed):
    """The per-arc common-neighbour table built on the device at upload
    equals (A @ A)[u, w] at every arc (u, w), with hub bitmaps (power-law
    graphs whose hubs take bit tests) and without the two-hop index's help."""
    import scipy.sparse as sp
    g = graph_for(9700 + seed) if seed < 2 else cuts_amd.Graph.chung_lu(3000, 24000, 2.1, 0, 9702)
    v = g.view()
    dg = ctx.upload(g)
    info = dg.info()
    assert info["arc_common"] == 1 and info["self_loops"] == 0
    if seed == 2:
        assert info["hubs"] > 0
    V = v["V"]
    src = np.repeat(np.arange(V), np.diff(v["out_off"]).astype(np.int64))
    A = sp.csr_matrix((np.ones(len(src), np.int64), (src, v["out_adj"].astype(np.int64))), shape=(V, V))
    want = np.asarray((A @ A)[src, v["out_adj"].astype(np.int64)]).ravel()
    assert np.array_equal(dg.arc_common(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_pair_level_from_the_arc_table(ctx, seed):
    """Pair-level orders whose two pendant owners are adjacent in the query
    take |N(u) & N(w)| from the per-arc table when there is no two-hop index
    (FLAG_NO_TWO_HOP stands in for a graph too large for one): no k_common
    pass runs, and every mode counts the oracle's number.  Covers the
    5-cycle + chord + 2 tails order with its pendants on positions 2 and 3,
    and the collapsing triangle + 2 pendants."""
    g = graph_for(9800 + seed)
    og = oracle.OGraph.from_view(g.view())
    dg = ctx.upload(g)
    F = cuts_amd
    for name, order in (("5cycle+chord+2tails", [0, 2, 4, 3, 1, 5, 6]),
                        ("5cycle+chord+2tails", [2, 3, 0, 4, 1, 5, 6]),
                        ("triangle+2pend-two", None)):
        edges = SHAPES[name]
        k = 1 + max(max(e) for e in edges)
        want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
        p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges), order)
        assert pair_eligible(p), (name, p.order)
        r = ctx.match(dg, p, flags=F.FLAG_NO_TWO_HOP | F.FLAG_PROFILE)
        assert r.count == want, (name, p.order)
        assert r.kernels["common"][1] == 0 and r.kernels["pair"][1] > 0, (name, r.kernels)
        r2 = ctx.match(dg, p, flags=F.FLAG_NO_TWO_HOP | F.FLAG_NO_ARC_COMMON | F.FLAG_PROFILE)
        assert r2.count == want
        assert r2.level_sizes == r.level_sizes
        for flags in (0, F.FLAG_NO_TWO_HOP | F.FLAG_NO_BITMAP, F.FLAG_NO_PAIR):
            assert ctx.match(dg, p, flags=flags).count == want, (name, flags)
        assert ctx.match(dg, p, flags=F.FLAG_NO_TWO_HOP, budget=2048).count == want


def _chain_row_base(N, deg, R, cu, ta, tb):
    """The candidate-free part of the chain closed form for the older
    positions R (k_sib_base in cuts_amd/csrc/match.hip)."""
    u = R[cu]
    cnt = lambda x, t: sum(1 for y in N[x] if deg[y] >= t)
    inu = {l for l in range(len(R)) if l != cu and R[l] in N[u] and deg[R[l]] >= ta}
    dold = {i for i in range(len(R)) if i != cu and deg[R[i]] >= tb}
    na_old = cnt(u, ta) - len(inu)
    c = sum(cnt(x, tb) for x in N[u] if deg[x] >= ta) - sum(cnt(R[l], tb) for l in inu)
    dtb = deg[u] >= tb
    if dtb:
        c -= na_old
    c += sum(1 for i in dold for j in inu if j != i and R[j] in N[R[i]])
    c -= sum(len(N[u] & N[R[i]]) for i in dold)  # the shared intersections
    return c, na_old, inu, dold, dtb


def _chain_item(N, deg, R, cu, v, base, ta, tb):
    """The candidate's part on top of the row base (chain_sib_item)."""
    c, na_old, inu, dold, dtb = base
    u = R[cu]
    cnt = lambda x, t: sum(1 for y in N[x] if deg[y] >= t)
    inv = deg[v] >= ta and v in N[u]
    if inv:
        c += -cnt(v, tb) + (1 if dtb else 0) + sum(1 for i in dold if v in N[R[i]])
    if deg[v] >= tb:
        c += sum(1 for j in inu if R[j] in N[v])
        c -= len(N[u] & N[v])
    return c


@pytest.mark.parametrize("seed", range(6))
def test_chain_row_base_plus_item_vs_brute_force(seed):
    """CPU pin of the split the shared-candidate chain kernel evaluates: the
    per-row base plus each candidate's part equals the brute-force
    completions of the row + candidate (loop-free symmetric graphs, ta <= 2)."""
    rng = random.Random(4700 + seed)
    n = rng.randint(12, 40)
    N = [set() for _ in range(n)]
    for _ in range(rng.randint(n, 4 * n)):
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b:
            N[a].add(b)
            N[b].add(a)
    deg = [len(x) for x in N]
    for _ in range(30):
        L = rng.randint(2, 5)
        R = rng.sample(range(n), L)
        cu = rng.randrange(L)
        for ta in (1, 2):
            for tb in (1, 2, 3):
                base = _chain_row_base(N, deg, R, cu, ta, tb)
                for v in range(n):
                    if v in R:
                        continue
                    assert _chain_item(N, deg, R, cu, v, base, ta, tb) == _chain_brute(N, deg, R + [v], cu, ta, tb)
