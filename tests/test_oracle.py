"""Pins the CPU oracle before anything is compared against it.

* networkx monomorphism counts (tests/golden/nx_counts.json), including the
  reference's 33 directed query graphs;
* brute force over all injective maps on tiny graphs with self-loops;
* closed forms: n!/(n-k)! into the complete digraph, trace(A^3) for
  triangles, sum d(d-1) for 2-paths, sum d!/(d-s)! for s-stars, and
  sum over ordered triangles of (d_c - 2) for a triangle with a pendant;
* the score-order restatement against the order the reference's own host
  code computes (tests/golden/ref_plans.json);
* the CPU trie baseline against the oracle count for every order, and its
  per-level table sizes against oracle counts of the prefix sub-queries.
"""
import itertools
import json
import math
import os
import random

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


NX = load("nx_counts.json")["cases"]
PLANS = load("ref_plans.json")["plans"]


@pytest.mark.parametrize("case", NX, ids=[c["name"] for c in NX])
def test_oracle_matches_networkx(case):
    g = oracle.OGraph.from_arcs(case["V"], case["arcs"])
    qa = oracle.qmatrix(case["k"], case["qarcs"])
    assert oracle.count(g, qa, nthreads=2) == case["count"]


def brute(V, arcs, k, qarcs):
    """Reference semantics: injective, every query arc u->v (u != v) kept,
    signature (in/out degree, self-loops included) dominated."""
    A = set(map(tuple, arcs))
    din = [0] * V
    dout = [0] * V
    for s, t in A:
        dout[s] += 1
        din[t] += 1
    Q = set(map(tuple, qarcs))
    qin = [sum(1 for s, t in Q if t == u) for u in range(k)]
    qout = [sum(1 for s, t in Q if s == u) for u in range(k)]
    rows = []
    for f in itertools.permutations(range(V), k):
        if any(din[f[u]] < qin[u] or dout[f[u]] < qout[u] for u in range(k)):
            continue
        if all((f[s], f[t]) in A for s, t in Q if s != t):
            rows.append(f)
    return rows


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_bruteforce(seed):
    rng = random.Random(seed)
    V = rng.randint(3, 7)
    arcs = [(s, t) for s in range(V) for t in range(V) if rng.random() < 0.45]  # self-loops allowed
    k = rng.randint(1, min(4, V))
    q = set()
    for v in range(1, k):
        u = rng.randrange(v)
        q.add((u, v) if rng.random() < 0.5 else (v, u))
        if rng.random() < 0.4:
            q.add((v, u) if (u, v) in q else (u, v))
    if rng.random() < 0.3:
        x = rng.randrange(k)
        q.add((x, x))  # query self-loop: raises thresholds only
    q = sorted(q)
    g = oracle.OGraph.from_arcs(V, arcs)
    qa = oracle.qmatrix(k, q)
    expect = brute(V, arcs, k, q)
    assert oracle.count(g, qa, nthreads=1) == len(expect)
    got = oracle.enumerate_embeddings(g, qa, cap=10000)
    assert sorted(map(tuple, got.tolist())) == sorted(expect)


def complete_digraph(n):
    return oracle.OGraph.from_arcs(n, [(s, t) for s in range(n) for t in range(n) if s != t])


@pytest.mark.parametrize("n,k", [(6, 3), (7, 5), (8, 4), (9, 6)])
def test_closed_form_complete_digraph(n, k):
    rng = random.Random(n * 31 + k)
    # any loop-free connected k-vertex query, directed arcs at random
    q = set()
    for v in range(1, k):
        u = rng.randrange(v)
        q.add((u, v) if rng.random() < 0.5 else (v, u))
    q.add((0, k - 1))
    g = complete_digraph(n)
    assert oracle.count(g, oracle.qmatrix(k, sorted(q))) == math.factorial(n) // math.factorial(n - k)


def sym_random(n, p, seed):
    rng = np.random.default_rng(seed)
    A = (rng.random((n, n)) < p).astype(np.int64)
    A = np.triu(A, 1)
    A = A + A.T
    arcs = np.argwhere(A > 0).tolist()
    return A, oracle.OGraph.from_arcs(n, arcs)


@pytest.mark.parametrize("seed", range(5))
def test_closed_forms_undirected(seed):
    A, g = sym_random(40, 0.2, seed)
    d = A.sum(1)
    tri = oracle.qmatrix(3, oracle.undirected([(0, 1), (1, 2), (2, 0)]))
    assert oracle.count(g, tri) == int(np.trace(A @ A @ A))
    path2 = oracle.qmatrix(3, oracle.undirected([(0, 1), (1, 2)]))
    assert oracle.count(g, path2) == int((d * (d - 1)).sum())
    for s in (2, 3, 4):
        star = oracle.qmatrix(s + 1, oracle.undirected([(0, i) for i in range(1, s + 1)]))
        expect = sum(math.perm(int(x), s) for x in d)
        assert oracle.count(g, star) == expect
    # triangle (0,1,2) + pendant 3 on 2: sum over ordered triangles of d_c - 2
    tp = oracle.qmatrix(4, oracle.undirected([(0, 1), (1, 2), (2, 0), (2, 3)]))
    T = (A @ A) * A  # T[b,c] = #a closing the triangle over edge (b,c)
    expect = int(sum(T[b, c] * (d[c] - 2) for b in range(40) for c in range(40)))
    assert oracle.count(g, tp) == expect


@pytest.mark.parametrize("plan", PLANS, ids=[p["name"] for p in PLANS])
def test_order_restatement_matches_reference(plan):
    qa = oracle.qmatrix(plan["k"], plan["qarcs"])
    assert oracle.heuristic_order(qa) == plan["order"]


def valid_orders(k, qarcs, rng, n):
    adj = {u: set() for u in range(k)}
    for s, t in qarcs:
        if s != t:
            adj[s].add(t)
            adj[t].add(s)
    out = []
    for _ in range(n):
        start = rng.randrange(k)
        order, seen = [start], {start}
        while len(order) < k:
            front = sorted({w for u in order for w in adj[u]} - seen)
            w = rng.choice(front)
            order.append(w)
            seen.add(w)
        out.append(order)
    return out


@pytest.mark.parametrize("case", NX[:40], ids=[c["name"] for c in NX[:40]])
def test_trie_cpu_counts_and_levels(case):
    g = oracle.OGraph.from_arcs(case["V"], case["arcs"])
    k = case["k"]
    qa = oracle.qmatrix(k, case["qarcs"])
    qin = qa.sum(0).astype(np.uint32)
    qout = qa.sum(1).astype(np.uint32)
    rng = random.Random(case["name"])
    for order in valid_orders(k, case["qarcs"], rng, 3):
        r = oracle.trie_cpu(g, qa, order, max_rows=64, nthreads=2)  # tiny tables force chunking
        assert r.count == case["count"]
        for l in range(k):
            pre = order[: l + 1]
            sub = qa[np.ix_(pre, pre)]
            want = oracle.count_ex(g, sub, thr_in=qin[pre], thr_out=qout[pre], nthreads=1)
            assert r.level[l] == want, (order, l)


def brute_merge_bytes(V, arcs, k, qarcs, order):
    """The merge-model bytes from their definition, over trie levels built
    by brute force (every injective extension checked against every arc):
    8 per data vertex, per prefix of T_(l-1) 8 + 4*len for each distinct
    backward list of position l (a symmetric graph's in- and out-list of a
    vertex are one list), 16 per node of T_0..T_(k-2)."""
    A = set(map(tuple, arcs))
    outd, ind = [0] * V, [0] * V
    for a, b in A:
        outd[a] += 1
        ind[b] += 1
    sym = all((b, a) in A for a, b in A)
    Q = set(map(tuple, qarcs))
    qout, qin = [0] * k, [0] * k
    for a, b in Q:
        qout[a] += 1
        qin[b] += 1

    def fits(q, x):
        return ind[x] >= qin[q] and outd[x] >= qout[q]

    T = [[(x,) for x in range(V) if fits(order[0], x)]]
    for l in range(1, k):
        v = order[l]
        nxt = []
        for p in T[-1]:
            for x in range(V):
                if x in p or not fits(v, x):
                    continue
                if all(((order[j], v) not in Q or (p[j], x) in A) and ((v, order[j]) not in Q or (x, p[j]) in A)
                       for j in range(l)):
                    nxt.append(p + (x,))
        T.append(nxt)
    B = 8 * V
    for l in range(1, k):
        v = order[l]
        lists = []
        for j in range(l):
            dirs = ([0] if (order[j], v) in Q else []) + ([1] if (v, order[j]) in Q else [])
            if sym and dirs:
                dirs = [0]
            lists += [(j, d) for d in dirs]
        for p in T[l - 1]:
            for j, d in lists:
                B += 8 + 4 * (outd[p[j]] if d == 0 else ind[p[j]])
    B += 16 * sum(len(T[l]) for l in range(k - 1))
    return B, [len(t) for t in T]


@pytest.mark.parametrize("seed", range(16))
def test_trie_cpu_merge_bytes_from_definition(seed):
    rng = random.Random(seed)
    V = rng.randint(5, 9)
    directed = seed % 2 == 1
    arcs = set()
    for a in range(V):
        for b in range(V):
            if a != b and (directed or a < b) and rng.random() < 0.45:
                arcs.add((a, b))
                if not directed:
                    arcs.add((b, a))
    k = rng.randint(2, 4)
    q = set()
    for v in range(1, k):
        u = rng.randrange(v)
        q.add((u, v))
        if not directed:
            q.add((v, u))
        elif rng.random() < 0.4:
            q.add((v, u))
    qarcs = sorted(q)
    g = oracle.OGraph.from_arcs(V, sorted(arcs))
    qa = oracle.qmatrix(k, qarcs)
    for order in valid_orders(k, qarcs, rng, 3):
        want, levels = brute_merge_bytes(V, arcs, k, qarcs, order)
        r = oracle.trie_cpu(g, qa, order, nthreads=2)
        assert [r.level[i] for i in range(k)] == levels
        assert r.bytes == want, (order, r.bytes, want)


def test_oracle_root_restriction_partitions_count():
    _, g = sym_random(50, 0.15, 11)
    qa = oracle.qmatrix(4, oracle.undirected([(0, 1), (1, 2), (2, 3), (3, 0)]))
    total = oracle.count(g, qa)
    parts = sum(oracle.count_ex(g, qa, root_q=2, roots=(r, 50, 3)) for r in range(3))
    assert parts == total > 0
