"""Learned matching order: the FP64 reference network is pinned by finite
differences; the HIP forward/backward kernels are checked against it; the
order DP is checked against brute force; training reduces the loss; and a
network trained on the solver's own logged tables proposes orders that count
exactly what the score order counts, touch fewer trie rows than random
orders, and -- once the score order has judged them -- never examine more
candidates than the score order."""
import itertools
import math
import os
import random

import numpy as np
import pytest

import cuts_amd
from cuts_amd import learned as LN
from cuts_amd import workloads as W

torch = pytest.importorskip("torch")
from oracle import mpnn_ref  # noqa: E402


def random_samples(n, seed):
    rng = np.random.default_rng(seed)
    samples = []
    for _ in range(n):
        k = int(rng.integers(2, 13))
        edges = [(int(rng.integers(0, v)), v) for v in range(1, k)]
        edges += [tuple(map(int, rng.choice(k, 2, replace=False))) for _ in range(int(rng.integers(0, k)))]
        adj = LN.adjacency(k, edges)
        sm = int(rng.integers(1, 1 << k))
        samples.append((k, adj, sm))
    gfeat = rng.random(16).astype(np.float32) * 0.3
    x, A, meta = LN.encode(samples, gfeat)
    t = rng.random(n).astype(np.float32) * 3
    return x, A, meta, t


def test_reference_gradient_matches_finite_differences():
    x, A, meta, t = random_samples(3, 1)
    p = LN.init_params(3).astype(np.float64)
    p = mpnn_ref.bf16_round(p)  # exactly representable start point
    _, g, _ = mpnn_ref.loss_and_grad(p, x, A, meta, t)
    rng = np.random.default_rng(0)
    idx = list(rng.choice(len(p), 25, replace=False)) + [len(p) - 1, len(p) - 2]
    tp = torch.tensor(p)
    for i in idx:
        h = 1e-6
        a, b = tp.clone(), tp.clone()
        a[i] += h
        b[i] -= h
        la = float(((mpnn_ref.forward(a, x, A, meta) - torch.tensor(t, dtype=torch.float64)) ** 2).mean())
        lb = float(((mpnn_ref.forward(b, x, A, meta) - torch.tensor(t, dtype=torch.float64)) ** 2).mean())
        fd = (la - lb) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def brute_best(k, adj, f, first=None, eng=LN.LAST_LEVEL_ONLY):
    best = None
    for perm in itertools.permutations(range(k)):
        if first is not None and perm[0] != first:
            continue
        ok, sm = True, 1 << perm[0]
        for v in perm[1:]:
            if not adj[v] & sm:
                ok = False
                break
            sm |= 1 << v
        if ok:
            c = LN.order_cost(list(perm), adj, f, eng)
            best = c if best is None else min(best, c)
    return best


ENGINES = [LN.LAST_LEVEL_ONLY, LN.Engine(), LN.Engine(two_hop=True), LN.Engine(arc_common=True),
           LN.Engine(two_hop=True, arc_common=True), LN.Engine(symmetric=False)]


@pytest.mark.parametrize("seed", range(24))
def test_order_dp_is_exact(seed):
    """The DP (connected subsets + the last three positions enumerated) finds
    the minimum of the engine-aware cost over all connected orders, with and
    without a fixed first vertex, for every engine configuration."""
    rng = random.Random(seed)
    k = rng.randint(2, 7)
    edges = [(rng.randrange(v), v) for v in range(1, k)] + [tuple(rng.sample(range(k), 2)) for _ in range(rng.randint(0, 4))]
    adj = LN.adjacency(k, edges)
    f = {sm: rng.uniform(1, 1000) for sm in LN.connected_subsets(k, adj)}
    for eng in ENGINES:
        order, cost = LN.best_order(k, adj, f, eng=eng)
        assert sorted(order) == list(range(k))
        assert math.isclose(cost, LN.order_cost(order, adj, f, eng))
        assert math.isclose(cost, brute_best(k, adj, f, eng=eng))
        first = rng.randrange(k)
        order, cost = LN.best_order(k, adj, f, first=first, eng=eng)
        assert order[0] == first and sorted(order) == list(range(k))
        assert math.isclose(cost, LN.order_cost(order, adj, f, eng))
        assert math.isclose(cost, brute_best(k, adj, f, first, eng))


@pytest.mark.parametrize("seed", range(12))
def test_top_orders_are_valid_distinct_and_sorted(seed):
    """best_orders: the first is best_order's; every proposal is a
    connected order (with the fixed first vertex) whose stated cost is its
    engine cost; costs never decrease; the last three positions differ."""
    rng = random.Random(100 + seed)
    k = rng.randint(4, 8)
    edges = [(rng.randrange(v), v) for v in range(1, k)] + [tuple(rng.sample(range(k), 2)) for _ in range(rng.randint(0, 5))]
    adj = LN.adjacency(k, edges)
    f = {sm: rng.uniform(1, 1000) for sm in LN.connected_subsets(k, adj)}
    for eng in ENGINES:
        first = rng.randrange(k)
        top = LN.best_orders(k, adj, f, first=first, eng=eng, top=4)
        assert top[0] == LN.best_order(k, adj, f, first=first, eng=eng)
        costs = [c for _, c in top]
        assert costs == sorted(costs)
        tails = set()
        for order, c in top:
            assert order[0] == first and sorted(order) == list(range(k))
            seen = 1 << order[0]
            for v in order[1:]:
                assert adj[v] & seen  # every prefix connected
                seen |= 1 << v
            assert math.isclose(c, LN.order_cost(order, adj, f, eng))
            tails.add(tuple(order[-3:]))
        assert len(tails) == len(top)


def test_engine_modes_follow_the_planner_rules():
    """The cost model's view of which positions the engine expands agrees with
    the planner rules restated in tests/test_pair.py (pair, collapse and
    chain eligibility of a plan), on every connected order of the shapes
    there and of random walk-like queries."""
    import test_pair as TP
    shapes = list(TP.SHAPES.values())
    rng = random.Random(5)
    for _ in range(30):
        k = rng.randint(4, 7)
        shapes.append([(rng.randrange(v), v) for v in range(1, k)] +
                      [tuple(rng.sample(range(k), 2)) for _ in range(rng.randint(0, 3))])
    seen = set()
    for edges in shapes:
        k = 1 + max(max(e) for e in edges)
        adj = LN.adjacency(k, edges)
        qg = cuts_amd.Graph.from_edges(k, edges)
        orders = [o for o in itertools.permutations(range(k))
                  if all(adj[o[i]] & sum(1 << o[j] for j in range(i)) for i in range(1, k))]
        for o in rng.sample(orders, min(40, len(orders))):
            p = cuts_amd.Plan(qg, list(o))
            T = min(3, k - 1)
            s0 = sum(1 << v for v in o[:k - T])
            for eng in (LN.Engine(two_hop=True), LN.Engine(), LN.Engine(arc_common=True)):
                E, extra = LN._tail_mode(adj, s0, list(o[k - T:]), k, eng)
                b = TP.backward(p)
                sc = len(b[k - 1]) == 1
                pair = TP.pair_eligible(p)
                chain = not pair and TP.chain_eligible(p)
                coll = pair and TP.collapse_eligible(p)
                if coll and not eng.two_hop:  # without the two-hop index only adjacent owners collapse
                    (a,), (bb,) = b[k - 2], b[k - 1]
                    coll = eng.arc_common and bool(adj[o[a]] >> o[bb] & 1)
                want = k - 4 if coll else k - 3 if (pair or chain) else k - 2 if sc else k - 1
                assert E == want, (edges, o, eng, E, want)
                seen.add(want - k)
    assert seen == {-4, -3, -2, -1}


def test_connected_subset_counts():
    for k in range(1, 9):
        path = LN.adjacency(k, [(i, i + 1) for i in range(k - 1)])
        assert len(LN.connected_subsets(k, path)) == k * (k + 1) // 2
        full = LN.adjacency(k, [(i, j) for i in range(k) for j in range(i + 1, k)])
        assert len(LN.connected_subsets(k, full)) == 2 ** k - 1


@pytest.mark.gpu
def test_mpnn_forward_and_gradient_match_fp64_reference(ctx):
    x, A, meta, t = random_samples(40, 2)
    p = LN.init_params(5)
    net = LN.Mpnn(ctx, p)
    y = net.forward(x, A, meta)
    loss, g = net.gradient(x, A, meta, t)
    rl, rg, ry = mpnn_ref.loss_and_grad(p, x, A, meta, t)
    # FP32 accumulation against FP64: tolerances relative to the magnitudes
    assert np.allclose(y, ry, rtol=1e-4, atol=1e-4)
    assert abs(loss - rl) <= 1e-4 * max(1.0, rl)
    assert np.max(np.abs(g - rg)) <= 1e-3 * np.max(np.abs(rg)) + 1e-6


@pytest.mark.gpu
def test_mpnn_training_reduces_loss(ctx):
    x, A, meta, _ = random_samples(2000, 3)
    # a learnable target: a fixed function of the subset size and degrees
    t = (np.array([bin(int(m) >> 16).count("1") for m in meta]) * 0.3 + x[:, :, 2].sum(1) * 0.1).astype(np.float32)
    net = LN.Mpnn(ctx, seed=1)
    hist = net.fit(x, A, meta, t, epochs=25, batch=500, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0]


@pytest.mark.gpu
def test_learned_orders_count_the_same_beat_random_orders_and_never_lose_after_judging(ctx):
    """A network trained on the solver's own logged tables proposes orders
    (DP over its predicted table sizes, no fallback).  Every proposal counts
    exactly what the score order counts; the proposals touch far fewer trie
    rows than random connected orders (the network carries real information
    about table sizes); and after the score order judges them, no query
    examines more candidates than with the score order.  On these walk
    queries the score order is usually already optimal, so the judged gain
    over it is reported, not asserted."""
    g = cuts_amd.Graph.chung_lu(4096, 32768, 2.3, 0, 11)
    dg = ctx.upload(g)
    gfeat = LN.graph_features(g)
    rng = np.random.default_rng(0)
    # one execution mode throughout (the last two levels expanded, not
    # counted together): work -- which admits queries and judges orders --
    # depends on the mode, and this one keeps the admitted workload and the
    # uncapped random-order runs below at their intended size
    F = cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN

    def rows(k, r):
        return sum(r.level_sizes[: k - 1])

    def queries(seed, n):
        out = []
        for i, k, ne, edges in W.batch_candidates(g, seed):
            if not 5 <= k <= 8:
                continue
            p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
            r = ctx.match(dg, p, max_work=1 << 27, flags=F)
            if not r.work_capped:
                out.append((k, edges, p, r))
            if len(out) == n:
                return out

    train = queries(1, 200)
    samples, targets = [], []
    for k, edges, p, r in train:
        runs = [(p.order, r)]
        for _ in range(2):
            o = LN.random_connected_order(k, LN.adjacency(k, edges), rng)
            rr = ctx.match(dg, cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges), o), max_work=1 << 27, flags=F)
            if not rr.work_capped:
                runs.append((o, rr))
        for o, rr in runs:
            for kk, adj, sm, size in LN.table_samples(k, edges, o, rr.level_sizes):
                samples.append((kk, adj, sm))
                targets.append(LN.target(size))
    x, A, meta = LN.encode(samples, gfeat)
    net = LN.Mpnn(ctx, seed=2)
    hist = net.fit(x, A, meta, np.asarray(targets, np.float32), epochs=40, batch=1024, lr=2e-3)
    assert hist[-1] < hist[0]

    test = queries(2, 60)
    qs = [(k, e) for k, e, _, _ in test]
    choices = LN.choose_orders(net, qs, gfeat, [p.order for _, _, p, _ in test], margin=-1e9)  # never fall back
    rows_h = rows_l = rows_r = 0
    for (k, edges, p, r), ch in zip(test, choices):
        rl = ctx.match(dg, cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges), ch.order), flags=F)
        assert rl.count == r.count
        o = LN.random_connected_order(k, LN.adjacency(k, edges), rng)
        rr = ctx.match(dg, cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges), o), flags=F)
        assert rr.count == r.count
        rows_h += rows(k, r)
        rows_l += rows(k, rl)
        rows_r += rows(k, rr)
    print(f"rows: score order {rows_h}, network proposals {rows_l}, random orders {rows_r}", flush=True)
    assert rows_l < rows_r

    judged = LN.judge(ctx, dg, qs, [p for _, _, p, _ in test], [r for _, _, _, r in test], choices, flags=F)
    for j, (k, edges, p, r) in zip(judged, test):
        assert j.count == r.count
        assert j.work <= r.work
    print(f"judged: {sum(j.learned for j in judged)} of {len(test)} proposals kept; candidates "
          f"{sum(j.work for j in judged)} vs score order {sum(r.work for _, _, _, r in test)}", flush=True)


@pytest.mark.gpu
def test_config5_train_on_rmat18_evaluate_on_rmat20(ctx):
    """Config 5 at test scale.  A network is trained with the HIP kernels on
    tables the solver logs for a small admitted set of walk queries on an
    R-MAT scale-18 graph; its engine-aware orders for held-out queries of the
    R-MAT scale-20 fixture (kept or fallen back after the score order's
    timed judgement) count exactly the fixture's score-order counts.  With
    the committed weights, the learned order of the config-3 query is the
    pair order and at least 10x faster than the score order, and over the
    judged evaluation queries the learned orders take less total time."""
    import json
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat20_queries.json")))
    g18 = W.rmat18_graph(1)
    dg18 = ctx.upload(g18)
    qs18, _ = W.admit(ctx, dg18, g18, 40, seed=101, kmin=5, kmax=8, max_work=1 << 22, max_tries_per_k=400)
    assert len(qs18) >= 20
    x, A, meta, t = LN.log_tables(ctx, dg18, LN.graph_features(g18), qs18, np.random.default_rng(0))
    net = LN.Mpnn(ctx, seed=3)
    hist = net.fit(x, A, meta, t, epochs=30, batch=512, lr=2e-3)
    assert hist[-1] < 0.5 * hist[0]
    del dg18
    g = W.config3_graph()
    assert W.csr_fingerprint(g) == fx["fingerprint"]
    dg = ctx.upload(g)
    info = dg.info()
    eng = LN.Engine(two_hop=info["two_hop_entries"] > 0, arc_common=bool(info["arc_common"]))
    items = [fx["config3"]] + fx["queries"][:60]
    qs = [(q["k"], [tuple(e) for e in q["edges"]]) for q in items]
    roots = [tuple(q["roots"]) for q in items]
    want = [q["count"] for q in items]
    heur = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    work = [q["work"] for q in items]
    gf = LN.graph_features(g)
    committed, _ = LN.load_weights(ctx, os.path.join(os.path.dirname(LN.__file__), "data", "mpnn_rmat18.npy"))
    for model in (net, committed):
        choices = LN.choose_orders(model, qs, gf, [p.order for p in heur], keep_first=True, eng=eng)
        judged = LN.judge_by_time(ctx, dg, qs, heur, want, choices, roots=roots, heuristic_work=work)
        assert any(not j.learned for j in judged)  # the fallback path is exercised
        for j, rt, w in zip(judged, roots, want):
            assert ctx.match(dg, j.plan, roots=rt).count == w
    # the committed network on the config-3 query and the judged set
    c3 = judged[0]
    assert c3.learned and c3.ms * 10 < c3.heuristic_ms, (c3.ms, c3.heuristic_ms)
    assert sum(j.ms for j in judged[1:]) < sum(j.heuristic_ms for j in judged[1:])
