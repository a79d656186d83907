"""Concurrent single-query matching (cuts_match_many): many queries against
one resident graph on several worker streams must give every query exactly
the result of a sequential cuts_match -- counts, level sizes, work -- for
every worker count, with per-query root samples and under every closed-form
mode (the lazily built degree and chain tables are shared across workers)."""
import pytest

import cuts_amd
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_match_many_equals_sequential(ctx):
    # small enough that every query finishes in milliseconds (sizes as in
    # the oracle-checked random cases)
    g = cuts_amd.Graph.chung_lu(400, 1600, 2.4, 0, 31)
    dg = ctx.upload(g)
    shapes = [W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2, W.C5CHORD_2TAIL]
    plans, roots = [], []
    for i in range(48):
        k = 4 + i % 3
        e, _ = g.walk_query(k, k + i % 3, 700 + i)
        plans.append(cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e.tolist())))
        roots.append((i % 7, g.V, 7 + i % 5))
    plans += [cuts_amd.Plan(W.query_graph(s)) for s in shapes]
    roots += [(0, g.V, 13)] * len(shapes)
    for flags in (0, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN):
        want = [ctx.match(dg, p, roots=rt, flags=flags) for p, rt in zip(plans, roots)]
        for workers in (1, 4, 8):
            got, wall = ctx.match_many(dg, plans, roots=roots, flags=flags, workers=workers)
            assert wall > 0
            for a, b in zip(got, want):
                assert (a.count, a.level_sizes, a.work) == (b.count, b.level_sizes, b.work), (flags, workers)
            print(f"match_many flags={flags} workers={workers}: {len(got)} queries equal, {wall:.1f} ms", flush=True)
    # a fresh device graph: the degree and chain tables are first built under concurrency
    dg2 = ctx.upload(g)
    got, _ = ctx.match_many(dg2, plans, roots=roots, workers=8)
    assert [r.count for r in got] == [ctx.match(dg2, p, roots=rt).count for p, rt in zip(plans, roots)]
