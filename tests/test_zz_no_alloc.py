"""The no-allocation entry point (cuts_match_noalloc): every buffer of a run
comes from the device's static scratch array.  Its results must equal
cuts_match's on the same inputs -- counts, level sizes and written
embeddings -- on two calls in a row, the second with a smaller input, and
it must refuse inputs beyond its scratch with a clear error."""
import numpy as np
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu


def plan(shape):
    return cuts_amd.Plan(W.query_graph(shape))


def check_same(ctx, dg, p, **kw):
    got = ctx.match(dg, p, noalloc=True, **kw)  # first: degree tables not cached yet -> built in the scratch
    want = ctx.match(dg, p, **kw)
    assert got.count == want.count, kw
    assert got.level_sizes == want.level_sizes, kw
    assert got.complete == want.complete
    if kw.get("write"):
        assert got.rows.shape == want.rows.shape
        assert sorted(map(tuple, got.rows.tolist())) == sorted(map(tuple, want.rows.tolist()))
    return got


def test_noalloc_equals_cuts_match_twice_second_smaller(ctx):
    assert cuts_amd.noalloc_scratch_bytes() >= 1 << 20
    # call 1: the 36,692-vertex Chung-Lu graph (two-hop index, hub bitmaps):
    # collapsed pair level, chain level with shared candidates, and the
    # plain expansion with a degree-threshold table
    big = W.config2_graph()
    dg = ctx.upload(big)
    r1 = check_same(ctx, dg, plan(W.K4E_TAIL2), roots=(0, big.V, 97))
    assert r1.count > 0
    check_same(ctx, dg, plan(W.C4_2PEND), roots=(0, big.V, 97))
    check_same(ctx, dg, plan(W.K4E_TAIL2), roots=(0, big.V, 997),
               flags=cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN | cuts_amd.FLAG_NO_FUSE)
    # call 2, smaller: a 2,048-vertex graph, fresh device graph, small budget
    # (many chunks) and written embeddings checked against the oracle's
    small = cuts_amd.Graph.chung_lu(2048, 9000, 2.4, 0, 3)
    dg2 = ctx.upload(small)
    p = plan(W.TRI_TAIL2)
    got = check_same(ctx, dg2, p, budget=1 << 16)
    og = oracle.OGraph.from_view(small.view())
    assert got.count == oracle.count(og, oracle.qmatrix(5, W.shape_arcs(W.TRI_TAIL2)))
    tiny = cuts_amd.Graph.chung_lu(300, 900, 2.5, 0, 4)
    dg3 = ctx.upload(tiny)
    w = check_same(ctx, dg3, p, write=True, emb_cap=1 << 20)
    assert w.rows.shape[0] == w.count
    emb = oracle.enumerate_embeddings(oracle.OGraph.from_view(tiny.view()),
                                      oracle.qmatrix(5, W.shape_arcs(W.TRI_TAIL2)))
    assert sorted(map(tuple, w.rows.tolist())) == sorted(map(tuple, emb[:, p.order].tolist()))


def test_noalloc_directed_graph(ctx):
    rng = np.random.default_rng(8)
    V = 400
    src = rng.integers(0, V, 4000)
    dst = rng.integers(0, V, 4000)
    g = cuts_amd.Graph.from_arcs(V, src, dst)
    dg = ctx.upload(g)
    q = cuts_amd.Graph.from_arcs(4, [0, 1, 2, 0], [1, 2, 0, 3])
    p = cuts_amd.Plan(q)
    got = check_same(ctx, dg, p)
    og = oracle.OGraph.from_view(g.view())
    assert got.count == oracle.count(og, oracle.qmatrix(4, [(0, 1), (1, 2), (2, 0), (0, 3)]))


def test_noalloc_refuses_graphs_beyond_its_scratch(ctx):
    V = cuts_amd.noalloc_scratch_bytes() // 128 + 10
    g = cuts_amd.Graph.from_arcs(V, [0, 1, 2], [1, 2, 0], symmetrize=True)
    dg = ctx.upload(g)
    p = plan(W.TRI_TAIL2)
    with pytest.raises(cuts_amd.CutsError) as e:
        ctx.match(dg, p, noalloc=True)
    assert e.value.code == -1 and "no-allocation scratch" in str(e.value)
    # the context stays usable for both entry points
    small = cuts_amd.Graph.chung_lu(256, 800, 2.5, 0, 5)
    dgs = ctx.upload(small)
    assert ctx.match(dgs, p, noalloc=True).count == ctx.match(dgs, p).count
