"""Levels expanded without the readback of their work (the default) against
levels that read it back first (CUTS_FLAG_SYNC_LEVELS): identical counts,
level sizes and work, on a fresh context -- whose buffers are still small,
so writing levels outgrow them and are redone with the readback, and
counting levels outgrow the tile index and make the run repeat itself -- and
again once the buffers have grown, in every closed-form mode, for single
queries and for concurrent ones."""
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu

SYNC = cuts_amd.FLAG_SYNC_LEVELS
MODES = [0, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_TWO_HOP | cuts_amd.FLAG_NO_ARC_COMMON,
         cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN, cuts_amd.FLAG_NO_FUSE, cuts_amd.FLAG_NO_SHORTCUT]


def same(a, b):
    return (a.count, a.level_sizes, a.work) == (b.count, b.level_sizes, b.work)


def test_speculative_levels_equal_synchronous_on_a_fresh_context():
    g = cuts_amd.Graph.chung_lu(4096, 32768, 2.3, 0, 5)
    shapes = [W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2, W.C5CHORD_2TAIL]
    plans = [cuts_amd.Plan(W.query_graph(s)) for s in shapes]
    roots = [None, None, None, (0, g.V, 61)]
    ctx = cuts_amd.Context(0)  # fresh: level buffers and tile index start small
    dg = ctx.upload(g)
    first = [ctx.match(dg, p, roots=rt) for p, rt in zip(plans, roots)]
    for flags in MODES:
        for p, rt, f in zip(plans, roots, first):
            if flags & (cuts_amd.FLAG_NO_SHORTCUT | cuts_amd.FLAG_NO_FUSE):
                # every embedding enumerated: a root sample (about 10^10 candidates
                # at full size), and not the 7-vertex query (far more)
                if p.k > 6:
                    continue
                rt = (0, g.V, 5)
            print(f"flags={flags} k={p.k} roots={rt}", flush=True)
            want = ctx.match(dg, p, roots=rt, flags=flags | SYNC)
            got = ctx.match(dg, p, roots=rt, flags=flags)
            assert same(got, want), (flags, p.order, got, want)
            if flags == 0:
                assert same(f, want)
            print(f"  count {got.count}, {got.device_ms:.1f} ms", flush=True)
    # writing levels beyond the first 64K rows allocated
    assert max(first[1].level_sizes[1:-2]) > 1 << 16
    many, _ = ctx.match_many(dg, plans, roots=roots, workers=4)
    assert all(same(a, b) for a, b in zip(many, first))


def test_counting_level_beyond_the_first_tile_index_repeats_the_run():
    """Without closed forms the last level is a counting expansion of about
    10^9 candidates: far beyond the 4,097-entry tile index a fresh context
    starts with (and beyond one launch's 2^31 items on the full graph), so
    the speculative pass flags it and the run is repeated with readbacks."""
    g = cuts_amd.Graph.chung_lu(4096, 32768, 2.3, 0, 5)
    p = cuts_amd.Plan(W.query_graph(W.TRI_TAIL2))
    for rt in ((0, g.V, 5), None):
        ctx = cuts_amd.Context(0)
        dg = ctx.upload(g)
        print(f"fresh context, roots={rt}", flush=True)
        got = ctx.match(dg, p, roots=rt, flags=cuts_amd.FLAG_NO_SHORTCUT)
        want = ctx.match(dg, p, roots=rt, flags=cuts_amd.FLAG_NO_SHORTCUT | SYNC)
        assert same(got, want), (got, want)
        assert want.work > 4097 * 1024
        assert got.count == ctx.match(dg, p, roots=rt).count
        print(f"  count {got.count}, work {got.work}", flush=True)


def test_speculative_levels_vs_oracle_small_graphs(ctx):
    for seed in range(6):
        g = cuts_amd.Graph.chung_lu(120, 420, 2.4, 0, 700 + seed)
        og = oracle.OGraph.from_view(g.view())
        dg = ctx.upload(g)
        for shape in (W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2):
            qg = W.query_graph(shape)
            p = cuts_amd.Plan(qg)
            want = oracle.count(og, oracle.qmatrix(shape.k, W.shape_arcs(shape)))
            for flags in MODES:
                assert ctx.match(dg, p, flags=flags).count == want, (seed, flags)
