"""The launch-shape autotuner on small workloads: every point it sweeps
(blocks per CU, XCD tile mapping and block size of the counting expansions
for single queries; queue kind, blocks per CU, item work and level-0 chunk
for the batch) must count the same, the chosen point is one of the swept
points, and the cache returns the stored result."""
import pytest

import cuts_amd
from cuts_amd import autotune
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_tune_single_and_batch(ctx, tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    g = cuts_amd.Graph.chung_lu(3000, 15000, 2.3, 0, 21)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(W.query_graph(s)) for s in (W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2)]
    want = [ctx.match(dg, p).count for p in plans]
    print("tune single", flush=True)
    res = autotune.tune_single(ctx, dg, plans, "t-single", reps=1, grid=(2, 6), threads=(256, 128, 64))
    assert res["counts"] == want
    assert res["best"] in res["points"] and len(res["points"]) == 12
    again = autotune.tune_single(ctx, dg, plans, "t-single", reps=1, grid=(2, 6), threads=(256, 128, 64))
    assert again["best"] == res["best"]
    best = res["best"]
    assert [ctx.match(dg, p, blocks_per_cu=best["blocks_per_cu"], expand_threads=best["threads"],
                      flags=0 if best["xcd"] else cuts_amd.FLAG_NO_XCD).count for p in plans] == want
    with pytest.raises(cuts_amd.CutsError, match="expand_threads"):
        ctx.match(dg, plans[0], expand_threads=96)
    print("tune batch", flush=True)
    rb = autotune.tune_batch(ctx, dg, plans, "t-batch", reps=1, blocks=(2, 6), item_works=(4096, 65536),
                             vchunks=(0, 64), queues=("level", "persistent"), arena_bytes=1 << 30,
                             queue_cap=1 << 20)
    assert rb["best"] in rb["points"] and len(rb["points"]) == 16
    bb = rb["best"]
    b = cuts_amd.Batch(ctx, dg, plans, blocks_per_cu=bb["blocks_per_cu"], item_work=bb["item_work"],
                       vchunk=bb["vchunk"], queue=bb["queue"], arena_bytes=1 << 30, queue_cap=1 << 20)
    assert b.run().counts == want


def test_tune_classes(ctx, tmp_path, monkeypatch):
    """Blocks per CU per kernel class: the profiled sweep counts the same at
    every point, the table names only tiled classes with swept values, and a
    run with the table counts the same as the default run."""
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    g = cuts_amd.Graph.chung_lu(3000, 15000, 2.3, 0, 21)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(W.query_graph(s)) for s in (W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2, W.C5CHORD_2TAIL)]
    roots = [None, None, None, (0, g.V, 11)]
    want = [ctx.match(dg, p, roots=rt).count for p, rt in zip(plans, roots)]
    print("tune classes", flush=True)
    res = autotune.tune_classes(ctx, dg, plans, "t-classes", roots=roots, grid=(1, 3, 6), reps=1)
    assert res["counts"] == want
    assert res["blocks_per_cu"] in (1, 3, 6)
    assert set(res["class_blocks_per_cu"]) <= set(autotune.TILED_CLASSES)
    assert all(b in (1, 3, 6) for b in res["class_blocks_per_cu"].values())
    table = {"common": 1, "expand_write": 2, "expand_fused": 5, "pair": 1, "chain": 7, "expand_count": 9}
    assert [ctx.match(dg, p, roots=rt, class_blocks_per_cu=table).count for p, rt in zip(plans, roots)] == want
    got, _ = ctx.match_many(dg, plans, roots=roots, class_blocks_per_cu=table)
    assert [r.count for r in got] == want
    with pytest.raises(cuts_amd.CutsError, match="class_blocks_per_cu"):
        ctx.match(dg, plans[0], class_blocks_per_cu={"common": 65})
