"""The launch-shape autotuner on small workloads: every point it sweeps
(blocks per CU and XCD tile mapping for single queries; blocks per CU and
item work for the batch) must count the same, the chosen point is one of
the swept points, and the cache returns the stored result."""
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
    res = autotune.tune_single(ctx, dg, plans, "t-single", reps=1, grid=(2, 6))
    assert res["counts"] == want
    assert res["best"] in res["points"] and len(res["points"]) == 4
    again = autotune.tune_single(ctx, dg, plans, "t-single", reps=1, grid=(2, 6))
    assert again["best"] == res["best"]
    rb = autotune.tune_batch(ctx, dg, plans, "t-batch", reps=1, blocks=(2, 6), item_works=(4096, 65536),
                             arena_bytes=1 << 30, queue_cap=1 << 20)
    assert rb["best"] in rb["points"] and len(rb["points"]) == 4
    b = cuts_amd.Batch(ctx, dg, plans, blocks_per_cu=rb["best"]["blocks_per_cu"],
                       item_work=rb["best"]["item_work"], arena_bytes=1 << 30, queue_cap=1 << 20)
    assert b.run().counts == want
