"""Writes tests/golden/batch_queries.json: the 256 queries of the batched
workload (random walks on the mid-size Chung-Lu graph; sizes 4..12 in equal
quotas), admitted by cuts_amd.workloads.admit, with their exact counts and
per-level table sizes from single-query GPU runs.  Needs a GPU; run from the repository root."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import cuts_amd  # noqa: E402
from cuts_amd import workloads as W  # noqa: E402

if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "batch_queries.json")
    g = W.config2_graph()
    ctx = cuts_amd.Context(0)
    dg = ctx.upload(g)
    qs, tried = W.admit(ctx, dg, g, 256, seed=0, log=lambda m: print(m, flush=True))
    with open(out, "w") as f:
        json.dump(dict(graph="chung-lu n=36692 m=183831 gamma=2.1 maxdeg=1400 seed=0", seed=0,
                       tried={str(k): v for k, v in tried.items()}, quotas={str(k): v for k, v in W.quotas(256).items()},
                       max_work=W.MAX_WORK, max_rows=W.MAX_ROWS, queries=qs), f, separators=(",", ":"))
    print(f"wrote {out}: {len(qs)} admitted; tries per size {tried}", flush=True)
