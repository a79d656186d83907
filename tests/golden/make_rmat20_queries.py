"""Writes tests/golden/rmat20_queries.json: the queries matched on the R-MAT
scale-20 graph (rmat scale=20 ef=16 a,b,c=.57,.19,.19 seed=0):

* the 7-vertex 5-cycle + chord + 2 tails query on the root sample
  v = 127 (mod 128) of order position 0 (its full count exceeds 2^64);
* the held-out evaluation set of the learned order: 6-10-vertex random-walk
  queries (seed 7), sizes in equal quotas, each admitted on its densest
  nested root sample of position 0 within 2^28 candidates
  (cuts_amd.workloads.admit).

Counts and per-level sizes come from single-query GPU runs with the score
order; the config-3 count is also checked in a second execution mode (pair
and chain levels off).  Needs a GPU; run from the repository root."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import cuts_amd  # noqa: E402
from cuts_amd import workloads as W  # noqa: E402

N_EVAL = 500
TRIES = 8000

if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "rmat20_queries.json")
    t0 = time.time()
    g = W.config3_graph()
    ctx = cuts_amd.Context(0)
    dg = ctx.upload(g)
    shape = W.C5CHORD_2TAIL
    p = cuts_amd.Plan(W.query_graph(shape))
    roots = W.root_sample(g.V, 128)
    r = ctx.match(dg, p, roots=roots)
    r2 = ctx.match(dg, p, roots=roots, flags=cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN)
    assert r.count == r2.count and r.level_sizes == r2.level_sizes, (r.count, r2.count)
    c3 = dict(k=shape.k, edges=[list(e) for e in shape.edges], name=shape.name, roots=list(roots), count=r.count,
              level_sizes=r.level_sizes, work=r2.work, order=p.order)
    print(f"config-3 query: count {r.count} ({time.time() - t0:.0f}s)", flush=True)
    qs, tried = W.admit(ctx, dg, g, N_EVAL, seed=7, kmin=6, kmax=10, max_tries_per_k=TRIES,
                        log=lambda m: print(m, flush=True) if m.startswith("tried") or
                        m.split()[1].split("/")[0].endswith("0") else None)
    with open(out, "w") as f:
        json.dump(dict(graph="rmat scale=20 ef=16 a,b,c=.57,.19,.19 seed=0", fingerprint=W.csr_fingerprint(g),
                       config3=c3, eval_seed=7, tried={str(k): v for k, v in tried.items()},
                       quotas={str(k): v for k, v in W.quotas(N_EVAL, 6, 10).items()},
                       max_work=W.MAX_WORK, max_rows=W.MAX_ROWS, max_tries_per_k=TRIES, queries=qs),
                  f, separators=(",", ":"))
    print(f"wrote {out}: {len(qs)} admitted; tries per size {tried} ({time.time() - t0:.0f}s)", flush=True)
