"""Trains the table-size network on the solver's own logged tables.

    python -m cuts_amd.train --out cuts_amd/data/mpnn_rmat18.npy [--queries 5000]

Data graphs: four R-MAT scale-18 graphs (edge factor 16, a,b,c = .57,.19,.19,
seeds 1-4).  Queries: 5-10-vertex random walks on each, admitted on nested
root samples of position 0 (cuts_amd.workloads.admit, capped at 2^24
candidates so that admission stays fast); each is run with the score order
and one random connected order, and every prefix's |T| becomes a sample.
The network (forward, backward and Adam in HIP) is fitted on all samples and
written with a JSON description of the run next to it.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import Context
from . import learned as LN
from . import workloads as W

TRAIN_MAX_WORK = 1 << 24
TRAIN_TRIES = 10000  # candidates per size and graph before a quota is given up


def train(ctx: Context, n_queries: int, seeds=(1, 2, 3, 4), epochs: int = 60, log=print):
    per = max(1, n_queries // len(seeds))
    rng = np.random.default_rng(0)
    parts, admitted, tried_all = [], 0, {}
    t0 = time.perf_counter()
    for s in seeds:
        g = W.rmat18_graph(s)
        dg = ctx.upload(g)
        qs, tried = W.admit(ctx, dg, g, per, seed=100 + s, kmin=5, kmax=10, max_work=TRAIN_MAX_WORK,
                            max_tries_per_k=TRAIN_TRIES)
        admitted += len(qs)
        tried_all[s] = tried
        parts.append(LN.log_tables(ctx, dg, LN.graph_features(g), qs, rng))
        log(f"R-MAT 18 seed {s}: {len(qs)} queries admitted (tries per size {tried}), "
            f"{sum(len(p[3]) for p in parts)} samples so far, {time.perf_counter() - t0:.0f}s")
        del dg
    x = np.concatenate([p[0] for p in parts])
    A = np.concatenate([p[1] for p in parts])
    meta = np.concatenate([p[2] for p in parts])
    t = np.concatenate([p[3] for p in parts])
    t1 = time.perf_counter()
    net = LN.Mpnn(ctx, seed=0)
    hist = net.fit(x, A, meta, t, epochs=epochs, batch=2048, lr=2e-3)
    t2 = time.perf_counter()
    log(f"trained on {len(t)} samples: mse {hist[0]:.4f} -> {hist[-1]:.4f} ({t2 - t1:.0f}s)")
    info = {"graphs": [f"rmat scale=18 ef=16 a,b,c=.57,.19,.19 seed={s}" for s in seeds],
            "queries_requested": n_queries, "queries_admitted": admitted,
            "admission": f"walk queries k 5-10, nested root samples, at most {TRAIN_MAX_WORK} candidates, "
                         f"at most {TRAIN_TRIES} candidate queries per size and graph",
            "tries_per_size": {str(s): {str(k): v for k, v in tr.items()} for s, tr in tried_all.items()},
            "samples": int(len(t)), "epochs": epochs, "mse_first": hist[0], "mse_last": hist[-1],
            "log_s": t1 - t0, "train_s": t2 - t1}
    return net, info


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--queries", type=int, default=5000)
    ap.add_argument("--epochs", type=int, default=60)
    a = ap.parse_args(argv)
    ctx = Context(0)
    net, info = train(ctx, a.queries, epochs=a.epochs, log=lambda m: print(m, flush=True))
    LN.save_weights(a.out, net, info)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
