"""Launch-shape autotuner.

Sweeps the tunable launch parameters of the matcher on a given workload and
keeps the fastest point; every point's counts must equal every other
point's (counts do not depend on launch shape), so a tuning run is also a
consistency check.  Results are cached per (device, graph fingerprint,
workload) in ``$CUTS_TUNE_CACHE`` (default ``~/.cache/cuts_amd/tune.json``).

Tunables:
  single query -- blocks per CU of the expansion kernels (grid = CUs x n),
                  the XCD-contiguous tile mapping, and the block size of the
                  counting expansions (64, 128 or 256 threads; a tile stays
                  1,024 work items, so the work items each lane carries --
                  16, 8 or 4 -- change with it).  The reference's sub-wave
                  groups (8-lane virtual warps sharing one intersection,
                  src/device_funcs.cu:618-733) have no counterpart to tune:
                  every lane here owns whole candidates of the flattened
                  work list, and items per lane is the width that varies;
  batch        -- blocks per CU of the level launches (2..8), the candidate
                  work at which a new work item is started (the chunk size),
                  the level-0 chunk (roots per initial item) and the queue
                  (one launch per level, or the persistent kernel).
"""
from __future__ import annotations

import json
import os
import statistics
import time
from typing import Dict, List, Optional, Sequence

from . import FLAG_NO_XCD, Batch, Context, DeviceGraph, Plan

CACHE = os.environ.get("CUTS_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "cuts_amd", "tune.json"))


def _load() -> Dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: Dict):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


class CountMismatch(AssertionError):
    pass


def tune_single(ctx: Context, dg: DeviceGraph, plans: Sequence[Plan], key: str, reps: int = 2,
                grid: Sequence[int] = (2, 4, 8, 16), xcd: Sequence[bool] = (True, False),
                threads: Sequence[int] = (256, 128, 64), roots=None, log=None, use_cache: bool = True) -> Dict:
    cache = _load()
    if use_cache and key in cache:
        return cache[key]
    roots = list(roots) if roots is not None else [None] * len(plans)
    points: List[Dict] = []
    ref_counts = None
    for bpc in grid:
        for x in xcd:
            for nt in threads:
                flags = 0 if x else FLAG_NO_XCD
                ms = []
                counts = []
                for _ in range(reps):
                    t = 0.0
                    counts = []
                    for p, rt in zip(plans, roots):
                        r = ctx.match(dg, p, flags=flags, blocks_per_cu=bpc, expand_threads=nt, roots=rt)
                        t += r.device_ms
                        counts.append(r.count)
                    ms.append(t)
                if ref_counts is None:
                    ref_counts = counts
                elif counts != ref_counts:
                    raise CountMismatch(f"counts differ at blocks_per_cu={bpc} xcd={x} threads={nt}")
                pt = dict(blocks_per_cu=bpc, xcd=x, threads=nt, ms=min(ms))
                points.append(pt)
                if log:
                    log(f"tune single: {pt}")
    best = min(points, key=lambda p: p["ms"])
    res = dict(best=best, points=points, counts=ref_counts)
    cache[key] = res
    _store(cache)
    return res


def tune_batch(ctx: Context, dg: DeviceGraph, plans: Sequence[Plan], key: str, reps: int = 2,
               blocks: Sequence[int] = (2, 4, 6, 8), item_works: Sequence[int] = (4096, 16384, 65536),
               vchunks: Sequence[int] = (0,), queues: Sequence[str] = ("level",),
               arena_bytes: int = 32 << 30, queue_cap: int = 0, roots=None, log=None,
               use_cache: bool = True) -> Dict:
    """vchunk 0 is the default level-0 chunk (4,096 roots per item)."""
    cache = _load()
    if use_cache and key in cache:
        return cache[key]
    points: List[Dict] = []
    ref = None
    for queue in queues:
        for bpc in blocks:
            for iw in item_works:
                for vc in vchunks:
                    b = Batch(ctx, dg, plans, blocks_per_cu=bpc, item_work=iw, vchunk=vc, arena_bytes=arena_bytes,
                              queue_cap=queue_cap, roots=roots, queue=queue)
                    runs = [b.run() for _ in range(reps)]
                    for r in runs:
                        if ref is None:
                            ref = r.counts
                        elif r.counts != ref:
                            raise CountMismatch(f"batch counts differ at queue={queue} blocks_per_cu={bpc} "
                                                f"item_work={iw} vchunk={vc}")
                    pt = dict(queue=queue, blocks_per_cu=bpc, item_work=iw, vchunk=vc,
                              ms=min(r.device_ms for r in runs), p50_latency_ms=statistics.median(runs[-1].latency_ms))
                    points.append(pt)
                    if log:
                        log(f"tune batch: {pt}")
                    del b
    best = min(points, key=lambda p: p["ms"])
    res = dict(best=best, points=points)
    cache[key] = res
    _store(cache)
    return res
