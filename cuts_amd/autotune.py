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
  per class    -- blocks per CU of each tiled kernel class (tune_classes):
                  one profiled sweep of uniform values, each class taking its
                  fastest, kept if the table beats the best uniform value;
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

from . import FLAG_NO_XCD, FLAG_PROFILE, Batch, Context, DeviceGraph, Plan

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


TILED_CLASSES = ("expand_write", "expand_count", "expand_fused", "pair", "common", "chain")


def tune_classes(ctx: Context, dg: DeviceGraph, plans: Sequence[Plan], key: str, roots=None,
                 grid: Sequence[int] = (2, 3, 4, 6, 8), expand_threads: int = 0, reps: int = 2, log=None,
                 use_cache: bool = True) -> Dict:
    """Blocks per CU for each tiled kernel class from one profiled sweep.

    Every plan runs at each uniform blocks-per-CU value with per-launch HIP
    events (FLAG_PROFILE); each tiled class (expansions, intersection walks,
    chain level) takes the value at which its summed time was least -- list
    walks at random addresses often prefer fewer resident blocks (less L2
    thrashing) than the counting expansions.  The table is then timed on the
    plain path against the best uniform value and kept only if faster.
    Every run must count the same.  The default grid starts at 2: a class
    tuned to 1 block per CU on a sample of small launches starved the large
    launches of other queries (the R-MAT 20 config-3 query: 3.3 ms against
    2.1 ms at 3 blocks per CU)."""
    cache = _load()
    if use_cache and key in cache:
        return cache[key]
    roots = list(roots) if roots is not None else [None] * len(plans)
    ref_counts = None

    def check(counts, what):
        nonlocal ref_counts
        if ref_counts is None:
            ref_counts = counts
        elif counts != ref_counts:
            raise CountMismatch(f"counts differ at {what}")

    def plain_ms(**kw):
        best = None
        for _ in range(reps):
            t, counts = 0.0, []
            for p, rt in zip(plans, roots):
                r = ctx.match(dg, p, roots=rt, expand_threads=expand_threads, **kw)
                t += r.device_ms
                counts.append(r.count)
            check(counts, str(kw))
            best = t if best is None else min(best, t)
        return best

    per: Dict[int, Dict[str, float]] = {}
    for b in grid:
        cls: Dict[str, float] = {}
        for _ in range(reps):
            acc: Dict[str, float] = {}
            counts = []
            for p, rt in zip(plans, roots):
                r = ctx.match(dg, p, roots=rt, flags=FLAG_PROFILE, blocks_per_cu=b, expand_threads=expand_threads)
                counts.append(r.count)
                for name, (ms, n, _) in r.kernels.items():
                    if n:
                        acc[name] = acc.get(name, 0.0) + ms
            check(counts, f"blocks_per_cu={b} (profiled)")
            for name, ms in acc.items():
                cls[name] = min(cls.get(name, ms), ms)
        per[b] = cls
        if log:
            log(f"tune classes: blocks_per_cu {b}: " + ", ".join(f"{k} {v:.1f}" for k, v in sorted(cls.items())))
    table = {}
    for name in TILED_CLASSES:
        seen = [(per[b][name], b) for b in grid if name in per[b]]
        # a class the sample barely exercised keeps the default
        if seen and max(t for t, _ in seen) > 0.05:
            table[name] = min(seen)[1]
    uniform = {b: plain_ms(blocks_per_cu=b) for b in grid}
    ub = min(uniform, key=uniform.get)
    t_table = plain_ms(blocks_per_cu=ub, class_blocks_per_cu=table)
    keep = t_table < uniform[ub]
    res = dict(blocks_per_cu=ub, class_blocks_per_cu=table if keep else {}, uniform_ms=uniform,
               table_ms=t_table, per_class_ms={str(b): v for b, v in per.items()}, counts=ref_counts)
    if log:
        log(f"tune classes: uniform best {ub} ({uniform[ub]:.1f} ms), per-class table {table} "
            f"({t_table:.1f} ms){' kept' if keep else ' not kept'}")
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
