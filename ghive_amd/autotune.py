"""Autotuner of the batched form: tile size (rows per tile), workgroups per CU, and the probe
structure of the shared dimension (LDS bitmap with or without ranked group codes, HBM bitmap, or
the open-addressing hash table at load factor 0.5 / 0.7 / 0.9). Every candidate is timed on the actual batch (HIP graph replay,
best of `reps`), the fastest wins, and the whole search is reported (time spent and per-candidate
results) and cached in a JSON file keyed by the batch signature.
"""
import json
import os
import time

from . import workloads as W
from ._lib import GhiveError

TILE_ROWS = (2048, 4096, 8192, 16384, 32768)
BLOCKS_PER_CU = (2, 4, 8)
# (probe_mode, load_factor): 1 LDS bitmap, 2 HBM bitmap, 3 hash table, 4 LDS bitmap + ranked group codes
PROBES = ((4, 0.5), (1, 0.5), (2, 0.5), (3, 0.5), (3, 0.7), (3, 0.9))


def signature(ctx, B):
    info = ctx.device_info()
    return f"{info['name']}|{info['arch']}|frags={len(B.rows)}|rows={sum(B.rows)}|seed={B.seed}"


def tune_batch(ctx, B, tile_rows=TILE_ROWS, blocks_per_cu=BLOCKS_PER_CU, probes=PROBES, reps=3,
               cache_path=None, log=None):
    """Returns (best, results, search_seconds)."""
    sig = signature(ctx, B)
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cached = json.load(fh)
        if cached.get("signature") == sig:
            return cached["best"], cached["results"], cached["search_s"]
    t0 = time.time()
    results = []
    tables = {}
    for mode, lf in probes:
        if lf not in tables:
            tables[lf] = W.batch_dim_table(ctx, B, load_factor=lf)
        ht = tables[lf]
        for tr in tile_rows:
            for bpc in blocks_per_cu:
                try:
                    bt = W.batch_create(ctx, B, ht, tile_rows=tr, blocks_per_cu=bpc, probe_mode=mode)
                except GhiveError:  # a probe mode this shape does not support (e.g. tables too large for LDS)
                    continue
                ms = min(bt.run() for _ in range(reps))
                info = bt.info()
                bt.free()
                r = dict(tile_rows=tr, blocks_per_cu=bpc, probe_mode=mode, load_factor=lf, ms=ms,
                         grid=info["grid"])
                results.append(r)
                if log:
                    log(f"tune {r}")
    for h in tables.values():
        h.free()
    best = min(results, key=lambda r: r["ms"])
    search_s = time.time() - t0
    if cache_path:
        os.makedirs(os.path.dirname(cache_path) or ".", exist_ok=True)
        with open(cache_path, "w") as fh:
            json.dump(dict(signature=sig, best=best, results=results, search_s=search_s), fh, indent=1)
    return best, results, search_s
