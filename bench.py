"""Benchmark of the star-join fragment on one MI355X (plain form).

Workload (one "step"): fact table of 600M rows resident in HBM (7 int32 columns), three
dimension tables (date 2556, supplier 200k, part 1.4M rows); per step the dimension filters run,
the three hash tables are built, and the fused scan -> probe x3 -> price*discount -> GROUP BY
year, brand kernel makes one pass over the fact table, then the groups are written out.

Prints ONE JSON line on stdout. `value` = fact rows/s of the whole step. `roofline` is for the
fused kernel: algorithmic bytes per launch (the 5 referenced int32 fact columns, 20 B/row;
hash tables are cache-resident and not counted) / its average duration from HIP events on the
stream it runs on. `cpu_baseline` times the oracle's C restatement of Hive 3.1's vectorized
map-join pipeline on the host cores over a bounded sample of the same fact rows.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fact rows/s + %HBM peak, star join+agg @600M rows; batched p99 ms"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """The host threads the CPU baselines use: the cores this process may run on (its affinity
    mask), capped by a cgroup CPU quota and by OMP_NUM_THREADS where the environment sets the
    process's CPU share that way; with the CPU model name."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
        if q != "max":
            quota = max(1, -(-int(q) // int(per)))
    except (OSError, ValueError):
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    share = int(env) if env.isdigit() and int(env) > 0 else None
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": min(x for x in (aff, quota, share) if x), "affinity_cpus": aff,
            "cgroup_quota_cpus": quota, "omp_num_threads": share, "model": model}


def median_of(fn, runs=5):
    """One warm-up call, then the median wall time of `runs` calls (seconds)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline(T, sample_rows, threads, one_thread_rows):
    """Hive-restatement baseline (oracle) on the first fact rows: T map tasks over contiguous
    slices (each with its own dimension tables and partial groups) + the mergepartial step, and
    the same pipeline on one thread over a smaller sample. Median of 5 after a warm-up."""
    import oracle
    from ghive_amd import abi, workloads as W
    fact = [c[:sample_rows].cpu().numpy() for c in T.fact.columns]
    d, s, p = ([c.cpu().numpy() for c in t.columns] for t in (T.date, T.supplier, T.part))
    dims = [
        dict(columns=d, key_col=0, payload_col=1, fact_key_col=0,
             filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
        dict(columns=s, key_col=0, payload_col=1, fact_key_col=2,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
        dict(columns=p, key_col=0, payload_col=2, fact_key_col=3,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
    ]

    def spec(rows):
        return oracle.FragmentSpec([c[:rows] for c in fact], dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)],
                                   aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 5, 6)])

    def med(sp, nt):
        sp.run(nthreads=nt)  # warm-up
        return float(np.median([sp.run(nthreads=nt)["seconds"] for _ in range(5)]))

    secs = med(spec(sample_rows), threads)
    secs1 = med(spec(one_thread_rows), 1)
    rate = sample_rows / secs
    return {"value": rate, "unit": "fact rows/s", "cores": threads, "kind": "port",
            "gbs": rate * W.STAR_REFERENCED_BYTES_PER_ROW / 1e9, "seconds": secs,
            "sample": f"first {sample_rows} fact rows of the same tables; C restatement of Hive 3.1 vectorized "
                      f"map-join + hash group-by as {threads} map tasks on contiguous slices, each with its own "
                      f"dimension hash tables and partial groups, + the mergepartial step; median of 5 after a warm-up",
            "one_thread": {"value": one_thread_rows / secs1, "seconds": secs1,
                           "sample": f"first {one_thread_rows} fact rows, one map task"}}


def cpu_batch_baseline(B, threads, max_rows=200_000_000, one_thread_rows=20_000_000):
    """Per-fragment CPU restatement (oracle FragmentSpec, one thread per fragment = one map task)
    over a pool of `threads` workers, on the first fragments of the batch up to max_rows fact rows;
    and the first fragments up to one_thread_rows on one thread. Median of 5 after a warm-up."""
    import concurrent.futures as cf
    import oracle
    from ghive_amd import abi, workloads as W
    dim = [c.cpu().numpy() for c in B.dim.columns[:3]]

    def first(limit):
        take, total = 0, 0
        while take < len(B.rows) and total + B.rows[take] <= limit:
            total += B.rows[take]
            take += 1
        return take, total

    take, total = first(max_rows)
    end = B.starts[take - 1] + B.rows[take - 1] if take else 0
    fk, me, pr = (c[:end].cpu().numpy() for c in (B.fk, B.measure, B.price))
    specs = []
    for f in range(take):
        s, r, c = B.starts[f], B.rows[f], B.attr_const[f]
        specs.append(oracle.FragmentSpec(
            [fk[s:s + r], me[s:s + r], pr[s:s + r]],
            [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                  filter=([abi.make_pred(abi.FILTER_EQ, 1, [c], data_type=2)], 0))],
            keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)]))
    with cf.ThreadPoolExecutor(threads) as ex:
        secs = median_of(lambda: list(ex.map(lambda sp: sp.run(nthreads=1), specs)))
    take1, total1 = first(one_thread_rows)
    secs1 = median_of(lambda: [sp.run(nthreads=1) for sp in specs[:take1]])
    rate = total / secs
    return {"value": rate, "unit": "fact rows/s", "cores": threads, "kind": "port",
            "gbs": rate * W.SMALL_REFERENCED_BYTES_PER_ROW / 1e9,
            "sample": f"first {take} fragments ({total} fact rows), oracle restatement of Hive 3.1's vectorized "
                      f"map-join + hash group-by, one map task per fragment on a pool of {threads} threads; "
                      f"median of 5 after a warm-up",
            "one_thread": {"value": total1 / secs1, "seconds": secs1,
                           "sample": f"first {take1} fragments ({total1} fact rows) one after another on one thread"}}


def batched_section(ctx, reps, fixed=None, with_cpu=True):
    """4096 lognormal fragments of the small query shape: autotuned persistent kernel replayed from
    a HIP graph vs. 4096 plain per-fragment launches; identical outputs are checked. `fixed` (a
    dict of tile_rows, blocks_per_cu, probe_mode, load_factor) replaces the autotuner's search."""
    import torch
    import numpy as np
    from ghive_amd import autotune, workloads as W
    t0 = time.time()
    B = W.batch_tables(ctx, nfrags=4096)
    torch.cuda.synchronize()
    total = int(sum(B.rows))
    log(f"batched: {len(B.rows)} fragments, {total} rows generated in {time.time() - t0:.1f}s")
    if fixed:
        best, results, search_s = dict(fixed), [], 0.0
    else:
        best, results, search_s = autotune.tune_batch(
            ctx, B, cache_path=os.path.join(ROOT, "gpurun_out", "batch_tuning.json"))
    log(f"batched: autotuner picked {best} after {len(results)} candidates in {search_s:.1f}s")
    ht = W.batch_dim_table(ctx, B, load_factor=best["load_factor"])
    bt = W.batch_create(ctx, B, ht, tile_rows=best["tile_rows"], blocks_per_cu=best["blocks_per_cu"],
                        probe_mode=best["probe_mode"])
    bt.run()
    lat = []
    gc.collect()
    gc.disable()  # as for the plain steps: no collection of the harness's heap inside a timed run
    for _ in range(reps):
        t1 = time.perf_counter()
        bt.run()
        lat.append(1e3 * (time.perf_counter() - t1))
    gc.enable()
    dev_ms = bt.info()["last_ms"]
    sums, counts = bt.results()
    got = [{g: int(sums[f, g]) for g in np.nonzero(counts[f])[0].tolist()} for f in range(len(B.rows))]
    plain = W.PlainBatch(ctx, B)
    plain.run()  # warm-up
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plain.run()
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t1
    identical = plain.run(collect=True) == got
    cpu = None
    if with_cpu:
        cpu = cpu_batch_baseline(B, host_cores()["threads"])
        log(f"batched cpu baseline: {cpu['value']:.3e} rows/s ({cpu['sample']})")
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    out = {
        "workload": "4096 map-task fragments, rows lognormal(median 256k, sigma 1.32) clipped to "
                    "[1k, 16M], per-fragment filter dim.attr == c, join, GROUP BY category SUM(measure*price)",
        "fragments": len(B.rows), "rows": total, "p50_ms": p50, "p99_ms": p99, "device_ms": dev_ms,
        "rows_per_s": total / (p50 * 1e-3), "plain_ms": 1e3 * plain_s,
        "plain_rows_per_s": total / plain_s, "speedup_vs_plain": plain_s / (p50 * 1e-3),
        "identical_to_plain": identical, "cpu_baseline": cpu, "tuning": best, "tuning_candidates": len(results),
        "tuning_search_s": search_s,
        "roofline": {"achieved": W.SMALL_REFERENCED_BYTES_PER_ROW * total / (dev_ms * 1e-3) / 1e9,
                     "unit": "GB/s", "peak": HBM_PEAK_GBS},
    }
    out["roofline"]["frac"] = out["roofline"]["achieved"] / HBM_PEAK_GBS
    log(f"batched: p50 {p50:.3f} ms p99 {p99:.3f} ms (device {dev_ms:.3f} ms), plain {1e3 * plain_s:.1f} ms, "
        f"speedup {out['speedup_vs_plain']:.1f}x, identical={identical}")
    bt.free()
    ht.free()
    del B, plain
    return out


def cpu_skew_baseline(seed, threads):
    """The oracle's restatement of Hive's CPU operators on bounded samples of the skewed-join stages,
    on one thread and on `threads` threads as Tez runs them:
    * map join: 10M Zipf(1.0) probe rows (10% misses) into 1M build keys; the threads probe
      contiguous slices of one shared, read-only table (as LLAP's broadcast-table cache does);
    * ORDER BY (key, tiebreak) of 2M joined rows: one thread sorts everything; at T threads each map
      task sorts its slice and the single ORDER BY reducer merges the T sorted runs;
    * GROUP BY of 4M rows over 400k distinct keys (10 rows a key, as at full size), COUNT(*) and
      SUM(double) in Hive's double arithmetic: at T threads, T map tasks each aggregate a slice into
      their own hash table, their partial groups are partitioned by key over T reducers, and each
      reducer merges its share (mergepartial).
    Each figure is the median of 5 after a warm-up; build time of the probe's table is reported
    apart."""
    import concurrent.futures as cf
    import numpy as np
    import torch
    import oracle
    from ghive_amd import abi, workloads as W
    nb, npr, ns, ng, nk = 1_000_000, 10_000_000, 2_000_000, 4_000_000, 400_000
    T = threads
    split = lambda n: [(n * t // T, n * (t + 1) // T) for t in range(T)]
    bk = W.key62_torch(seed, torch.arange(1, nb + 1, dtype=torch.int64)).numpy()
    bp = np.arange(1, nb + 1, dtype=np.int64)
    pk, _ = oracle.gen_zipf_keys(seed, 0, 0, npr, nb, 1.0, 0.1)
    t0 = time.perf_counter()
    mj = oracle.MapJoinTable(bk, bp)
    t_build = time.perf_counter() - t0
    rows, _ = oracle.mapjoin_inner(bk, pk[: 4 * ns], payload=bp)
    tb = oracle.gen_uniform_f32(seed, 1, 0, 4 * ns)
    sel = rows[:ns].astype(np.int64)
    jk, jt = np.ascontiguousarray(pk[sel]), np.ascontiguousarray(tb[sel])
    keys = oracle.gen_uniform_i64(seed, 2, 0, ng, 1, nk)
    vals = oracle.gen_normal_f64(seed, 3, 0, ng)
    aggs = [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 1)]
    out = {"cores": T, "kind": "port"}
    with cf.ThreadPoolExecutor(T) as ex:
        probe1 = median_of(lambda: mj.probe(pk))
        probeT = median_of(lambda: list(ex.map(lambda r: mj.probe(pk[r[0]:r[1]]), split(npr))))
        sort1 = median_of(lambda: oracle.sort_perm([jk, jt], [0, 1], "++"))
        phases = []
        sortT = median_of(lambda: phases.append(oracle.sort_runs([jk, jt], [0, 1], "++", T)[1:]))
        gby1 = median_of(lambda: oracle.groupby([keys, vals], [0], aggs, exact_double_sum=False))

        def tez_groupby():
            parts = [None] * T

            def map_task(t):
                b, e = split(ng)[t]
                k = keys[b:e]
                rep, (cnt, sm), _ = oracle.groupby([k, vals[b:e]], [0], aggs, exact_double_sum=False)
                gk = k[rep]
                dest = gk % T
                order = np.argsort(dest, kind="stable")
                parts[t] = (gk[order], cnt[order], sm[order], np.searchsorted(dest[order], np.arange(T + 1)))

            def reduce_task(r):
                cat = lambda i: np.concatenate([p[i][p[3][r]:p[3][r + 1]] for p in parts])
                oracle.groupby([cat(0), cat(1), cat(2)], [0], [(abi.AGG_SUM, 1), (abi.AGG_SUM, 2)],
                               exact_double_sum=False)

            list(ex.map(map_task, range(T)))
            list(ex.map(reduce_task, range(T)))

        gbyT = median_of(tez_groupby)
    mj.free()
    ph = np.median(np.array(phases), axis=0)
    out.update({
        "probe_rows_per_s": npr / probeT, "probe_rows_per_s_one_thread": npr / probe1, "build_ms_one_thread": 1e3 * t_build,
        "sort_rows_per_s": ns / sortT, "sort_rows_per_s_one_thread": ns / sort1,
        "sort_phases_ms": {"map_side_run_sorts": 1e3 * float(ph[0]), "reducer_merge": 1e3 * float(ph[1])},
        "groupby_rows_per_s": ng / gbyT, "groupby_rows_per_s_one_thread": ng / gby1,
        "sample": "oracle (C restatement of Hive 3.1 VectorMapJoinFastLongHashTable, ORDER BY as a total order, "
                  f"hash group-by) at {T} threads as Tez map tasks and at one thread: 10M Zipf probe rows x 1M "
                  "build keys (probe only); 2M-row ORDER BY (map-side run sorts + one merging reducer); 4M rows / "
                  "400k keys GROUP BY COUNT(*), SUM(double) (map-side hash aggregation + mergepartial over "
                  f"{T} reducers); median of 5 after a warm-up",
    })
    return out


def skew_section(ctx, reps, with_cpu=True, gby_form="auto"):
    """Config 3 at full size: 64M-key build, 1e9 Zipf(1.0) probe rows with 10% misses, ORDER BY
    (key, float tiebreak) over the joined rows, and GROUP BY over 1e9 rows with 100M distinct int64
    keys (COUNT(*), exact SUM(double)). Wall time around each synchronous call, median of reps;
    GB/s over the algorithmic bytes of each stage."""
    import torch
    import numpy as np
    from ghive_amd import abi, ops, workloads as W

    reps_ms = {}  # every repetition's wall time, per stage

    def timed(fn, name):
        ts, out = [], None
        for _ in range(reps):
            out = None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        reps_ms[name] = [round(1e3 * t, 3) for t in ts]
        return float(np.median(ts)), out

    S = W.skew_tables(ctx)
    nb, npr = S.build_keys.numel(), S.probe_keys.numel()
    t_build, ht = timed(lambda: ops.HashTable(ctx, S.build_keys, payload=S.build_pay, load_factor=0.5), "build")
    t_probe, (rows, pay) = timed(lambda: ht.probe(S.probe_keys, capacity=npr), "probe")
    matches = rows.numel()
    ht.free()
    r64 = rows.to(torch.int64) & 0xFFFFFFFF
    jk = S.probe_keys[r64].contiguous()
    jt = S.tiebreak[r64].contiguous()
    del rows, pay, r64, S
    t_sort, perm = timed(lambda: ops.sort(ctx, ops.Table([jk, jt]), [0, 1], ["+", "+"]), "sort")
    si = ops.sort_info(ctx)
    del perm, jk, jt
    torch.cuda.empty_cache()
    keys, vals = W.highcard_tables(ctx)
    ng = [0]

    def gby():
        outs = ops.GroupByPredicate([0], [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 1)]).gpu_execute(
            ctx, ops.Table([keys, vals]), capacity=110_000_000)
        ng[0] = outs[0].numel()
        return None

    # the group-by form: chosen by the library (auto), or forced to the global-table form (direct)
    ctx.set_option("groupby.partition", {"auto": 1, "direct": 0, "partitioned": 2}[gby_form])
    fb0 = ctx.get_option("groupby.partition_fallbacks")
    t_gby, _ = timed(gby, "groupby")
    fallbacks = ctx.get_option("groupby.partition_fallbacks") - fb0
    ctx.set_option("groupby.partition", 1)
    nrows = keys.numel()
    del keys, vals
    torch.cuda.empty_cache()
    b_build = nb * 16
    b_probe = npr * 8 + matches * 12          # probe keys in; (probe row id, payload) out
    b_sort = matches * 16                      # lower bound: 12 B of keys in, 4 B of permutation out
    b_gby = nrows * 16 + ng[0] * 24            # key + value in; key, COUNT, SUM out
    gbs = lambda b, t: b / t / 1e9
    out = {
        "workload": "64M unique int64 build keys (int64 payload) x 1e9 probe rows, Zipf(1.0) over build "
                    "ranks + 10% misses; ORDER BY key, float32 tiebreak over the joined rows; GROUP BY "
                    "1e9 rows, 100M distinct int64 keys, COUNT(*), SUM(double N(0,1)) exact",
        "reps_ms": reps_ms,  # the stage times above are the medians of these
        "build": {"rows": nb, "ms": 1e3 * t_build, "rows_per_s": nb / t_build, "gbs": gbs(b_build, t_build)},
        "probe": {"rows": npr, "matches": matches, "ms": 1e3 * t_probe, "rows_per_s": npr / t_probe,
                  "gbs": gbs(b_probe, t_probe), "frac": gbs(b_probe, t_probe) / HBM_PEAK_GBS},
        # build + probe as one stage: probe keys in, build (key, payload) in, matches out as this
        # build writes them (4-byte probe row id + 8-byte payload); bytes_16 counts each match as a
        # 16-byte (key, payload) pair, the form of the join's algorithmic bytes in its specification
        "join": {"ms": 1e3 * (t_build + t_probe), "bytes": b_build + b_probe,
                 "gbs": gbs(b_build + b_probe, t_build + t_probe),
                 "frac": gbs(b_build + b_probe, t_build + t_probe) / HBM_PEAK_GBS,
                 "bytes_16": npr * 8 + nb * 16 + matches * 16},
        "sort": {"rows": matches, "ms": 1e3 * t_sort, "rows_per_s": matches / t_sort,
                 "gbs_lower_bound_bytes": gbs(b_sort, t_sort),
                 # P x 2 x 16 B x N: P passes, each reading and writing a 16-byte (key, row id) record
                 "passes": si["passes"], "key_words": si["key_words"], "embedded_ids": bool(si["embedded_ids"]),
                 "gbs_passes": gbs(si["passes"] * 32 * matches, t_sort),
                 "frac": gbs(si["passes"] * 32 * matches, t_sort) / HBM_PEAK_GBS,
                 "scatter_bytes_moved": si["scatter_bytes"],
                 "frac_bytes_moved": gbs(si["scatter_bytes"], t_sort) / HBM_PEAK_GBS},
        "groupby": {"rows": nrows, "groups": ng[0], "form": gby_form, "ms": 1e3 * t_gby, "rows_per_s": nrows / t_gby,
                    "gbs": gbs(b_gby, t_gby), "frac": gbs(b_gby, t_gby) / HBM_PEAK_GBS,
                    "partition_fallbacks": fallbacks},  # repetitions that ran the direct form after the partitioned one
    }
    out["cpu_baseline"] = None
    if with_cpu:
        out["cpu_baseline"] = cpu_skew_baseline(W.SEED_BASE + 3, host_cores()["threads"])
        log(f"skew cpu baseline: {out['cpu_baseline']}")
    log(f"skew: build {1e3 * t_build:.1f} ms, probe {1e3 * t_probe:.1f} ms ({out['probe']['gbs']:.0f} GB/s), "
        f"sort {1e3 * t_sort:.1f} ms of {matches} rows ({si['passes']} passes, frac {out['sort']['frac']:.3f}), group-by {1e3 * t_gby:.1f} ms "
        f"({out['groupby']['gbs']:.0f} GB/s, {ng[0]} groups)")
    return out


def _digest(kind, outs):
    """A 64-bit digest of a request's result, computed on the device: for a sort the permutation in
    order; for group rows (fragment, group-by) a sum over the rows of a mix of each row's values, so
    that rows in any order give the same digest."""
    import torch
    m = (1 << 63) - 1
    c1, c2, c3 = 0x1E3779B97F4A7C15, 0x3F58476D1CE4E5B9, 0x14D049BB133111EB
    if kind == "sort":
        p = outs[0].to(torch.int64)
        i = torch.arange(p.numel(), device=p.device, dtype=torch.int64)
        return int(((p * c1) ^ (i * c2)).sum().item()) & m, int(p.numel())
    h = torch.zeros(outs[0].numel(), dtype=torch.int64, device=outs[0].device)
    for j, o in enumerate(outs):
        x = (o.view(torch.int64) if o.dtype == torch.float64 else o.to(torch.int64)) * (c1 + 2 * j)
        h = (h ^ x) * c2
        h = h ^ ((h >> 29) & ((1 << 35) - 1))
    h = h * c3
    return int(h.sum().item()) & m, int(outs[0].numel())


def cpu_serving_baseline(I, reqs, max_sort_rows=16_000_000):
    """The same requests run one after another on the CPU restatement (oracle, one thread each):
    throughput only. Sorts above max_sort_rows are left out of the sample (stated)."""
    import oracle
    from ghive_amd import abi
    fact = [c.cpu().numpy() for c in I.small.fact.columns]
    dim = [c.cpu().numpy() for c in I.small.dim.columns]
    gk, gv = I.gb_keys.cpu().numpy(), I.gb_vals.cpu().numpy()
    sk = I.sort_keys[:max_sort_rows].cpu().numpy()
    rows, secs, done, skipped = 0, 0.0, 0, 0
    for kind, p in reqs:
        if kind == "sort" and p["rows"] > max_sort_rows:
            skipped += 1
            continue
        t0 = time.perf_counter()
        if kind == "fragment":
            s, r = p["start"], p["rows"]
            oracle.FragmentSpec([c[s:s + r] for c in fact],
                                [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                      filter=([abi.make_pred(abi.FILTER_EQ, 1, [p["attr"]], data_type=2)], 0))],
                                keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)]).run(nthreads=1)
        elif kind == "groupby":
            n = p["rows"]
            oracle.groupby([gk[:n], gv[:n]], [0], [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)])
        else:
            oracle.sort_perm([sk[:p["rows"]]], [0], "+")
        secs += time.perf_counter() - t0
        rows += p["rows"]
        done += 1
    return {"value": rows / secs if secs else None, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"{done} requests of the mix run sequentially on the oracle restatement, one thread"
                      f" ({skipped} sorts above {max_sort_rows} rows left out)"}


def serving_section(ctx, rates, seconds, slots):
    """Serving form: requests of the mix (60% config-1 fragments, 25% group-bys, 15% sorts) arrive
    as a Poisson process at each rate and go through ghive_server_*; latency = completion - submit
    (host steady clock), p50/p99 per rate, sustained fact rows/s; the first requests are checked
    against the plain entry points."""
    import threading
    import torch
    import numpy as np
    from ghive_amd import workloads as W
    t0 = time.time()
    I = W.serving_inputs(ctx)
    torch.cuda.synchronize()
    log(f"serving: resident inputs ready in {time.time() - t0:.1f}s")
    srv = ops_server(slots)
    per_rate = []
    served = []  # (kind, params, digest of the served result) of every request at every rate
    for rate in rates:
        n = max(int(rate * seconds), 1)
        reqs = W.serving_requests(n, 1000 + rate, I)
        gaps = np.random.default_rng(rate).exponential(1.0 / rate, n)
        tickets = [None] * n
        rows = [0] * n

        def submitter():
            t_next = time.perf_counter()
            for i, (k, p) in enumerate(reqs):
                t_next += gaps[i]
                d = t_next - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i], rows[i] = W.serving_submit(srv, I, k, p)

        th = threading.Thread(target=submitter)
        t_start = time.perf_counter()
        th.start()
        lat = {"fragment": [], "groupby": [], "sort": []}
        first_submit, last_end = None, None
        for i in range(n):
            while tickets[i] is None:
                time.sleep(0.0002)
            outs = srv.wait(tickets[i])
            served.append((reqs[i][0], reqs[i][1], _digest(reqs[i][0], outs)))
            del outs
            r = tickets[i].response
            lat[reqs[i][0]].append((r.end_us - r.submit_us) * 1e-3)
            first_submit = r.submit_us if first_submit is None else min(first_submit, r.submit_us)
            last_end = r.end_us if last_end is None else max(last_end, r.end_us)
            tickets[i] = True  # drop the outputs
        th.join()
        wall = time.perf_counter() - t_start
        span_s = (last_end - first_submit) * 1e-6
        all_lat = sorted(x for v in lat.values() for x in v)
        st = srv.stats()
        entry = {"offered_req_per_s": rate, "requests": n, "seconds": wall,
                 "p50_ms": float(np.percentile(all_lat, 50)), "p99_ms": float(np.percentile(all_lat, 99)),
                 "served_req_per_s": n / span_s, "rows_per_s": sum(rows) / span_s,
                 "p99_ms_by_kind": {k: float(np.percentile(v, 99)) for k, v in lat.items() if v},
                 "p50_ms_by_kind": {k: float(np.percentile(v, 50)) for k, v in lat.items() if v},
                 "peak_inflight": st["peak_inflight"], "peak_inflight_gb": st["peak_inflight_bytes"] / 1e9}
        log(f"serving @{rate} req/s: p50 {entry['p50_ms']:.2f} ms p99 {entry['p99_ms']:.2f} ms, "
            f"{entry['rows_per_s']:.3e} rows/s, peak in flight {st['peak_inflight']}")
        per_rate.append(entry)
    srv.close()
    # every served result against the same request run alone through the plain entry point
    mismatches = 0
    for kind, p, dg in served:
        mismatches += _digest(kind, W.serving_plain(ctx, I, kind, p)) != dg
    identical = mismatches == 0
    log(f"serving: {len(served)} served results checked against the plain form, {mismatches} differ")
    cpu = cpu_serving_baseline(I, W.serving_requests(40, 7, I))
    log(f"serving cpu baseline: {cpu}")
    del I
    torch.cuda.empty_cache()
    return {"workload": "Poisson arrivals; 60% config-1 fragments (1k..16M rows, lognormal median 256k), "
                        "25% group-by SUM/COUNT (1k..16M rows, <= 10M distinct int64 keys), 15% sort "
                        "(1M..256M int64 keys, lognormal median 8M); dimension tables resident",
            "slots": slots, "identical_to_plain": identical, "requests_checked": len(served),
            "rates": per_rate, "cpu_baseline": cpu}


def ops_server(slots):
    from ghive_amd import ops
    return ops.Server(slots=slots, hbm_budget=96 << 30, slot_mem_limit=32 << 30)


def dropin_section(ctx, T, want_groups, reps):
    """The same star-join query through the drop-in boundary: the authored Hive plan of the query
    (tests/golden/plans_authored/star3*), its three dimension vertices and then its fact vertex run
    by ghive_gpu_process on GTables host buffers -- every integer column LONG, as Hive's vectorized
    reader hands them over. The fact vertex runs as one fused fragment; its LONG columns are sent as
    32-bit integers. Reported: staging (host copies + PCIe), execution on the device (hash builds,
    fused kernel, write-out), result marshalling, and the fused kernel alone."""
    import torch
    from ghive_amd import ops, workloads as W
    d = os.path.join(ROOT, "tests", "golden", "plans_authored")
    with open(os.path.join(d, "star3.txt")) as a, open(os.path.join(d, "star3_extended.txt")) as b:
        plan = ops.Plan(a.read(), b.read())
    host = lambda t: t.to(torch.int64).cpu().numpy()
    t0 = time.time()
    dims = {}
    for v, tbl, ncols in (("Map 2", T.date, 2), ("Map 3", T.supplier, 2), ("Map 4", T.part, 3)):
        cols, _, _ = ops.gpu_process(ctx, plan, [(v, [host(c) for c in tbl.columns[:ncols]])], v)
        dims[v] = cols
    fact = [host(T.fact.columns[c]) for c in (W.F_DATE, W.F_SUPP, W.F_PART, W.F_PRICE, W.F_DISC)]
    rows = len(fact[0])
    log(f"drop-in: LONG host columns ready in {time.time() - t0:.1f}s")
    tables = [("Map 1", fact)] + [(v, dims[v]) for v in ("Map 2", "Map 3", "Map 4")]
    runs = []
    for r in range(reps + 1):
        cols, _, _ = ops.gpu_process(ctx, plan, tables, "Map 1")
        pt, info = ops.process_times(ctx), ctx.plan_info()
        if r:
            runs.append(dict(pt, fused_kernel_ms=info["fused_kernel_ms"]))
    got = {(int(y), int(b)): int(v) for y, b, v in zip(*cols)}
    med = lambda k: float(np.median([x[k] for x in runs]))
    out = {"workload": "the 600M-row star join as Hive's plan vertex 'Map 1' through ghive_gpu_process: 5 LONG "
                       "fact columns + 3 filtered dimension tables in GTables host buffers",
           "rows": rows, "fused": bool(runs[-1]["fused"]), "narrowed_columns": runs[-1]["narrowed_columns"],
           "h2d_bytes": runs[-1]["h2d_bytes"], "stage_ms": med("stage_ms"), "execute_ms": med("execute_ms"),
           "result_ms": med("result_ms"), "total_ms": med("total_ms"), "fused_kernel_ms": med("fused_kernel_ms"),
           "execute_rows_per_s": rows / (med("execute_ms") * 1e-3),
           "pcie_inclusive_rows_per_s": rows / (med("total_ms") * 1e-3),
           "identical_to_plain": got == want_groups, "reps": reps}
    log(f"drop-in: {out}")
    del fact, tables
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=600_000_000)
    ap.add_argument("--late", action="store_true", help="late-materialising kernel variant")
    ap.add_argument("--no-dense", action="store_true", help="probe every dimension through its hash table")
    ap.add_argument("--no-lds-bits", action="store_true", help="keep dimension bitmaps in HBM/L2")
    ap.add_argument("--rows-per-thread", type=int, default=0, help="fused kernel rows/thread (4 or 8)")
    ap.add_argument("--generic", action="store_true", help="general fused kernel (no dense-probe form)")
    ap.add_argument("--no-pipe", action="store_true", help="dense-probe kernel without the pipelined tile loop")
    ap.add_argument("--grid", type=int, default=0, help="fused kernel workgroups (0 = resident wave)")
    ap.add_argument("--block", type=int, default=0, help="fused kernel workgroup size (256/512/1024)")
    ap.add_argument("--cpu-sample", type=int, default=600_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 4096-fragment batched form")
    ap.add_argument("--no-dropin", action="store_true", help="skip the run through ghive_gpu_process")
    ap.add_argument("--dropin-reps", type=int, default=2)
    ap.add_argument("--batched-reps", type=int, default=50)
    ap.add_argument("--batched-config", default="",
                    help="tile_rows,blocks_per_cu,probe_mode,load_factor instead of the autotuner's search")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving form")
    ap.add_argument("--no-skew", action="store_true", help="skip the skewed join / sort / group-by stages")
    ap.add_argument("--skew-reps", type=int, default=3)
    ap.add_argument("--sort-shape", type=int, default=-1, help="LSD tile of the sort (option sort.shape)")
    ap.add_argument("--sort-pack", type=int, default=-1, help="16-byte LDS records in the sort (option sort.pack)")
    ap.add_argument("--sort-occupancy", type=int, default=0,
                    help="workgroups per CU of the sort's 8-byte-staging passes (option sort.occupancy)")
    ap.add_argument("--part-threads", type=int, default=0, help="group-by partition workgroup size (256/512)")
    ap.add_argument("--groupby-form", default="auto", choices=["auto", "direct", "partitioned"],
                    help="form of the 1e9-row / 100M-key group-by (direct = one global hash table)")
    ap.add_argument("--serving-rates", default="50,200,1000,2000", help="offered request rates (req/s)")
    ap.add_argument("--serving-seconds", type=float, default=1.5)
    ap.add_argument("--serving-slots", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the fragment shards "
              "by fact row range with replicated dimension tables", file=sys.stderr)
        sys.exit(2)

    import torch
    from ghive_amd import ops, workloads as W

    ctx = ops.Context(0)
    if args.late:
        ctx.set_option("fragment.late", 1)
    if args.no_dense:
        ctx.set_option("fragment.dense", 0)
    if args.no_lds_bits:
        ctx.set_option("fragment.lds_bits", 0)
    if args.rows_per_thread:
        ctx.set_option("fragment.rows_per_thread", args.rows_per_thread)
    if args.grid:
        ctx.set_option("fragment.grid", args.grid)
    if args.block:
        ctx.set_option("fragment.block", args.block)
    if args.generic:
        ctx.set_option("fragment.generic", 1)
    if args.no_pipe:
        ctx.set_option("fragment.pipe", 0)
    if args.sort_shape >= 0:
        ctx.set_option("sort.shape", args.sort_shape)
    if args.sort_pack >= 0:
        ctx.set_option("sort.pack", args.sort_pack)
    if args.sort_occupancy:
        ctx.set_option("sort.occupancy", args.sort_occupancy)
    if args.part_threads:
        ctx.set_option("groupby.part_threads", args.part_threads)
    variant = ("late" if args.late else "early") + ("/hash" if args.no_dense else
                                                    ("/hbm-bitmap" if args.no_lds_bits else "/lds-bitmap"))
    if args.generic or args.late or args.no_dense:
        variant += "/generic-kernel"
    elif args.no_pipe or args.rows_per_thread == 8:
        variant += "/dense-kernel"
    else:
        variant += "/pipe-kernel"
    info = ctx.device_info()
    log(f"device {info['name']} {info['arch']} CUs={info['compute_units']}")
    t0 = time.time()
    T = W.star_tables(ctx, args.rows)
    torch.cuda.synchronize()
    log(f"generated {args.rows} fact rows in {time.time() - t0:.1f}s")

    host = {"filters_builds": [], "fragment": [], "free": []}
    # prepared once, like a compiled plan: the three build sides' filters, the fragment descriptor
    # (rebound to each step's tables) and the result buffers; every step still filters, builds,
    # runs the fused kernel and reads its result back
    builds = W.star_builds(T)
    frag, outs = None, None

    def step():
        nonlocal frag, outs
        t_a = time.perf_counter()
        hd, hs, hp = ops.filtered_hash_tables(ctx, builds)
        t_b = time.perf_counter()
        if frag is None:
            frag = W.star_fragment(T.fact, hd, hs, hp)
            outs = [torch.empty(8192, dtype=torch.int32, device="cuda") for _ in range(2)] + \
                   [torch.empty(8192, dtype=torch.int64, device="cuda")]
        else:
            frag.rebind([hd, hs, hp])
        out = frag.execute(ctx, capacity=8192, outs=outs)
        t_c = time.perf_counter()
        for h in (hd, hs, hp):
            h.free()
        t_d = time.perf_counter()
        for k, v in zip(host, (t_b - t_a, t_c - t_b, t_d - t_c)):
            host[k].append(1e3 * v)
        return out

    # The harness's own Python objects (torch's included) are moved out of the collector's reach and
    # collection is paused while steps are timed: a full collection of that heap takes ~40 ms and
    # would otherwise land inside some step. The library itself has no Python on its path.
    gc.collect()
    gc.freeze()
    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    for v in host.values():
        v.clear()
    gc.disable()
    kernel_ms, fin_ms, step_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        out = step()
        stt = ctx.stage_times()
        kernel_ms.append(stt["main_kernel_ms"])
        fin_ms.append(stt["finalize_ms"])
        step_ms.append(1e3 * (time.perf_counter() - ts))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gc.enable()
    breakdown = {"host_ms_median": {k: float(np.median(v)) for k, v in host.items()},
                 "host_ms_max": {k: float(np.max(v)) for k, v in host.items()},
                 "step_ms": [round(x, 3) for x in step_ms],
                 "device_ms_mean": {"fused_kernel": float(np.mean(kernel_ms)),
                                    "group_write_out": float(np.mean(fin_ms))},
                 "note": "Python's collector paused during the timed steps (a full collection of the harness's "
                         "heap takes ~40 ms); each step: the three dimension tables filtered and built by one "
                         "ghive_hash_build_filtered call (predicates prepared once), the fragment descriptor "
                         "rebound to them, the fused fragment into result buffers allocated once; host_ms: wall "
                         "time of each call of the step (the builds are submitted without a host wait; the "
                         "fragment call holds the step's host waits: one read-back of the three builds' "
                         "statistics and one of the group count); device_ms: HIP events on the context stream"}
    log(f"step breakdown: {breakdown}")
    st = ctx.stage_times()
    groups = int(out[0].numel())
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.rows * args.steps / elapsed
    k_ms = float(np.mean(kernel_ms))
    alg_bytes = W.STAR_REFERENCED_BYTES_PER_ROW * args.rows
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    log(f"step {ms_per_step:.3f} ms, fused kernel {k_ms:.3f} ms, joined {st['rows_joined']}, groups {groups}")

    read_gbs, _ = ops.bench_stream(ctx, 8 << 30, mode=0, iters=10)
    copy_gbs, _ = ops.bench_stream(ctx, 4 << 30, mode=1, iters=10)
    # the fused kernel's own access pattern with no work: its five fact columns streamed side by side
    cols_gbs, _ = ops.bench_stream(ctx, alg_bytes - alg_bytes % 80, mode=2, iters=10)
    log(f"stream probe: read {read_gbs:.0f} GB/s, copy {copy_gbs:.0f} GB/s (read+write), "
        f"5 columns side by side {cols_gbs:.0f} GB/s")

    traffic = None
    prof = load_traffic("star600m_fragment_traffic.json")
    if prof and prof.get("rows") == args.rows and prof.get("variant") == variant:
        traffic = prof.get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        hc = host_cores()
        sample = min(args.cpu_sample, args.rows)
        cpu = cpu_baseline(T, sample, hc["threads"], min(sample, 60_000_000))
        cpu["host"] = hc
        log(f"cpu baseline: {cpu}")

    plain_groups = {(int(y), int(b)): int(v) for y, b, v in zip(*[o.cpu().numpy() for o in out])}
    dropin = None
    if not args.no_dropin:
        dropin = dropin_section(ctx, T, plain_groups, args.dropin_reps)

    batched = None
    T = None  # the 600M-row tables are not needed past this point
    torch.cuda.empty_cache()
    if not args.no_batched:
        fixed = None
        if args.batched_config:
            tr, bpc, mode, lf = args.batched_config.split(",")
            fixed = dict(tile_rows=int(tr), blocks_per_cu=int(bpc), probe_mode=int(mode), load_factor=float(lf))
        batched = batched_section(ctx, args.batched_reps, fixed, with_cpu=not args.no_cpu_baseline)

    skew = None
    if not args.no_skew:
        skew = skew_section(ctx, args.skew_reps, with_cpu=not args.no_cpu_baseline, gby_form=args.groupby_form)

    serving = None
    if not args.no_serving:
        serving = serving_section(ctx, [int(r) for r in args.serving_rates.split(",")], args.serving_seconds,
                                  args.serving_slots)

    line = {
        "metric": METRIC, "value": value, "unit": "fact rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": "star-join fragment, plain form: 600M-row fact x date(2556) x "
                               "supplier(200k) x part(1.4M), GROUP BY year, brand SUM(price*discount)",
                   "fact_rows": args.rows, "groups": groups, "rows_joined": st["rows_joined"],
                   "kernel_variant": variant,
                   "fused_kernel_ms": k_ms, "step_breakdown": breakdown},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "measured_read_peak": read_gbs, "frac_of_measured_read_peak": achieved / read_gbs,
                     "measured_copy_peak": copy_gbs, "measured_5col_read_peak": cols_gbs,
                     "frac_of_measured_5col_read_peak": achieved / cols_gbs},
        "cpu_baseline": cpu,
        "dropin": dropin,
        "batched": batched,
        "skew": skew,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
