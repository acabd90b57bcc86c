"""GPU parity of the fused fragment (filter -> star joins -> int64 expression -> GROUP BY) against
the oracle's restatement of Hive's vectorized map-join pipeline: group keys and int64 SUMs
bit-exact. At 600M rows (too large for the oracle) the checks are size-independent: slice
linearity of SUM, the dense-key closed form, and agreement of the two kernel variants."""
import numpy as np
import pytest

from conftest import to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu


def _gpu_dict(outs, nkeys):
    keys = [to_np(o) for o in outs[:nkeys]]
    aggs = [to_np(o) for o in outs[nkeys:]]
    return {tuple(int(k[g]) for k in keys): tuple(int(a[g]) for a in aggs) for g in range(len(aggs[0]))}


def _orc_dict(res):
    return {tuple(int(x) for x in k): tuple(int(x) for x in v) for k, v in zip(res["keys"], res["aggs"])}


def test_small_fragment_full_size_vs_oracle(ctx, orc):
    from ghive_amd import workloads as W
    T = W.small_tables(ctx, 6_000_000)
    ht = W.small_dim_table(ctx, T.dim)
    frag = W.small_fragment(T.fact, ht)
    got = _gpu_dict(frag.execute(ctx, capacity=64), 1)
    joined = ctx.stage_times()["rows_joined"]
    got_late = _gpu_dict(frag.execute(ctx, capacity=64, late=True), 1)
    fact = [to_np(c) for c in T.fact.columns]
    dim = [to_np(c) for c in T.dim.columns]
    spec = orc.FragmentSpec(fact, [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                        filter=([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0))],
                            keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)])
    want = spec.run(nthreads=8)
    assert got == _orc_dict(want) == got_late
    assert joined == want["joined"]
    assert len(got) == 25


def test_small_fragment_fact_filter_count_min_max(ctx, orc):
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 1_000_003, seed=99)
    ht = W.small_dim_table(ctx, T.dim, attr_value=5)
    pred = ops.FilterOperator("(c1 > 50 or c2 < 200000)", ["c0", "c1", "c2"]).generatePredicate()
    aggs = [(abi.AGG_SUM, abi.EXPR_COL, 2, 0), (abi.AGG_CNT, abi.EXPR_ONE, 0, 0),
            (abi.AGG_MIN, abi.EXPR_COL, 1, 0), (abi.AGG_MAX, abi.EXPR_MUL, 1, 2)]
    frag = ops.Fragment(T.fact, joins=[(0, ht)], keys=[(abi.SRC_JOIN, 0)], aggs=aggs, fact_filter=pred)
    got = _gpu_dict(frag.execute(ctx, capacity=64), 1)
    fact = [to_np(c) for c in T.fact.columns]
    dim = [to_np(c) for c in T.dim.columns]
    spec = orc.FragmentSpec(fact, [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                        filter=([abi.make_pred(abi.FILTER_EQ, 1, [5], data_type=2)], 0))],
                            keys=[(abi.SRC_JOIN, 0)], aggs=aggs, fact_filter=(pred.nodes, pred.root))
    assert got == _orc_dict(spec.run(nthreads=4))


def _star_oracle(orc, T, rows=None):
    fact = [to_np(c[:rows] if rows else c) for c in T.fact.columns]
    d, s, p = ([to_np(c) for c in t.columns] for t in (T.date, T.supplier, T.part))
    dims = [
        dict(columns=d, key_col=0, payload_col=1, fact_key_col=0,
             filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
        dict(columns=s, key_col=0, payload_col=1, fact_key_col=2,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
        dict(columns=p, key_col=0, payload_col=2, fact_key_col=3,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
    ]
    spec = orc.FragmentSpec(fact, dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)],
                            aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 5, 6)])
    return spec.run(nthreads=8, capacity=1 << 14)


_DEFAULTS = {"fragment.dense": 1, "fragment.lds_bits": 1, "fragment.generic": 0,
             "fragment.rows_per_thread": 4, "fragment.block": 256, "fragment.pipe": 1}


@pytest.mark.parametrize("opts", [
    {},                                                        # pipelined kernel, LDS bitmaps
    {"fragment.lds_bits": 0},                                  # pipelined kernel, HBM bitmaps
    {"fragment.pipe": 0},                                      # dense-probe kernel, LDS bitmaps
    {"fragment.pipe": 0, "fragment.lds_bits": 0},              # dense-probe kernel, HBM bitmaps
    {"fragment.generic": 1},                                   # general kernel, LDS bitmaps
    {"fragment.generic": 1, "fragment.lds_bits": 0},           # general kernel, HBM bitmaps
    {"fragment.dense": 0},                                     # hash-table probes
    {"fragment.rows_per_thread": 8, "fragment.block": 512},
    {"fragment.block": 1024, "fragment.pipe": 0},
], ids=["pipe-lds", "pipe-hbm", "dense-lds", "dense-hbm", "generic-lds", "generic-hbm", "hash", "r8-b512",
        "b1024"])
def test_star_fragment_vs_oracle_20m(ctx, orc, opts):
    """Every fused kernel, every probe mode (LDS bitmap, HBM bitmap, hash table), both load orders
    and several tile shapes agree with the oracle bit for bit; 20M+1 rows leaves a partial tile."""
    from ghive_amd import workloads as W
    T = W.star_tables(ctx, 20_000_001)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    frag = W.star_fragment(T.fact, hd, hs, hp)
    for k, v in opts.items():
        ctx.set_option(k, v)
    try:
        got = _gpu_dict(frag.execute(ctx, capacity=8192), 2)
        joined = ctx.stage_times()["rows_joined"]
        got_late = _gpu_dict(frag.execute(ctx, capacity=8192, late=True), 2)
    finally:
        for k in opts:
            ctx.set_option(k, _DEFAULTS[k])
    want = _star_oracle(orc, T)
    assert got == _orc_dict(want) == got_late
    assert joined == want["joined"]


@pytest.mark.parametrize("rows", [1, 3, 1023, 1024, 1025, 2048, 3072, 4096, 262_144, 262_147])
def test_star_fragment_tile_boundaries(ctx, orc, rows):
    """Tables smaller than, equal to and just past whole tiles (1024 rows at the default shape):
    the partial last tile goes through the staging copy, whole tiles do not; odd and even numbers
    of whole tiles per workgroup end the pipelined loop in either register set."""
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 300_000)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    part = ops.Table([c[:rows] for c in T.fact.columns])
    got = _gpu_dict(W.star_fragment(part, hd, hs, hp).execute(ctx, capacity=8192), 2)
    res = {}
    for opt in ("fragment.generic", "fragment.pipe"):
        ctx.set_option(opt, 1 - _DEFAULTS[opt])
        try:
            res[opt] = _gpu_dict(W.star_fragment(part, hd, hs, hp).execute(ctx, capacity=8192), 2)
        finally:
            ctx.set_option(opt, _DEFAULTS[opt])
    assert got == res["fragment.generic"] == res["fragment.pipe"] == _orc_dict(_star_oracle(orc, T, rows))


@pytest.mark.parametrize("grid", [1, 3, 7])
def test_star_fragment_few_workgroups(ctx, orc, grid):
    """A handful of workgroups, each streaming hundreds of tiles through the pipelined loop."""
    from ghive_amd import workloads as W
    T = W.star_tables(ctx, 1_000_037)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    ctx.set_option("fragment.grid", grid)
    try:
        got = _gpu_dict(W.star_fragment(T.fact, hd, hs, hp).execute(ctx, capacity=8192), 2)
        joined = ctx.stage_times()["rows_joined"]
    finally:
        ctx.set_option("fragment.grid", 0)
    want = _star_oracle(orc, T)
    assert got == _orc_dict(want)
    assert joined == want["joined"]


@pytest.mark.parametrize("generic,pipe", [(0, 1), (0, 0), (1, 0)])
def test_star_fragment_count_min_max(ctx, orc, generic, pipe):
    """COUNT(*), SUM, MIN and MAX of single columns (unused aggregate inputs) in every kernel."""
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 2_000_003)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    aggs = [(abi.AGG_CNT, abi.EXPR_ONE, 0, 0), (abi.AGG_SUM, abi.EXPR_COL, 5, 0),
            (abi.AGG_MIN, abi.EXPR_COL, 6, 0), (abi.AGG_MAX, abi.EXPR_COL, 4, 0)]
    frag = ops.Fragment(T.fact, joins=[(W.F_DATE, hd), (W.F_SUPP, hs), (W.F_PART, hp)],
                        keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)], aggs=aggs)
    ctx.set_option("fragment.generic", generic)
    ctx.set_option("fragment.pipe", pipe)
    try:
        got = _gpu_dict(frag.execute(ctx, capacity=8192), 2)
    finally:
        ctx.set_option("fragment.generic", 0)
        ctx.set_option("fragment.pipe", 1)
    fact = [to_np(c) for c in T.fact.columns]
    d, s_, p = ([to_np(c) for c in t.columns] for t in (T.date, T.supplier, T.part))
    dims = [
        dict(columns=d, key_col=0, payload_col=1, fact_key_col=0,
             filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
        dict(columns=s_, key_col=0, payload_col=1, fact_key_col=2,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
        dict(columns=p, key_col=0, payload_col=2, fact_key_col=3,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
    ]
    want = orc.FragmentSpec(fact, dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)], aggs=aggs).run(
        nthreads=8, capacity=1 << 14)
    assert got == _orc_dict(want)


def test_star_fragment_600m_size_independent(ctx):
    """Full C2 size: slice linearity, dense-key closed form and both kernel variants agree."""
    import torch
    from ghive_amd import ops, workloads as W
    n = 600_000_000
    T = W.star_tables(ctx, n)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    full = _gpu_dict(W.star_fragment(T.fact, hd, hs, hp).execute(ctx, capacity=8192), 2)
    assert ctx.stage_times()["rows_in"] == n
    late = _gpu_dict(W.star_fragment(T.fact, hd, hs, hp).execute(ctx, capacity=8192, late=True), 2)
    assert late == full
    # SUM is additive: uneven slices (16-byte aligned starts) add up to the whole
    bounds = [0, 100_000_004, 377_777_776, n]
    acc = {}
    for a, b in zip(bounds[:-1], bounds[1:]):
        part = ops.Table([c[a:b] for c in T.fact.columns])
        for k, v in _gpu_dict(W.star_fragment(part, hd, hs, hp).execute(ctx, capacity=8192), 2).items():
            acc[k] = (acc.get(k, (0,))[0] + v[0],)
    assert acc == full
    # dense keys 1..N: each join is an array lookup; closed form with torch on the same data
    year = T.date.columns[1]
    region = T.supplier.columns[1]
    cat, brand = T.part.columns[1], T.part.columns[2]
    sums = torch.zeros(7 * 1000, dtype=torch.int64, device=year.device)
    step = 50_000_000
    for a in range(0, n, step):
        fd, fs, fp, pr, di = (T.fact.columns[i][a:a + step].to(torch.int64) for i in (0, 2, 3, 5, 6))
        y = year[fd - 1].to(torch.int64)
        ok = (y >= 1992) & (y <= 1993) & (region[fs - 1] == 1) & (cat[fp - 1] == 1)
        g = (y[ok] - 1992) * 1000 + brand[fp - 1][ok].to(torch.int64)
        sums.index_add_(0, g, pr[ok] * di[ok])
    counts = torch.zeros(7 * 1000, dtype=torch.int64, device=year.device)
    for a in range(0, n, step):
        fd, fs, fp = (T.fact.columns[i][a:a + step].to(torch.int64) for i in (0, 2, 3))
        y = year[fd - 1].to(torch.int64)
        ok = (y >= 1992) & (y <= 1993) & (region[fs - 1] == 1) & (cat[fp - 1] == 1)
        counts.index_add_(0, (y[ok] - 1992) * 1000 + brand[fp - 1][ok].to(torch.int64),
                          torch.ones(int(ok.sum()), dtype=torch.int64, device=year.device))
    closed = {(1992 + g // 1000, g % 1000): (int(sums[g]),) for g in range(7000) if counts[g] > 0}
    assert closed == full


def test_fragment_edges(ctx):
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 1000)
    ht = W.small_dim_table(ctx, T.dim)
    empty = ops.Table([c[:0] for c in T.fact.columns])
    assert _gpu_dict(W.small_fragment(empty, ht).execute(ctx, capacity=64), 1) == {}
    # keyless: one group even over no rows; COUNT is 0 and SUM is NULL (LONG_MAX)
    f = ops.Fragment(empty, joins=[(0, ht)], keys=[],
                     aggs=[(abi.AGG_CNT, abi.EXPR_ONE, 0, 0), (abi.AGG_SUM, abi.EXPR_COL, 1, 0)])
    outs = f.execute(ctx, capacity=4)
    assert [int(o[0]) for o in outs] == [0, abi.LONG_NULL]
    # an empty dimension joins nothing (probed through its empty hash table)
    none = W.small_dim_table(ctx, T.dim, attr_value=99)
    assert _gpu_dict(W.small_fragment(T.fact, none).execute(ctx, capacity=64), 1) == {}
    # a misaligned column is rejected, not read
    from ghive_amd._lib import GhiveError
    mis = ops.Table([c[1:] for c in T.fact.columns])
    with pytest.raises(GhiveError):
        W.small_fragment(mis, ht).execute(ctx, capacity=64)


def _orc_dict_nulls(res):
    return {tuple(int(x) for x in k): tuple(abi.LONG_NULL if nl else int(x) for x, nl in zip(v, nn))
            for k, v, nn in zip(res["keys"], res["aggs"], res["nulls"])}


@pytest.mark.parametrize("pipe", [1, 0])
def test_small_fragment_lds_slots_all_aggregates(ctx, orc, pipe):
    """25 groups (the workgroup-private LDS table, so the dense-probe kernel whatever the pipelined
    kernel's setting): COUNT(*), SUM, MIN and MAX of a product."""
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 3_000_017, seed=7)
    ht = W.small_dim_table(ctx, T.dim)
    aggs = [(abi.AGG_CNT, abi.EXPR_ONE, 0, 0), (abi.AGG_SUM, abi.EXPR_COL, 2, 0),
            (abi.AGG_MIN, abi.EXPR_COL, 1, 0), (abi.AGG_MAX, abi.EXPR_MUL, 1, 2)]
    frag = ops.Fragment(T.fact, joins=[(0, ht)], keys=[(abi.SRC_JOIN, 0)], aggs=aggs)
    ctx.set_option("fragment.pipe", pipe)
    try:
        got = _gpu_dict(frag.execute(ctx, capacity=64), 1)
        joined = ctx.stage_times()["rows_joined"]
    finally:
        ctx.set_option("fragment.pipe", 1)
    fact = [to_np(c) for c in T.fact.columns]
    dim = [to_np(c) for c in T.dim.columns]
    spec = orc.FragmentSpec(fact, [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                        filter=([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0))],
                            keys=[(abi.SRC_JOIN, 0)], aggs=aggs)
    want = spec.run(nthreads=8)
    assert got == _orc_dict_nulls(want)
    assert joined == want["joined"]
    assert len(got) == 25


@pytest.mark.parametrize("pipe", [1, 0])
@pytest.mark.parametrize("neg", ["one", "both"])
def test_fragment_negative_keys_and_nulls(ctx, orc, pipe, neg):
    """Group keys from two joins with negative attribute values and a domain too large for the
    LDS table (direct slots offset by each key's minimum; when both ranges hold -1 a key could pack
    to the empty-slot pattern and the hashed table is used), NULL foreign keys and NULL aggregate
    inputs, keys outside the dimension ranges. Most rows survive, so tiles overflow the pipelined
    kernel's survivor queues and go straight to the global table."""
    import torch
    from ghive_amd import ops
    rng = np.random.default_rng(11)
    n = 700_001
    ka = np.arange(1, 301, dtype=np.int32)
    pa = rng.integers(-3, 41, 300).astype(np.int32)
    kb = np.arange(10, 210, dtype=np.int32)
    pb = rng.integers(-1 if neg == "both" else 0, 31, 200).astype(np.int32)
    cols = [rng.integers(0, 320, n), rng.integers(0, 225, n), rng.integers(-50, 51, n), rng.integers(1, 1001, n)]
    cols = [c.astype(np.int32) for c in cols]
    for c in cols:
        c[rng.random(n) < 0.01] = abi.INT_NULL
    dev = f"cuda:{ctx.device}"
    tcols = [torch.from_numpy(c).to(dev) for c in cols]
    fact = ops.Table(tcols)
    dka, dpa, dkb, dpb = (torch.from_numpy(x).to(dev) for x in (ka, pa, kb, pb))
    ha = ops.HashTable(ctx, dka, payload=dpa)
    hb = ops.HashTable(ctx, dkb, payload=dpb)
    aggs = [(abi.AGG_SUM, abi.EXPR_MUL, 2, 3), (abi.AGG_CNT, abi.EXPR_ONE, 0, 0),
            (abi.AGG_MIN, abi.EXPR_COL, 2, 0), (abi.AGG_MAX, abi.EXPR_COL, 3, 0)]
    frag = ops.Fragment(fact, joins=[(0, ha), (1, hb)], keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 1)], aggs=aggs)
    ctx.set_option("fragment.pipe", pipe)
    try:
        got = _gpu_dict(frag.execute(ctx, capacity=4096), 2)
        joined = ctx.stage_times()["rows_joined"]
    finally:
        ctx.set_option("fragment.pipe", 1)
    dims = [dict(columns=[ka, pa], key_col=0, payload_col=1, fact_key_col=0),
            dict(columns=[kb, pb], key_col=0, payload_col=1, fact_key_col=1)]
    want = orc.FragmentSpec(cols, dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 1)], aggs=aggs).run(nthreads=8)
    assert got == _orc_dict_nulls(want)
    assert joined == want["joined"]
    assert any(k[0] < 0 for k in got)


@pytest.mark.parametrize("pipe", [1, 0])
def test_fragment_keyless_star(ctx, orc, pipe):
    """No group keys: one group (key 0) holds the global aggregates."""
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 1_500_001)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    aggs = [(abi.AGG_SUM, abi.EXPR_MUL, 5, 6), (abi.AGG_CNT, abi.EXPR_ONE, 0, 0)]
    frag = ops.Fragment(T.fact, joins=[(W.F_DATE, hd), (W.F_SUPP, hs), (W.F_PART, hp)], keys=[], aggs=aggs)
    ctx.set_option("fragment.pipe", pipe)
    try:
        outs = frag.execute(ctx, capacity=4)
    finally:
        ctx.set_option("fragment.pipe", 1)
    got = tuple(int(to_np(o)[0]) for o in outs)
    fact = [to_np(c) for c in T.fact.columns]
    d, s_, p = ([to_np(c) for c in t.columns] for t in (T.date, T.supplier, T.part))
    dims = [
        dict(columns=d, key_col=0, payload_col=1, fact_key_col=0,
             filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
        dict(columns=s_, key_col=0, payload_col=1, fact_key_col=2,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
        dict(columns=p, key_col=0, payload_col=2, fact_key_col=3,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
    ]
    want = orc.FragmentSpec(fact, dims, keys=[], aggs=aggs).run(nthreads=8)
    assert len(outs[0]) == 1
    assert got == tuple(int(x) for x in want["aggs"][0])


@pytest.mark.parametrize("naggs", [1, 2, 4])
def test_star_fragment_queue_sizes(ctx, orc, naggs):
    """The pipelined kernel's per-wave survivor queues shrink as aggregates are added (the LDS the
    bitmaps leave is fixed); every size gives the oracle's groups."""
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 3_000_011)
    hd, hs, hp = W.star_hash_tables(ctx, T)
    pool = [(abi.AGG_SUM, abi.EXPR_MUL, 5, 6), (abi.AGG_CNT, abi.EXPR_ONE, 0, 0),
            (abi.AGG_MIN, abi.EXPR_COL, 4, 0), (abi.AGG_MAX, abi.EXPR_MUL, 4, 6)]
    aggs = pool[:naggs]
    frag = ops.Fragment(T.fact, joins=[(W.F_DATE, hd), (W.F_SUPP, hs), (W.F_PART, hp)],
                        keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)], aggs=aggs)
    got = _gpu_dict(frag.execute(ctx, capacity=8192), 2)
    fact = [to_np(c) for c in T.fact.columns]
    d, s_, p = ([to_np(c) for c in t.columns] for t in (T.date, T.supplier, T.part))
    dims = [
        dict(columns=d, key_col=0, payload_col=1, fact_key_col=0,
             filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
        dict(columns=s_, key_col=0, payload_col=1, fact_key_col=2,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
        dict(columns=p, key_col=0, payload_col=2, fact_key_col=3,
             filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
    ]
    want = orc.FragmentSpec(fact, dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)], aggs=aggs).run(
        nthreads=8, capacity=1 << 14)
    assert got == _orc_dict_nulls(want)


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 3 * (1 << 20) + 16 * 257])
def test_bandwidth_probe_copies_and_reads(ctx, nbytes):
    """The bandwidth probe behind roofline peaks: the copy form copies every 16-byte word of odd-
    sized ranges (workgroup ranges of unequal length, partial unrolled steps), the read form runs."""
    import ctypes as C
    import torch
    from ghive_amd._lib import lib
    src = torch.randint(-2**31, 2**31 - 1, (nbytes // 4,), dtype=torch.int32, device=f"cuda:{ctx.device}")
    dst = torch.zeros_like(src)
    ms = C.c_float()
    ctx.check(lib().ghive_bench_stream(ctx.h, C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()), nbytes, 1, 2,
                                       C.byref(ms)))
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    ctx.check(lib().ghive_bench_stream(ctx.h, None, C.c_void_p(src.data_ptr()), nbytes, 0, 2, C.byref(ms)))
    assert ms.value > 0
    # the five-column form takes whole 80-byte rows of its five columns, and rejects the rest
    st = lib().ghive_bench_stream(ctx.h, None, C.c_void_p(src.data_ptr()), nbytes, 2, 2, C.byref(ms))
    if nbytes % 80:
        assert st == abi.ERR_INVALID
    else:
        ctx.check(st)
    ctx.check(lib().ghive_bench_stream(ctx.h, None, C.c_void_p(src.data_ptr()), nbytes - nbytes % 80, 2, 2,
                                       C.byref(ms)) if nbytes >= 80 else abi.OK)


def test_async_filter_and_build_match_the_waiting_forms(ctx, orc):
    """Filters that leave their count on the device and builds that read their statistics on
    first use give the same tables and fragment results as the forms that wait after each step;
    an empty selection builds an empty table."""
    import torch
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 2_000_003)
    sync_tabs = W.star_hash_tables(ctx, T, wait_each=True)
    async_tabs = W.star_hash_tables(ctx, T)
    for a, b in zip(sync_tabs, async_tabs):
        ia, ib = a.info(), b.info()
        assert (ia["entries"], ia["unique"]) == (ib["entries"], ib["unique"])
    got_sync = _gpu_dict(W.star_fragment(T.fact, *sync_tabs).execute(ctx, capacity=8192), 2)
    got_async = _gpu_dict(W.star_fragment(T.fact, *W.star_hash_tables(ctx, T)).execute(ctx, capacity=8192), 2)
    assert got_sync == got_async == _orc_dict(_star_oracle(orc, T))
    # the filter's device count equals the waiting filter's count
    pred = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)
    sel_a, cnt = pred.process_async(ctx, T.part)
    sel_s = pred.process(ctx, T.part)
    assert int(cnt.item()) == sel_s.numel()
    assert torch.equal(sel_a[: sel_s.numel()], sel_s)
    # probe through an asynchronously built table
    sel_d, cnt_d = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)],
                                       0).process_async(ctx, T.date)
    hd = ops.HashTable(ctx, T.date.columns[0], sel=sel_d, payload=T.date.columns[1], count=cnt_d)
    rows, pay = hd.probe(T.fact.columns[0][:100_000])
    hs = ops.HashTable(ctx, T.date.columns[0], sel=sel_d[: int(cnt_d.item())], payload=T.date.columns[1])
    rows2, pay2 = hs.probe(T.fact.columns[0][:100_000])
    key = lambda r, p: sorted(zip(r.cpu().tolist(), p.cpu().tolist()))
    assert key(rows, pay) == key(rows2, pay2)
    # nothing selected
    sel_e, cnt_e = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [99], data_type=2)], 0).process_async(ctx, T.part)
    he = ops.HashTable(ctx, T.part.columns[0], sel=sel_e, payload=T.part.columns[2], count=cnt_e)
    assert he.info()["entries"] == 0
    frag = W.star_fragment(T.fact, async_tabs[0], async_tabs[1], he)
    assert _gpu_dict(frag.execute(ctx, capacity=8192), 2) == {}
    for h in (*sync_tabs, *async_tabs, hd, hs, he):
        h.free()
