"""GPU parity of the hash group-by against the oracle: integer aggregates bit-exact, double
SUM/AVG within 1e-12 relative of the correctly rounded oracle sum (BASELINE tolerance)."""
import math
from collections import defaultdict

import numpy as np
import pytest

from conftest import np_col, to_dev, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
INT_NULL, LONG_NULL = abi.INT_NULL, abi.LONG_NULL
REL = 1e-12


def _dev(c):
    """numpy column, or a ('string', blob, idx) fixture column -> device column."""
    from ghive_amd import ops
    if isinstance(c, tuple) and c[0] == "string":
        import torch
        blob = torch.frombuffer(bytearray(c[1]), dtype=torch.uint8) if len(c[1]) else torch.zeros(0, dtype=torch.uint8)
        return ops.StringColumn(blob.cuda(), to_dev(np.asarray(c[2], dtype=np.int32)))
    return to_dev(c)


def _host(o):
    """Device output -> numpy array, or a list of str / None for a STRING column."""
    from ghive_amd import ops
    if isinstance(o, ops.StringColumn):
        return [None if v is None else v.decode() for v in o.to_list()]
    return to_np(o)


def _run(ctx, cols, keys, aggs, sel=None, capacity=None):
    from ghive_amd import ops
    t = ops.Table([_dev(c) for c in cols])
    outs = ops.GroupByPredicate(keys, aggs).gpu_execute(
        ctx, t, sel=None if sel is None else to_dev(sel.astype(np.int32)), capacity=capacity)
    return [_host(o) for o in outs]


def _item(x):
    return x.item() if hasattr(x, "item") else x


def _as_dict(outs, nkeys):
    res = {}
    for g in range(len(outs[0]) if outs else 0):
        key = tuple(_item(o[g]) for o in outs[:nkeys])
        res[key] = [_item(o[g]) for o in outs[nkeys:]]
    return res


def _is_null(v):
    return v is None or v == LONG_NULL or v == INT_NULL or (isinstance(v, float) and math.isnan(v))


def _same(a, b):
    if _is_null(a) or _is_null(b):
        return _is_null(a) and _is_null(b)
    if isinstance(a, float) or isinstance(b, float):
        return a == b or abs(a - b) <= REL * max(abs(a), abs(b))
    return a == b


def _oracle_key(col, r):
    if isinstance(col, tuple) and col[0] == "string":
        s, e = col[2][2 * r], col[2][2 * r + 1]
        return None if s < 0 else col[1][s:e].decode()
    return col[r].item()


def _strings(rng, n, vocab, null_frac=0.02):
    """A ('string', blob, idx) column drawn from `vocab` distinct strings (empty ones included)."""
    words = [bytes(rng.integers(97, 123, int(rng.integers(0, 12))).astype(np.uint8)) for _ in range(vocab)]
    words[0] = b""
    pick = rng.integers(0, vocab, n)
    blob, idx, o = bytearray(), np.empty(2 * n, dtype=np.int32), 0
    for i, w in enumerate(pick):
        if rng.random() < null_frac:
            idx[2 * i], idx[2 * i + 1] = -1, -1
            continue
        blob += words[w]
        idx[2 * i], idx[2 * i + 1] = o, o + len(words[w])
        o += len(words[w])
    return ("string", bytes(blob), idx)


def _oracle_dict(orc, cols, keys, aggs, sel=None):
    rep, vals, nulls = orc.groupby(cols, keys, aggs, sel=sel)
    res = {}
    for g, r in enumerate(rep):
        key = tuple(_oracle_key(cols[k], r) for k in keys)
        key = tuple(float("nan") if isinstance(x, float) and math.isnan(x) else x for x in key)
        res[key] = [None if nulls[a][g] else vals[a][g].item() for a in range(len(aggs))]
    return res


def _compare(got, want):
    def norm(k):
        return tuple("nan" if isinstance(x, float) and math.isnan(x) else x for x in k)
    g = {norm(k): v for k, v in got.items()}
    w = {norm(k): v for k, v in want.items()}
    assert g.keys() == w.keys()
    for k in w:
        for a, b in zip(g[k], w[k]):
            assert _same(a, b), (k, g[k], w[k])


def test_gpu_impl_kats(ctx, kats):
    """Every group-by KAT of GroupByPredicateTest.cu, STRING keys included (dictionary path)."""
    for k in kats["groupby_gpu_impl"]:
        cols = [np_col(c) for c in k["columns"]]
        outs = _run(ctx, cols, k["keys"], [tuple(a) for a in k["aggs"]])
        got = _as_dict(outs, len(k["keys"]))
        want = {tuple(e["key"]): e["aggs"] for e in k["expect"]}
        assert got == want, k["cite"]


_AGG = {"sum": abi.AGG_SUM, "min": abi.AGG_MIN, "max": abi.AGG_MAX, "avg": abi.AGG_AVG,
        "count": abi.AGG_CNT}


def test_hive_aggregate_kats(ctx, kats):
    for k in kats["hive_groupby"]:
        dbl = k.get("dtype") == "DOUBLE"
        vals = np.array([(np.nan if dbl else LONG_NULL) if v is None else v for v in k["values"]],
                        dtype=np.float64 if dbl else np.int64)
        cols, keys = [vals], []
        if k["keys"] is not None:
            cols.append(np.array([LONG_NULL if v is None else v for v in k["keys"]], dtype=np.int64))
            keys = [1]
        outs = _run(ctx, cols, keys, [(_AGG[k["agg"]], 0)])
        if k["keys"] is None:
            assert len(outs[0]) == 1
            v = outs[0][0].item()
            assert (None if _is_null(v) else v) == k["expect"], k["cite"]
        else:
            got = {(None if kk == LONG_NULL else kk): (None if _is_null(v) else v)
                   for kk, v in zip(outs[0].tolist(), outs[1].tolist())}
            assert got == {a: b for a, b in k["expect"]}, k["cite"]


@pytest.mark.parametrize("partition", [1, 2], ids=["direct", "partitioned"])
@pytest.mark.parametrize("card", [7, 3000, 200_000])
@pytest.mark.parametrize("ktype", ["INT", "LONG", "DOUBLE", "INT2"])
def test_groupby_all_aggs_vs_oracle(ctx, orc, card, ktype, partition):
    """Every aggregate kind and key type against the oracle, in the direct form and in the
    high-cardinality form (hash partitions aggregated in LDS); LONG keys include -1, whose bits
    equal the table's EMPTY pattern."""
    ctx.set_option("groupby.partition", partition)
    try:
        _all_aggs_case(ctx, orc, card, ktype)
    finally:
        ctx.set_option("groupby.partition", 1)


def _all_aggs_case(ctx, orc, card, ktype):
    rng = np.random.default_rng(card + len(ktype))
    n = 400_000
    if ktype == "INT2":
        k1 = rng.integers(0, max(card // 50, 2), n).astype(np.int32)
        k2 = rng.integers(-25, 25, n).astype(np.int32)
        kcols, keys = [k1, k2], [0, 1]
    elif ktype == "DOUBLE":
        k1 = rng.integers(0, card, n).astype(np.float64) / 4
        k1[::211] = np.nan
        kcols, keys = [k1], [0]
    else:
        dt = np.int32 if ktype == "INT" else np.int64
        k1 = rng.integers(-card // 2, card - card // 2, n).astype(dt)
        k1[::101] = INT_NULL if ktype == "INT" else LONG_NULL
        kcols, keys = [k1], [0]
    v_long = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    v_long[::37] = LONG_NULL
    v_int = rng.integers(-1000, 1000, n).astype(np.int32)
    v_dbl = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    v_dbl[::41] = np.nan
    v_flt = rng.standard_normal(n).astype(np.float32)
    cols = kcols + [v_long, v_int, v_dbl, v_flt]
    b = len(kcols)
    aggs = [(abi.AGG_SUM, b), (abi.AGG_MIN, b), (abi.AGG_MAX, b + 1), (abi.AGG_CNT, b),
            (abi.AGG_AVG, b + 1), (abi.AGG_SUM, b + 2), (abi.AGG_CNT, abi.COUNT_STAR),
            (abi.AGG_MAX, b + 2)]
    outs = _run(ctx, cols, keys, aggs)
    _compare(_as_dict(outs, len(keys)), _oracle_dict(orc, cols, keys, aggs))
    aggs2 = [(abi.AGG_MIN, b + 3), (abi.AGG_AVG, b + 2), (abi.AGG_SUM, b + 3)]
    outs = _run(ctx, cols, keys, aggs2)
    _compare(_as_dict(outs, len(keys)), _oracle_dict(orc, cols, keys, aggs2))


@pytest.mark.parametrize("partition", [1, 2], ids=["direct", "partitioned"])
@pytest.mark.parametrize("kset", ["INT+LONG", "LONG3", "DOUBLE+LONG", "INT5", "STRING", "STRING+INT"])
def test_groupby_wide_keys_vs_oracle(ctx, orc, kset, partition):
    """Keys wider than 64 bits, more than four key columns, and STRING keys (the dictionary path)
    against the oracle, with NULLs in every key column, in both aggregation forms."""
    rng = np.random.default_rng(len(kset) * 7 + partition)
    n = 300_000

    def ints(dt, lo, hi, null):
        x = rng.integers(lo, hi, n).astype(dt)
        x[rng.random(n) < 0.01] = null
        return x
    if kset == "INT+LONG":
        kcols = [ints(np.int32, 1992, 1999, INT_NULL), ints(np.int64, -(2**40), -(2**40) + 300, LONG_NULL)]
    elif kset == "LONG3":
        kcols = [ints(np.int64, 0, 9, LONG_NULL), ints(np.int64, 2**50, 2**50 + 30, LONG_NULL),
                 ints(np.int64, -1, 2, LONG_NULL)]
    elif kset == "DOUBLE+LONG":
        d = rng.integers(0, 50, n) / 8.0
        d[rng.random(n) < 0.01] = np.nan
        kcols = [d, ints(np.int64, 0, 40, LONG_NULL)]
    elif kset == "INT5":
        kcols = [ints(np.int32, 0, 3, INT_NULL) for _ in range(5)]
    elif kset == "STRING":
        kcols = [_strings(rng, n, 5000)]
    else:
        kcols = [_strings(rng, n, 300), ints(np.int32, 0, 20, INT_NULL)]
    v_long = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    v_long[::37] = LONG_NULL
    v_dbl = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    v_dbl[::41] = np.nan
    cols = kcols + [v_long, v_dbl]
    b = len(kcols)
    keys = list(range(b))
    aggs = [(abi.AGG_SUM, b), (abi.AGG_MIN, b), (abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, b + 1),
            (abi.AGG_MAX, b + 1), (abi.AGG_AVG, b)]
    ctx.set_option("groupby.partition", partition)
    try:
        outs = _run(ctx, cols, keys, aggs)
        _compare(_as_dict(outs, len(keys)), _oracle_dict(orc, cols, keys, aggs))
        sel = np.nonzero(rng.random(n) < 0.4)[0].astype(np.uint32)
        outs = _run(ctx, cols, keys, aggs, sel=sel)
        _compare(_as_dict(outs, len(keys)), _oracle_dict(orc, cols, keys, aggs, sel=sel))
    finally:
        ctx.set_option("groupby.partition", 1)


def test_groupby_wide_keys_edges(ctx, orc):
    from ghive_amd._lib import GhiveError
    rng = np.random.default_rng(5)
    k1 = np.arange(1000, dtype=np.int64)
    k2 = np.arange(1000, dtype=np.int64) % 7
    with pytest.raises(GhiveError) as e:  # 1000 groups, room for 100
        _run(ctx, [k1, k2, k1], [0, 1], [(abi.AGG_SUM, 2)], capacity=100)
    assert e.value.status == abi.ERR_CAPACITY
    outs = _run(ctx, [k1[:0], k2[:0], k1[:0]], [0, 1], [(abi.AGG_SUM, 2)])  # no rows, no groups
    assert len(outs[0]) == 0
    # every string NULL or empty, with a high-cardinality second key in the partitioned form
    s = ("string", b"", np.array([-1, -1, 0, 0] * 50_000, dtype=np.int32))
    k = rng.integers(0, 60_000, 100_000).astype(np.int64)
    cols = [s, k, k]
    outs = _run(ctx, cols, [0, 1], [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 2)])
    _compare(_as_dict(outs, 2), _oracle_dict(orc, cols, [0, 1], [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 2)]))


@pytest.mark.parametrize("threads,occ", [(256, 1), (512, 2), (512, 1), (1024, 1)])
@pytest.mark.parametrize("aggset", ["count_star", "one_column", "two_columns"])
def test_groupby_partitioned_value_forms(ctx, orc, aggset, threads, occ):
    """The high-cardinality form reads aggregate inputs three ways: none (COUNT(*) only), one column
    carried through the partition passes, several columns read at the row id; its partition passes
    at every workgroup size, the 512-thread ones compiled for one or two workgroups per CU (option
    groupby.part_occupancy). LONG keys include -1, whose bits equal the EMPTY pattern (aggregated in
    the small global table), and NULL."""
    rng = np.random.default_rng(len(aggset))
    n = 500_000
    k = rng.integers(-1000, 60_000, n).astype(np.int64)
    k[::97] = -1
    k[::89] = LONG_NULL
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    x[::41] = np.nan
    y = rng.integers(-10**12, 10**12, n).astype(np.int64)
    aggs = {"count_star": [(abi.AGG_CNT, abi.COUNT_STAR)],
            "one_column": [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 1), (abi.AGG_MIN, 1), (abi.AGG_AVG, 1),
                           (abi.AGG_CNT, 1)],
            "two_columns": [(abi.AGG_SUM, 1), (abi.AGG_MAX, 2), (abi.AGG_CNT, abi.COUNT_STAR)]}[aggset]
    cols = [k, x, y]
    ctx.set_option("groupby.partition", 2)
    ctx.set_option("groupby.part_threads", threads)
    ctx.set_option("groupby.part_occupancy", occ)
    try:
        outs = _run(ctx, cols, [0], aggs)
        _compare(_as_dict(outs, 1), _oracle_dict(orc, cols, [0], aggs))
        sel = np.nonzero(rng.random(n) < 0.5)[0].astype(np.uint32)
        outs = _run(ctx, cols, [0], aggs, sel=sel)
        _compare(_as_dict(outs, 1), _oracle_dict(orc, cols, [0], aggs, sel=sel))
        from ghive_amd._lib import GhiveError
        with pytest.raises(GhiveError) as e:  # more groups than the bound
            _run(ctx, cols, [0], aggs, capacity=1000)
        assert e.value.status == abi.ERR_CAPACITY
    finally:
        ctx.set_option("groupby.partition", 1)
        ctx.set_option("groupby.part_threads", 512)
        ctx.set_option("groupby.part_occupancy", 2)


@pytest.mark.parametrize("threads", [256, 512])
def test_groupby_partitioned_double_sum_without_nulls(ctx, orc, threads):
    """A carried DOUBLE column with no NULL: the first partition pass records that no row is NULL,
    and the emit keeps no non-NULL count for SUM (every group has a value). Against the oracle:
    SUM, COUNT(*), MIN and COUNT(x) of values over seven decades, keys with -1 (the EMPTY pattern's
    bits, aggregated in the small global table) and NULL; then the same keys with one NULL value
    added (the count-keeping form) on a selection."""
    rng = np.random.default_rng(threads + 7)
    n = 400_000
    k = rng.integers(-1000, 50_000, n).astype(np.int64)
    k[::97] = -1
    k[::89] = LONG_NULL
    x = rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)
    aggs = [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_MIN, 1), (abi.AGG_CNT, 1)]
    ctx.set_option("groupby.partition", 2)
    ctx.set_option("groupby.part_threads", threads)
    try:
        cols = [k, x]
        _compare(_as_dict(_run(ctx, cols, [0], aggs), 1), _oracle_dict(orc, cols, [0], aggs))
        x2 = x.copy()
        x2[12345] = np.nan
        cols = [k, x2]
        sel = np.nonzero(rng.random(n) < 0.7)[0].astype(np.uint32)
        _compare(_as_dict(_run(ctx, cols, [0], aggs, sel=sel), 1), _oracle_dict(orc, cols, [0], aggs, sel=sel))
    finally:
        ctx.set_option("groupby.partition", 1)
        ctx.set_option("groupby.part_threads", 512)


@pytest.mark.parametrize("threads", [256, 512])
def test_groupby_two_partition_passes(ctx, threads):
    """A group bound of 3M puts 2^11 partitions behind two 8-bit partition passes, the second of
    which counts its digits from the byte plane the first one wrote. 5M rows over 1.5M keys:
    COUNT(*) and SUM(LONG) exact against numpy, every key present once."""
    rng = np.random.default_rng(threads)
    n = 5_000_000
    k = rng.integers(-2**40, 2**40, 1_500_000)[rng.integers(0, 1_500_000, n)].astype(np.int64)
    y = rng.integers(-10**9, 10**9, n).astype(np.int64)
    ctx.set_option("groupby.part_threads", threads)
    try:
        outs = _run(ctx, [k, y], [0], [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 1)], capacity=3_000_000)
    finally:
        ctx.set_option("groupby.part_threads", 512)
    uk, inv = np.unique(k, return_inverse=True)
    cnt = np.bincount(inv)
    sm = np.zeros(len(uk), dtype=np.int64)
    np.add.at(sm, inv, y)
    order = np.argsort(outs[0])
    assert np.array_equal(outs[0][order], uk)
    assert np.array_equal(outs[1][order], cnt)
    assert np.array_equal(outs[2][order], sm)


@pytest.mark.parametrize("partition", [1, 2], ids=["direct", "partitioned"])
def test_double_sum_rejects_infinity(ctx, partition):
    """SUM over a DOUBLE column holding +-Inf has no fixed-point scale: both forms reject it with
    UNSUPPORTED (the partitioned form finds the magnitude inside its first pass), and the context
    keeps working."""
    from ghive_amd._lib import GhiveError
    n = 50_000
    k = np.arange(n, dtype=np.int64) % 7919
    x = np.random.default_rng(3).standard_normal(n)
    x[12345] = -np.inf
    ctx.set_option("groupby.partition", partition)
    try:
        with pytest.raises(GhiveError) as e:
            _run(ctx, [k, x], [0], [(abi.AGG_SUM, 1)], capacity=n)
        assert e.value.status == abi.ERR_UNSUPPORTED
        x[12345] = 1.0
        outs = _run(ctx, [k, x], [0], [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)], capacity=n)
        assert len(outs[0]) == 7919 and int(outs[2].sum()) == n
    finally:
        ctx.set_option("groupby.partition", 1)


def test_groupby_with_selection_and_empty(ctx, orc):
    rng = np.random.default_rng(2)
    n = 100_000
    k = rng.integers(0, 40, n).astype(np.int32)
    v = rng.integers(0, 100, n).astype(np.int64)
    sel = np.nonzero(rng.random(n) < 0.3)[0].astype(np.uint32)
    outs = _run(ctx, [k, v], [0], [(abi.AGG_SUM, 1)], sel=sel)
    _compare(_as_dict(outs, 1), _oracle_dict(orc, [k, v], [0], [(abi.AGG_SUM, 1)], sel=sel))
    # keyless over no rows: one group, SUM NULL, COUNT 0
    outs = _run(ctx, [k[:0], v[:0]], [], [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)])
    assert len(outs[0]) == 1 and outs[0][0] == LONG_NULL and outs[1][0] == 0
    # keyed over no rows: no groups
    outs = _run(ctx, [k[:0], v[:0]], [0], [(abi.AGG_SUM, 1)])
    assert len(outs[0]) == 0


def test_double_sum_order_independent_and_exact(ctx):
    """Cancelling sums: every group's SUM equals math.fsum of its values (correct rounding)."""
    rng = np.random.default_rng(12)
    g = 5000
    base = rng.standard_normal(g * 10) * 10.0 ** rng.integers(-8, 8, g * 10)
    vals = np.concatenate([base, -base[: g * 5]])
    keys = np.concatenate([np.repeat(np.arange(g), 10), np.repeat(np.arange(g), 10)[: g * 5]]).astype(np.int64)
    perm = rng.permutation(len(vals))
    outs1 = _run(ctx, [keys[perm], vals[perm]], [0], [(abi.AGG_SUM, 1)])
    perm2 = rng.permutation(len(vals))
    outs2 = _run(ctx, [keys[perm2], vals[perm2]], [0], [(abi.AGG_SUM, 1)])
    d1, d2 = _as_dict(outs1, 1), _as_dict(outs2, 1)
    assert d1 == d2  # bit-identical under a different arrival order
    groups = defaultdict(list)
    for kk, vv in zip(keys.tolist(), vals.tolist()):
        groups[kk].append(vv)
    worst = 0.0
    for (kk,), (s,) in d1.items():
        exact = math.fsum(groups[kk])
        if exact != 0:
            worst = max(worst, abs(s - exact) / abs(exact))
    assert worst <= REL


def test_groupby_capacity_error(ctx):
    from ghive_amd._lib import GhiveError
    k = np.arange(1000, dtype=np.int64)
    with pytest.raises(GhiveError) as e:
        _run(ctx, [k, k], [0], [(abi.AGG_SUM, 1)], capacity=100)
    assert e.value.status == abi.ERR_CAPACITY


def test_groupby_high_cardinality_1e7(ctx):
    """10M rows, ~5M distinct LONG keys: Σcount = rows and ΣSUM = global SUM (int64, exact)."""
    import torch
    from ghive_amd import ops
    n = 10_000_000
    k = ctx.gen_uniform_i64(5, 0, n, 0, 5_000_000)
    v = ctx.gen_uniform_i64(5, 1, n, -1000, 1000)
    outs = ops.GroupByPredicate([0], [(abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_SUM, 1)]).gpu_execute(
        ctx, ops.Table([k, v]), capacity=6_000_000)
    assert int(outs[1].sum().item()) == n
    assert int(outs[2].sum().item()) == int(v.sum().item())
    assert outs[0].numel() == torch.unique(k).numel()
