"""GPU parity of hash build/probe and the N-way join against the oracle (row-pair multisets)."""
import numpy as np
import pytest

from conftest import to_dev, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
INT_NULL, LONG_NULL = abi.INT_NULL, abi.LONG_NULL


def _pairs(rows, pay):
    return sorted(zip(rows.astype(np.int64).tolist(), pay.astype(np.int64).tolist()))


@pytest.mark.parametrize("ktype", ["INT", "LONG"])
@pytest.mark.parametrize("payload", [None, "INT", "LONG"])
def test_probe_duplicates_nulls_vs_oracle(ctx, orc, ktype, payload):
    from ghive_amd import ops
    rng = np.random.default_rng(11)
    dt = np.int32 if ktype == "INT" else np.int64
    null = INT_NULL if ktype == "INT" else LONG_NULL
    bk = rng.integers(-500, 500, 30000).astype(dt)
    bk[::50] = null
    pk = rng.integers(-700, 700, 200000).astype(dt)
    pk[::31] = null
    pay = None
    if payload == "INT":
        pay = rng.integers(-2**31, 2**31 - 1, 30000).astype(np.int32)
    elif payload == "LONG":
        pay = rng.integers(-2**62, 2**62, 30000).astype(np.int64)
    ht = ops.HashTable(ctx, to_dev(bk), payload=None if pay is None else to_dev(pay))
    info = ht.info()
    assert not info["unique"] and info["entries"] == int((bk != null).sum())
    rows, vals = ht.probe(to_dev(pk))
    want_rows, want_vals = orc.mapjoin_inner(bk, pk, payload=pay)
    assert _pairs(to_np(rows), to_np(vals)) == _pairs(want_rows, want_vals)
    ht.free()


def test_probe_unique_with_selections(ctx, orc):
    from ghive_amd import ops
    rng = np.random.default_rng(3)
    bk = (rng.permutation(400000)[:150000] * 3 + 1).astype(np.int32)
    pk = rng.integers(0, 1200000, 1_000_000).astype(np.int32)
    bsel = np.nonzero(rng.random(len(bk)) < 0.4)[0].astype(np.uint32)
    psel = np.nonzero(rng.random(len(pk)) < 0.6)[0].astype(np.uint32)
    for lf in (0.25, 0.5, 0.9):
        ht = ops.HashTable(ctx, to_dev(bk), sel=to_dev(bsel.astype(np.int32)), load_factor=lf)
        assert ht.info()["unique"]
        rows, vals = ht.probe(to_dev(pk), sel=to_dev(psel.astype(np.int32)))
        want = orc.mapjoin_inner(bk, pk, build_sel=bsel, probe_sel=psel)
        assert _pairs(to_np(rows), to_np(vals)) == _pairs(*want)
        ht.free()


def test_probe_edge_cases(ctx):
    from ghive_amd import ops
    empty = to_dev(np.zeros(0, dtype=np.int64))
    ht = ops.HashTable(ctx, empty)
    rows, vals = ht.probe(to_dev(np.arange(1000, dtype=np.int64)))
    assert rows.numel() == 0
    ht2 = ops.HashTable(ctx, to_dev(np.array([INT_NULL, 5], dtype=np.int32)))
    rows, vals = ht2.probe(to_dev(np.array([INT_NULL, 5, 6], dtype=np.int32)))  # NULL never joins
    assert to_np(rows).tolist() == [1] and to_np(vals).tolist() == [1]
    # a LONG probe key outside the INT range never matches an INT table
    rows, _ = ht2.probe(to_dev(np.array([5 + 2**32, 5], dtype=np.int64)))
    assert to_np(rows).tolist() == [1]


def test_probe_capacity_error_reports_need(ctx):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    ht = ops.HashTable(ctx, to_dev(np.zeros(100, dtype=np.int64)))  # one key, 100 entries
    with pytest.raises(GhiveError) as e:
        ht.probe(to_dev(np.zeros(10, dtype=np.int64)), capacity=50)
    assert e.value.status == abi.ERR_CAPACITY
    rows, _ = ht.probe(to_dev(np.zeros(10, dtype=np.int64)))
    assert rows.numel() == 1000


def test_join_reference_kat(ctx, kats):
    from ghive_amd import ops
    for k in kats["join"]:
        tables = [ops.Table([to_dev(np.array(t, dtype=np.int64))]) for t in k["tables"]]
        jp = ops.JoinPredicate(k["join_cols"], [tuple(c) for c in k["conds"]])
        got = to_np(jp.process(ctx, tables))
        want = np.array(k["expect_offsets"], dtype=np.int32).reshape(-1, k["width"])
        assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, want.tolist())), k["cite"]


@pytest.mark.parametrize("types", [(0, 0), (1, 0), (4, 1), (0, 1), (1, 4)])
def test_join_nway_vs_oracle(ctx, orc, types):
    from ghive_amd import ops
    rng = np.random.default_rng(sum(types) + 8)
    t0 = rng.integers(0, 300, 20000).astype(np.int64)
    t0[::101] = LONG_NULL
    t1 = rng.integers(0, 300, 600).astype(np.int32)
    t2 = rng.integers(100, 400, 500).astype(np.int64)
    conds = [(types[0], 0, 1), (types[1], 0, 2)]
    jp = ops.JoinPredicate([0, 0, 0], conds)
    got = to_np(jp.process(ctx, [ops.Table([to_dev(t)]) for t in (t0, t1, t2)]))
    want = orc.join([[t0], [t1], [t2]], [0, 0, 0], conds)
    assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, want.tolist()))


def test_tables_outlive_their_context(ctx):
    """A hash table (or batch) whose context was shut down is orphaned, not left dangling: using it
    is an INVALID error, and freeing it afterwards only releases the handle."""
    from ghive_amd import ops, workloads as W
    from ghive_amd._lib import GhiveError
    T = W.small_tables(ctx, 10_000, seed=21)
    other = ops.Context(0)
    ht = W.small_dim_table(other, T.dim)
    empty = W.small_dim_table(other, T.dim, attr_value=99)  # a selection that keeps nothing
    assert empty.info()["entries"] == 0
    other.close()
    with pytest.raises(GhiveError) as e:
        W.small_fragment(T.fact, ht).execute(ctx, capacity=64)
    assert e.value.status == abi.ERR_INVALID and "shut down" in str(e.value)
    with pytest.raises(GhiveError):
        ht.probe(T.fact.columns[0])
    ht.free()
    empty.free()
    # the session context is unaffected
    live = W.small_dim_table(ctx, T.dim)
    assert len(W.small_fragment(T.fact, live).execute(ctx, capacity=64)[0]) > 0


def _gather_expected(cols, offsets, null):
    out = np.empty(len(offsets), dtype=cols.dtype)
    ok = offsets >= 0
    out[ok] = cols[offsets[ok]]
    out[~ok] = null
    return out


def test_generate_result_reference_kat(ctx, kats):
    """generate_result over the reference's own join KAT (LEFT SEMI + LEFT OUTER): the maintained key
    columns gathered by the KAT's result_offsets, LONG_MAX where the offset is -1."""
    from ghive_amd import ops
    for k in kats["join"]:
        tables = [ops.Table([to_dev(np.array(t, dtype=np.int64))]) for t in k["tables"]]
        conds = [tuple(c) for c in k["conds"]]
        jp = ops.JoinPredicate(k["join_cols"], conds, maintain_cols=[[0] for _ in tables])
        jp.process(ctx, tables)
        got_cols = [to_np(c) for c in jp.generate_result(ctx, tables)]
        offs = to_np(jp.result_offsets)
        semi = {c[2] for c in conds if c[0] == abi.LEFT_SEMI_JOIN}
        result_tables = [t for t in range(len(tables)) if t not in semi]
        want_rows = sorted(zip(*[_gather_expected(np.array(k["tables"][t], dtype=np.int64),
                                                  np.array(k["expect_offsets"], dtype=np.int32).reshape(-1, k["width"])[:, w],
                                                  LONG_NULL)
                                 for w, t in enumerate(result_tables)]))
        got_rows = sorted(zip(*got_cols))
        assert got_rows == want_rows, k["cite"]
        # and the gather follows this run's own offsets column by column
        for w, t in enumerate(result_tables):
            assert np.array_equal(got_cols[w], _gather_expected(np.array(k["tables"][t], dtype=np.int64),
                                                               offs[:, w], LONG_NULL))


def test_generate_result_types_and_maintained_columns(ctx, orc):
    """Maintained columns of every type in maintain order, NULL sentinels of each type on the outer
    side, the first table's column 0 when it maintains nothing, semi inputs emitting nothing."""
    import torch
    from ghive_amd import ops
    rng = np.random.default_rng(41)
    n0, n1, n2 = 5000, 400, 300
    k0 = rng.integers(0, 500, n0).astype(np.int64)
    k1 = rng.integers(0, 500, n1).astype(np.int32)
    k2 = rng.integers(200, 700, n2).astype(np.int64)
    t0 = [k0, rng.standard_normal(n0), rng.integers(-9, 9, n0).astype(np.int32)]
    t1 = [k1, rng.standard_normal(n1).astype(np.float32), rng.integers(0, 1 << 40, n1).astype(np.int64)]
    t2 = [k2, rng.integers(0, 9, n2).astype(np.int32)]
    tables = [ops.Table([to_dev(c) for c in t]) for t in (t0, t1, t2)]
    for conds, maintain in ((((abi.LEFT_OUTER_JOIN, 0, 1), (abi.LEFT_OUTER_JOIN, 0, 2)), [[2, 1], [1, 2, 0], [1]]),
                            (((abi.INNER_JOIN, 0, 1), (abi.LEFT_SEMI_JOIN, 0, 2)), [[], [2], [1]])):
        jp = ops.JoinPredicate([0, 0, 0], conds, maintain_cols=maintain)
        jp.process(ctx, tables)
        offs = to_np(jp.result_offsets)
        want_offs = orc.join([[k0], [k1], [k2]], [0, 0, 0], list(conds))
        assert sorted(map(tuple, offs.tolist())) == sorted(map(tuple, want_offs.tolist()))
        got = [to_np(c) for c in jp.generate_result(ctx, tables)]
        semi = {c[2] for c in conds if c[0] == abi.LEFT_SEMI_JOIN}
        result_tables = [t for t in range(3) if t not in semi]
        want = []
        for w, t in enumerate(result_tables):
            cols = maintain[t] if (maintain[t] or w > 0) else [0]
            for c in cols:
                src = (t0, t1, t2)[t][c]
                null = {np.dtype(np.int64): LONG_NULL, np.dtype(np.int32): 2**31 - 1,
                        np.dtype(np.float64): np.nan, np.dtype(np.float32): np.float32(np.nan)}[src.dtype]
                want.append(_gather_expected(src, offs[:, w], null))
        assert len(got) == len(want)
        for g, x in zip(got, want):
            assert g.dtype == x.dtype
            assert np.array_equal(g, x, equal_nan=g.dtype.kind == "f")


@pytest.fixture
def regional(ctx):
    """Every build in the test takes the regional form (option hash.regional_rows 1)."""
    ctx.set_option("hash.regional_rows", 1)
    yield
    ctx.set_option("hash.regional_rows", 1 << 22)


@pytest.mark.parametrize("ktype", ["INT", "LONG"])
@pytest.mark.parametrize("payload", [None, "INT", "LONG"])
def test_regional_build_vs_oracle(ctx, orc, regional, ktype, payload):
    """The regional build (rows sorted by table region, each region of 8192 slots built in LDS):
    duplicate keys, NULL keys, every key and payload type; 65536 slots = 8 regions. The probe's
    (row, payload) multiset equals the oracle's and the statistics match the rows."""
    from ghive_amd import ops
    rng = np.random.default_rng(17)
    dt = np.int32 if ktype == "INT" else np.int64
    null = INT_NULL if ktype == "INT" else LONG_NULL
    bk = rng.integers(-3000, 3000, 30000).astype(dt)
    bk[::50] = null
    pk = rng.integers(-4000, 4000, 200000).astype(dt)
    pk[::31] = null
    pay = None
    if payload == "INT":
        pay = rng.integers(-2**31, 2**31 - 1, 30000).astype(np.int32)
    elif payload == "LONG":
        pay = rng.integers(-2**62, 2**62, 30000).astype(np.int64)
    ht = ops.HashTable(ctx, to_dev(bk), payload=None if pay is None else to_dev(pay))
    info = ht.info()
    live = bk[bk != null]
    assert not info["unique"] and info["entries"] == len(live) and info["slots"] == 65536
    rows, vals = ht.probe(to_dev(pk))
    want_rows, want_vals = orc.mapjoin_inner(bk, pk, payload=pay)
    assert _pairs(to_np(rows), to_np(vals)) == _pairs(want_rows, want_vals)
    ht.free()


def test_regional_build_unique_with_selections(ctx, orc, regional):
    """Unique INT keys through a build selection at load factors 0.25, 0.5 and 0.9 (the last fills
    regions enough that clusters run past a region's end into the overflow pass)."""
    from ghive_amd import ops
    rng = np.random.default_rng(5)
    bk = (rng.permutation(900000)[:400000] * 3 + 1).astype(np.int32)
    pk = rng.integers(0, 2700000, 1_000_000).astype(np.int32)
    bsel = np.nonzero(rng.random(len(bk)) < 0.7)[0].astype(np.uint32)
    psel = np.nonzero(rng.random(len(pk)) < 0.6)[0].astype(np.uint32)
    for lf in (0.25, 0.5, 0.9):
        ht = ops.HashTable(ctx, to_dev(bk), sel=to_dev(bsel.astype(np.int32)), load_factor=lf)
        info = ht.info()
        assert info["unique"] and info["entries"] == len(bsel)
        rows, vals = ht.probe(to_dev(pk), sel=to_dev(psel.astype(np.int32)))
        want = orc.mapjoin_inner(bk, pk, build_sel=bsel, probe_sel=psel)
        assert _pairs(to_np(rows), to_np(vals)) == _pairs(*want)
        ht.free()


def test_regional_build_overflow_and_wrap(ctx):
    """Keys made to share home slots at the end of region 3 and at the end of the last region (the
    table's last slots): their clusters run past the region, into the overflow pass, and past the
    table's end to slot 0. 65536 slots (20000 keys at load factor 0.5). Every key is found with its
    payload, absent keys are not, and the regional and global forms give the same probe output and
    statistics."""
    from ghive_amd import ops
    M = 1 << 64
    C = 0x9E3779B97F4A7C15
    inv = pow(C, -1, M)
    log2 = 16
    rng = np.random.default_rng(23)

    def key_at(slot):  # a key whose home slot is `slot`: its product's top log2 bits
        w = (slot << (64 - log2)) | int(rng.integers(0, 1 << (64 - log2)))
        k = (w * inv) % M
        return k - M if k >= 1 << 63 else k

    crafted = [key_at(4 * 8192 - 1 - int(rng.integers(0, 8))) for _ in range(300)]
    crafted += [key_at((1 << log2) - 1 - int(rng.integers(0, 8))) for _ in range(300)]
    rest = rng.integers(-2**62, 2**62, 20000 - len(crafted)).astype(np.int64)
    bk = np.unique(np.concatenate([np.array(crafted, dtype=np.int64), rest]))
    rng.shuffle(bk)
    pay = rng.integers(-2**62, 2**62, len(bk)).astype(np.int64)
    absent = rng.integers(-2**62, 2**62, 5000).astype(np.int64)
    absent = absent[~np.isin(absent, bk)]
    pk = np.concatenate([bk, absent])
    out = {}
    for rows_min in (1, 0):  # regional, then global
        ctx.set_option("hash.regional_rows", rows_min)
        try:
            ht = ops.HashTable(ctx, to_dev(bk), payload=to_dev(pay), load_factor=0.5)
            info = ht.info()
            rows, vals = ht.probe(to_dev(pk))
            out[rows_min] = (info, _pairs(to_np(rows), to_np(vals)))
            ht.free()
        finally:
            ctx.set_option("hash.regional_rows", 1 << 22)
    info, pairs = out[1]
    assert info["unique"] and info["entries"] == len(bk) and info["slots"] == 1 << log2
    assert pairs == sorted(zip(range(len(bk)), pay.tolist()))
    assert out[1] == out[0]


def test_filtered_build_equals_filter_then_build(ctx):
    """ghive_hash_build_filtered over four build sides at once (INT key with an INT payload, LONG key
    with a LONG payload, INT key with the row id, LONG key with no filter; NULL keys; AND / OR
    trees; a filter that keeps nothing) against filter_async + the asynchronous build of each: the
    same entries, slots and uniqueness, the same probe output, and the same dense companion (a
    fragment over the tables gives the same groups)."""
    from ghive_amd import ops
    rng = np.random.default_rng(41)
    n = 300_000
    ki = rng.permutation(n).astype(np.int32) + 1
    ki[::97] = INT_NULL
    kl = rng.integers(-2**60, 2**60, n).astype(np.int64)
    kl[::89] = LONG_NULL
    a = rng.integers(0, 7, n).astype(np.int32)
    b = rng.integers(-100, 100, n).astype(np.int64)
    pi = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    T = ops.Table([to_dev(ki), to_dev(kl), to_dev(a), to_dev(b), to_dev(pi)])
    P = ops.FilterPredicate
    eq = P([abi.make_pred(abi.FILTER_EQ, 2, [3], data_type=2)], 0)
    tree = P([abi.make_pred(abi.FILTER_EQ_RANGE, 2, [1, 2], data_type=2), abi.make_pred(abi.FILTER_GT, 3, [40]),
              abi.make_pred(abi.FILTER_AND, children=(0, 1)), abi.make_pred(abi.FILTER_LT, 3, [-90]),
              abi.make_pred(abi.FILTER_OR, children=(2, 3))], 4)
    none = P([abi.make_pred(abi.FILTER_EQ, 2, [99], data_type=2)], 0)
    builds = [(T, eq, 0, 4, 0.5), (T, tree, 1, 3, 0.5), (T, eq, 0, None, 0.75), (T, None, 1, None, 0.5),
              (T, none, 0, 4, 0.5)]
    fused = ops.filtered_hash_tables(ctx, builds)
    probe_i = to_dev(np.concatenate([ki, rng.integers(0, n + 1000, 50_000).astype(np.int32)]))
    probe_l = to_dev(np.concatenate([kl, rng.integers(-2**60, 2**60, 50_000).astype(np.int64)]))
    for (tab, pred, kc, pc, lf), hf in zip(builds, fused):
        if pred is None:
            sel, cnt = None, None
            hs = ops.HashTable(ctx, tab.columns[kc], payload=None if pc is None else tab.columns[pc], load_factor=lf)
        else:
            sel, cnt = pred.process_async(ctx, tab)
            hs = ops.HashTable(ctx, tab.columns[kc], sel=sel, payload=None if pc is None else tab.columns[pc],
                               load_factor=lf, count=cnt)
        fi, si = hf.info(), hs.info()
        assert fi["entries"] == si["entries"] and fi["unique"] == si["unique"]
        if pred is not None:
            assert fi["slots"] == si["slots"]
        probe = probe_i if kc == 0 else probe_l
        rf, vf = hf.probe(probe)
        rs, vs = hs.probe(probe)
        assert _pairs(to_np(rf), to_np(vf)) == _pairs(to_np(rs), to_np(vs))
        hs.free()
        hf.free()


def test_star_fragment_over_filtered_builds(ctx):
    """The star-join fragment over the dimension tables of one filtered build (the default) and over
    filter-then-build tables: the same groups and sums; then one fragment descriptor rebound to
    fresh tables for each of three more runs (as the benchmark's step does), into the same result
    buffers, with the same result each time."""
    import torch
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, fact_rows=3_000_000, seed=77)

    def canon(out):
        keys = np.stack([to_np(o).astype(np.int64) for o in out[:-1]], axis=1)
        order = np.lexsort(keys.T[::-1])
        return [to_np(o)[order].tolist() for o in out]

    res = []
    for fused in (True, False):
        hd, hs, hp = W.star_hash_tables(ctx, T, fused=fused)
        res.append(canon(W.star_fragment(T.fact, hd, hs, hp).execute(ctx, capacity=8192)))
        for h in (hd, hs, hp):
            h.free()
    assert res[0] == res[1] and len(res[0][0]) > 0
    builds = W.star_builds(T)
    frag = None
    outs = [torch.empty(8192, dtype=torch.int32, device="cuda") for _ in range(2)] + \
           [torch.empty(8192, dtype=torch.int64, device="cuda")]
    for _ in range(3):
        tabs = ops.filtered_hash_tables(ctx, builds)
        if frag is None:
            frag = W.star_fragment(T.fact, *tabs)
        else:
            frag.rebind(tabs)
        assert canon(frag.execute(ctx, capacity=8192, outs=outs)) == res[0]
        for h in tabs:
            h.free()
