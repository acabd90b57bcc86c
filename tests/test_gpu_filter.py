"""GPU parity of the generator and of filter + compaction against the CPU oracle (bit-exact row ids)."""
import numpy as np
import pytest

from conftest import str_col, to_dev, to_dev_col, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
INT_NULL, LONG_NULL = abi.INT_NULL, abi.LONG_NULL


def test_generator_matches_oracle(ctx, orc):
    seed = 0x6768697667
    for col, (lo, hi) in enumerate([(1, 200000), (0, 6), (-2**31, 2**31 - 1), (100000, 10000000)]):
        g = to_np(ctx.gen_uniform_i32(seed, col, 100003, lo, hi, row_begin=17))
        w = orc.gen_uniform_i32(seed, col, 17, 100003, lo, hi)
        bad = np.nonzero(g != w)[0]
        assert len(bad) == 0, (col, lo, hi, len(bad), bad[:8].tolist(), g[bad[:8]].tolist(), w[bad[:8]].tolist())
    g = to_np(ctx.gen_uniform_i64(seed, 9, 50001, -10**17, 10**17))
    assert np.array_equal(g, orc.gen_uniform_i64(seed, 9, 0, 50001, -10**17, 10**17))
    g = to_np(ctx.gen_uniform_f32(seed, 5, 50001))
    assert np.array_equal(g, orc.gen_uniform_f32(seed, 5, 0, 50001))
    z = to_np(ctx.gen_normal_f64(seed, 6, 50001))
    assert np.allclose(z, orc.gen_normal_f64(seed, 6, 0, 50001), rtol=1e-13, atol=1e-14)
    s = to_np(ctx.gen_sequence_i32(1000, 1, row_begin=5))
    assert np.array_equal(s, np.arange(6, 1006, dtype=np.int32))


def _modes_case(ctype, n, seed):
    rng = np.random.default_rng(seed)
    if ctype == "DOUBLE":
        x = rng.integers(-20, 21, n).astype(np.float64) / 2
        x[rng.random(n) < 0.05] = np.nan
        return x, (-3.5, 4.0)
    if ctype == "FLOAT":
        x = (rng.integers(-20, 21, n) / 2).astype(np.float32)
        x[rng.random(n) < 0.05] = np.nan
        return x, (-3.5, 4.0)
    dt = np.int32 if ctype == "INT" else np.int64
    x = rng.integers(-20, 21, n).astype(dt)
    x[rng.random(n) < 0.05] = INT_NULL if ctype == "INT" else LONG_NULL
    return x, (-3, 4)


@pytest.mark.parametrize("ctype", ["INT", "LONG", "DOUBLE", "FLOAT"])
def test_filter_every_mode_vs_oracle(ctx, orc, ctype):
    from ghive_amd import ops
    for n in (0, 1, 63, 64, 65, 4095, 4096, 4097, 300001):
        x, params = _modes_case(ctype, n, n + 7)
        t = ops.Table([to_dev(x)])
        for mode in range(11):
            node = abi.make_pred(mode, 0, params if mode != abi.FILTER_NOT_NULL else ())
            got = to_np(ops.FilterPredicate([node], 0).process(ctx, t)).astype(np.uint32)
            want = orc.filter_rows([x], [node], 0)
            assert np.array_equal(got, want), (ctype, n, mode)


def test_filter_tree_and_columns_vs_oracle(ctx, orc):
    from ghive_amd import ops
    rng = np.random.default_rng(5)
    n = 1_000_003
    a = rng.integers(0, 10, n).astype(np.int32)
    b = rng.integers(0, 10, n).astype(np.int64)
    c = rng.random(n)
    a[::97] = INT_NULL
    c[::89] = np.nan
    nodes = [
        abi.make_pred(abi.FILTER_OR, children=(1, 6)),
        abi.make_pred(abi.FILTER_AND, children=(2, 5)),
        abi.make_pred(abi.FILTER_OR, children=(3, 4)),
        abi.make_pred(abi.FILTER_EQ, 0, [3], data_type=2),
        abi.make_pred(abi.FILTER_EQ_RANGE, 1, [2, 4]),
        abi.make_pred(abi.FILTER_LT, 2, [0.5]),
        abi.make_pred(abi.FILTER_AND, children=(7, 8)),
        abi.make_pred(abi.FILTER_GT, 0, [7], data_type=2),
        abi.make_pred(abi.FILTER_GE, 0, col2=1),  # column vs column
    ]
    t = ops.Table([to_dev(a), to_dev(b), to_dev(c)])
    got = to_np(ops.FilterPredicate(nodes, 0).process(ctx, t)).astype(np.uint32)
    assert np.array_equal(got, orc.filter_rows([a, b, c], nodes, 0))


@pytest.mark.parametrize("n", [4095, 4096, 12288, 12289, 1_000_003])
def test_filter_int_tree_whole_tiles_vs_oracle(ctx, orc, n):
    """AND/OR trees over INT columns only: whole 4096-row tiles go through the 16-byte-load form,
    the partial tile through the general one; a column slice that is not 16-byte aligned sends
    everything through the general form. Both give the oracle's rows."""
    from ghive_amd import ops
    rng = np.random.default_rng(n)
    a = rng.integers(0, 10, n + 1).astype(np.int32)
    b = rng.integers(-5, 5, n + 1).astype(np.int32)
    a[::61] = INT_NULL
    b[::53] = INT_NULL
    nodes = [
        abi.make_pred(abi.FILTER_OR, children=(1, 4)),
        abi.make_pred(abi.FILTER_AND, children=(2, 3)),
        abi.make_pred(abi.FILTER_EQ_RANGE, 0, [2, 6], data_type=2),
        abi.make_pred(abi.FILTER_NOT_EQ, 1, [0], data_type=2),
        abi.make_pred(abi.FILTER_NOT_EQ_RANGE, 1, [-4, 3], data_type=2),
    ]
    da, db = to_dev(a), to_dev(b)
    for sl in (slice(0, n), slice(1, n + 1)):
        t = ops.Table([da[sl], db[sl]])
        got = to_np(ops.FilterPredicate(nodes, 0).process(ctx, t)).astype(np.uint32)
        assert np.array_equal(got, orc.filter_rows([a[sl], b[sl]], nodes, 0)), (n, sl)


def test_filter_hive_kats_on_gpu(ctx, kats):
    from ghive_amd import ops
    for k in kats["hive_filter"]:
        dt = np.int64 if k["type"] == "LONG" else np.float64
        x = np.array(k["values"], dtype=dt)
        for i in k["nulls"]:
            x[i] = LONG_NULL if dt == np.int64 else np.nan
        cols = [x]
        p = k["pred"]
        if p.get("col2"):
            cols.append(np.array(k["values2"], dtype=dt))
            node = abi.make_pred(p["mode"], 0, col2=1)
        else:
            params = [float(v) for v in p["params"]] if k["type"] == "DOUBLE" else p["params"]
            node = abi.make_pred(p["mode"], 0, params)
        got = to_np(ops.FilterPredicate([node], 0).process(ctx, ops.Table([to_dev(c) for c in cols]))).tolist()
        if "sel" in k:
            got = [r for r in got if r in k["sel"]]
        assert got == k["expect"], k["cite"]


def test_clean_data_compacts_every_column(ctx, orc):
    from ghive_amd import ops
    rng = np.random.default_rng(9)
    n = 200_000
    cols = [rng.integers(0, 7, n).astype(np.int32), rng.integers(-5, 5, n).astype(np.int64),
            rng.random(n), rng.random(n).astype(np.float32)]
    op = ops.FilterOperator("(c0 = 3 and c1 >= 0)", ["c0", "c1", "c2", "c3"])
    out = op.clean_data(ctx, ops.Table([to_dev(c) for c in cols]))
    sel = orc.filter_rows(cols, op.generatePredicate().nodes, op.generatePredicate().root)
    assert out.row_num == len(sel)
    for got, c in zip(out.columns, cols):
        assert np.array_equal(to_np(got), c[sel])


def test_filter_large_size_independent_properties(ctx):
    """At 200M rows: ids strictly ascending, every id passes, count = closed-form count."""
    import torch
    from ghive_amd import ops
    n = 200_000_000
    x = ctx.gen_uniform_i32(77, 0, n, 0, 9)
    sel = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 0, [5], data_type=2)], 0).process(
        ctx, ops.Table([x]))
    assert sel.numel() == int((x == 5).sum().item())
    s64 = sel.to(torch.int64) & 0xFFFFFFFF
    assert bool((s64[1:] > s64[:-1]).all().item())
    assert bool((x[s64] == 5).all().item())


def test_filter_rejects_bad_descriptors(ctx):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    t = ops.Table([to_dev(np.arange(10, dtype=np.int32))])
    with pytest.raises(GhiveError):
        ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 3, [1])], 0).process(ctx, t)  # no column 3
    cyc = [abi.make_pred(abi.FILTER_AND, children=(1, 0)), abi.make_pred(abi.FILTER_EQ, 0, [1])]
    with pytest.raises(GhiveError):
        ops.FilterPredicate(cyc, 0).process(ctx, t)


# ---------------------------------------------------------------- STRING columns --------------
def _rand_strings(rng, n, null=0.08):
    alphabet = [b"a", b"b", b"#", b"1", b"\x7f", b"\x80", b"\xff"]
    return [None if rng.random() < null else b"".join(alphabet[i] for i in rng.integers(0, 7, rng.integers(0, 5)))
            for _ in range(n)]


def test_string_filter_every_mode_vs_oracle(ctx, orc):
    """STRING column against string literals in every mode, and against another STRING column, in
    Hive's order (unsigned bytes, then length), NULLs and empty strings included."""
    from ghive_amd import ops
    rng = np.random.default_rng(31)
    lo, hi = b"a\x80", b"b#"
    for n in (0, 1, 65, 4097, 120_001):
        vals, vals2 = _rand_strings(rng, n), _rand_strings(rng, n)
        cols = [str_col(vals), str_col(vals2)]
        t = ops.Table([to_dev_col(c) for c in cols])
        for mode in range(11):
            two = mode in (abi.FILTER_RANGE, abi.FILTER_EQ_RANGE, abi.FILTER_NOT_RANGE, abi.FILTER_NOT_EQ_RANGE)
            params = () if mode == abi.FILTER_NOT_NULL else ([lo, hi] if two else [lo])
            node = abi.make_pred(mode, 0, params)
            got = to_np(ops.FilterPredicate([node], 0).process(ctx, t)).astype(np.uint32)
            assert np.array_equal(got, orc.filter_rows(cols, [node], 0)), (n, mode)
        for mode in (abi.FILTER_EQ, abi.FILTER_LT, abi.FILTER_GE, abi.FILTER_NOT_EQ):
            node = abi.make_pred(mode, 0, col2=1)
            got = to_np(ops.FilterPredicate([node], 0).process(ctx, t)).astype(np.uint32)
            assert np.array_equal(got, orc.filter_rows(cols, [node], 0)), (n, "col-col", mode)


def test_string_filter_reference_kat_on_gpu(ctx):
    """FilterPredicateTest.cu:99-204: clean_data of (INT, STRING, DOUBLE) rows by
    brand BETWEEN 'MFGR#1221' AND 'MFGR#1223' keeps the six rows its expectation lists."""
    from ghive_amd import ops
    brands = [b"MFGR#1221", b"MFGR#1223", b"MFGR#1238", b"MFGR#1221", b"MFGR#1238", b"MFGR#122",
              b"MFGR#1232", b"MFGR#1228", b"MFGR#123", b"MFGR#1224", b"MFGR#1248", b"MFGR#12223",
              b"MFGR#1237", b"MFGR#1222", b"MFGR#1230", b"MFGR#1221", b"MFGR#1234", b"MFGR#1220",
              b"MFGR#1239", b"MFGR#12234"]
    years = np.array([1998] * 13 + [1997] * 7, dtype=np.int32)
    price = np.array([30000.0, 80000.0, 75000.0, 27000.0, 14660.0, 23333.0, 77000.0, 38438.0, 10395.0,
                      22422.0, 73300.0, 27999.0, 69897.0, 12941.0, 20719.0, 75246.5, 13932.4, 83299.1,
                      49278.3, 46527.6])
    t = ops.Table([to_dev(years), to_dev_col(str_col(brands)), to_dev(price)])
    op = ops.FilterOperator("b BETWEEN 'MFGR#1221' AND 'MFGR#1223'", ["y", "b", "p"])
    out = op.clean_data(ctx, t)
    rows = list(zip(to_np(out.columns[0]).tolist(), out.columns[1].to_list(), to_np(out.columns[2]).tolist()))
    assert rows == [(1998, b"MFGR#1221", 30000.0), (1998, b"MFGR#1223", 80000.0), (1998, b"MFGR#1221", 27000.0),
                    (1998, b"MFGR#12223", 27999.0), (1997, b"MFGR#1222", 12941.0), (1997, b"MFGR#1221", 75246.5)]


def test_clean_data_with_string_columns_vs_oracle(ctx, orc):
    """Predicate text mixing string and numeric leaves; every column compacted, STRING ones as
    (start,end) pairs into the shared blob."""
    from ghive_amd import ops
    rng = np.random.default_rng(8)
    n = 300_000
    region = [[b"AMERICA", b"ASIA", b"EUROPE", b"AFRICA", None][i] for i in rng.integers(0, 5, n)]
    nation = _rand_strings(rng, n)
    qty = rng.integers(0, 50, n).astype(np.int64)
    cols = [str_col(region), str_col(nation), qty]
    op = ops.FilterOperator("((c_region = 'AMERICA') or ((c_nation >= 'b') and (q < 10L)))", ["c_region", "c_nation", "q"])
    out = op.clean_data(ctx, ops.Table([to_dev_col(c) for c in cols]))
    pred = op.generatePredicate()
    sel = orc.filter_rows(cols, pred.nodes, pred.root)
    assert out.row_num == len(sel) > 0
    assert out.columns[0].to_list() == [region[i] for i in sel]
    assert out.columns[1].to_list() == [nation[i] for i in sel]
    assert np.array_equal(to_np(out.columns[2]), qty[sel])
