"""GPU parity of the generator and of filter + compaction against the CPU oracle (bit-exact row ids)."""
import numpy as np
import pytest

from conftest import to_dev, to_np
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
