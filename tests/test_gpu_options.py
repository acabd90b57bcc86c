"""Tuning options: defaults, round trips and rejected values; the filtered build's argument checks
(no kernel runs for a rejected call)."""
import numpy as np
import pytest

from conftest import to_dev
from ghive_amd import abi

pytestmark = pytest.mark.gpu


def test_option_defaults_round_trips_and_rejections(ctx):
    from ghive_amd._lib import GhiveError
    defaults = {"sort.pack": 0, "sort.occupancy": 2, "sort.merge": 1, "groupby.part_occupancy": 2,
                "hash.regional_rows": 1 << 22, "sort.shape": 1, "groupby.part_threads": 512}
    for name, want in defaults.items():
        assert ctx.get_option(name) == want, name
    for name, value in (("sort.occupancy", 1), ("groupby.part_occupancy", 1), ("hash.regional_rows", 0),
                        ("sort.merge", 0), ("sort.pack", 1)):
        ctx.set_option(name, value)
        assert ctx.get_option(name) == value
        ctx.set_option(name, defaults[name])
    for name, bad in (("sort.occupancy", 3), ("groupby.part_occupancy", 0), ("hash.regional_rows", -1),
                      ("sort.shape", 4), ("no.such.option", 1), ("groupby.partition_fallbacks", 0)):
        with pytest.raises(GhiveError) as e:
            ctx.set_option(name, bad)
        assert e.value.status == abi.ERR_INVALID
    for name, want in defaults.items():
        assert ctx.get_option(name) == want, name


def test_filtered_build_rejects_bad_arguments(ctx):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    k = to_dev(np.arange(1000, dtype=np.int32))
    f = to_dev(np.linspace(0, 1, 1000).astype(np.float32))
    T = ops.Table([k, f])
    pred = ops.FilterPredicate([abi.make_pred(abi.FILTER_GT, 0, [10], data_type=2)], 0)
    cases = [[],                                   # no build side
             [(T, pred, 0, None, 0.5)] * 9,       # more than GHIVE_MAX_FILTERED_BUILDS
             [(T, pred, 2, None, 0.5)],           # key column out of range
             [(T, pred, 1, None, 0.5)],           # FLOAT key
             [(T, pred, 0, 1, 0.5)],              # FLOAT payload
             [(T, pred, 0, 5, 0.5)],              # payload column out of range
             [(T, pred, 0, None, 0.99)]]          # load factor above 0.95
    for builds in cases:
        with pytest.raises((GhiveError, ValueError)) as e:
            ops.filtered_hash_tables(ctx, builds)
        if isinstance(e.value, GhiveError):
            assert e.value.status == abi.ERR_INVALID
    bad_pred = ops.FilterPredicate([abi.make_pred(abi.FILTER_GT, 7, [10], data_type=2)], 0)  # no column 7
    with pytest.raises(GhiveError):
        ops.filtered_hash_tables(ctx, [(T, bad_pred, 0, None, 0.5)])
    (h,) = ops.filtered_hash_tables(ctx, [(T, pred, 0, None, 0.5)])  # and the valid call
    assert h.info()["entries"] == 989
    h.free()


def test_partition_overflow_counts_one_fallback(ctx):
    """The partitioned group-by sizes its partitions from the group bound: with a bound far below
    the real group count its one partition overflows the LDS table, the direct form runs instead
    (and reports the bound exceeded), and the read-only counter records it; a bound that holds
    every group runs partitioned and adds nothing."""
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    n = 300_000
    keys = to_dev(np.arange(n, dtype=np.int64) * 7919 + 11)
    gb = ops.GroupByPredicate([0], [(abi.AGG_CNT, abi.COUNT_STAR)])
    ctx.set_option("groupby.partition", 2)
    try:
        f0 = ctx.get_option("groupby.partition_fallbacks")
        with pytest.raises(GhiveError) as e:
            gb.gpu_execute(ctx, ops.Table([keys]), capacity=1000)
        assert e.value.status == abi.ERR_CAPACITY
        assert ctx.get_option("groupby.partition_fallbacks") == f0 + 1
        out = gb.gpu_execute(ctx, ops.Table([keys]), capacity=n)
        assert out[0].numel() == n and int(out[1].sum()) == n
        assert ctx.get_option("groupby.partition_fallbacks") == f0 + 1
    finally:
        ctx.set_option("groupby.partition", 1)
