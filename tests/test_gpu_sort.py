"""GPU parity of the LSD radix ORDER BY: the permutation equals the oracle's total order
(keys, then row id) exactly."""
import numpy as np
import pytest

from conftest import str_col, to_dev, to_dev_col, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu


def _check(ctx, orc, cols, keys, orders):
    from ghive_amd import ops
    perm = to_np(ops.sort(ctx, ops.Table([to_dev_col(c) for c in cols]), keys, orders)).astype(np.uint32)
    want = orc.sort_perm(cols, keys, orders)
    assert np.array_equal(perm, want)


@pytest.mark.parametrize("n", [0, 1, 2, 4095, 4096, 4097, 250_001])
def test_sort_int64_float32(ctx, orc, n):
    rng = np.random.default_rng(n)
    k = rng.integers(-50, 50, n).astype(np.int64)
    f = rng.random(n).astype(np.float32)
    f[::53] = np.nan
    _check(ctx, orc, [k, f], [0, 1], ["+", "+"])


@pytest.mark.parametrize("ctype", ["INT", "LONG", "DOUBLE", "FLOAT"])
@pytest.mark.parametrize("order", ["+", "-"])
def test_sort_single_key(ctx, orc, ctype, order):
    rng = np.random.default_rng(7)
    n = 300_000
    if ctype == "INT":
        x = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
        x[::99] = abi.INT_NULL
    elif ctype == "LONG":
        x = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
        x[::7] = x[::7] % 1000  # many duplicates
    elif ctype == "DOUBLE":
        x = rng.standard_normal(n) * 1e300 ** rng.random(n)
        x[::31] = np.nan
        x[::37] = -0.0
        x[::41] = 0.0
        x[::43] = np.inf
        x[::47] = -np.inf
    else:
        x = (rng.standard_normal(n) * 100).astype(np.float32)
        x[::31] = np.nan
    _check(ctx, orc, [x], [0], [order])


def test_sort_composite_mixed_orders(ctx, orc):
    rng = np.random.default_rng(3)
    n = 200_000
    a = rng.integers(0, 5, n).astype(np.int32)
    b = rng.integers(-3, 3, n).astype(np.int64)
    c = (rng.integers(0, 4, n) / 2).astype(np.float64)
    _check(ctx, orc, [a, b, c], [0, 1, 2], ["-", "+", "-"])
    _check(ctx, orc, [a, b, c], [2, 0], ["+", "+"])
    # three and four LONG keys: 192 and 256 bits, three and four key words with the row ids beside
    # them (no spare bits), the first pass building every word
    d = rng.integers(-2**62, 2**62, n).astype(np.int64)
    e = rng.integers(0, 3, n).astype(np.int64)
    _check(ctx, orc, [b, e, d], [0, 1, 2], ["+", "-", "+"])
    _check(ctx, orc, [e, b, d, e], [0, 1, 2, 3], ["-", "+", "-", "+"])


@pytest.mark.parametrize("n", [1, 4097, 200_003])
def test_sort_string_keys(ctx, orc, n):
    """STRING keys (ranked among their distinct values, then sorted as INT) against the oracle:
    shared prefixes, empty strings, NUL and high bytes, NULLs, long strings spanning several
    8-byte chunks, both orders, with a numeric tiebreak key."""
    rng = np.random.default_rng(n)
    alphabet = [b"a", b"b", b"\x00", b"\x80", b"\xff", b"MFGR#1", b"UNITED KI"]
    vals = [None if rng.random() < 0.03 else b"".join(alphabet[i] for i in rng.integers(0, 7, rng.integers(0, 6)))
            for _ in range(n)]
    k2 = rng.integers(0, 4, n).astype(np.int64)
    col = str_col(vals)
    for orders in (["+", "+"], ["-", "+"], ["+", "-"]):
        _check(ctx, orc, [col, k2], [0, 1], orders)
    _check(ctx, orc, [k2, col], [0, 1], ["+", "-"])


def test_sort_large_sorted_and_permutation(ctx):
    """20M rows: keys non-decreasing along the permutation, and it is a permutation."""
    import torch
    from ghive_amd import ops
    n = 20_000_000
    k = ctx.gen_uniform_i64(9, 0, n, -10**12, 10**12)
    f = ctx.gen_uniform_f32(9, 1, n)
    perm = ops.sort(ctx, ops.Table([k, f]), [0, 1], ["+", "+"]).to(torch.int64) & 0xFFFFFFFF
    ks, fs = k[perm], f[perm]
    assert bool(((ks[1:] > ks[:-1]) | ((ks[1:] == ks[:-1]) & (fs[1:] >= fs[:-1]))).all().item())
    assert bool((torch.bincount(perm, minlength=n) == 1).all().item())


def test_sort_info_counts_the_passes_it_ran(ctx, orc):
    """ghive_last_sort_info: the first pass sorts on the lowest key byte while it builds the
    composite keys from the key columns, and any later digit byte that is the same in every row
    costs no pass. Keys in [0, 256) vary in that byte only: one pass. LONG: the 8-byte column in,
    an 8-byte key word and a 4-byte id out, plus the 1-byte digit plane of the next byte; INT: the
    4-byte column in, one 8-byte word out (the id in its low half) and the plane. Keys that vary
    in a higher byte (LONG, multiples of 256): that byte's pass too, 8 + 4 bytes each way reading
    the plane. One value in every row: the first pass alone. The permutations equal the oracle's."""
    from ghive_amd import ops
    n = 100_003
    rng = np.random.default_rng(5)
    for dtype, key_words, embedded, per_row in ((np.int64, 1, False, 8 + 12 + 1), (np.int32, 1, True, 4 + 8 + 1)):
        k = rng.integers(0, 256, n).astype(dtype)
        _check(ctx, orc, [k], [0], ["+"])
        info = ops.sort_info(ctx)
        assert info == {"rows": n, "key_words": key_words, "passes": 1, "embedded_ids": int(embedded),
                        "scatter_bytes": per_row * n}
    k = rng.integers(0, 256, n).astype(np.int64) * 256
    _check(ctx, orc, [k], [0], ["+"])
    info = ops.sort_info(ctx)
    assert info["passes"] == 2 and info["scatter_bytes"] == (8 + 12 + 1) * n + (12 + 12) * n
    _check(ctx, orc, [np.full(n, 7, dtype=np.int64)], [0], ["-"])
    assert ops.sort_info(ctx)["passes"] == 1


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("n", [1, 8191, 8193, 16383, 16385, 300_007])
def test_sort_tile_shapes(ctx, orc, shape, n):
    """Every LSD tile shape (option sort.shape), with and without 16-byte LDS records (sort.pack),
    with and without the pass that merges a key word's high half and its row id into one word
    (sort.merge), compiled for one or two workgroups per CU (sort.occupancy), gives the oracle's
    permutation, on sizes around both tile lengths: the row id carried
    beside one key word (INT + FLOAT) and beside two (LONG + FLOAT + INT), and inside two key words
    through the pass that drops word 0 (LONG + FLOAT)."""
    rng = np.random.default_rng(shape * 1000 + n % 997)
    k = rng.integers(-3000, 3000, n).astype(np.int64)
    f = rng.random(n).astype(np.float32)
    f[::61] = np.nan
    i32 = rng.integers(-5, 5, n).astype(np.int32)
    ctx.set_option("sort.shape", shape)
    try:
        for pack, merge, occ in ((0, 1, 2), (1, 1, 2), (0, 1, 1), (0, 0, 2)):
            ctx.set_option("sort.pack", pack)
            ctx.set_option("sort.merge", merge)
            ctx.set_option("sort.occupancy", occ)
            _check(ctx, orc, [i32, f], [0, 1], ["+", "-"])
            _check(ctx, orc, [k, f], [0, 1], ["-", "+"])
            _check(ctx, orc, [k, f, i32], [0, 1, 2], ["+", "+", "-"])
    finally:
        ctx.set_option("sort.shape", 1)
        ctx.set_option("sort.pack", 0)
        ctx.set_option("sort.merge", 1)
        ctx.set_option("sort.occupancy", 2)


def test_sort_info_counts_digit_planes(ctx, orc):
    """Two varying bytes: the first pass (the 8-byte key column in, key word and id out) writes the
    second pass's digit as a one-byte plane (one byte a row more on write-out), the second histogram
    reads it; the order is the oracle's."""
    from ghive_amd import ops
    n = 70_001
    k = np.random.default_rng(8).integers(0, 65536, n).astype(np.int64)
    _check(ctx, orc, [k], [0], ["+"])
    info = ops.sort_info(ctx)
    assert info["passes"] == 2 and info["scatter_bytes"] == (8 + 12 + 1) * n + 24 * n


def test_sort_merges_the_row_id_into_the_key_word(ctx, orc):
    """LONG keys over the whole range: the first pass (byte 0, the 8-byte column in, key word + id
    out, 21 bytes a row with the next plane), bytes 1 and 2 (key word + id both ways, 25), byte 3 as
    the merging pass (12 in, one 8-byte word out: 21), bytes 4-6 on the merged words (17) and byte 7
    (the merged words in, the 4-byte ids out as the permutation: 12): 155 bytes a row over 8 passes
    against 195 without the merge (option sort.merge 0). Both orders equal the oracle's, also with
    duplicates and a second key (INT + FLOAT: one key word, the id beside it)."""
    from ghive_amd import ops
    n = 100_003
    rng = np.random.default_rng(11)
    k = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    k[::5] = k[::5] % 977
    try:
        for merge, per_row in ((1, 155), (0, 195)):
            ctx.set_option("sort.merge", merge)
            for order in ("+", "-"):
                _check(ctx, orc, [k], [0], [order])
                info = ops.sort_info(ctx)
                assert info["passes"] == 8 and info["scatter_bytes"] == per_row * n
            i = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
            f = rng.standard_normal(n).astype(np.float32)
            f[::13] = np.nan
            _check(ctx, orc, [i, f], [0, 1], ["-", "+"])
    finally:
        ctx.set_option("sort.merge", 1)


def test_sort_info_is_left_by_builds_and_group_bys(ctx, orc):
    """ghive_last_sort_info describes the last ORDER BY: a regional hash build and a partitioned
    group-by after it (both run radix passes of their own) leave it unchanged."""
    from ghive_amd import ops
    from conftest import to_dev
    rng = np.random.default_rng(2)
    k = rng.integers(-2**40, 2**40, 100_003).astype(np.int64)
    _check(ctx, orc, [k], [0], ["+"])
    before = ops.sort_info(ctx)
    ctx.set_option("hash.regional_rows", 1)
    try:
        ht = ops.HashTable(ctx, to_dev(rng.integers(0, 2**50, 50_000).astype(np.int64)))
        ht.info()
        ht.free()
    finally:
        ctx.set_option("hash.regional_rows", 1 << 22)
    ctx.set_option("groupby.partition", 2)
    try:
        ops.GroupByPredicate([0], [(abi.AGG_CNT, abi.COUNT_STAR)]).gpu_execute(
            ctx, ops.Table([to_dev(k)]), capacity=200_000)
    finally:
        ctx.set_option("groupby.partition", 1)
    assert ops.sort_info(ctx) == before
