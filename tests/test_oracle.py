"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against (a) the reference's own known-answer tests transcribed in
tests/golden/reference_kats.json, (b) published vectors (Random123 Philox KATs), and (c) brute force
or closed forms computed here with independent numpy/Python code, on inputs chosen so that a
plausible slip (wrong bound inclusivity, dropped NULL check, swapped operands, lost duplicate
matches, wrong wrap-around) fails at least one case.
"""
import math
from collections import defaultdict

import numpy as np
import pytest

from conftest import np_col
from ghive_amd import abi

INT_NULL, LONG_NULL = abi.INT_NULL, abi.LONG_NULL


# ---------------------------------------------------------------- generator ------------------
def test_philox_random123_kats(orc):
    # Random123 kat_vectors, philox4x32 10 rounds
    assert list(orc.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(orc.philox([0xffffffff] * 4, [0xffffffff] * 2)) == [
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                           [0xa4093822, 0x299f31d0])) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_uniform_int_mapping(orc):
    seed, col = 0x6768697666, 3
    a = orc.gen_uniform_i32(seed, col, 0, 200000, 1, 7)
    assert a.min() == 1 and a.max() == 7
    counts = np.bincount(a)[1:]
    assert np.all(np.abs(counts / len(a) - 1 / 7) < 0.01)
    # value of row r is a pure function of (seed, col, r): slices agree
    b = orc.gen_uniform_i32(seed, col, 1000, 50, 1, 7)
    assert np.array_equal(a[1000:1050], b)
    # closed form of the mapping lo + floor(x * range / 2^32) with x = philox word 0
    x = orc.philox([1234, 0, col, 0], [seed & 0xffffffff, seed >> 32])[0]
    assert orc.gen_uniform_i32(seed, col, 1234, 1, -5, 5)[0] == -5 + (int(x) * 11 >> 32)
    full = orc.gen_uniform_i32(seed, 9, 0, 10000, -2**31, 2**31 - 1)
    assert full.min() < -2**30 and full.max() > 2**30
    c = orc.gen_uniform_i64(seed, 4, 0, 100000, -10**15, 10**15)
    assert c.min() >= -10**15 and c.max() <= 10**15 and abs(c.mean()) < 2e13


def test_uniform_f32_and_normal(orc):
    u = orc.gen_uniform_f32(7, 1, 0, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    z = orc.gen_normal_f64(7, 2, 0, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


# ---------------------------------------------------------------- filter ---------------------
def _null_col(values, nulls, dtype):
    a = np.array(values, dtype=dtype)
    sentinel = {np.int64: LONG_NULL, np.int32: INT_NULL}.get(dtype, np.nan)
    for i in nulls:
        a[i] = sentinel
    return a


def test_filter_hive_kats(orc, kats):
    for k in kats["hive_filter"]:
        dt = np.int64 if k["type"] == "LONG" else np.float64
        cols = [_null_col(k["values"], k["nulls"], dt)]
        p = k["pred"]
        if p.get("col2"):
            cols.append(np.array(k["values2"], dtype=dt))
            node = abi.make_pred(p["mode"], 0, col2=1)
        else:
            params = [float(v) for v in p["params"]] if k["type"] == "DOUBLE" else p["params"]
            node = abi.make_pred(p["mode"], 0, params)
        got = orc.filter_rows(cols, [node], 0)
        if "sel" in k:  # the batch arrives with selectedInUse: restrict to the given rows
            got = [r for r in got if r in k["sel"]]
        assert list(got) == k["expect"], k["cite"]


def _brute_leaf(mode, x, a, b):
    return {0: x == a, 1: x <= a, 2: x >= a, 3: x < a, 4: x > a, 5: (x > a) & (x < b),
            6: (x >= a) & (x <= b), 7: x != a, 8: (x < a) | (x > b), 9: (x <= a) | (x >= b),
            10: np.ones_like(x, dtype=bool)}[mode]


@pytest.mark.parametrize("mode", list(range(11)))
@pytest.mark.parametrize("ctype", ["INT", "LONG", "DOUBLE"])
def test_filter_modes_brute_force(orc, mode, ctype):
    rng = np.random.default_rng(mode * 7 + len(ctype))
    n = 5000
    if ctype == "DOUBLE":
        x = rng.integers(-20, 21, n).astype(np.float64) / 2
        x[rng.random(n) < 0.05] = np.nan
        null = np.isnan(x)
        params = (-3.5, 4.0)
    else:
        dt = np.int32 if ctype == "INT" else np.int64
        x = rng.integers(-20, 21, n).astype(dt)
        x[rng.random(n) < 0.05] = INT_NULL if ctype == "INT" else LONG_NULL
        null = x == (INT_NULL if ctype == "INT" else LONG_NULL)
        params = (-3, 4)
    node = abi.make_pred(mode, 0, params if mode not in (10,) else ())
    got = orc.filter_rows([x], [node], 0)
    with np.errstate(invalid="ignore"):
        want = np.nonzero(_brute_leaf(mode, x, params[0], params[1]) & ~null)[0]
    assert np.array_equal(got, want)


def test_filter_and_or_tree_brute_force(orc):
    rng = np.random.default_rng(5)
    n = 10000
    a = rng.integers(0, 10, n).astype(np.int32)
    b = rng.integers(0, 10, n).astype(np.int64)
    c = rng.random(n)
    a[::97] = INT_NULL
    c[::89] = np.nan
    # ((a = 3 or b between 2 and 4) and c < 0.5) or (a > 7 and b <> 1)
    nodes = [
        abi.make_pred(abi.FILTER_OR, children=(1, 6)),
        abi.make_pred(abi.FILTER_AND, children=(2, 5)),
        abi.make_pred(abi.FILTER_OR, children=(3, 4)),
        abi.make_pred(abi.FILTER_EQ, 0, [3], data_type=2),
        abi.make_pred(abi.FILTER_EQ_RANGE, 1, [2, 4]),
        abi.make_pred(abi.FILTER_LT, 2, [0.5]),
        abi.make_pred(abi.FILTER_AND, children=(7, 8)),
        abi.make_pred(abi.FILTER_GT, 0, [7], data_type=2),
        abi.make_pred(abi.FILTER_NOT_EQ, 1, [1]),
    ]
    got = orc.filter_rows([a, b, c], nodes, 0)
    an = a == INT_NULL
    cn = np.isnan(c)
    with np.errstate(invalid="ignore"):
        left = (((a == 3) & ~an) | ((b >= 2) & (b <= 4))) & (c < 0.5) & ~cn
        right = (a > 7) & ~an & (b != 1)
    assert np.array_equal(got, np.nonzero(left | right)[0])


def test_filter_empty_and_all(orc):
    x = np.arange(0, dtype=np.int64)
    assert len(orc.filter_rows([x], [abi.make_pred(abi.FILTER_GE, 0, [0])], 0)) == 0
    y = np.arange(3000, dtype=np.int32)
    assert np.array_equal(orc.filter_rows([y], [abi.make_pred(abi.FILTER_GE, 0, [0])], 0), y)


def test_filter_literal_conversion(orc):
    # an INT column against a double literal compares values; a LONG column against an int literal widens
    x = np.array([1, 2, 3, 4], dtype=np.int32)
    assert list(orc.filter_rows([x], [abi.make_pred(abi.FILTER_LT, 0, [2.5])], 0)) == [0, 1]
    y = np.array([2**40, 5, -2**40], dtype=np.int64)
    assert list(orc.filter_rows([y], [abi.make_pred(abi.FILTER_GT, 0, [4], data_type=2)], 0)) == [0, 1]


# ---------------------------------------------------------------- map join -------------------
def _brute_join(bk, pk, payload=None, bsel=None, psel=None, null=LONG_NULL):
    bsel = range(len(bk)) if bsel is None else bsel
    psel = range(len(pk)) if psel is None else psel
    idx = defaultdict(list)
    for r in bsel:
        if bk[r] != null:
            idx[int(bk[r])].append(int(payload[r]) if payload is not None else int(r))
    out = []
    for r in psel:
        if pk[r] == null:
            continue
        for v in idx.get(int(pk[r]), []):
            out.append((int(r), v))
    return sorted(out)


def test_mapjoin_brute_force_duplicates_nulls(orc):
    rng = np.random.default_rng(11)
    bk = rng.integers(-500, 500, 3000).astype(np.int64)  # many duplicates, negatives
    bk[::50] = LONG_NULL
    pk = rng.integers(-700, 700, 20000).astype(np.int64)
    pk[::31] = LONG_NULL
    pay = rng.integers(0, 1 << 40, 3000).astype(np.int64)
    rows, vals = orc.mapjoin_inner(bk, pk, payload=pay)
    assert sorted(zip(rows.tolist(), vals.tolist())) == _brute_join(bk, pk, pay)
    # with selections on both sides
    bsel = np.nonzero(rng.random(3000) < 0.3)[0].astype(np.uint32)
    psel = np.nonzero(rng.random(20000) < 0.5)[0].astype(np.uint32)
    rows, vals = orc.mapjoin_inner(bk, pk, build_sel=bsel, probe_sel=psel)
    assert sorted(zip(rows.tolist(), vals.tolist())) == _brute_join(bk, pk, None, bsel, psel)
    # one table probed from several threads at once: each slice's match count is its brute count
    import concurrent.futures as cf
    mj = orc.MapJoinTable(bk, pay)
    cuts = [0, 1, 777, 9000, 20000]
    with cf.ThreadPoolExecutor(4) as ex:
        got = list(ex.map(lambda i: mj.probe(pk[cuts[i]:cuts[i + 1]]), range(4)))
    mj.free()
    assert got == [len(_brute_join(bk, pk[cuts[i]:cuts[i + 1]], pay)) for i in range(4)]


def test_mapjoin_int_keys_and_expansion(orc):
    rng = np.random.default_rng(3)
    bk = rng.permutation(200000).astype(np.int32)[:50000] * 7
    pk = rng.integers(0, 1400000, 100000).astype(np.int32)
    rows, vals = orc.mapjoin_inner(bk, pk)
    want = _brute_join(bk, pk, null=INT_NULL)
    assert sorted(zip(rows.tolist(), vals.tolist())) == want
    st = orc.mapjoin_stats(bk)
    assert st["keys"] == 50000 and st["buckets"] & (st["buckets"] - 1) == 0
    assert st["keys"] <= 0.75 * st["buckets"]
    # a table built at a tiny load factor still answers exactly (triangular probing + rehash)
    rows2, vals2 = orc.mapjoin_inner(bk[:5000], pk, load_factor=0.9)
    assert sorted(zip(rows2.tolist(), vals2.tolist())) == _brute_join(bk[:5000], pk, null=INT_NULL)


def test_hive_long_hash_known_values(orc):
    # key ^= (key >>> 7) ^ (key >>> 4); key ^= (key >>> 20) ^ (key >>> 12); (int) key
    assert orc.hive_long_hash(0) == 0
    assert orc.hive_long_hash(1) == 1
    assert orc.hive_long_hash(16) == 17            # 16 ^ (16>>>4)
    assert orc.hive_long_hash(128) == 128 ^ 1 ^ 8  # 128 ^ (128>>>7) ^ (128>>>4)
    assert orc.hive_long_hash(-1) == orc.hive_long_hash(-1)


# ---------------------------------------------------------------- N-way join -----------------
def test_join_reference_kat(orc, kats):
    for k in kats["join"]:
        tables = [[np.array(t, dtype=np.int64)] for t in k["tables"]]
        got = orc.join(tables, k["join_cols"], [tuple(c) for c in k["conds"]])
        want = np.array(k["expect_offsets"], dtype=np.int32).reshape(-1, k["width"])
        assert sorted(map(tuple, got.tolist())) == sorted(map(tuple, want.tolist())), k["cite"]


def test_join_inner_chain_brute_force(orc):
    rng = np.random.default_rng(8)
    t0 = rng.integers(0, 30, 400).astype(np.int64)
    t1 = rng.integers(0, 30, 50).astype(np.int64)
    t2 = rng.integers(0, 30, 40).astype(np.int64)
    got = orc.join([[t0], [t1], [t2]], [0, 0, 0],
                   [(abi.INNER_JOIN, 0, 1), (abi.INNER_JOIN, 0, 2)])
    want = sorted((i, j, k) for i in range(400) for j in range(50) if t0[i] == t1[j]
                  for k in range(40) if t0[i] == t2[k])
    assert sorted(map(tuple, got.tolist())) == want


# ---------------------------------------------------------------- group-by -------------------
def _kat_key(cols, keys, r):
    out = []
    for k in keys:
        c = cols[k]
        if isinstance(c, tuple):
            s, e = c[2][2 * r], c[2][2 * r + 1]
            out.append(c[1][s:e].decode())
        else:
            out.append(c[r].item())
    return out


def test_groupby_gpu_impl_kats(orc, kats):
    for k in kats["groupby_gpu_impl"]:
        cols = [np_col(c) for c in k["columns"]]
        rep, vals, nulls = orc.groupby(cols, k["keys"], [tuple(a) for a in k["aggs"]])
        got = sorted((tuple(_kat_key(cols, k["keys"], r)), [v[g].item() for v in vals])
                     for g, r in enumerate(rep))
        want = sorted((tuple(e["key"]), e["aggs"]) for e in k["expect"])
        assert got == want, k["cite"]


_AGG = {"sum": abi.AGG_SUM, "min": abi.AGG_MIN, "max": abi.AGG_MAX, "avg": abi.AGG_AVG,
        "count": abi.AGG_CNT}


def test_groupby_hive_kats(orc, kats):
    for k in kats["hive_groupby"]:
        dbl = k.get("dtype") == "DOUBLE"
        vals = np.array([(np.nan if dbl else LONG_NULL) if v is None else v for v in k["values"]],
                        dtype=np.float64 if dbl else np.int64)
        cols = [vals]
        keys = []
        if k["keys"] is not None:
            cols.append(np.array([LONG_NULL if v is None else v for v in k["keys"]], dtype=np.int64))
            keys = [1]
        rep, out, nulls = orc.groupby(cols, keys, [(_AGG[k["agg"]], 0)])
        if k["keys"] is None:
            assert len(rep) == 1
            got = None if nulls[0][0] else out[0][0].item()
            assert got == k["expect"], k["cite"]
        else:
            got = {}
            for g, r in enumerate(rep):
                kv = cols[1][r].item()
                got[None if kv == LONG_NULL else kv] = None if nulls[0][g] else out[0][g].item()
            assert got == {a: b for a, b in k["expect"]}, k["cite"]


def test_groupby_brute_force_all_aggs(orc):
    rng = np.random.default_rng(21)
    n = 20000
    k1 = rng.integers(0, 50, n).astype(np.int32)
    k1[::101] = INT_NULL  # NULL keys form one group
    v_long = rng.integers(-2**62, 2**62, n, dtype=np.int64)  # sums wrap like a Java long
    v_long[::37] = LONG_NULL
    v_dbl = rng.standard_normal(n)
    v_dbl[::41] = np.nan
    cols = [k1, v_long, v_dbl]
    aggs = [(abi.AGG_SUM, 1), (abi.AGG_MIN, 1), (abi.AGG_MAX, 1), (abi.AGG_CNT, 1),
            (abi.AGG_AVG, 1), (abi.AGG_SUM, 2), (abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_MAX, 2)]
    rep, vals, nulls = orc.groupby(cols, [0], aggs)
    groups = defaultdict(list)
    for r in range(n):
        groups[int(k1[r])].append(r)
    assert len(rep) == len(groups)
    for g, r0 in enumerate(rep):
        rows = groups[int(k1[r0])]
        ls = [int(v_long[r]) for r in rows if v_long[r] != LONG_NULL]
        ds = [float(v_dbl[r]) for r in rows if not np.isnan(v_dbl[r])]
        wrap = (sum(ls) + 2**63) % 2**64 - 2**63
        assert vals[0][g] == wrap
        assert vals[1][g] == min(ls) and vals[2][g] == max(ls) and vals[3][g] == len(ls)
        assert vals[4][g] == pytest.approx(wrap / len(ls), rel=1e-15)
        assert vals[5][g] == math.fsum(ds)  # correctly rounded
        assert vals[6][g] == len(rows)
        assert vals[7][g] == max(ds)


def test_exact_sum_against_fsum(orc):
    rng = np.random.default_rng(2)
    for trial in range(20):
        x = rng.standard_normal(1000) * 10.0 ** rng.integers(-5, 5, 1000)
        x = np.concatenate([x, -x[: 500]])  # heavy cancellation
        rng.shuffle(x)
        assert orc.exact_sum(x) == math.fsum(x.tolist())
    assert orc.exact_sum(np.array([1e16, 1.0, -1e16])) == 1.0
    assert orc.exact_sum(np.array([0.1] * 10)) == math.fsum([0.1] * 10)


# ---------------------------------------------------------------- STRING filters -------------
def _str_col(values):
    blob, idx, o = bytearray(), [], 0
    for v in values:
        if v is None:
            idx += [-1, -1]
            continue
        blob += v
        idx += [o, o + len(v)]
        o += len(v)
    return ("string", bytes(blob), np.array(idx, dtype=np.int32))


def test_string_filter_reference_kat(orc):
    """FilterPredicateTest.cu:99-204: BETWEEN 'MFGR#1221' AND 'MFGR#1223' over 20 brands keeps the
    six rows its (commented) expectation lists."""
    brands = [b"MFGR#1221", b"MFGR#1223", b"MFGR#1238", b"MFGR#1221", b"MFGR#1238", b"MFGR#122",
              b"MFGR#1232", b"MFGR#1228", b"MFGR#123", b"MFGR#1224", b"MFGR#1248", b"MFGR#12223",
              b"MFGR#1237", b"MFGR#1222", b"MFGR#1230", b"MFGR#1221", b"MFGR#1234", b"MFGR#1220",
              b"MFGR#1239", b"MFGR#12234"]
    node = abi.make_pred(abi.FILTER_EQ_RANGE, 0, [b"MFGR#1221", b"MFGR#1223"])
    sel = orc.filter_rows([_str_col(brands)], [node], 0)
    assert [brands[i] for i in sel] == [b"MFGR#1221", b"MFGR#1223", b"MFGR#1221", b"MFGR#12223",
                                        b"MFGR#1222", b"MFGR#1221"]


def test_string_filter_brute_force(orc):
    """Every mode against Python's bytes order, which is Hive's (unsigned bytes, then length):
    prefixes, empty strings, bytes above 0x7f, NULLs, and STRING column against STRING column."""
    rng = np.random.default_rng(17)
    alphabet = [b"a", b"b", b"\x7f", b"\x80", b"\xff"]
    vals = [None if rng.random() < 0.1 else b"".join(alphabet[i] for i in rng.integers(0, 5, rng.integers(0, 4)))
            for _ in range(3000)]
    vals2 = [None if rng.random() < 0.1 else b"".join(alphabet[i] for i in rng.integers(0, 5, rng.integers(0, 4)))
             for _ in range(3000)]
    col, col2 = _str_col(vals), _str_col(vals2)
    lo, hi = b"a\x80", b"\x7f"
    lo, hi = min(lo, hi), max(lo, hi)
    ref = {abi.FILTER_EQ: lambda v: v == lo, abi.FILTER_LE: lambda v: v <= lo, abi.FILTER_GE: lambda v: v >= lo,
           abi.FILTER_LT: lambda v: v < lo, abi.FILTER_GT: lambda v: v > lo, abi.FILTER_NOT_EQ: lambda v: v != lo,
           abi.FILTER_RANGE: lambda v: lo < v < hi, abi.FILTER_EQ_RANGE: lambda v: lo <= v <= hi,
           abi.FILTER_NOT_RANGE: lambda v: not (lo <= v <= hi), abi.FILTER_NOT_EQ_RANGE: lambda v: not (lo < v < hi)}
    for mode, f in ref.items():
        two = mode in (abi.FILTER_RANGE, abi.FILTER_EQ_RANGE, abi.FILTER_NOT_RANGE, abi.FILTER_NOT_EQ_RANGE)
        node = abi.make_pred(mode, 0, [lo, hi] if two else [lo])
        want = [i for i, v in enumerate(vals) if v is not None and f(v)]
        assert orc.filter_rows([col], [node], 0).tolist() == want, mode
    node = abi.make_pred(abi.FILTER_NOT_NULL, 0)
    assert orc.filter_rows([col], [node], 0).tolist() == [i for i, v in enumerate(vals) if v is not None]
    node = abi.make_pred(abi.FILTER_LT, 0, col2=1)
    want = [i for i, (u, v) in enumerate(zip(vals, vals2)) if u is not None and v is not None and u < v]
    assert orc.filter_rows([col, col2], [node], 0).tolist() == want


def test_cast_string_to_int_known_answers(orc):
    """Hive 3.1 CAST(string AS INT) (UDFToInteger of Text -> LazyInteger.parseInt with trim)."""
    cases = {b"123": 123, b" 42 ": 42, b"-7": -7, b"+8": 8, b"1.9": 1, b"-1.9": -1, b"1.": 1, b".5": 0,
             b"abc": None, b"": None, b"-": None, b"2147483648": None, b"-2147483648": -2147483648,
             b"12a": None, b"1.2.3": None, b"1e5": None, b"N123": None, b"x": None, b"7": 7, b" 5": 5,
             b"\t12\n": 12, b"1 2": None, b"0x1F": None, b"--1": None, b"1.x": None, b"  ": None}
    got = orc.cast_string_to_int(list(cases))
    assert dict(zip(cases, got)) == cases
    assert orc.cast_string_to_int([b"9223372036854775807", b"9223372036854775808"], bits=64) == [2**63 - 1, None]
    assert orc.cast_string_to_int([None]) == [None]


def test_substr_known_answers(orc):
    """Hive's documented SUBSTR behaviour (UDFSubstr.makeIndex): 1-based, negative from the end."""
    s = [b"Facebook"]
    assert orc.substr(s, 5, 2**31 - 1) == [b"book"]
    assert orc.substr(s, 5, 1) == [b"b"]
    assert orc.substr(s, -4, 2**31 - 1) == [b"book"]
    assert orc.substr(s, 0, 2) == [b"Fa"]
    assert orc.substr([b"abc"], 4, 5) == [b""] and orc.substr([b"abc"], 3, 5) == [b"c"]
    assert orc.substr([b"abc"], -3, 2) == [b"ab"] and orc.substr([b"abc"], -4, 1) == [b""]
    assert orc.substr([b"abc"], 1, 0) == [b""] and orc.substr([None], 1, 1) == [None]
    assert orc.substr(["h\u00e9llo".encode()], 2, 2) == ["\u00e9l".encode()]  # a 2-byte character is one unit
    assert orc.substr(["a\U0001F600b".encode()], 2, 2) == ["\U0001F600".encode()]  # a 4-byte one is two


# ---------------------------------------------------------------- projection arithmetic -------
def _java_int(v):
    return (v + 2**31) % 2**32 - 2**31


def _java_long(v):
    return (v + 2**63) % 2**64 - 2**63


def test_select_arith_known_answers(orc):
    i32, i64 = np.int32, np.int64
    a = np.array([2147483000, -5, 7, INT_NULL, 0, -2147483648], dtype=i32)
    b = np.array([1000, 3, -2, 1, 0, -1], dtype=i32)
    # INT op INT wraps like a Java int; NULL in, NULL out
    for op, f in (("+", lambda x, y: x + y), ("-", lambda x, y: x - y), ("*", lambda x, y: x * y)):
        r = orc.select_arith(a, b, op)
        assert r.dtype == i32
        want = [INT_NULL if x == INT_NULL else _java_int(f(int(x), int(y))) for x, y in zip(a, b)]
        assert r.tolist() == want
    assert orc.select_arith(a, b, "+")[0] == -2147483296
    # '/' is DOUBLE; x / 0 is NULL (NaN)
    r = orc.select_arith(a, b, "/")
    assert r.dtype == np.float64
    assert r[1] == -5 / 3 and r[2] == -3.5 and math.isnan(r[3]) and math.isnan(r[4])
    assert r[5] == 2147483648.0
    # INT op LONG widens to LONG and wraps like a Java long
    la = np.array([2**62, -(2**63), 5, LONG_NULL], dtype=i64)
    lb = np.array([4, 5, 6, 7], dtype=i32)
    r = orc.select_arith(la, lb, "*")
    assert r.dtype == i64
    assert r.tolist() == [_java_long(2**64), _java_long(-(2**63) * 5), 30, LONG_NULL]
    assert orc.select_arith(la, lb, "-").tolist() == [2**62 - 4, _java_long(-(2**63) - 5), -1, LONG_NULL]
    # a floating operand makes DOUBLE; NaN is NULL; a 1-row operand broadcasts
    d = np.array([1.5, np.nan, -2.0], dtype=np.float64)
    r = orc.select_arith(d, np.array([2], dtype=i32), "*")
    assert r[0] == 3.0 and math.isnan(r[1]) and r[2] == -4.0
    r = orc.select_arith(np.array([10.0], dtype=np.float32), np.array([4, 0, INT_NULL], dtype=i32), "/")
    assert r[0] == 2.5 and math.isnan(r[1]) and math.isnan(r[2])
    assert len(orc.select_arith(np.zeros(0, i32), np.zeros(0, i32), "+")) == 0


# ---------------------------------------------------------------- sort -----------------------
def test_sort_total_order(orc):
    rng = np.random.default_rng(4)
    n = 30000
    k = rng.integers(-50, 50, n).astype(np.int64)
    f = rng.integers(0, 20, n).astype(np.float32) / 4
    f[::53] = np.nan
    perm = orc.sort_perm([k, f], [0, 1], ["+", "+"])
    fkey = np.where(np.isnan(f), np.inf, f)  # NULL (NaN) last ascending; ties by row id
    want = np.lexsort((np.arange(n), fkey, k))
    assert np.array_equal(perm, want)
    perm = orc.sort_perm([k, f], [0, 1], ["-", "+"])
    want = np.lexsort((np.arange(n), fkey, -k))
    assert np.array_equal(perm, want)


@pytest.mark.parametrize("nruns", [1, 2, 7, 16])
def test_sort_runs_merge_is_the_same_total_order(orc, nruns):
    """Run sorts + one k-way merge (the multi-core CPU baseline's ORDER BY) give the permutation of
    the one-pass sort and of numpy's lexsort, ties broken by row id across run boundaries."""
    rng = np.random.default_rng(40 + nruns)
    n = 20011
    k = rng.integers(-30, 30, n).astype(np.int64)
    f = rng.integers(0, 8, n).astype(np.float32)
    f[::41] = np.nan
    fkey = np.where(np.isnan(f), np.inf, f)
    for orders, kk in ((["+", "+"], k), (["-", "+"], -k)):
        perm, run_s, merge_s = orc.sort_runs([k, f], [0, 1], orders, nruns)
        assert np.array_equal(perm, np.lexsort((np.arange(n), fkey, kk)))
        assert np.array_equal(perm, orc.sort_perm([k, f], [0, 1], orders))
        assert run_s >= 0 and merge_s >= 0
    # fewer rows than runs, and no rows
    assert np.array_equal(orc.sort_runs([k[:3]], [0], "-", nruns)[0], np.lexsort((np.arange(3), -k[:3])))
    assert len(orc.sort_runs([k[:0]], [0], "+", nruns)[0]) == 0


def test_sort_string_keys_against_python_order(orc):
    """STRING keys sort in Python's bytes order (unsigned bytes, then length = StringExpr.compare),
    NULL last ascending, ties by the next key and then by row id."""
    rng = np.random.default_rng(14)
    alphabet = [b"a", b"b", b"\x00", b"\x80", b"\xff"]
    n = 4000
    vals = [None if rng.random() < 0.05 else b"".join(alphabet[i] for i in rng.integers(0, 5, rng.integers(0, 4)))
            for _ in range(n)]
    k2 = rng.integers(0, 3, n).astype(np.int64)
    col = _str_col(vals)
    for orders in (["+", "+"], ["-", "+"], ["+", "-"]):
        perm = orc.sort_perm([col, k2], [0, 1], orders)
        s0 = 1 if orders[0] == "+" else -1
        s1 = 1 if orders[1] == "+" else -1

        def key(i):
            v = vals[i]
            v = (1, b"") if v is None else (0, v)
            return v, k2[i] * s1, i
        want = sorted(range(n), key=key)
        if s0 < 0:  # descending on the string (NULL first), ascending ties
            from functools import cmp_to_key

            def cmp(i, j):
                a, b = key(i)[0], key(j)[0]
                if a != b:
                    return -1 if a > b else 1
                return -1 if (key(i)[1:] < key(j)[1:]) else (1 if key(i)[1:] > key(j)[1:] else 0)
            want = sorted(range(n), key=cmp_to_key(cmp))
        assert perm.tolist() == want, orders


# ---------------------------------------------------------------- skewed-join generator ------
def test_key62_is_injective_and_below_null(orc):
    ranks = np.arange(1, 200001, dtype=np.uint64)
    keys = np.array([orc.key62(99, int(r)) for r in ranks[:20000]], dtype=np.uint64)
    assert len(np.unique(keys)) == len(keys)
    assert keys.max() < 2**62


@pytest.mark.parametrize("theta", [1.0, 0.5, 1.5])
def test_zipf_rejection_inversion_matches_pmf(orc, theta):
    n, m = 1000, 400_000
    keys, ranks = orc.gen_zipf_keys(5, 3, 0, m, n, theta, 0.0)
    assert ranks.min() >= 1 and ranks.max() <= n
    pmf = 1.0 / np.arange(1, n + 1) ** theta
    pmf /= pmf.sum()
    counts = np.bincount(ranks.astype(np.int64), minlength=n + 1)[1:]
    for k in (1, 2, 3, 10, 100):  # frequent ranks within 5 standard errors
        se = np.sqrt(m * pmf[k - 1] * (1 - pmf[k - 1]))
        assert abs(counts[k - 1] - m * pmf[k - 1]) < 5 * se + 2, (k, counts[k - 1], m * pmf[k - 1])
    # keys are the rank bijection
    assert all(int(keys[i]) == orc.key62(5, int(ranks[i])) for i in range(0, m, 9973))


def test_zipf_misses_are_outside_the_build_ranks(orc):
    n, m = 10_000, 200_000
    keys, ranks = orc.gen_zipf_keys(8, 1, 0, m, n, 1.0, 0.1)
    miss = ranks > n
    assert abs(miss.mean() - 0.1) < 0.005
    build = {orc.key62(8, r) for r in range(1, n + 1)}
    assert not any(int(k) in build for k in keys[miss][:2000])
    assert all(int(k) in build for k in keys[~miss][:2000])


# ---------------------------------------------------------------- fragment -------------------
def c1_tables(orc, fact_rows, seed=0x6768697666):
    """C1 shape: fact(fk, measure, price), dim(key, attr, category)."""
    fk = orc.gen_uniform_i32(seed, 0, 0, fact_rows, 1, 200000)
    measure = orc.gen_uniform_i32(seed, 1, 0, fact_rows, 1, 100)
    price = orc.gen_uniform_i32(seed, 2, 0, fact_rows, 100000, 10000000)
    key = np.arange(1, 200001, dtype=np.int32)
    attr = orc.gen_uniform_i32(seed, 10, 0, 200000, 0, 6)
    cat = orc.gen_uniform_i32(seed, 11, 0, 200000, 0, 24)
    return (fk, measure, price), (key, attr, cat)


def test_fragment_c1_closed_form(orc):
    (fk, measure, price), (key, attr, cat) = c1_tables(orc, 300000)
    spec = orc.FragmentSpec(
        [fk, measure, price],
        [dict(columns=[key, attr, cat], key_col=0, payload_col=2, fact_key_col=0,
              filter=([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0))],
        keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)])
    r1 = spec.run(nthreads=1)
    r4 = spec.run(nthreads=4)
    # dense keys 1..N: the join is the array lookup dim[fk-1]
    ok = attr[fk - 1] == 3
    g = cat[fk - 1][ok]
    prod = measure[ok].astype(np.int64) * price[ok].astype(np.int64)
    want = {int(c): int(prod[g == c].sum()) for c in np.unique(g)}
    for r in (r1, r4):
        got = {int(k[0]): int(v[0]) for k, v in zip(r["keys"], r["aggs"])}
        assert got == want
        assert r["joined"] == int(ok.sum())


def test_fragment_star_three_joins_two_keys_brute_force(orc):
    """The fragment path of the 600M-row star query (three filtered map joins, group keys from the
    first and third dimension payloads, SUM(price * discount)) plus a fact filter, COUNT(*), MAX,
    MIN and NULLs, on sparse shuffled dimension keys, against a dictionary join written here."""
    rng = np.random.default_rng(7)
    n = 6000

    def dim(rows, key_space):
        key = (rng.choice(key_space, size=rows, replace=False) + 1).astype(np.int32)
        return key

    dk, sk, pk = dim(300, 5000), dim(800, 40000), dim(1500, 90000)
    year = rng.integers(1992, 1999, dk.size).astype(np.int32)
    region = rng.integers(0, 5, sk.size).astype(np.int32)
    pcat = rng.integers(0, 25, pk.size).astype(np.int32)
    brand = rng.integers(0, 40, pk.size).astype(np.int32)
    pcat[:200] = 1                        # enough part rows pass the category filter

    def fks(keys, space):
        f = np.where(rng.random(n) < 0.85, rng.choice(keys, n), rng.integers(1, space + 2, n)).astype(np.int32)
        f[rng.random(n) < 0.03] = INT_NULL
        return f

    f_date, f_supp, f_part = fks(dk, 5000), fks(sk, 40000), fks(pk, 90000)
    qty = rng.integers(1, 51, n).astype(np.int32)
    price = rng.integers(100_000, 10_000_001, n).astype(np.int32)
    disc = rng.integers(0, 11, n).astype(np.int32)
    price[rng.random(n) < 0.02] = INT_NULL
    disc[rng.random(n) < 0.02] = INT_NULL
    fact = [f_date, f_supp, f_part, qty, price, disc]
    P = abi.make_pred
    spec = orc.FragmentSpec(
        fact,
        [dict(columns=[dk, year], key_col=0, payload_col=1, fact_key_col=0,
              filter=([P(abi.FILTER_EQ_RANGE, 1, [1993, 1996], data_type=2)], 0)),
         dict(columns=[sk, region], key_col=0, payload_col=1, fact_key_col=1,
              filter=([P(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
         dict(columns=[pk, pcat, brand], key_col=0, payload_col=2, fact_key_col=2,
              filter=([P(abi.FILTER_EQ, 1, [1], data_type=2)], 0))],
        keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)],
        aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 4, 5), (abi.AGG_CNT, abi.EXPR_ONE, 0, 0),
              (abi.AGG_MAX, abi.EXPR_COL, 4, 0), (abi.AGG_MIN, abi.EXPR_COL, 5, 0)],
        fact_filter=([P(abi.FILTER_LT, 3, [25], data_type=2)], 0))
    dmap = {int(k): int(y) for k, y in zip(dk, year) if 1993 <= y <= 1996}
    smap = {int(k): int(r) for k, r in zip(sk, region) if r == 1}
    pmap = {int(k): int(b) for k, c, b in zip(pk, pcat, brand) if c == 1}
    groups, joined = {}, 0
    for i in range(n):
        if not qty[i] < 25:
            continue
        a, b, c = int(f_date[i]), int(f_supp[i]), int(f_part[i])
        if a not in dmap or b not in smap or c not in pmap:
            continue
        joined += 1
        g = groups.setdefault((dmap[a], pmap[c]), [0, 0, None, None])
        if price[i] != INT_NULL and disc[i] != INT_NULL:
            g[0] += int(price[i]) * int(disc[i])
        g[1] += 1
        if price[i] != INT_NULL:
            g[2] = int(price[i]) if g[2] is None else max(g[2], int(price[i]))
        if disc[i] != INT_NULL:
            g[3] = int(disc[i]) if g[3] is None else min(g[3], int(disc[i]))
    assert joined > 20 and len(groups) > 10
    for nt in (1, 3):
        r = spec.run(nthreads=nt)
        assert r["joined"] == joined
        got = {}
        for k, v, z in zip(r["keys"], r["aggs"], r["nulls"]):
            got[(int(k[0]), int(k[1]))] = [int(v[0]), int(v[1]), None if z[2] else int(v[2]),
                                           None if z[3] else int(v[3])]
        assert got == groups
