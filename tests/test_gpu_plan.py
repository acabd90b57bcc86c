"""GPU parity of projection arithmetic (ghive_select_arith) against the oracle, and of whole SSB
queries run vertex by vertex from their Hive plans (ghive_plan_execute) against the query's SQL
restated in numpy. The plans are the reference's own EXPLAIN fixtures (tests/golden/plans/); the
tables are synthetic, seeded."""
import math
import os

import numpy as np
import pytest

from conftest import to_dev, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
INT_NULL, LONG_NULL = abi.INT_NULL, abi.LONG_NULL
PLANS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans")


def _same(got, want):
    """Bit-exact equality, NaN (DOUBLE NULL) matching NaN."""
    assert got.dtype == want.dtype and got.shape == want.shape
    if want.dtype.kind == "f":
        assert np.array_equal(np.isnan(got), np.isnan(want))
        m = ~np.isnan(want)
        assert np.array_equal(got[m].view(np.uint64 if want.dtype == np.float64 else np.uint32),
                              want[m].view(np.uint64 if want.dtype == np.float64 else np.uint32))
    else:
        assert np.array_equal(got, want)


def _operand(rng, dtype, n):
    if dtype == np.float64:
        x = rng.standard_normal(n) * 1e3
        x[rng.random(n) < 0.02] = np.nan
        x[rng.random(n) < 0.05] = 0.0
    elif dtype == np.float32:
        x = (rng.standard_normal(n) * 1e2).astype(np.float32)
        x[rng.random(n) < 0.02] = np.nan
    elif dtype == np.int32:
        x = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
        small = rng.random(n) < 0.3  # small values too, so that / gives exact small quotients
        x[small] = rng.integers(-3, 4, int(small.sum()))
        x[rng.random(n) < 0.02] = INT_NULL
        x[rng.random(n) < 0.05] = 0
    else:
        x = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
        x[rng.random(n) < 0.02] = LONG_NULL
        x[rng.random(n) < 0.05] = 0
    return x.astype(dtype)


@pytest.mark.parametrize("ta,tb", [(np.int32, np.int32), (np.int64, np.int32), (np.int32, np.int64),
                                   (np.int64, np.int64), (np.float64, np.int64), (np.float32, np.int32),
                                   (np.float64, np.float64)])
@pytest.mark.parametrize("op", ["+", "-", "*", "/"])
def test_select_arith_vs_oracle(ctx, orc, ta, tb, op):
    from ghive_amd import ops
    rng = np.random.default_rng(hash((ta.__name__, tb.__name__, op)) % 2**32)
    n = 300_001
    a, b = _operand(rng, ta, n), _operand(rng, tb, n)
    _same(to_np(ops.select_arith(ctx, to_dev(a), to_dev(b), op)), orc.select_arith(a, b, op))
    # a literal on either side broadcasts
    for lit in (b[:1], np.array([0], dtype=tb)):
        _same(to_np(ops.select_arith(ctx, to_dev(a), to_dev(lit), op)), orc.select_arith(a, lit, op))
        _same(to_np(ops.select_arith(ctx, to_dev(lit), to_dev(a[:1000]), op)), orc.select_arith(lit, a[:1000], op))


def test_select_arith_edges(ctx, orc):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    e = to_dev(np.zeros(0, dtype=np.int32))
    assert ops.select_arith(ctx, e, e, "+").numel() == 0
    a = to_dev(np.array([2147483000, INT_NULL, 4], dtype=np.int32))
    b = to_dev(np.array([1000, 1, 0], dtype=np.int32))
    assert to_np(ops.select_arith(ctx, a, b, "+")).tolist() == [-2147483296, INT_NULL, 4]
    r = to_np(ops.select_arith(ctx, a, b, "/"))
    assert r[0] == 2147483.0 and math.isnan(r[1]) and math.isnan(r[2])
    with pytest.raises(GhiveError):
        ops.select_arith(ctx, a, b, "%")
    with pytest.raises(GhiveError):  # ragged operands
        ops.select_arith(ctx, a, to_dev(np.array([1, 2], dtype=np.int32)), "+")


# ---------------------------------------------------------------- whole queries -----------------
def _read(name):
    with open(os.path.join(PLANS, name)) as fh:
        return fh.read()


def _plan(q, maintain=None):
    from ghive_amd import ops
    return ops.Plan(_read(f"ssb_10_orc_transformed_{q}.txt"), _read(f"ssb_10_orc_transformed_{q}_extended.txt"),
                    maintain)


def _dates():
    days = np.arange(np.datetime64("1992-01-01"), np.datetime64("1999-01-01"))
    y = days.astype("datetime64[Y]").astype(np.int64) + 1970
    m = days.astype("datetime64[M]").astype(np.int64) % 12 + 1
    d = (days - days.astype("datetime64[M]")).astype(np.int64) + 1
    return (y * 10000 + m * 100 + d).astype(np.int64), y.astype(np.int32)


def _fk(rng, keys, n, miss=0.05, null=0.01):
    """Foreign keys drawn from `keys`, with some that match nothing and some NULL."""
    x = keys[rng.integers(0, len(keys), n)].astype(np.int64)
    x[rng.random(n) < miss] = -7
    x[rng.random(n) < null] = LONG_NULL
    return x


def _run(plan, ctx, vertex, inputs):
    from ghive_amd import ops
    ins = {name: (ops.Table([to_dev(c) for c in cols]), names) for name, (cols, names) in inputs.items()}
    names, cols = plan.execute(ctx, vertex, ins)
    return names, [to_np(c) for c in cols]


def _concat(parts):
    return [np.concatenate([p[i] for p in parts]) for i in range(len(parts[0]))]


def test_ssb_q1_1_three_vertices(ctx, orc):
    """SELECT sum(lo_extendedprice*lo_discount) FROM lineorder, dates WHERE lo_orderdate = d_datekey
    AND d_year = 1993 AND lo_discount BETWEEN 1 AND 3 AND lo_quantity < 25 -- Map 3 (dates), Map 1
    (lineorder, as two map tasks), Reducer 2 (merges the partial sums)."""
    rng = np.random.default_rng(101)
    dk, dy = _dates()
    n = 400_000
    od = _fk(rng, dk, n)
    qty = rng.integers(1, 51, n).astype(np.float64)
    qty[rng.random(n) < 0.01] = np.nan
    price = np.round(rng.random(n) * 1e5, 2)
    disc = rng.integers(0, 11, n).astype(np.float64)
    disc[rng.random(n) < 0.01] = np.nan
    p = _plan("1_1")
    lo_names = ["lo_orderdate", "lo_quantity", "lo_extendedprice", "lo_discount"]
    n3, c3 = _run(p, ctx, "Map 3", {"ssb_10_orc_transformed@dates": ([dk, dy], ["d_datekey", "d_year"])})
    assert n3 == ["_col0"] and np.array_equal(np.sort(c3[0]), np.sort(dk[dy == 1993]))
    parts = []
    for lo, hi in ((0, n // 3), (n // 3, n)):  # two map tasks over splits of lineorder
        names, cols = _run(p, ctx, "Map 1", {
            "ssb_10_orc_transformed@lineorder": ([od[lo:hi], qty[lo:hi], price[lo:hi], disc[lo:hi]], lo_names),
            "Map 3": (c3, n3)})
        assert names == ["_col0"] and len(cols[0]) == 1
        parts.append(cols)
    names, cols = _run(p, ctx, "Reducer 2", {"Map 1": (_concat(parts), ["_col0"])})
    assert names == ["_col0"] and len(cols[0]) == 1
    year93 = np.isin(od, dk[dy == 1993])
    m = year93 & (disc >= 1) & (disc <= 3) & (qty < 25)
    rev = price * disc
    want = math.fsum([math.fsum(rev[lo:hi][m[lo:hi]].tolist()) for lo, hi in ((0, n // 3), (n // 3, n))])
    assert cols[0][0] == want


def test_ssb_q2_1_seven_vertices_with_order_by(ctx, orc):
    """SELECT sum(lo_revenue), d_year, p_brand1 FROM lineorder, dates, part, supplier WHERE
    lo_orderdate = d_datekey AND lo_partkey = p_partkey AND lo_suppkey = s_suppkey AND p_category = 1
    AND s_region = 1 GROUP BY d_year, p_brand1 ORDER BY d_year, p_brand1 -- map joins in Map 1 and
    Reducer 2, a merge join in Reducer 2 over two shuffled inputs, partial and final group-by, and
    the final ORDER BY. Column types are the plan's: d_year int, p_brand1 bigint (96-bit group keys)."""
    rng = np.random.default_rng(202)
    dk, dy = _dates()
    ns, npart, n = 2000, 20000, 300_000
    s_key = rng.permutation(10 * ns)[:ns].astype(np.int64) + 1
    s_region = rng.integers(0, 5, ns).astype(np.int64)
    p_key = rng.permutation(10 * npart)[:npart].astype(np.int64) + 1
    p_cat = rng.integers(0, 5, npart).astype(np.int64)
    p_brand = rng.integers(0, 40, npart).astype(np.int64) + 1000
    lo_part, lo_supp, lo_date = _fk(rng, p_key, n), _fk(rng, s_key, n), _fk(rng, dk, n)
    lo_rev = np.round(rng.random(n) * 1e6, 2)
    maintain = ('{"MAPJOIN_52":{"0":["_col0","_col2","_col3"],"1":[]},'
                '"MERGEJOIN_43":{"0":["VALUE._col1","VALUE._col2"],"1":["VALUE._col1"]},'
                '"MAPJOIN_44":{"0":["_col3","_col8"],"1":["_col1"]}}')
    p = _plan("2_1", maintain)
    v5 = _run(p, ctx, "Map 5", {"ssb_10_orc_transformed@supplier": ([s_key, s_region], ["s_suppkey", "s_region"])})
    v6 = _run(p, ctx, "Map 6", {"ssb_10_orc_transformed@part": ([p_key, p_cat, p_brand],
                                                                 ["p_partkey", "p_category", "p_brand1"])})
    v7 = _run(p, ctx, "Map 7", {"ssb_10_orc_transformed@dates": ([dk, dy], ["d_datekey", "d_year"])})
    assert v6[0] == ["_col0", "_col2"] and v7[0] == ["_col0", "_col1"]
    lo_names = ["lo_partkey", "lo_suppkey", "lo_orderdate", "lo_revenue"]
    v1 = _run(p, ctx, "Map 1", {"ssb_10_orc_transformed@lineorder": ([lo_part, lo_supp, lo_date, lo_rev], lo_names),
                                "Map 5": (v5[1], v5[0])})
    assert v1[0] == ["_col0", "_col2", "_col3"]
    v2 = _run(p, ctx, "Reducer 2", {"Map 1": (v1[1], v1[0]), "Map 6": (v6[1], v6[0]), "Map 7": (v7[1], v7[0])})
    v3 = _run(p, ctx, "Reducer 3", {"Reducer 2": (v2[1], v2[0])})
    assert v3[0] == ["_col1", "_col2", "_col0"]  # keys first
    names, (s, year, brand) = _run(p, ctx, "Reducer 4", {"Reducer 3": (v3[1], v3[0])})
    assert names == ["_col0", "_col1", "_col2"]
    # the query restated
    s_ok = dict(zip(s_key.tolist(), (s_region == 1).tolist()))
    p_of = {k: b for k, c, b in zip(p_key.tolist(), p_cat.tolist(), p_brand.tolist()) if c == 1}
    y_of = dict(zip(dk.tolist(), dy.tolist()))
    groups = {}
    for pk, sk, d, r in zip(lo_part.tolist(), lo_supp.tolist(), lo_date.tolist(), lo_rev.tolist()):
        if s_ok.get(sk) and pk in p_of and d in y_of:
            groups.setdefault((y_of[d], p_of[pk]), []).append(r)
    want = sorted(groups)
    assert list(zip(year.tolist(), brand.tolist())) == want
    assert year.dtype == np.int32 and brand.dtype == np.int64
    # one partial sum per group from the map side, summed exactly in the reducer
    assert s.tolist() == [math.fsum(groups[k]) for k in want]


def test_plan_execute_errors(ctx):
    from ghive_amd._lib import GhiveError
    p = _plan("1_1")
    with pytest.raises(GhiveError) as e:
        _run(p, ctx, "Map 9", {})
    assert e.value.status == abi.ERR_INVALID and "Map 9" in str(e.value)
    with pytest.raises(GhiveError) as e:  # the dates table is not given
        _run(p, ctx, "Map 3", {})
    assert "ssb_10_orc_transformed@dates" in str(e.value)
    with pytest.raises(GhiveError) as e:  # a column missing from the input
        _run(p, ctx, "Map 3", {"ssb_10_orc_transformed@dates": ([np.zeros(3, np.int64)], ["d_datekey"])})
    assert "d_year" in str(e.value)


# ---------------------------------------------------------------- string functions ---------------
def _rand_texts(rng, n):
    pool = [b"123", b" 42 ", b"-7", b"+8", b"1.9", b"-1.9", b".5", b"abc", b"", b"-", b"2147483648",
            b"-2147483648", b"12a", b"1.2.3", b"N123", b"x", b" 5", b"\t12\n", b"9223372036854775807",
            b"9223372036854775808", "héllo wörld".encode(), "中文字".encode()]
    out = []
    for _ in range(n):
        u = rng.random()
        if u < 0.05:
            out.append(None)
        elif u < 0.5:
            out.append(pool[rng.integers(0, len(pool))])
        else:
            digits = b"".join(bytes([48 + int(d)]) for d in rng.integers(0, 10, rng.integers(1, 21)))
            sign = [b"", b"-", b"+"][rng.integers(0, 3)]
            frac = b"" if rng.random() < 0.7 else b"." + digits[:3]
            out.append(sign + digits + frac)
    return out


def _dev_strings(vals):
    from ghive_amd import ops
    return ops.StringColumn.from_list(vals)


@pytest.mark.parametrize("to_type", [abi.INT, abi.LONG])
def test_cast_string_to_integer_vs_oracle(ctx, orc, to_type):
    from ghive_amd import ops
    rng = np.random.default_rng(to_type)
    vals = _rand_texts(rng, 50_000)
    got = to_np(ops.select_cast(ctx, _dev_strings(vals), to_type)).tolist()
    bits = 32 if to_type == abi.INT else 64
    null = INT_NULL if to_type == abi.INT else LONG_NULL
    want = [null if v is None else v for v in orc.cast_string_to_int(vals, bits=bits)]
    assert got == want


def test_cast_numeric_like_java(ctx):
    from ghive_amd import ops
    i64 = np.array([5, -(2**40) + 7, 2**31 + 3, LONG_NULL, -1], dtype=np.int64)
    assert to_np(ops.select_cast(ctx, to_dev(i64), abi.INT)).tolist() == [5, 7, -(2**31) + 3, INT_NULL, -1]
    i32 = np.array([5, -3, INT_NULL], dtype=np.int32)
    assert to_np(ops.select_cast(ctx, to_dev(i32), abi.LONG)).tolist() == [5, -3, LONG_NULL]
    d = np.array([1.9, -1.9, 3e10, -3e10, np.nan, 1e300], dtype=np.float64)
    assert to_np(ops.select_cast(ctx, to_dev(d), abi.INT)).tolist() == [1, -1, 2**31 - 1, -(2**31), INT_NULL,
                                                                       2**31 - 1]
    assert to_np(ops.select_cast(ctx, to_dev(d), abi.LONG)).tolist()[:4] == [1, -1, 30000000000, -30000000000]
    r = to_np(ops.select_cast(ctx, to_dev(i32), abi.DOUBLE))
    assert r[0] == 5.0 and r[1] == -3.0 and math.isnan(r[2])


@pytest.mark.parametrize("pos,length", [(1, 3), (5, 2**31 - 1), (-4, 2**31 - 1), (0, 2), (3, 1), (-3, 2),
                                        (2, 0), (30, 5), (-30, 1)])
def test_substr_vs_oracle(ctx, orc, pos, length):
    from ghive_amd import ops
    rng = np.random.default_rng(abs(pos) * 31 + length % 97)
    words = [b"Facebook", b"abc", b"", "héllo".encode(), "中文字符".encode(),
             b"MFGR#1221", None]
    vals = [words[i] for i in rng.integers(0, len(words), 20_000)]
    out = ops.select_substr(ctx, _dev_strings(vals), pos, length)
    assert out.to_list() == orc.substr(vals, pos, length)


def test_substr_reference_functor_case(ctx):
    """SelectPredicateTest.cu substrTest: substr_functor(0, 3) (the first three bytes) of the 20
    strings, then FILTER_EQ "122" -- as Hive's SUBSTR(s, 1, 3)."""
    from ghive_amd import ops
    strs = [b"1221", b"1223", b"1238", b"1221", b"1238", b"122", b"1232", b"1228", b"123", b"1224", b"1248",
            b"12223", b"1237", b"1222", b"1230", b"1221", b"1234", b"1220", b"1239", b"12234"]
    sub = ops.select_substr(ctx, _dev_strings(strs), 1, 3)
    assert sub.to_list() == [s[:3] for s in strs]
    t = ops.Table([sub])
    sel = to_np(ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 0, [b"122"])], 0).process(ctx, t)).tolist()
    assert sel == [i for i, s in enumerate(strs) if s[:3] == b"122"]
    # string_convert_int_functor over the same strings (stringConversionTest), as CAST(s AS INT)
    assert to_np(ops.select_cast(ctx, _dev_strings(strs), abi.INT)).tolist() == [int(s) for s in strs]


_TINY_PLAN = """Plan optimized by CBO.

Stage-0
  Fetch Operator
    limit:-1
    Stage-1
      Map 1
      File Output Operator [FS_2]
        Select Operator [SEL_1] (rows=10 width=100)
          Output:["_col0","_col1","_col2"]
          TableScan [TS_0] (rows=10 width=100)
            db@t,t,Tbl:COMPLETE,Col:NONE,Output:["s","n"]
"""
_TINY_EXTENDED = """STAGE PLANS:
  Stage: Stage-1
    Tez
      Vertices:
        Map 1 
            Map Operator Tree:
                TableScan
                  alias: t
                  Select Operator
                    expressions: substr(s, 2, 3) (type: string), UDFToInteger(s) (type: int), UDFToDouble(n) (type: double)
                    outputColumnNames: _col0, _col1, _col2
                    File Output Operator
                      compressed: false
"""


def test_plan_projection_functions(ctx, orc):
    """A projection calling substr, UDFToInteger and UDFToDouble (a hand-written single-vertex plan
    in EXPLAIN / EXPLAIN EXTENDED form) against the oracle's SUBSTR and CAST."""
    from ghive_amd import ops
    rng = np.random.default_rng(77)
    vals = _rand_texts(rng, 5000)
    n = rng.integers(-10**6, 10**6, 5000).astype(np.int64)
    n[::13] = LONG_NULL
    plan = ops.Plan(_TINY_PLAN, _TINY_EXTENDED)
    names, cols = plan.execute(ctx, "Map 1", {"db@t": (ops.Table([_dev_strings(vals), to_dev(n)]), ["s", "n"])})
    assert names == ["_col0", "_col1", "_col2"]
    assert cols[0].to_list() == orc.substr(vals, 2, 3)
    assert to_np(cols[1]).tolist() == [INT_NULL if v is None else v for v in orc.cast_string_to_int(vals)]
    d = to_np(cols[2])
    assert np.array_equal(np.isnan(d), n == LONG_NULL) and np.array_equal(d[n != LONG_NULL], n[n != LONG_NULL])


_LIMIT_PLAN = """Plan optimized by CBO.

Vertex dependency in root stage
Reducer 2 <- Map 1 (SIMPLE_EDGE)

Stage-0
  Fetch Operator
    limit:10
    Stage-1
      Reducer 2
      File Output Operator [FS_6]
        Limit [LIM_5] (rows=10 width=16)
          Number of rows:10
          Select Operator [SEL_4] (rows=1000 width=16)
            Output:["_col0","_col1"]
          <-Map 1 [SIMPLE_EDGE]
            SHUFFLE [RS_3]
              Select Operator [SEL_2] (rows=1000 width=16)
                Output:["_col0","_col1"]
                TableScan [TS_0] (rows=1000 width=16)
                  db@t,t,Tbl:COMPLETE,Col:NONE,Output:["k","v"]
"""
_LIMIT_EXTENDED = """STAGE PLANS:
  Stage: Stage-1
    Tez
      Vertices:
        Map 1 
            Map Operator Tree:
                TableScan
                  alias: t
                  Select Operator
                    expressions: k (type: bigint), v (type: bigint)
                    outputColumnNames: _col0, _col1
                    Reduce Output Operator
                      key expressions: _col0 (type: bigint)
                      null sort order: a
                      sort order: +
                      tag: -1
                      value expressions: _col1 (type: bigint)
        Reducer 2 
            Reduce Operator Tree:
              Select Operator
                expressions: KEY.reducesinkkey0 (type: bigint), VALUE._col0 (type: bigint)
                outputColumnNames: _col0, _col1
                Limit
                  Number of rows: 10
                  File Output Operator
                    compressed: false
"""


def test_plan_limit_applies_after_order_by(ctx):
    """ORDER BY k LIMIT 10 over a reduce input in hash order (a hand-written two-vertex plan in
    EXPLAIN / EXPLAIN EXTENDED form): the ten smallest keys, not a sort of ten arbitrary rows."""
    from ghive_amd import ops
    rng = np.random.default_rng(91)
    k = rng.permutation(1000).astype(np.int64) * 3 - 500
    v = k * 7 + 1
    plan = ops.Plan(_LIMIT_PLAN, _LIMIT_EXTENDED)
    # the map side passes its rows through (keys first)
    names, cols = plan.execute(ctx, "Map 1", {"db@t": (ops.Table([to_dev(k), to_dev(v)]), ["k", "v"])})
    assert names == ["_col0", "_col1"]
    names, (rk, rv) = plan.execute(ctx, "Reducer 2", {"Map 1": (ops.Table(cols), None)})
    order = np.lexsort((np.arange(len(k)), k))[:10]
    assert to_np(rk).tolist() == k[order].tolist()
    assert to_np(rv).tolist() == v[order].tolist()
