"""The drop-in boundary end to end: GTables host buffers (LONG integers, as Hive's vectorized path
delivers them) through ghive_gpu_process -- staging, one plan vertex, GPUResult arrays -- with the
map vertex of the CS1 shape running as one fused fragment. Checked against the oracle (the C
restatement of Hive's vectorized map-join + group-by), the query restated in numpy, and the same
vertex run as the operator chain (option plan.fuse = 0)."""
import math
import os

import numpy as np
import pytest

from conftest import to_dev, to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PLANS = os.path.join(HERE, "golden", "plans")
AUTHORED = os.path.join(HERE, "golden", "plans_authored")
LONG_NULL = abi.LONG_NULL


def _read(d, name):
    with open(os.path.join(d, name)) as fh:
        return fh.read()


def _star_plan():
    from ghive_amd import ops
    return ops.Plan(_read(AUTHORED, "star3.txt"), _read(AUTHORED, "star3_extended.txt"))


def _star_data(n, seed=5):
    """Fact (date, supplier, part foreign keys, price, discount) and the three dimensions, dense
    keys; every integer LONG, as the vectorized reader hands them over."""
    rng = np.random.default_rng(seed)
    nd, ns, npart = 2556, 200_000, 1_400_000
    d_key = np.arange(1, nd + 1, dtype=np.int64)
    d_year = np.minimum(1992 + np.arange(nd) // 365, 1998).astype(np.int64)
    s_key = np.arange(1, ns + 1, dtype=np.int64)
    s_region = rng.integers(0, 5, ns).astype(np.int64)
    p_key = np.arange(1, npart + 1, dtype=np.int64)
    p_cat = rng.integers(0, 25, npart).astype(np.int64)
    p_brand = rng.integers(0, 1000, npart).astype(np.int64)
    fact = [rng.integers(1, nd + 1, n), rng.integers(1, ns + 1, n), rng.integers(1, npart + 1, n),
            rng.integers(100_000, 10_000_001, n), rng.integers(0, 11, n)]
    fact = [f.astype(np.int64) for f in fact]
    return fact, (d_key, d_year), (s_key, s_region), (p_key, p_cat, p_brand)


def _oracle_star(orc, fact, dates, supp, part):
    """The oracle's fragment (Hive's vectorized map joins + hash group-by) over the same rows."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    dims = [dict(columns=[i32(c) for c in dates], key_col=0, payload_col=1, fact_key_col=0,
                 filter=([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [1992, 1993], data_type=2)], 0)),
            dict(columns=[i32(c) for c in supp], key_col=0, payload_col=1, fact_key_col=1,
                 filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0)),
            dict(columns=[i32(c) for c in part], key_col=0, payload_col=2, fact_key_col=2,
                 filter=([abi.make_pred(abi.FILTER_EQ, 1, [1], data_type=2)], 0))]
    spec = orc.FragmentSpec([i32(c) for c in fact], dims, keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)],
                            aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 3, 4)])
    res = spec.run(nthreads=4)
    return {(int(k[0]), int(k[1])): int(v[0]) for k, v in zip(res["keys"], res["aggs"])}


def _dims_through_boundary(ctx, plan, dates, supp, part):
    from ghive_amd import ops
    out = {}
    for v, cols in (("Map 2", list(dates)), ("Map 3", list(supp)), ("Map 4", list(part))):
        cols_out, seq, counts = ops.gpu_process(ctx, plan, [(v, cols)], v)
        assert ops.process_times(ctx)["fused"] == 0  # filter vertices: the operator chain
        out[v] = cols_out
    return out


@pytest.mark.parametrize("n", [1, 4097, 3_000_000])
def test_star_map_vertex_fused_through_gpu_process(ctx, orc, n):
    from ghive_amd import ops
    plan = _star_plan()
    fact, dates, supp, part = _star_data(n)
    dims = _dims_through_boundary(ctx, plan, dates, supp, part)
    assert len(dims["Map 2"][0]) == 730 and dims["Map 2"][0].dtype == np.int64
    tables = [("Map 1", fact)] + [(v, dims[v]) for v in ("Map 2", "Map 3", "Map 4")]
    cols, seq, counts = ops.gpu_process(ctx, plan, tables, "Map 1")
    pt = ops.process_times(ctx)
    assert pt["fused"] == 1, ctx.plan_info()["not_fused"]
    info = ctx.plan_info()
    assert info["joins"] == 3 and not info["has_filter"]  # the IS NOT NULL conjuncts are implied
    assert pt["narrowed_columns"] == 10  # 5 fact columns + date (2), supplier (1), part (2)
    assert counts == [3, 0, 0, 0] and seq == [0, 1, 2]  # year, brand, sum: LONG, keys first
    got = {(int(y), int(b)): int(s) for y, b, s in zip(*cols)}
    assert got == _oracle_star(orc, fact, dates, supp, part)
    # the operator chain on the same inputs gives the same rows
    ctx.set_option("plan.fuse", 0)
    try:
        cols2, _, _ = ops.gpu_process(ctx, plan, tables, "Map 1")
        assert ops.process_times(ctx)["fused"] == 0
    finally:
        ctx.set_option("plan.fuse", 1)
    assert {(int(y), int(b)): int(s) for y, b, s in zip(*cols2)} == got
    # dimension tables at the regional-build size are built one by one, not in one batched call
    ctx.set_option("hash.regional_rows", 1)
    try:
        cols3, _, _ = ops.gpu_process(ctx, plan, tables, "Map 1")
        assert ops.process_times(ctx)["fused"] == 1
    finally:
        ctx.set_option("hash.regional_rows", 1 << 22)
    assert {(int(y), int(b)): int(s) for y, b, s in zip(*cols3)} == got
    # and the reduce vertex finishes the query
    final, _, _ = ops.gpu_process(ctx, plan, [("Map 1", cols)], "Reducer 5")
    assert {(int(y), int(b)): int(s) for y, b, s in zip(*final)} == got


def test_star_values_beyond_32_bits_run_unfused(ctx, orc):
    """A LONG fact column with a value that does not fit 32 bits: the staging keeps it LONG, the
    fused form is ruled out, and the operator chain gives the answer."""
    from ghive_amd import ops
    plan = _star_plan()
    fact, dates, supp, part = _star_data(200_000, seed=9)
    fact[3][17] = 2**40 + 3  # a price that needs 64 bits
    dims = _dims_through_boundary(ctx, plan, dates, supp, part)
    tables = [("Map 1", fact)] + [(v, dims[v]) for v in ("Map 2", "Map 3", "Map 4")]
    cols, _, _ = ops.gpu_process(ctx, plan, tables, "Map 1")
    assert ops.process_times(ctx)["fused"] == 0
    assert "32 bits" in ctx.plan_info()["not_fused"]
    d_year, s_reg, p_cat, p_brand = dates[1], supp[1], part[1], part[2]
    fd, fs, fp, pr, di = fact
    m = (d_year[fd - 1] >= 1992) & (d_year[fd - 1] <= 1993) & (s_reg[fs - 1] == 1) & (p_cat[fp - 1] == 1)
    want = {}
    for y, b, v in zip(d_year[fd[m] - 1], p_brand[fp[m] - 1], pr[m] * di[m]):
        want[(int(y), int(b))] = want.get((int(y), int(b)), 0) + int(v)
    assert {(int(y), int(b)): int(s) for y, b, s in zip(*cols)} == want


def test_star_plan_execute_narrows_device_long_columns(ctx, orc):
    """ghive_plan_execute on device LONG columns (no staging): the fused form narrows them on the
    device after a range check."""
    from ghive_amd import ops
    plan = _star_plan()
    fact, dates, supp, part = _star_data(500_000, seed=11)
    dims = _dims_through_boundary(ctx, plan, dates, supp, part)
    ins = {"star@lineorder": (ops.Table([to_dev(c) for c in fact]),
                              ["lo_orderdate", "lo_suppkey", "lo_partkey", "lo_price", "lo_discount"])}
    for v in ("Map 2", "Map 3", "Map 4"):
        ins[v] = (ops.Table([to_dev(c) for c in dims[v]]), None)
    names, cols = plan.execute(ctx, "Map 1", ins)
    info = ctx.plan_info()
    assert info["fused"] and info["narrowed_columns"] == 10 and info["fused_kernel_ms"] > 0
    got = {(int(y), int(b)): int(s) for y, b, s in zip(*[to_np(c) for c in cols])}
    assert got == _oracle_star(orc, fact, dates, supp, part)


def _dates():
    days = np.arange(np.datetime64("1992-01-01"), np.datetime64("1999-01-01"))
    y = days.astype("datetime64[Y]").astype(np.int64) + 1970
    m = days.astype("datetime64[M]").astype(np.int64) % 12 + 1
    d = (days - days.astype("datetime64[M]")).astype(np.int64) + 1
    return (y * 10000 + m * 100 + d).astype(np.int64), y.astype(np.int64)


@pytest.mark.parametrize("schema", ["ssb_10_orc_transformed", "ssb_2_orc"])
def test_ssb_q1_1_map_vertex_fused(ctx, schema):
    """SSB Q1.1's map vertex (the reference's own plan fixture): a fact filter on DOUBLE columns,
    a map join to the filtered dates, SUM(lo_extendedprice * lo_discount) over DOUBLE -- fused,
    and exactly the correctly rounded sum of the rounded products."""
    from ghive_amd import ops
    plan = ops.Plan(_read(PLANS, f"{schema}_1_1.txt"), _read(PLANS, f"{schema}_1_1_extended.txt"))
    rng = np.random.default_rng(31)
    dk, dy = _dates()
    n = 600_000
    od = dk[rng.integers(0, len(dk), n)]
    od[rng.random(n) < 0.03] = -5
    od[rng.random(n) < 0.01] = LONG_NULL
    qty = rng.integers(1, 51, n).astype(np.float64)
    qty[rng.random(n) < 0.01] = np.nan
    price = np.round(rng.random(n) * 1e5, 2)
    disc = rng.integers(0, 11, n).astype(np.float64)
    disc[rng.random(n) < 0.01] = np.nan
    v3, _, _ = ops.gpu_process(ctx, plan, [("Map 3", [dk, dy])], "Map 3")
    tables = [("Map 1", [od, qty, price, disc]), ("Map 3", v3)]
    cols, seq, counts = ops.gpu_process(ctx, plan, tables, "Map 1")
    pt, info = ops.process_times(ctx), ctx.plan_info()
    assert pt["fused"] == 1, info["not_fused"]
    assert info["has_filter"] and info["joins"] == 1 and pt["narrowed_columns"] == 2
    assert counts == [0, 1, 0, 0] and len(cols[0]) == 1
    m = np.isin(od, dk[dy == 1993]) & (disc >= 1) & (disc <= 3) & (qty < 25)
    assert cols[0][0] == math.fsum((price * disc)[m].tolist())
    ctx.set_option("plan.fuse", 0)
    try:
        cols2, _, _ = ops.gpu_process(ctx, plan, tables, "Map 1")
    finally:
        ctx.set_option("plan.fuse", 1)
    assert cols2[0].view(np.uint64)[0] == cols[0].view(np.uint64)[0]


def test_gpu_process_missing_input_gives_no_result(ctx):
    from ghive_amd import ops
    plan = _star_plan()
    fact, dates, supp, part = _star_data(1000)
    assert ops.gpu_process(ctx, plan, [("Map 1", fact), (None, [dates[0]])], "Map 1") is None


_PROJ_PLAN = """Plan optimized by CBO.

Stage-0
  Fetch Operator
    limit:-1
    Stage-1
      Map 1
      File Output Operator [FS_2]
        Select Operator [SEL_1] (rows=10 width=100)
          Output:["_col0","_col1","_col2","_col3"]
          TableScan [TS_0] (rows=10 width=100)
            db@t,t,Tbl:COMPLETE,Col:NONE,Output:["s","n"]
"""
_PROJ_EXTENDED = """STAGE PLANS:
  Stage: Stage-1
    Tez
      Vertices:
        Map 1 
            Map Operator Tree:
                TableScan
                  alias: t
                  Select Operator
                    expressions: s (type: string), UDFToInteger(n) (type: int), UDFToDouble(n) (type: double), n (type: bigint)
                    outputColumnNames: _col0, _col1, _col2, _col3
                    File Output Operator
                      compressed: false
"""


def test_gpu_process_string_and_mixed_result(ctx):
    """GPUResult's grouping (generate_tbl_result, ghive_JNIInterface.cu:223-342): a projection with
    STRING, INT, DOUBLE and LONG outputs comes back grouped LONG, DOUBLE, INT, STRING with
    sequence[c] = the position of output column c; strings as blob + (start,end) pairs; and a
    vertex with no group-by (SSB Q2.1 Map 5) equals the plan executed on device inputs."""
    from ghive_amd import ops
    words = [b"MFGR#12", b"", None, "h\xe9llo".encode(), b"x" * 40]
    rng = np.random.default_rng(8)
    vals = [words[i] for i in rng.integers(0, len(words), 3000)]
    blob, idx, o = bytearray(), [], 0
    for v in vals:
        if v is None:
            idx += [-1, -1]
            continue
        blob += v
        idx += [o, o + len(v)]
        o += len(v)
    n = rng.integers(-10**6, 10**6, 3000).astype(np.int64)
    plan = ops.Plan(_PROJ_PLAN, _PROJ_EXTENDED)
    cols, seq, counts = ops.gpu_process(ctx, plan, [("Map 1", [(bytes(blob), np.array(idx, np.int32)), n])], "Map 1")
    assert counts == [1, 1, 1, 1] and seq == [3, 2, 1, 0]
    sblob, pairs = cols[0]
    got = [None if pairs[2 * r] < 0 else sblob[pairs[2 * r]:pairs[2 * r + 1]] for r in range(len(vals))]
    assert got == vals
    assert cols[1].dtype == np.int32 and cols[1].tolist() == n.astype(np.int32).tolist()
    assert cols[2].dtype == np.float64 and np.array_equal(cols[2], n.astype(np.float64))
    assert cols[3].dtype == np.int64 and np.array_equal(cols[3], n)
    plan = ops.Plan(_read(PLANS, "ssb_10_orc_transformed_2_1.txt"),
                    _read(PLANS, "ssb_10_orc_transformed_2_1_extended.txt"))
    key = rng.permutation(5000).astype(np.int64) + 1
    reg = rng.integers(0, 5, 5000).astype(np.int64)
    cols, seq, counts = ops.gpu_process(ctx, plan, [("Map 5", [key, reg])], "Map 5")
    assert counts == [1, 0, 0, 0] and sorted(cols[0].tolist()) == sorted(key[reg == 1].tolist())
    names, dev = plan.execute(ctx, "Map 5", {"ssb_10_orc_transformed@supplier": (
        ops.Table([to_dev(key), to_dev(reg)]), ["s_suppkey", "s_region"])})
    assert sorted(to_np(dev[0]).tolist()) == sorted(cols[0].tolist())
