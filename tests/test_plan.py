"""Plan parsing (ghive_plan_*): the operator tree of every vertex and the operators' resolved
descriptors, checked against values read off the reference's own plan files (Hive EXPLAIN output
of SSB queries, copied into tests/golden/plans/ by tests/golden/make_fixtures.py). Host code only:
runs without a GPU."""
import os

import pytest

from ghive_amd import abi

PLANS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans")
INNER, LEFT_OUTER, LEFT_SEMI = abi.INNER_JOIN, abi.LEFT_OUTER_JOIN, abi.LEFT_SEMI_JOIN
SUM = abi.AGG_SUM


def _plan(name, extended=True, maintain=None):
    from ghive_amd import ops
    with open(os.path.join(PLANS, name + ".txt")) as fh:
        text = fh.read()
    ext = None
    if extended:
        with open(os.path.join(PLANS, name + "_extended.txt")) as fh:
            ext = fh.read()
    return ops.Plan(text, ext, maintain)


def test_ssb_q1_1_tree_and_descriptors():
    p = _plan("ssb_10_orc_transformed_1_1")
    d = p.describe()
    assert d["root"] == "FS_32"
    assert d["vertices"] == {"Reducer 2": "FS_32", "Map 1": "RS_30", "Map 3": "RS_24"}
    assert p.chain("Map 1") == ["RS_30", "GBY_29", "SEL_28", "MAPJOIN_27", "SEL_26", "FIL_25", "TS_0"]
    assert p.chain("Map 3") == ["RS_24", "SEL_23", "FIL_22", "TS_3"]
    assert p.chain("Reducer 2") == ["FS_32", "GBY_31"]
    assert all(o["linked"] for o in d["operators"].values())
    ts = p.op("TS_0")
    assert ts["table"] == "ssb_10_orc_transformed@lineorder"
    assert ts["outputs"] == ["lo_orderdate", "lo_quantity", "lo_extendedprice", "lo_discount"]
    assert p.op("FIL_25")["predicate"] == \
        "(lo_discount BETWEEN 1.0D AND 3.0D and (lo_quantity < 25.0D) and lo_orderdate is not null)"
    sel = p.op("SEL_26")
    assert sel["outputs"] == ["_col0", "_col2", "_col3"]
    assert sel["expressions"] == ["lo_orderdate", "lo_extendedprice", "lo_discount"]
    assert p.op("SEL_28")["expressions"] == ["(_col2 * _col3)"]
    mj = p.op("MAPJOIN_27")
    assert mj["join_inputs"] == ["SEL_26", "RS_24"]          # big table first, as Conds names them
    assert mj["join_cols"] == [[0], [0]]                    # SEL_26._col0 = RS_24._col0
    assert mj["join_types"] == [[INNER, 0, 1]]
    assert mj["outputs"] == ["_col2", "_col3"]
    assert p.op("GBY_29")["keys"] == [] and p.op("GBY_29")["aggs"] == [[SUM, 0]]
    assert p.op("GBY_29")["mode"] == "hash"
    # the reduce side: sum(VALUE._col0) over a sink with no keys -> column 0
    assert p.op("GBY_31")["aggs"] == [[SUM, 0]] and p.op("GBY_31")["mode"] == "mergepartial"
    rs = p.op("RS_24")
    assert rs["key_expressions"] == ["_col0"] and rs["key_offsets"] == [0] and rs["sort_orders"] == "+"
    assert p.op("FS_32")["require_sort"] and p.op("FS_32")["sort_orders"] == ""


def test_ssb_q2_1_reduce_side_joins_and_order():
    p = _plan("ssb_10_orc_transformed_2_1")
    # keys are resolved against the child's outputs: _col10, _col8 of [_col3, _col8, _col10]
    g = p.op("GBY_22")
    assert g["keys"] == [2, 1] and g["aggs"] == [[SUM, 0]]
    mj = p.op("MAPJOIN_44")
    assert mj["join_inputs"] == ["MERGEJOIN_43", "RS_59"] and mj["join_cols"] == [[0], [0]]
    mg = p.op("MERGEJOIN_43")
    assert mg["join_inputs"] == ["RS_53", "RS_56"] and mg["join_cols"] == [[0], [0]]
    assert p.op("MAPJOIN_52")["join_cols"] == [[1], [0]]  # SEL_51._col1 = RS_49._col0
    # a reduce sink lists its key columns first
    rs = p.op("RS_62")
    assert rs["key_offsets"] == [1, 2] and rs["outputs"] == ["_col1", "_col2", "_col0"]
    assert rs["original_outputs"] == ["_col0", "_col1", "_col2"]
    # KEY._colN -> N, VALUE._colN -> N + the sink's key count
    assert p.op("GBY_60")["keys"] == [0, 1] and p.op("GBY_60")["aggs"] == [[SUM, 2]]
    assert p.op("FS_64")["require_sort"] and p.op("FS_64")["sort_orders"] == "++"


def test_ssb_q3_1_descending_order():
    p = _plan("ssb_10_orc_transformed_3_1")
    assert p.op("GBY_22")["keys"] == [3, 1, 2]
    assert p.op("GBY_59")["keys"] == [0, 1, 2] and p.op("GBY_59")["aggs"] == [[SUM, 3]]
    assert p.op("RS_61")["key_offsets"] == [2, 3] and p.op("RS_61")["sort_orders"] == "+-"
    assert p.op("FS_63")["sort_orders"] == "+-"


def test_ssb_q4_1_nine_vertices():
    p = _plan("ssb_10_orc_transformed_4_1")
    d = p.describe()
    assert set(d["vertices"]) == {f"Map {i}" for i in (1, 6, 7, 8, 9)} | {f"Reducer {i}" for i in (2, 3, 4, 5)}
    assert all(o["linked"] for o in d["operators"].values())
    assert p.chain("Reducer 3") == ["RS_30", "GBY_29", "SEL_27", "MAPJOIN_58", "MERGEJOIN_57"]
    assert p.op("MAPJOIN_66")["join_cols"] == [[2], [0]]
    assert p.op("MERGEJOIN_57")["join_inputs"] == ["RS_21", "RS_73"]
    assert p.op("MERGEJOIN_57")["join_cols"] == [[0], [0]]
    assert p.op("MAPJOIN_58")["join_inputs"] == ["MERGEJOIN_57", "RS_76"]
    assert p.op("SEL_27")["expressions"] == ["_col14", "_col9", "(_col4 - _col5)"]
    assert p.op("GBY_29")["keys"] == [0, 1] and p.op("GBY_29")["aggs"] == [[SUM, 2]]
    assert p.op("GBY_77")["keys"] == [0, 1] and p.op("GBY_77")["aggs"] == [[SUM, 2]]
    assert p.op("FIL_71")["predicate"] == "(((p_mfgr = 0L) or (p_mfgr = 1L)) and p_partkey is not null)"
    assert p.op("FS_80")["sort_orders"] == "++"


def test_maintained_columns_from_the_java_map():
    # map-side join: names resolve against the input's outputs
    p = _plan("ssb_10_orc_transformed_1_1",
              maintain='{"MAPJOIN_27":{"0":["_col2","_col3"],"1":[]}}')
    assert p.op("MAPJOIN_27")["maintain_cols"] == [[1, 2], []]
    # reduce-side join: VALUE._colN is the sink's column _col(N+1) before the keys-first reordering
    p = _plan("ssb_10_orc_transformed_2_1",
              maintain='{"MERGEJOIN_43":{"0":["VALUE._col1"],"1":["VALUE._col1"]}}')
    assert p.op("MERGEJOIN_43")["maintain_cols"] == [[1], [1]]  # _col2 of [_col0,_col2,_col3] / [_col0,_col2]


def test_concise_plan_alone_and_limit():
    # this fixture's EXPLAIN EXTENDED comes from another optimizer run (other vertices), so only
    # the tree of the concise plan is checked
    p = _plan("tpcds_orc_100_transformed_query3", extended=False)
    assert p.describe()["root"] == "FS_47"
    lim = p.op("LIM_46")
    assert lim["kind"] == "Limit" and lim["limit"] == 100
    assert p.chain("Reducer 5") == ["FS_47", "LIM_46", "SEL_45"]
    assert p.op("GBY_16")["keys"] == [3, 1, 2]  # _col8, _col4, _col5 of MERGEJOIN_33's outputs
    assert p.op("MERGEJOIN_32")["join_inputs"] == ["RS_36", "RS_39"]
    assert p.op("TS_6")["table"] == "tpcds_100_transformed@date_dim"


@pytest.mark.parametrize("text,maintain", [("", None), ("Plan optimized by CBO.\n\nStage-0\n", None),
                                           (None, '{"MAPJOIN_1":{"0":["_col0"]'), (None, "[1, 2]")])
def test_plan_errors(text, maintain):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    if text is None:
        with open(os.path.join(PLANS, "ssb_10_orc_transformed_1_1.txt")) as fh:
            text = fh.read()
    with pytest.raises(GhiveError) as e:
        ops.Plan(text, None, maintain)
    assert e.value.status == abi.ERR_INVALID


def test_fused_form_recognised_from_the_plan():
    """The map vertex of the CS1 shape (scan, filter, map joins, group-by, sink) has a fused form;
    reduce vertices and chains through a merge join do not (host-only: the plan decides the shape)."""
    from ghive_amd import ops
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

    def gbys(d, a, b):
        p = ops.Plan(open(os.path.join(golden, d, a)).read(), open(os.path.join(golden, d, b)).read())
        return {k: (o["vertex"], o["fused_form"], o["not_fused"]) for k, o in p.describe()["operators"].items()
                if o["kind"] == "GroupBy"}

    g = gbys("plans_authored", "star3.txt", "star3_extended.txt")
    assert g["GBY_42"][:2] == ("Map 1", True) and g["GBY_44"][1] is False and "mergepartial" in g["GBY_44"][2]
    for schema in ("ssb_10_orc_transformed", "ssb_2_orc"):
        g = gbys("plans", f"{schema}_1_1.txt", f"{schema}_1_1_extended.txt")
        assert g["GBY_29"][:2] == ("Map 1", True)
    g = gbys("plans", "ssb_10_orc_transformed_2_1.txt", "ssb_10_orc_transformed_2_1_extended.txt")
    assert g["GBY_22"][1] is False and "MERGEJOIN_43" in g["GBY_22"][2]


def test_tpcds_q3_from_the_extended_plan_alone():
    """TPC-DS q3's EXPLAIN and EXPLAIN EXTENDED fixtures come from two optimizer runs (four
    reducers and merge joins only in one, three reducers and a broadcast map join in the other),
    so the extended text alone gives the tree: its Edges and operator trees, ids numbered in
    order of appearance. Expected values read off tpcds_orc_100_transformed_query3_extended.txt."""
    from ghive_amd import ops
    with open(os.path.join(PLANS, "tpcds_orc_100_transformed_query3_extended.txt")) as fh:
        p = ops.Plan(None, fh.read())
    d = p.describe()
    O = d["operators"]
    assert d["root"] == "FS_20"
    assert sorted(d["vertices"]) == ["Map 1", "Map 5", "Map 6", "Reducer 2", "Reducer 3", "Reducer 4"]
    assert p.chain("Map 1") == ["RS_3", "SEL_2", "FIL_1", "TS_0"]
    assert O["TS_0"]["table"] == "tpcds_orc_100_transformed@store_sales"
    assert O["TS_0"]["outputs"] == ["ss_sold_date_sk", "ss_item_sk", "ss_sales_price"]
    assert O["TS_4"]["outputs"] == ["i_item_sk", "i_brand_id", "i_brand", "i_manufact_id"]
    assert O["FIL_9"]["predicate"] == "((d_moy = 11) and d_date_sk is not null)"
    # Reducer 2: merge join of Map 1 (tag 0) and Map 5 (tag 1) on ss_item_sk = i_item_sk, then the
    # broadcast map join with Map 6 on ss_sold_date_sk = d_date_sk, then the hash group-by
    mj = O["MERGEJOIN_12"]
    assert mj["vertex"] == "Reducer 2" and mj["join_inputs"] == ["RS_3", "RS_7"]
    assert mj["join_cols"] == [[0], [0]] and mj["join_types"] == [[INNER, 0, 1]]
    assert O["RS_3"]["outputs"] == ["_col1", "_col0", "_col2"]  # the reduce key first
    mp = O["MAPJOIN_13"]
    assert mp["join_inputs"] == ["MERGEJOIN_12", "RS_11"] and mp["join_cols"] == [[0], [0]]
    assert mp["outputs"] == ["_col2", "_col4", "_col5", "_col8"]
    g = O["GBY_14"]
    assert g["keys"] == [3, 1, 2] and g["aggs"] == [[SUM, 0]] and g["mode"] == "hash"
    assert O["GBY_16"]["keys"] == [0, 1, 2] and O["GBY_16"]["aggs"] == [[SUM, 3]]
    assert O["RS_17"]["sort_orders"] == "+-+" and O["RS_17"]["key_offsets"] == [0, 3, 1]
    assert O["LIM_19"]["limit"] == 100
    assert O["FS_20"]["require_sort"] and O["FS_20"]["sort_orders"] == "+-+"


@pytest.mark.parametrize("q,why", [(1, "Map 7 has 2 sink operators"), (2, "UNION vertices are not supported")])
def test_tpcds_extended_plans_outside_the_vertex_model(q, why):
    """q1 runs one map vertex into two reducers (two sinks, while a vertex run returns one result)
    and q2 reads UNION vertices: both are refused with the reason, not parsed into a wrong tree."""
    from ghive_amd import ops
    with open(os.path.join(PLANS, f"tpcds_orc_100_transformed_query{q}_extended.txt")) as fh:
        text = fh.read()
    with pytest.raises(ops.GhiveError) as e:
        ops.Plan(None, text)
    assert e.value.status == abi.ERR_UNSUPPORTED and why in str(e.value)
