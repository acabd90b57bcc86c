"""Writes tests/golden/reference_kats.json: known-answer vectors transcribed from the reference's
own tests (inputs and expected outputs only -- data, no source text), and copies a few of the
reference's query-plan data files (Hive EXPLAIN output, gpu_impl/plan/*.txt) into
tests/golden/plans/ as inputs of the plan-parser tests.

Run in a container where the reference tree is mounted (it is read, never executed):
    python tests/golden/make_fixtures.py /root/reference

Sources (paths relative to the GHive tree; hive/ = hive-rel-release-3.1.0/):
  gpu_impl/test/FilterPredicateTest.cu:17-36          predicate parse-tree dumps
  gpu_impl/test/JoinPredicateTest.cu:153-191          LEFT SEMI + 2x LEFT OUTER result_offsets
  gpu_impl/test/GroupByPredicateTest.cu:20-57,63-150,154-189,194-2775   group-by tables
  hive/ql/src/test/.../exec/vector/expressions/TestVectorFilterExpressions.java  filter KATs
  hive/ql/src/test/.../exec/vector/TestVectorGroupByOperator.java               aggregate KATs
The small vectors are written out literally below with their line citations; the 2,550-row
bigTableTestTable input is extracted from the test file by a regex over its assignment lines.
"""
import json
import os
import re
import sys

INT_MAX = 2**31 - 1
LONG_MAX = 2**63 - 1
LONG_MIN = -2**63
INT_MIN = -2**31


def filter_parse_kats():
    # gpu_impl/test/FilterPredicateTest.cu:19-35
    return [
        {
            "cite": "gpu_impl/test/FilterPredicateTest.cu:19-28",
            "text": "(lo_discount BETWEEN 1.0D AND 3.0D and (lo_quantity < 25.0D) and lo_orderdate is not null)",
            "cols": ["lo_orderdate", "lo_quantity", "lo_extendedprice", "lo_discount"],
            "expect": "predicate: 11; filterCol: x;paramNum: 0;params: x;\n"
                      "  predicate: 11; filterCol: x;paramNum: 0;params: x;\n"
                      "    predicate: 10; filterCol: 0;paramNum: 0;params: x;\n"
                      "    predicate: 3; filterCol: 1;paramNum: 1;params: 25.000000 ;\n"
                      "  predicate: 6; filterCol: 3;paramNum: 2;params: 1.000000 3.000000 ;\n",
        },
        {
            "cite": "gpu_impl/test/FilterPredicateTest.cu:31-35",
            "text": "d_datekey is not null",
            "cols": ["d_datekey", "d_year"],
            "expect": "predicate: 10; filterCol: 0;paramNum: 0;params: x;\n",
        },
    ]


def join_kats():
    # gpu_impl/test/JoinPredicateTest.cu:153-191
    return [{
        "cite": "gpu_impl/test/JoinPredicateTest.cu:153-191",
        "tables": [[4, 1, 3, 4], [1, 4, 4, 3], [1, 1, 3, 3], [5, 2, 4, 3]],
        "type": "LONG",
        "join_cols": [0, 0, 0, 0],
        "conds": [[4, 0, 1], [1, 0, 2], [1, 0, 3]],  # LEFT_SEMI 0-1, LEFT_OUTER 0-2, 0-3
        "expect_offsets": [1, 0, -1, 1, 1, -1, 2, 2, 3, 2, 3, 3, 0, -1, 2, 3, -1, 2],
        "width": 3,
    }]


def groupby_gpu_impl_kats(ref_root):
    kats = []
    # GroupByPredicateTest.cu:20-57 -- STRING key, AVG(double)
    kats.append({
        "cite": "gpu_impl/test/GroupByPredicateTest.cu:20-57",
        "columns": [
            {"type": "LONG", "values": [1, 2, 2]},
            {"type": "DOUBLE", "values": [2.0, 3.5, 3.0]},
            {"type": "INT", "values": [2, 3, 2]},
            {"type": "STRING", "blob": "zxz", "idx": [0, 1, 1, 2, 2, 3]},
        ],
        "keys": [3], "aggs": [[3, 1]],
        "expect": [{"key": ["x"], "aggs": [3.5]}, {"key": ["z"], "aggs": [2.5]}],
    })
    # GroupByPredicateTest.cu:63-150 -- INT + STRING keys, SUM(double)
    d = [3794438.0, 832702.0, 6759001.0, 275248.0, 1.1119494E7, 4597888.0, 2.6017672E7,
         542696.0, 1.4258494E7, 3150352.0, 7865033.0, 2795390.0, 6987585.0, 1.1092823E7,
         2071920.0, 1.7524655E7, 1.003082E7, 8327856.0, 4927500.0, 4659127.0]
    years = [1998] * 13 + [1997] * 7
    strs = ["MFGR#1221", "MFGR#1221", "MFGR#1238", "MFGR#1221", "MFGR#1238", "MFGR#1221",
            "MFGR#1238", "MFGR#1221", "MFGR#1238", "MFGR#1221", "MFGR#1238", "MFGR#1221",
            "MFGR#1238", "MFGR#1221", "MFGR#1238", "MFGR#1221", "MFGR#1238", "MFGR#1221",
            "MFGR#1238", "MFGR#1221"]
    blob, idx, off = "", [], 0
    for s in strs:
        idx += [off, off + len(s)]
        off += len(s)
        blob += s
    kats.append({
        "cite": "gpu_impl/test/GroupByPredicateTest.cu:63-150",
        "columns": [{"type": "DOUBLE", "values": d}, {"type": "INT", "values": years},
                    {"type": "STRING", "blob": blob, "idx": idx}],
        "keys": [1, 2], "aggs": [[0, 0]],
        "expect": [{"key": [1997, "MFGR#1221"], "aggs": [41604461.0]},
                   {"key": [1997, "MFGR#1238"], "aggs": [17030240.0]},
                   {"key": [1998, "MFGR#1221"], "aggs": [15988714.0]},
                   {"key": [1998, "MFGR#1238"], "aggs": [73007279.0]}],
    })
    # GroupByPredicateTest.cu:154-189 -- LONG key, SUM(double)
    kats.append({
        "cite": "gpu_impl/test/GroupByPredicateTest.cu:154-189",
        "columns": [{"type": "DOUBLE", "values": [716471960.0, 659973086.0, 707313389.0,
                                                  704901122.0, 662317475.0, 624687821.0]},
                    {"type": "LONG", "values": [1997, 1993, 1992, 1997, 1993, 1992]}],
        "keys": [1], "aggs": [[0, 0]],
        "expect": [{"key": [1992], "aggs": [1332001210.0]}, {"key": [1993], "aggs": [1322290561.0]},
                   {"key": [1997], "aggs": [1421373082.0]}],
    })
    # GroupByPredicateTest.cu:194-2775 -- 2,550 rows, INT key, SUM(double), 7 groups
    path = os.path.join(ref_root, "gpu_impl/test/GroupByPredicateTest.cu")
    keys, vals = {}, {}
    pat = re.compile(r"key\[(\d+)\]\s*=\s*(-?\d+);\s*value\[(\d+)\]\s*=\s*([-+0-9.eE]+);")
    with open(path) as fh:
        for ln, line in enumerate(fh, 1):
            if 200 <= ln <= 2760:
                m = pat.search(line)
                if m:
                    keys[int(m.group(1))] = int(m.group(2))
                    vals[int(m.group(3))] = float(m.group(4))
    assert len(keys) == 2550 and len(vals) == 2550, (len(keys), len(vals))
    kats.append({
        "cite": "gpu_impl/test/GroupByPredicateTest.cu:194-2775",
        "columns": [{"type": "INT", "values": [keys[i] for i in range(2550)]},
                    {"type": "DOUBLE", "values": [vals[i] for i in range(2550)]}],
        "keys": [0], "aggs": [[0, 1]],
        "expect": [{"key": [1992], "aggs": [1332001226.0]}, {"key": [1993], "aggs": [1320663638.0]},
                   {"key": [1994], "aggs": [1429210220.0]}, {"key": [1995], "aggs": [1310278322.0]},
                   {"key": [1996], "aggs": [1442958982.0]}, {"key": [1997], "aggs": [1421373031.0]},
                   {"key": [1998], "aggs": [855955531.0]}],
    })
    return kats


def hive_filter_kats():
    c = "hive/ql/src/test/org/apache/hadoop/hive/ql/exec/vector/expressions/TestVectorFilterExpressions.java"
    v5 = [5, 20, 17, 15, 10]
    return [
        # mode numbers are PredicateMode; null rows listed by index (become sentinels)
        {"cite": c + ":336-357", "type": "LONG", "values": v5, "nulls": [],
         "pred": {"mode": 6, "params": [15, 17]}, "expect": [2, 3]},
        {"cite": c + ":359-380", "type": "LONG", "values": v5, "nulls": [0, 2],
         "pred": {"mode": 6, "params": [15, 17]}, "expect": [3]},
        {"cite": c + ":382-409", "type": "LONG", "values": [5, 20, 17, 15, 10, 19, 21],
         "nulls": [0, 2, 5], "sel": [1, 2, 4], "pred": {"mode": 6, "params": [15, 17]},
         "expect": []},
        {"cite": c + ":432-453", "type": "LONG", "values": v5, "nulls": [],
         "pred": {"mode": 8, "params": [10, 20]}, "expect": [0]},
        {"cite": c + ":455-477", "type": "DOUBLE", "values": v5, "nulls": [],
         "pred": {"mode": 6, "params": [20.0, 21.0]}, "expect": [1]},
        {"cite": c + ":479-500", "type": "DOUBLE", "values": v5, "nulls": [],
         "pred": {"mode": 8, "params": [10.0, 20.0]}, "expect": [0]},
        # FilterLongScalarLessLongColumn(15, 0): 15 < x, i.e. x > 15
        {"cite": c + ":232-251", "type": "LONG", "values": v5, "nulls": [],
         "pred": {"mode": 4, "params": [15]}, "expect": [1, 2]},
        # then FilterLongScalarGreaterLongColumn(18, 0) on that selection: x < 18
        {"cite": c + ":253-257", "type": "LONG", "values": v5, "nulls": [], "sel": [1, 2],
         "pred": {"mode": 3, "params": [18]}, "expect": [2]},
        # FilterLongColGreaterLongColumn(0, 1), size 3: col0 = 10,10,10 ; col1 = 20,1,7
        {"cite": c + ":78-98", "type": "LONG", "values": [10, 10, 10], "values2": [20, 1, 7],
         "nulls": [], "pred": {"mode": 4, "col2": True}, "expect": [1, 2]},
        {"cite": c + ":100-105", "type": "LONG", "values": [10, 10, 10], "values2": [20, 1, 7],
         "nulls": [1], "sel": [1, 2], "pred": {"mode": 4, "col2": True}, "expect": [2]},
    ]


def hive_groupby_kats():
    c = "hive/ql/src/test/org/apache/hadoop/hive/ql/exec/vector/TestVectorGroupByOperator.java"
    N = None
    k4 = [1, 1, 2, 2]
    v4 = [13, 5, 7, 19]
    kn = [N, 2, N, 2]
    out = [
        # keyed: (agg, keys, values, {key: expected})
        dict(cite=c + ":1015-1022", agg="min", keys=k4, values=v4, expect=[[1, 5], [2, 7]]),
        dict(cite=c + ":1035-1042", agg="min", keys=[1, 2, 1, 2], values=v4, expect=[[1, 7], [2, 5]]),
        dict(cite=c + ":1045-1052", agg="min", keys=kn, values=v4, expect=[[N, 7], [2, 5]]),
        dict(cite=c + ":1065-1072", agg="max", keys=kn, values=v4, expect=[[N, 13], [2, 19]]),
        dict(cite=c + ":1075-1082", agg="count", keys=kn, values=v4, expect=[[N, 2], [2, 2]]),
        dict(cite=c + ":1085-1092", agg="sum", keys=kn, values=v4, expect=[[N, 20], [2, 24]]),
        dict(cite=c + ":1095-1102", agg="avg", keys=kn, values=v4, expect=[[N, 10.0], [2, 12.0]]),
        dict(cite=c + ":1115-1122", agg="min", keys=kn, values=[N, N, N, N], expect=[[N, N], [2, N]]),
        # keyless: (agg, values, expected)
        dict(cite=c + ":1125-1131", agg="min", keys=None, values=v4, expect=5),
        dict(cite=c + ":1414-1420", agg="sum", keys=None, values=v4, expect=44),
        dict(cite=c + ":1423-1429", agg="sum", keys=None, values=[], expect=N),
        dict(cite=c + ":1432-1437", agg="sum", keys=None, values=[N], expect=N),
        dict(cite=c + ":1438-1442", agg="sum", keys=None, values=[N, N, N], expect=N),
        dict(cite=c + ":1443-1447", agg="sum", keys=None, values=[N, 5, 7, 19], expect=31),
        dict(cite=c + ":1448-1452", agg="sum", keys=None, values=[13, N, 7, 19], expect=39),
        dict(cite=c + ":1490-1496", agg="sum", keys=None, values=[-INT_MAX, INT_MAX], expect=0),
        dict(cite=c + ":1499-1505", agg="sum", keys=None, values=[INT_MAX, INT_MAX], expect=4294967294),
        dict(cite=c + ":1508-1514", agg="sum", keys=None, values=[INT_MIN, INT_MIN], expect=-4294967296),
        # Long.MIN_VALUE + Long.MIN_VALUE wraps to 0 (silent overflow, Java long)
        dict(cite=c + ":1526-1532", agg="sum", keys=None, values=[LONG_MIN, LONG_MIN], expect=0),
        dict(cite=c + ":1544-1550", agg="avg", keys=None, values=v4, expect=44 / 4),
        dict(cite=c + ":1553-1559", agg="avg", keys=None, values=[], expect=N),
        dict(cite=c + ":1395-1401", agg="sum", keys=None, values=[13.0, 5.0, 7.0, 19.0],
             expect=44.0, dtype="DOUBLE"),
    ]
    # Cases whose literals equal the LONG NULL sentinel are not transcribed: in the reference's
    # data model (Table.hpp:29-52) Long.MAX_VALUE *is* NULL, so testSumLong2MaxLong (:1517) and
    # testSumLongMinMaxLong (:1535) have a different (sentinel-model) answer.
    return out


PLAN_FIXTURES = ["ssb_10_orc_transformed_1_1", "ssb_10_orc_transformed_2_1", "ssb_10_orc_transformed_3_1",
                 "ssb_10_orc_transformed_4_1", "tpcds_orc_100_transformed_query1",
                 "tpcds_orc_100_transformed_query2", "tpcds_orc_100_transformed_query3"] + \
                ["ssb_2_orc_" + q for q in ("1_1", "1_2", "1_3", "1_4", "2_1", "2_2", "2_3", "3_1", "3_2", "3_3",
                                            "3_4", "4_1", "4_2", "4_3")] + \
                ["ssb_10_orc_transformed_" + q for q in ("1_2", "1_3", "2_2", "2_3", "3_2", "3_3", "3_4", "4_2",
                                                         "4_3")]


def copy_plans(ref_root):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plans")
    os.makedirs(out, exist_ok=True)
    for name in PLAN_FIXTURES:
        for suffix in (".txt", "_extended.txt"):
            with open(os.path.join(ref_root, "gpu_impl", "plan", name + suffix)) as fh:
                text = fh.read()
            with open(os.path.join(out, name + suffix), "w") as fh:
                fh.write(text)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    copy_plans(ref)
    data = {
        "filter_parse": filter_parse_kats(),
        "join": join_kats(),
        "groupby_gpu_impl": groupby_gpu_impl_kats(ref),
        "hive_filter": hive_filter_kats(),
        "hive_groupby": hive_groupby_kats(),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
