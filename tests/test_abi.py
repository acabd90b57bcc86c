"""CPU-only checks of the C-ABI library: it loads, exports every entry point include/ghive/*.h
declares, its struct layouts match the ctypes mirror, and the host-only entry points
(predicate parser and dump) reproduce the reference's known answers. No GPU call is made."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

from conftest import ROOT
from ghive_amd import abi

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "ghive", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ghive_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def L():
    from ghive_amd._lib import lib
    return lib()


def test_library_exports_every_declared_symbol(L):
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert L.ghive_abi_version() == abi.ABI_VERSION


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof from the C header (compiled with gcc) against the ctypes mirror."""
    structs = {
        "ghive_column": abi.Column, "ghive_table": abi.Table, "ghive_pred": abi.Pred,
        "ghive_join_cond": abi.JoinCond, "ghive_agg": abi.Agg,
        "ghive_fragment_join": abi.FragmentJoin, "ghive_fragment_key": abi.FragmentKey,
        "ghive_fragment_agg": abi.FragmentAgg, "ghive_fragment": abi.Fragment,
        "ghive_device_info": abi.DeviceInfo, "ghive_stage_times": abi.StageTimes,
        "ghive_batch_item": abi.BatchItem, "ghive_batch_spec": abi.BatchSpec,
        "ghive_batch_tuning": abi.BatchTuning, "ghive_plan_input": abi.PlanInput, "ghive_gtables": abi.GTables,
        "ghive_server_config": abi.ServerConfig,
        "ghive_request": abi.Request,
        "ghive_response": abi.Response,
        "ghive_server_stats": abi.ServerStats,
        "ghive_plan_exec_info": abi.PlanExecInfo,
        "ghive_process_times": abi.ProcessTimes,
        "ghive_sort_info": abi.SortInfo,
        "ghive_filtered_build": abi.FilteredBuild,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ghive/ghive.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l.strip())
    for cname, py in structs.items():
        assert int(got[f"{cname} size"]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_parse_predicate_reference_kats(kats):
    from ghive_amd import ops
    for k in kats["filter_parse"]:
        pred = ops.FilterOperator(k["text"], k["cols"]).generatePredicate()
        assert pred.toString() == k["expect"], k["cite"]


def test_parse_predicate_forms():
    from ghive_amd import ops
    cols = ["a", "b", "c"]
    P = lambda s: ops.FilterOperator(s, cols).generatePredicate()
    p = P("(a = 3 or (b <> 7L and c > 2.5D))")
    assert p.nodes[p.root].mode == abi.FILTER_OR
    p = P("a NOT BETWEEN 1 AND 5")  # not shadowed by BETWEEN
    n = p.nodes[p.root]
    assert n.mode == abi.FILTER_NOT_RANGE and n.int_params[0] == 1 and n.int_params[1] == 5
    p = P("b >= 3000000000")  # wider than int: kept as a long literal
    n = p.nodes[p.root]
    assert n.data_type == 0 and n.long_params[0] == 3000000000
    p = P("(a <= -4)")
    assert p.nodes[p.root].int_params[0] == -4
    # X1 and X2 and X3 folds from the right: AND(AND(X3, X2), X1)
    p = P("(a = 1 and b = 2 and c = 3)")
    top = p.nodes[p.root]
    inner = p.nodes[top.children[0]]
    assert p.nodes[top.children[1]].filter_col[0] == 0
    assert p.nodes[inner.children[0]].filter_col[0] == 2 and p.nodes[inner.children[1]].filter_col[0] == 1


def _lit(n, k):
    return bytes(n.string_params[k][: n.string_sizes[k]])


def test_parse_predicate_string_literals():
    """'...' literals (FilterOperator.cu:251-255: the text between the quotes) are one token, spaces,
    parentheses, '=' and ' AND ' inside included; the tree dump prints them as the reference does."""
    from ghive_amd import ops
    cols = ["p_brand1", "c_region", "s_nation", "x"]
    P = lambda s: ops.FilterOperator(s, cols).generatePredicate()
    p = P("(p_brand1 BETWEEN 'MFGR#2221' AND 'MFGR#2228')")
    n = p.nodes[p.root]
    assert n.mode == abi.FILTER_EQ_RANGE and n.data_type == 3 and n.param_num == 2
    assert _lit(n, 0) == b"MFGR#2221" and _lit(n, 1) == b"MFGR#2228"
    p = P("((c_region = 'AMERICA') and (s_nation = 'UNITED KI1'))")
    top = p.nodes[p.root]
    assert _lit(p.nodes[top.children[0]], 0) == b"UNITED KI1"
    assert _lit(p.nodes[top.children[1]], 0) == b"AMERICA"
    n = P("x = 'a (b) = c AND d'").nodes[0]
    assert n.mode == abi.FILTER_EQ and n.filter_col[0] == 3 and _lit(n, 0) == b"a (b) = c AND d"
    n = P("x <> ''").nodes[0]
    assert n.mode == abi.FILTER_NOT_EQ and n.string_sizes[0] == 0
    assert P("(c_region = 'ASIA')").toString() == "predicate: 0; filterCol: 1;paramNum: 1;params: ASIA ;\n"
    for bad in ["x = 'abc", "x = '" + "y" * 65 + "'", "x BETWEEN 'a' AND 3"]:
        with pytest.raises(Exception):
            P(bad)


@pytest.mark.parametrize("text", ["(a = 1", "z = 1", "a ~ 2", "a = xyz", "(a = 1) b"])
def test_parse_predicate_errors(text):
    from ghive_amd import ops
    from ghive_amd._lib import GhiveError
    with pytest.raises(GhiveError):
        ops.FilterOperator(text, ["a", "b"]).generatePredicate()


def test_gpuresult_sequence_numbering(L):
    """generate_tbl_result (ghive_JNIInterface.cu:223-342): GPUResult holds LONG columns first, then
    DOUBLE, INT, STRING, each in column order; sequence[c] is column c's place in that order."""
    import ctypes as C
    LONG, DOUBLE, INT, STRING, FLOAT = abi.LONG, abi.DOUBLE, abi.INT, abi.STRING, abi.FLOAT

    def seq(types):
        t = (C.c_int32 * max(len(types), 1))(*types)
        out = (C.c_int32 * max(len(types), 1))()
        counts = (C.c_uint32 * 4)()
        assert L.ghive_gpuresult_sequence(t, len(types), out, counts) == abi.OK
        return list(out)[: len(types)], list(counts)
    # 2 LONG (0, 1), 1 DOUBLE (2), 1 INT (3), 1 STRING (4)
    assert seq([INT, LONG, STRING, DOUBLE, LONG]) == ([3, 0, 4, 2, 1], [2, 1, 1, 1])
    assert seq([STRING, STRING, INT]) == ([1, 2, 0], [0, 0, 1, 2])
    assert seq([FLOAT, DOUBLE, LONG]) == ([1, 2, 0], [1, 2, 0, 0])  # FLOAT travels as double
    assert seq([]) == ([], [0, 0, 0, 0])
    bad = (C.c_int32 * 1)(9)
    assert L.ghive_gpuresult_sequence(bad, 1, (C.c_int32 * 1)(), (C.c_uint32 * 4)()) == abi.ERR_INVALID


def test_agg_output_types(L):
    assert L.ghive_agg_output_type(abi.AGG_SUM, abi.INT) == abi.LONG
    assert L.ghive_agg_output_type(abi.AGG_SUM, abi.DOUBLE) == abi.DOUBLE
    assert L.ghive_agg_output_type(abi.AGG_CNT, abi.DOUBLE) == abi.LONG
    assert L.ghive_agg_output_type(abi.AGG_AVG, abi.INT) == abi.DOUBLE
    assert L.ghive_agg_output_type(abi.AGG_MIN, abi.FLOAT) == abi.FLOAT
    for t in range(5):
        assert L.ghive_agg_output_type(t, abi.INT) == abi.agg_output_type(t, abi.INT)


def test_null_context_is_rejected(L):
    assert L.ghive_sync(None) == abi.ERR_INVALID
    assert L.ghive_last_error(None) == b"no context"
