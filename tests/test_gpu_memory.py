"""The device memory manager (the context's DeviceArena, read through ghive_mem_stats): repeated calls
of every operator give back all the memory they take -- the reference leaks every cudaMalloc
(gpu_impl/src/Operator/GroupByPredicate.cu:45 and the operators around it) -- and a call that needs
more than the context's limit fails with GHIVE_ERR_OOM, a message, and nothing left allocated, after
which the context keeps working."""
import os

import numpy as np
import pytest
import torch

from conftest import to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu
AUTHORED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plans_authored")


def _in_use(ctx):
    return ctx.mem_stats()["in_use"]


def _one_round(ctx, T, plan, star):
    """One call of each operator family; everything the library allocates is released by the end."""
    from ghive_amd import ops, workloads as W
    # filter + synchronous build + fused fragment
    ht = W.small_dim_table(ctx, T.dim)
    W.small_fragment(T.fact, ht).execute(ctx, capacity=64)
    ht.free()
    # asynchronous filter -> build -> fragment (the table is finished on its first use)
    sel, cnt = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0).process_async(ctx, T.dim)
    ha = ops.HashTable(ctx, T.dim.columns[0], sel=sel, payload=T.dim.columns[2], count=cnt)
    W.small_fragment(T.fact, ha).execute(ctx, capacity=64)
    ha.free()
    # group-by: the direct form and the hash-partitioned form
    k = T.fact.columns[0]
    ops.GroupByPredicate([0], [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)]).gpu_execute(
        ctx, ops.Table([k, T.fact.columns[1]]), capacity=300_000)
    ops.GroupByPredicate([0], [(abi.AGG_SUM, 1)]).gpu_execute(
        ctx, ops.Table([T.fact.columns[2], T.fact.columns[1]]), capacity=T.fact.row_num)
    # sort and an N-way join
    ops.sort(ctx, ops.Table([T.fact.columns[2], T.fact.columns[1]]), [0, 1], "+-")
    jp = ops.JoinPredicate([0, 0], [(abi.INNER_JOIN, 0, 1)], [[1], [2]])
    jp.process(ctx, [ops.Table([c[:200_000] for c in T.fact.columns]), T.dim])
    jp.generate_result(ctx, [ops.Table([c[:200_000] for c in T.fact.columns]), T.dim])
    # a fused plan vertex and the one-call boundary
    fact, dims = star
    ops.gpu_process(ctx, plan, [("Map 1", fact)] + [(v, dims[v]) for v in ("Map 2", "Map 3", "Map 4")], "Map 1")
    # the batched form (create, run, read, free)
    B = W.batch_tables(ctx, nfrags=24, seed=5)
    rows = [min(r, 40_000) for r in B.rows]
    starts, s = [], 0
    for r in rows:
        starts.append(s)
        s += r
    B.rows, B.starts = rows, starts
    hb = W.batch_dim_table(ctx, B)
    bt = W.batch_create(ctx, B, hb)
    bt.run()
    bt.results()
    bt.free()
    hb.free()
    torch.cuda.synchronize()


def test_repeated_calls_give_back_their_memory(ctx):
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 5_000_000)  # enough rows for the hash-partitioned group-by
    with open(os.path.join(AUTHORED, "star3.txt")) as a, open(os.path.join(AUTHORED, "star3_extended.txt")) as b:
        plan = ops.Plan(a.read(), b.read())
    rng = np.random.default_rng(4)
    n = 100_000
    fact = [rng.integers(1, 2557, n), rng.integers(1, 200_001, n), rng.integers(1, 1_400_001, n),
            rng.integers(100_000, 10_000_001, n), rng.integers(0, 11, n)]
    fact = [f.astype(np.int64) for f in fact]
    dims = {"Map 2": [np.arange(1, 731, dtype=np.int64), np.full(730, 1992, dtype=np.int64)],
            "Map 3": [np.arange(1, 40_001, dtype=np.int64) * 5],
            "Map 4": [np.arange(1, 56_001, dtype=np.int64) * 25, np.arange(56_000, dtype=np.int64) % 1000]}
    _one_round(ctx, T, plan, (fact, dims))  # first use: kernels load, the memory manager takes chunks
    base = _in_use(ctx)
    for _ in range(3):
        _one_round(ctx, T, plan, (fact, dims))
        assert _in_use(ctx) == base
    assert ctx.mem_stats()["peak"] >= base


def test_server_slots_give_back_their_memory(ctx):
    from ghive_amd import ops, workloads as W
    from ghive_amd._lib import lib
    T = W.small_tables(ctx, 3_000_000)
    ht = W.small_dim_table(ctx, T.dim)
    srv = ops.Server(slots=2, hbm_budget=8 << 30, slot_mem_limit=2 << 30)

    def slot_in_use():
        import ctypes as C
        out = []
        for i in range(srv.slots):
            vals = [C.c_uint64() for _ in range(4)]
            lib().ghive_mem_stats(lib().ghive_server_slot_context(srv.h, i), *[C.byref(x) for x in vals])
            out.append(vals[0].value)
        return out

    try:
        def round_():
            tickets = [srv.submit_fragment(W.small_fragment(T.fact, ht), capacity=64),
                       srv.submit_groupby(ops.Table([T.fact.columns[2], T.fact.columns[1]]), [0],
                                          [(abi.AGG_SUM, 1)], capacity=T.fact.row_num),
                       srv.submit_sort(ops.Table([T.fact.columns[1]]), [0], "+")]
            for t in tickets:
                srv.wait(t)
        round_()
        base = slot_in_use()
        for _ in range(3):
            round_()
            assert slot_in_use() == base
    finally:
        srv.close()
        ht.free()


def test_oom_is_a_clean_error():
    """A 96 MiB limit: a group-by whose table needs more and a sort whose key buffers need more each
    fail with GHIVE_ERR_OOM and leave nothing allocated; calls that fit still work."""
    from ghive_amd import ops, workloads as W
    from ghive_amd._lib import GhiveError
    c = ops.Context(0, mem_limit=96 << 20)
    try:
        base = c.mem_stats()["in_use"]
        k = c.gen_uniform_i64(9, 0, 12_000_000, 0, 8_000_000)
        v = c.gen_uniform_i64(9, 1, 12_000_000, -1000, 1000)
        with pytest.raises(GhiveError) as e:  # ~16M slots of keys and accumulators
            ops.GroupByPredicate([0], [(abi.AGG_SUM, 1)]).gpu_execute(c, ops.Table([k, v]), capacity=8_000_000)
        assert e.value.status == abi.ERR_OOM and "memory" in str(e.value)
        assert c.mem_stats()["in_use"] == base
        with pytest.raises(GhiveError) as e:  # two 16-byte key buffers of 12M rows
            ops.sort(c, ops.Table([k, v]), [0, 1], "++")
        assert e.value.status == abi.ERR_OOM
        assert c.mem_stats()["in_use"] == base
        # the same context still serves calls that fit
        small = ops.Table([k[:100_000], v[:100_000]])
        outs = ops.GroupByPredicate([0], [(abi.AGG_CNT, abi.COUNT_STAR)]).gpu_execute(c, small, capacity=100_000)
        assert int(to_np(outs[1]).sum()) == 100_000
        perm = to_np(ops.sort(c, small, [0], "+"))
        kk = to_np(k[:100_000])
        assert np.all(np.diff(kk[perm]) >= 0)
        assert c.mem_stats()["in_use"] == base
    finally:
        c.close()
