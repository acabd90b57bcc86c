"""Serving form (ghive_server_*): every request's result is identical whether it runs alone through
the plain entry point or inside a concurrent mix on the server (fragments, group-bys and sorts at
once, device or host I/O); admission by HBM footprint holds the in-flight bytes under the budget
and rejects requests that can never fit; a failing request reports its status without stopping
the loop."""
import numpy as np
import pytest
import torch

from conftest import to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu


def _groups(outs, nkeys):
    """Group outputs (unordered) -> dict key tuple -> aggregate tuple; doubles compared by bits."""
    cols = []
    for o in outs:
        a = to_np(o)
        cols.append(a.view(np.int64) if a.dtype == np.float64 else a.astype(np.int64))
    return {tuple(int(c[g]) for c in cols[:nkeys]): tuple(int(c[g]) for c in cols[nkeys:])
            for g in range(len(cols[0]))}


def _mix(ctx):
    """A request mix over resident inputs: (kind, args, plain-form result)."""
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 6_000_000)
    ht = W.small_dim_table(ctx, T.dim)
    reqs = []
    # slices start on 16-byte boundaries (the fused kernel's column alignment)
    for rows, begin in ((1_000, 0), (250_000, 8), (4_000_000, 124), (6_000_000, 0), (77_777, 4)):
        fact = ops.Table([c[begin: begin + rows] for c in T.fact.columns])
        frag = W.small_fragment(fact, ht)
        reqs.append(("fragment", (frag, 64), _groups(frag.execute(ctx, capacity=64), 1)))
    for n, distinct, seed in ((1_000_000, 1_000, 1), (3_000_000, 400_000, 2)):
        k = ctx.gen_uniform_i64(seed, 0, n, 0, distinct - 1)
        v = ctx.gen_uniform_i64(seed, 1, n, -10**12, 10**12)
        d = ctx.gen_normal_f64(seed, 2, n)
        t = ops.Table([k, v, d])
        aggs = [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR), (abi.AGG_MIN, 1), (abi.AGG_MAX, 1),
                (abi.AGG_SUM, 2), (abi.AGG_AVG, 2)]
        want = _groups(ops.GroupByPredicate([0], aggs).gpu_execute(ctx, t, capacity=distinct), 1)
        reqs.append(("groupby", (t, [0], aggs, distinct), want))
    for n, seed in ((2_000_000, 3), (123_457, 4)):
        k = ctx.gen_uniform_i64(seed, 0, n, -5000, 5000)
        f = ctx.gen_uniform_f32(seed, 1, n)
        t = ops.Table([k, f])
        reqs.append(("sort", (t, [0, 1], "+-"), to_np(ops.sort(ctx, t, [0, 1], "+-"))))
    torch.cuda.synchronize()
    return reqs


def _submit(srv, kind, args):
    if kind == "fragment":
        frag, cap = args
        return srv.submit_fragment(frag, capacity=cap)
    if kind == "groupby":
        t, keys, aggs, cap = args
        return srv.submit_groupby(t, keys, aggs, capacity=cap)
    t, keys, orders = args
    return srv.submit_sort(t, keys, orders)


def _check(kind, outs, want):
    if kind == "sort":
        assert np.array_equal(to_np(outs[0]), want)
    else:
        assert _groups(outs, 1) == want


def _host_args(kind, args):
    from ghive_amd import ops
    hostify = lambda t: ops.Table([c.cpu().pin_memory() for c in t.columns], host=True)
    if kind == "fragment":
        frag, cap = args
        return (ops.Fragment(hostify(frag.fact), frag.joins_py, [(abi.SRC_JOIN, 0)],
                             [(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)]), cap)
    t, *rest = args
    return (hostify(t), *rest)


@pytest.mark.parametrize("slots", [1, 4])
def test_serve_mix_identical_to_plain(ctx, slots):
    from ghive_amd import ops
    reqs = _mix(ctx)
    srv = ops.Server(slots=slots, hbm_budget=32 << 30, slot_mem_limit=8 << 30)
    try:
        for rep in range(2):  # twice: slots reuse their memory managers
            order = list(range(len(reqs)))[:: (1 if rep == 0 else -1)]
            tickets = [(i, _submit(srv, reqs[i][0], reqs[i][1])) for i in order]
            for i, tk in tickets:
                _check(reqs[i][0], srv.wait(tk), reqs[i][2])
                assert tk.response.status == abi.OK and tk.response.slot < slots
                assert tk.response.submit_us <= tk.response.start_us <= tk.response.end_us
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs) and st["queued"] == 0 and st["inflight"] == 0
        assert 1 <= st["peak_inflight"] <= slots
        if slots > 1:
            assert st["peak_inflight"] > 1  # requests were admitted while others ran
    finally:
        srv.close()


def test_serve_host_io_identical_to_plain(ctx):
    from ghive_amd import ops
    reqs = _mix(ctx)
    srv = ops.Server(slots=3, hbm_budget=32 << 30, slot_mem_limit=8 << 30)
    try:
        tickets = [(kind, _submit(srv, kind, _host_args(kind, args)), want) for kind, args, want in reqs]
        for kind, tk, want in tickets:
            outs = srv.wait(tk)
            assert all(not o.is_cuda for o in outs)
            _check(kind, outs, want)
    finally:
        srv.close()


def test_serve_admission_by_footprint(ctx):
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 2_000_000, seed=11)
    ht = W.small_dim_table(ctx, T.dim)
    frag = W.small_fragment(T.fact, ht)
    want = _groups(frag.execute(ctx, capacity=64), 1)
    torch.cuda.synchronize()
    budget = 100 << 20
    srv = ops.Server(slots=4, hbm_budget=budget, slot_mem_limit=1 << 30)
    try:
        with pytest.raises(ops.GhiveError) as e:
            srv.submit_fragment(frag, capacity=64, footprint=budget + 1)
        assert e.value.status == abi.ERR_ADMISSION
        # 60 MB each: only one fits the 100 MB budget at a time although four slots are free
        tickets = [srv.submit_fragment(frag, capacity=64, footprint=60 << 20) for _ in range(6)]
        for tk in tickets:
            assert _groups(srv.wait(tk), 1) == want
        st = srv.stats()
        assert st["rejected"] == 1 and st["completed"] == 6
        assert st["peak_inflight"] == 1 and st["peak_inflight_bytes"] == 60 << 20
        # admissions follow arrival order
        starts = [tk.response.start_us for tk in tickets]
        assert starts == sorted(starts)
    finally:
        srv.close()


def test_serve_error_does_not_stop_the_loop(ctx):
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 500_000, seed=12)
    ht = W.small_dim_table(ctx, T.dim)
    frag = W.small_fragment(T.fact, ht)
    want = _groups(frag.execute(ctx, capacity=64), 1)
    torch.cuda.synchronize()
    srv = ops.Server(slots=2, hbm_budget=8 << 30, slot_mem_limit=1 << 30)
    try:
        bad = srv.submit_fragment(frag, capacity=3)  # 25 groups do not fit 3 rows
        good = srv.submit_fragment(frag, capacity=64)
        with pytest.raises(ops.GhiveError) as e:
            srv.wait(bad)
        assert e.value.status == abi.ERR_CAPACITY and bad.response.groups > 3
        assert _groups(srv.wait(good), 1) == want
    finally:
        srv.close()


def test_async_built_table_used_first_by_a_server_slot(ctx):
    """A dimension table built asynchronously on one context (filter count left on the device) and
    used first by a server slot, another context on its own stream: the slot finishes the build
    from the owner's stream, and the fragment's groups equal those of a synchronously built table."""
    from ghive_amd import ops, workloads as W
    T = W.small_tables(ctx, 2_000_000)
    pred = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0)
    sel, cnt = pred.process_async(ctx, T.dim)
    ha = ops.HashTable(ctx, T.dim.columns[0], sel=sel, payload=T.dim.columns[2], count=cnt)
    hs = W.small_dim_table(ctx, T.dim)
    want = _groups(W.small_fragment(T.fact, hs).execute(ctx, capacity=64), 1)
    srv = ops.Server(slots=2, hbm_budget=8 << 30, slot_mem_limit=2 << 30)
    try:
        t = srv.submit_fragment(W.small_fragment(T.fact, ha), capacity=64)
        got = _groups(srv.wait(t), 1)
    finally:
        srv.close()
    assert got == want and len(want) > 1
    ha.free()
    hs.free()


def test_serve_full_size_requests_identical_to_plain(ctx):
    """The largest requests of the serving mix, all in flight together: a 256M-row sort, a group-by
    over 16M rows with up to 10M distinct int64 keys, and 16M-row fragments with two filter
    constants; each served result equals the same request run alone through the plain entry point
    (outputs compared on the device; groups after sorting by key)."""
    from ghive_amd import ops, workloads as W
    I = W.serving_inputs(ctx)
    reqs = [("sort", dict(rows=I.sort_keys.numel())),
            ("groupby", dict(rows=I.gb_keys.numel())),
            ("fragment", dict(start=0, rows=I.small.fact_rows, attr=3)),
            ("fragment", dict(start=4096, rows=I.small.fact_rows - 4096, attr=5)),
            ("sort", dict(rows=100_000_000))]
    torch.cuda.synchronize()
    srv = ops.Server(slots=4, hbm_budget=96 << 30, slot_mem_limit=32 << 30)
    try:
        tickets = [W.serving_submit(srv, I, k, p)[0] for k, p in reqs]
        served = [srv.wait(t) for t in tickets]
        assert srv.stats()["peak_inflight"] > 1
    finally:
        srv.close()

    def canon(kind, outs):  # groups come out in no fixed order; a sort's permutation is the result
        if kind == "sort":
            return outs
        order = torch.argsort(outs[0])
        return [o[order] for o in outs]

    for (kind, p), got in zip(reqs, served):
        want = W.serving_plain(ctx, I, kind, p)
        got, want = canon(kind, got), canon(kind, want)
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert g.shape == w.shape and torch.equal(g, w), (kind, p)
        if kind == "groupby":
            assert 5_000_000 < got[0].numel() <= W.SERVE_GB_DISTINCT
            assert int(got[2].sum()) == p["rows"]  # COUNT(*) over every group
        if kind == "sort":
            k = I.sort_keys[: p["rows"]][got[0].long()]
            assert bool((k[1:] >= k[:-1]).all())
