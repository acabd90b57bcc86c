"""Batched form: the persistent HIP-graph kernel over many uneven fragments returns, for every
fragment, exactly what the plain per-fragment operator returns, and agrees with the oracle."""
import numpy as np
import pytest

from conftest import to_np
from ghive_amd import abi

pytestmark = pytest.mark.gpu


def _small_batch(ctx, nfrags=48, seed=123):
    import torch
    from ghive_amd import workloads as W
    B = W.batch_tables(ctx, nfrags=nfrags, seed=seed)
    # rescale to a test-sized batch: keep the lognormal shape but cap each fragment at 60k rows
    rows = [min(r, 60_000) for r in B.rows]
    rows[0], rows[1] = 0, 4  # an empty fragment and a 4-row fragment
    starts, s = [], 0
    for r in rows:
        starts.append(s)
        s += r
    B.rows, B.starts = rows, starts
    B.fk, B.measure, B.price = (t[:s].contiguous() for t in (B.fk, B.measure, B.price))
    torch.cuda.synchronize()
    return B


def _batch_dicts(sums, counts):
    return [{g: int(sums[f, g]) for g in np.nonzero(counts[f])[0].tolist()} for f in range(sums.shape[0])]


@pytest.mark.parametrize("probe_mode,lf", [(1, 0.5), (2, 0.5), (3, 0.5), (3, 0.9), (4, 0.5)])
@pytest.mark.parametrize("tile_rows,use_graph", [(1024, True), (8192, True), (4096, False)])
def test_batch_equals_plain_per_fragment(ctx, probe_mode, lf, tile_rows, use_graph):
    from ghive_amd import workloads as W
    B = _small_batch(ctx)
    ht = W.batch_dim_table(ctx, B, load_factor=lf)
    bt = W.batch_create(ctx, B, ht, tile_rows=tile_rows, blocks_per_cu=2, probe_mode=probe_mode,
                        use_graph=use_graph)
    assert bt.info()["probe_mode"] == probe_mode
    bt.run()
    got = _batch_dicts(*bt.results())
    bt.run()  # a replay gives the same answer (the graph zeroes its accumulators)
    assert _batch_dicts(*bt.results()) == got
    plain = W.PlainBatch(ctx, B).run(collect=True)
    assert got == plain
    assert got[0] == {} and len(got) == len(B.rows)
    bt.free()
    ht.free()


def test_batch_ranked_codes_mode(ctx):
    """Probe mode 4 (filtered bitmap with ranked group codes in LDS) is the automatic choice for
    this shape; fragments whose filter passes no dimension key, and several filters, agree with
    the plain per-fragment operator."""
    from ghive_amd import workloads as W
    B = _small_batch(ctx, nfrags=40, seed=99)
    for f in (3, 4, 17):
        B.attr_const[f] = 99  # no dimension row has attr 99
    ht = W.batch_dim_table(ctx, B)
    bt = W.batch_create(ctx, B, ht, tile_rows=2048, blocks_per_cu=2)
    assert bt.info()["probe_mode"] == 4
    bt.run()
    got = _batch_dicts(*bt.results())
    assert got == W.PlainBatch(ctx, B).run(collect=True)
    assert got[3] == {} and got[4] == {} and got[17] == {}
    assert sum(1 for g in got if g) > 30
    bt.free()
    ht.free()


def test_batch_vs_oracle_some_fragments(ctx, orc):
    from ghive_amd import workloads as W
    B = _small_batch(ctx, nfrags=16, seed=7)
    ht = W.batch_dim_table(ctx, B)
    bt = W.batch_create(ctx, B, ht)
    bt.run()
    got = _batch_dicts(*bt.results())
    dim = [to_np(c) for c in B.dim.columns[:3]]
    for f in (2, 5, 15):
        s, r, c = B.starts[f], B.rows[f], B.attr_const[f]
        fact = [to_np(t[s:s + r]) for t in (B.fk, B.measure, B.price)]
        spec = orc.FragmentSpec(fact, [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                            filter=([abi.make_pred(abi.FILTER_EQ, 1, [c], data_type=2)], 0))],
                                keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)])
        res = spec.run(nthreads=2)
        want = {int(k[0]): int(v[0]) for k, v in zip(res["keys"], res["aggs"])}
        assert got[f] == want, f


def test_batch_rejects_bad_descriptors(ctx):
    import torch
    from ghive_amd import ops, workloads as W
    from ghive_amd._lib import GhiveError
    B = _small_batch(ctx, nfrags=4)
    ht = W.batch_dim_table(ctx, B)
    with pytest.raises(GhiveError):  # tile not a multiple of 1024
        W.batch_create(ctx, B, ht, tile_rows=1000)
    x = torch.zeros(9, dtype=torch.int32, device="cuda")
    with pytest.raises(GhiveError):  # misaligned column
        ops.Batch(ctx, ht, [(x[1:], x[1:], None, 0, 0)], 8, 3, 0, 5)


def test_full_4096_fragment_batch(ctx, orc):
    """The batched workload at full size: 4096 fragments with lognormal row counts (median 256k,
    clipped to 1k..16M, ~2.4e9 rows) in one persistent kernel replayed from a HIP graph, against the
    4096 plain per-fragment launches over the same columns, fragment by fragment, and the oracle on
    the smallest fragments and a few random ones."""
    import torch
    from ghive_amd import workloads as W
    B = W.batch_tables(ctx, nfrags=4096)
    assert len(B.rows) == 4096 and sum(B.rows) > 2_000_000_000
    ht = W.batch_dim_table(ctx, B)
    bt = W.batch_create(ctx, B, ht)
    bt.run()
    got = _batch_dicts(*bt.results())
    bt.free()
    plain = W.PlainBatch(ctx, B).run(collect=True)
    assert got == plain
    dim = [to_np(c) for c in B.dim.columns[:3]]
    order = np.argsort(B.rows)
    sample = list(order[:6]) + list(np.random.default_rng(1).choice(4096, 4, replace=False))
    for f in sample:
        s, r, c = B.starts[f], B.rows[f], B.attr_const[f]
        fact = [to_np(t[s:s + r]) for t in (B.fk, B.measure, B.price)]
        spec = orc.FragmentSpec(fact, [dict(columns=dim, key_col=0, payload_col=2, fact_key_col=0,
                                            filter=([abi.make_pred(abi.FILTER_EQ, 1, [c], data_type=2)], 0))],
                                keys=[(abi.SRC_JOIN, 0)], aggs=[(abi.AGG_SUM, abi.EXPR_MUL, 1, 2)])
        res = spec.run(nthreads=4)
        assert got[f] == {int(k[0]): int(v[0]) for k, v in zip(res["keys"], res["aggs"])}, f
    ht.free()
    del B
    torch.cuda.empty_cache()
