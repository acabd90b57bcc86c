"""ghive_fragment_execute_noalloc (the fused fragment with its scratch memory in a static device
array) against ghive_fragment_execute on the same inputs, through the C ABI: two calls in a row,
the second over a smaller fact table, the memory manager untouched, and the rejections."""
import numpy as np
import pytest
import torch

from ghive_amd import abi

pytestmark = pytest.mark.gpu


def _groups(outs):
    cols = [o.cpu().numpy() for o in outs]
    return {tuple(int(c[g]) for c in cols[:2]): int(cols[2][g]) for g in range(len(cols[0]))}


def _outs(dev, n=8192):
    return [torch.empty(n, dtype=torch.int32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
            torch.empty(n, dtype=torch.int64, device=dev)]


def test_noalloc_fragment_equals_fragment_execute(ctx):
    from ghive_amd import ops, workloads as W
    T = W.star_tables(ctx, 3_000_001)  # a partial last tile: the staging copy is used too
    hd, hs, hp = W.star_hash_tables(ctx, T, wait_each=True)  # synchronous builds: finished tables
    dev = T.fact.columns[0].device
    frag = W.star_fragment(T.fact, hd, hs, hp)
    want = _groups(frag.execute(ctx, capacity=8192))
    outs = _outs(dev)
    torch.cuda.synchronize()
    before = ctx.mem_stats()
    got = _groups(frag.execute(ctx, capacity=8192, outs=outs, noalloc=True))
    assert got == want and len(got) > 1500
    # the second call, over a smaller fact table (its first 1,000,003 rows)
    small = ops.Table([c[:1_000_003] for c in T.fact.columns])
    frag2 = W.star_fragment(small, hd, hs, hp)
    got2 = _groups(frag2.execute(ctx, capacity=8192, outs=outs, noalloc=True))
    after = ctx.mem_stats()
    assert after["in_use"] == before["in_use"] and after["peak"] == before["peak"]
    assert got2 == _groups(frag2.execute(ctx, capacity=8192)) and got2 != got
    for h in (hd, hs, hp):
        h.free()


def test_noalloc_fragment_general_kernel_and_rejections(ctx):
    """The general fused kernel (a fact filter), a table still being built, and a group table
    larger than the scratch array."""
    from ghive_amd import ops, workloads as W
    from ghive_amd._lib import GhiveError
    T = W.small_tables(ctx, 1_234_567)
    ht = W.small_dim_table(ctx, T.dim)
    flt = ops.FilterPredicate([abi.make_pred(abi.FILTER_LE, 1, [60], data_type=2)], 0)  # measure <= 60
    frag = ops.Fragment(T.fact, joins=[(0, ht)], keys=[(abi.SRC_JOIN, 0)],
                        aggs=[(abi.AGG_SUM, abi.EXPR_MUL, W.FACT_MEASURE, W.FACT_PRICE)], fact_filter=flt)
    dev = T.fact.columns[0].device
    outs = [torch.empty(64, dtype=torch.int32, device=dev), torch.empty(64, dtype=torch.int64, device=dev)]
    got = frag.execute(ctx, capacity=64, outs=outs, noalloc=True)
    g = {int(k): int(v) for k, v in zip(*[o.cpu().numpy() for o in got])}
    want = frag.execute(ctx, capacity=64)
    assert g == {int(k): int(v) for k, v in zip(*[o.cpu().numpy() for o in want])} and len(g) == 25
    # a group key taken from a fact column: the table sized for 2^22 groups exceeds the array
    wide = ops.Fragment(T.fact, joins=[(0, ht)], keys=[(abi.SRC_FACT, W.FACT_PRICE)],
                        aggs=[(abi.AGG_CNT, abi.EXPR_ONE, 0, 0)])
    big = [torch.empty(1 << 22, dtype=torch.int32, device=dev), torch.empty(1 << 22, dtype=torch.int64, device=dev)]
    with pytest.raises(GhiveError) as e:
        wide.execute(ctx, capacity=1 << 22, outs=big, noalloc=True)
    assert e.value.status == abi.ERR_UNSUPPORTED and "scratch" in str(e.value)
    # an asynchronously built table is finished on its first use, which allocates: rejected before it
    sel, cnt = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [3], data_type=2)], 0).process_async(ctx, T.dim)
    pending = ops.HashTable(ctx, T.dim.columns[0], sel=sel, payload=T.dim.columns[2], count=cnt)
    frag3 = W.small_fragment(T.fact, pending)
    with pytest.raises(GhiveError) as e:
        frag3.execute(ctx, capacity=64, outs=outs, noalloc=True)
    assert e.value.status == abi.ERR_INVALID and "still being built" in str(e.value)
    want3 = frag3.execute(ctx, capacity=64)  # its first use finishes it
    got3 = frag3.execute(ctx, capacity=64, outs=outs, noalloc=True)
    assert [o.cpu().tolist() for o in got3] == [o.cpu().tolist() for o in want3]
    for h in (ht, pending):
        h.free()
