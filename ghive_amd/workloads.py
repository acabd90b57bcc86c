"""Synthetic workloads of the star-join fragment, generated on the device with Philox4x32-10.

Column c of a table draws from counter (row, c) under the table's seed, so the host can reproduce
any slice. Dimension keys are dense and unique (1..N), as in SSB.

  small   : fact 6M rows (fk U[1,200000], measure U[1,100], price U[1e5,1e7]);
            dim 200k rows (key, attr U[0,6], category U[0,24]);
            filter dim.attr == 3 -> join on fk -> GROUP BY category SUM(measure*price) (int64)
  star600m: fact 600M rows (fk_date, fk_cust, fk_supp, fk_part, quantity U[1,50],
            price U[1e5,1e7], discount U[0,10]); date 2556 rows (year 1992..1998, month 1..12),
            customer 3M, supplier 200k (region U[0,4]), part 1.4M (category U[0,24],
            brand U[0,999]); filters year BETWEEN 1992 AND 1993, region = 1, category = 1;
            three joins -> GROUP BY year, brand SUM(price*discount) (int64)
"""
from dataclasses import dataclass, field

import torch

from . import abi
from . import ops

SEED_BASE = 0x6768697665

# column ids of the generator streams
FACT_FK, FACT_MEASURE, FACT_PRICE = 0, 1, 2
DIM_ATTR, DIM_CATEGORY = 10, 11
F_DATE, F_CUST, F_SUPP, F_PART, F_QTY, F_PRICE, F_DISC = 0, 1, 2, 3, 4, 5, 6
SUPP_REGION, PART_CATEGORY, PART_BRAND = 20, 30, 31

STAR_DIMS = dict(date=2556, customer=3_000_000, supplier=200_000, part=1_400_000)


@dataclass
class SmallTables:
    fact: ops.Table
    dim: ops.Table
    seed: int
    fact_rows: int


def small_tables(ctx, fact_rows=6_000_000, seed=SEED_BASE + 1, dim_rows=200_000, fact_begin=0):
    fk = ctx.gen_uniform_i32(seed, FACT_FK, fact_rows, 1, dim_rows, row_begin=fact_begin)
    measure = ctx.gen_uniform_i32(seed, FACT_MEASURE, fact_rows, 1, 100, row_begin=fact_begin)
    price = ctx.gen_uniform_i32(seed, FACT_PRICE, fact_rows, 100_000, 10_000_000, row_begin=fact_begin)
    key = ctx.gen_sequence_i32(dim_rows, 1)
    attr = ctx.gen_uniform_i32(seed, DIM_ATTR, dim_rows, 0, 6)
    cat = ctx.gen_uniform_i32(seed, DIM_CATEGORY, dim_rows, 0, 24)
    return SmallTables(ops.Table([fk, measure, price]), ops.Table([key, attr, cat]), seed, fact_rows)


def small_dim_table(ctx, dim, attr_value=3, load_factor=0.5):
    """Filter dim.attr == attr_value and build the (key -> category) hash table."""
    pred = ops.FilterPredicate([abi.make_pred(abi.FILTER_EQ, 1, [attr_value], data_type=2)], 0)
    sel = pred.process(ctx, dim)
    return ops.HashTable(ctx, dim.columns[0], sel=sel, payload=dim.columns[2], load_factor=load_factor)


def small_fragment(fact, ht):
    return ops.Fragment(fact, joins=[(0, ht)], keys=[(abi.SRC_JOIN, 0)],
                        aggs=[(abi.AGG_SUM, abi.EXPR_MUL, FACT_MEASURE, FACT_PRICE)])


@dataclass
class StarTables:
    fact: ops.Table
    date: ops.Table      # key, year, month
    supplier: ops.Table  # key, region
    part: ops.Table      # key, category, brand
    customer: ops.Table  # key
    seed: int
    fact_rows: int
    extra: dict = field(default_factory=dict)


def star_dims(ctx, seed=SEED_BASE + 2):
    dev = f"cuda:{ctx.device}"
    nd = STAR_DIMS["date"]
    dkey = ctx.gen_sequence_i32(nd, 1)
    day = torch.arange(nd, dtype=torch.int32, device=dev)
    year = torch.clamp(1992 + day // 365, max=1998).to(torch.int32)
    month = torch.clamp((day % 365) // 31 + 1, max=12).to(torch.int32)
    date = ops.Table([dkey, year.contiguous(), month.contiguous()])
    ns = STAR_DIMS["supplier"]
    supplier = ops.Table([ctx.gen_sequence_i32(ns, 1), ctx.gen_uniform_i32(seed, SUPP_REGION, ns, 0, 4)])
    npart = STAR_DIMS["part"]
    part = ops.Table([ctx.gen_sequence_i32(npart, 1), ctx.gen_uniform_i32(seed, PART_CATEGORY, npart, 0, 24),
                      ctx.gen_uniform_i32(seed, PART_BRAND, npart, 0, 999)])
    customer = ops.Table([ctx.gen_sequence_i32(STAR_DIMS["customer"], 1)])
    return date, supplier, part, customer


def star_fact(ctx, fact_rows, seed=SEED_BASE + 2, fact_begin=0):
    g = lambda col, lo, hi: ctx.gen_uniform_i32(seed, col, fact_rows, lo, hi, row_begin=fact_begin)
    cols = [g(F_DATE, 1, STAR_DIMS["date"]), g(F_CUST, 1, STAR_DIMS["customer"]),
            g(F_SUPP, 1, STAR_DIMS["supplier"]), g(F_PART, 1, STAR_DIMS["part"]),
            g(F_QTY, 1, 50), g(F_PRICE, 100_000, 10_000_000), g(F_DISC, 0, 10)]
    return ops.Table(cols)


def star_tables(ctx, fact_rows=600_000_000, seed=SEED_BASE + 2):
    date, supplier, part, customer = star_dims(ctx, seed)
    return StarTables(star_fact(ctx, fact_rows, seed), date, supplier, part, customer, seed, fact_rows)


def star_builds(T, year_lo=1992, year_hi=1993, region=1, category=1, load_factor=0.5):
    """The star query's three build sides as ops.filtered_hash_tables takes them: (dimension,
    filter, key column, payload column, load factor) for date (-> year), supplier (existence) and
    part (-> brand). Built once, run every step (a prepared plan)."""
    P = ops.FilterPredicate
    return [(T.date, P([abi.make_pred(abi.FILTER_EQ_RANGE, 1, [year_lo, year_hi], data_type=2)], 0), 0, 1, load_factor),
            (T.supplier, P([abi.make_pred(abi.FILTER_EQ, 1, [region], data_type=2)], 0), 0, 1, load_factor),
            (T.part, P([abi.make_pred(abi.FILTER_EQ, 1, [category], data_type=2)], 0), 0, 2, load_factor)]


def star_hash_tables(ctx, T, year_lo=1992, year_hi=1993, region=1, category=1, load_factor=0.5,
                     wait_each=False, fused=True):
    """Filtered dimension hash tables: date (-> year), supplier (existence), part (-> brand).
    By default the three filters and builds run as one filtered build (ghive_hash_build_filtered:
    a pass per table filters and inserts, one launch fills and one reduces all three tables'
    statistics); fused=False runs each filter then its build, the filter handing its count to the
    build on the device (no host round trip until the tables are first used); wait_each=True
    reads every count on the host first."""
    P = ops.FilterPredicate
    preds = [(T.date, [abi.make_pred(abi.FILTER_EQ_RANGE, 1, [year_lo, year_hi], data_type=2)], 1),
             (T.supplier, [abi.make_pred(abi.FILTER_EQ, 1, [region], data_type=2)], 1),
             (T.part, [abi.make_pred(abi.FILTER_EQ, 1, [category], data_type=2)], 2)]
    if fused and not wait_each:
        return tuple(ops.filtered_hash_tables(ctx, star_builds(T, year_lo, year_hi, region, category, load_factor)))
    tables = []
    for dim, nodes, pay_col in preds:
        if wait_each:
            sel = P(nodes, 0).process(ctx, dim)
            tables.append(ops.HashTable(ctx, dim.columns[0], sel=sel, payload=dim.columns[pay_col],
                                        load_factor=load_factor))
        else:
            sel, cnt = P(nodes, 0).process_async(ctx, dim)
            tables.append(ops.HashTable(ctx, dim.columns[0], sel=sel, payload=dim.columns[pay_col],
                                        load_factor=load_factor, count=cnt))
    return tuple(tables)


def star_fragment(fact, hd, hs, hp):
    return ops.Fragment(fact, joins=[(F_DATE, hd), (F_SUPP, hs), (F_PART, hp)],
                        keys=[(abi.SRC_JOIN, 0), (abi.SRC_JOIN, 2)],
                        aggs=[(abi.AGG_SUM, abi.EXPR_MUL, F_PRICE, F_DISC)])


STAR_REFERENCED_BYTES_PER_ROW = 20  # fk_date, fk_supp, fk_part, price, discount (int32)
SMALL_REFERENCED_BYTES_PER_ROW = 12  # fk, measure, price (int32)

# ---- skewed join + sort + high-cardinality group-by -------------------------------------------
@dataclass
class SkewTables:
    build_keys: torch.Tensor   # unique int64 keys of ranks 1..B
    build_pay: torch.Tensor    # int64 payload = rank
    probe_keys: torch.Tensor   # Zipf(theta) over the build ranks, miss_frac absent keys
    tiebreak: torch.Tensor     # float32 U[0,1) per probe row
    seed: int


def skew_tables(ctx, build_rows=64_000_000, probe_rows=1_000_000_000, seed=SEED_BASE + 3, theta=1.0,
                miss_frac=0.1):
    bk = ctx.gen_rank_keys(seed, build_rows, rank_begin=1)
    bp = torch.arange(1, build_rows + 1, dtype=torch.int64, device=bk.device)
    pk = ctx.gen_zipf_keys(seed, 0, probe_rows, build_rows, theta, miss_frac)
    tb = ctx.gen_uniform_f32(seed, 1, probe_rows)
    return SkewTables(bk, bp, pk, tb, seed)


def key62_torch(seed, rank):
    """The generator's rank -> key bijection with torch int64 ops (wraps mod 2^64, masked to 62 bits)."""
    M = (1 << 62) - 1
    off = (seed * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
    x = (rank + (off & M)) & M
    x = x ^ (x >> 31)
    x = (x * _i64(0x7fb5d329728ea185)) & M
    x = x ^ (x >> 27)
    x = (x * _i64(0x81dadef4bc2dd44d)) & M
    x = x ^ (x >> 33)
    return x


def _i64(u):
    return u - (1 << 64) if u >= (1 << 63) else u


def highcard_tables(ctx, rows=1_000_000_000, groups=100_000_000, seed=SEED_BASE + 5):
    """GROUP BY key COUNT(*), SUM(x): keys uniform over `groups` 62-bit values, x ~ N(0,1)."""
    ranks = ctx.gen_uniform_i64(seed, 0, rows, 1, groups)
    keys = key62_torch(seed, ranks)
    del ranks
    vals = ctx.gen_normal_f64(seed, 1, rows)
    return keys, vals


# ---- batched form: 4096 map-task fragments of the small query shape -------------------------
BATCH_ATTR_SHIFT, BATCH_ATTR_BITS = 8, 3  # dimension payload = attr << 8 | category
BATCH_CAT_SHIFT, BATCH_CAT_BITS = 0, 5


def lognormal_rows(nfrags, seed, median=256_000, sigma=1.32, lo=1_000, hi=16_000_000):
    """Fragment row counts: lognormal(median, sigma) clipped to [lo, hi], rounded to multiples of
    4 rows (16-byte aligned fragment starts in shared column buffers)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    r = np.exp(np.log(median) + sigma * rng.standard_normal(nfrags))
    r = np.clip(np.round(r / 4) * 4, lo, hi).astype(np.int64)
    return r


@dataclass
class BatchTables:
    fk: torch.Tensor       # all fragments' fact columns, concatenated
    measure: torch.Tensor
    price: torch.Tensor
    rows: list             # per fragment
    starts: list           # per fragment (multiples of 4)
    attr_const: list       # per-fragment filter constant (dim.attr == c)
    dim: ops.Table         # key, attr, category, packed payload
    seed: int


def batch_tables(ctx, nfrags=4096, seed=SEED_BASE + 4, dim_rows=200_000):
    import numpy as np
    rows = lognormal_rows(nfrags, seed)
    starts = np.concatenate([[0], np.cumsum(rows)[:-1]]).astype(np.int64)
    total = int(rows.sum())
    fk = ctx.gen_uniform_i32(seed, FACT_FK, total, 1, dim_rows)
    measure = ctx.gen_uniform_i32(seed, FACT_MEASURE, total, 1, 100)
    price = ctx.gen_uniform_i32(seed, FACT_PRICE, total, 100_000, 10_000_000)
    consts = np.random.default_rng(seed + 1).integers(0, 7, nfrags).tolist()
    key = ctx.gen_sequence_i32(dim_rows, 1)
    attr = ctx.gen_uniform_i32(seed, DIM_ATTR, dim_rows, 0, 6)
    cat = ctx.gen_uniform_i32(seed, DIM_CATEGORY, dim_rows, 0, 24)
    packed = ((attr << BATCH_ATTR_SHIFT) | (cat << BATCH_CAT_SHIFT)).contiguous()
    return BatchTables(fk, measure, price, rows.tolist(), starts.tolist(), consts,
                       ops.Table([key, attr, cat, packed]), seed)


def batch_dim_table(ctx, B, load_factor=0.5):
    """The resident, unfiltered dimension table shared by every fragment (payload = packed)."""
    return ops.HashTable(ctx, B.dim.columns[0], payload=B.dim.columns[3], load_factor=load_factor)


def batch_items(B):
    items = []
    for s, r, c in zip(B.starts, B.rows, B.attr_const):
        items.append((B.fk[s:s + r], B.measure[s:s + r], B.price[s:s + r], c, c))
    return items


def batch_create(ctx, B, dim_ht, tile_rows=0, blocks_per_cu=0, probe_mode=0, use_graph=True):
    return ops.Batch(ctx, dim_ht, batch_items(B), BATCH_ATTR_SHIFT, BATCH_ATTR_BITS, BATCH_CAT_SHIFT,
                     BATCH_CAT_BITS, tile_rows=tile_rows, blocks_per_cu=blocks_per_cu,
                     probe_mode=probe_mode, use_graph=use_graph)


class PlainBatch:
    """The plain form of the same batch: one ghive_fragment_execute per fragment, with the
    dimension filtered and its hash table built once per distinct filter constant."""

    def __init__(self, ctx, B, load_factor=0.5):
        self.ctx, self.B = ctx, B
        self.tables = {}
        for c in sorted(set(B.attr_const)):
            self.tables[c] = small_dim_table(ctx, B.dim, attr_value=c, load_factor=load_factor)
        dev = B.fk.device
        self.outs = [torch.empty(64, dtype=torch.int32, device=dev), torch.empty(64, dtype=torch.int64, device=dev)]
        self.frags = []
        for s, r, c in zip(B.starts, B.rows, B.attr_const):
            fact = ops.Table([B.fk[s:s + r], B.measure[s:s + r], B.price[s:s + r]])
            self.frags.append(small_fragment(fact, self.tables[c]))

    def run(self, collect=False):
        res = []
        for f in self.frags:
            k, v = f.execute(self.ctx, capacity=64, outs=self.outs)
            if collect:
                res.append(dict(zip(k.cpu().tolist(), v.cpu().tolist())))
        return res


# ---- serving: a request mix over resident inputs ----------------------------------------------
@dataclass
class ServingInputs:
    small: SmallTables          # fact rows sliced by fragment requests (config-1 shape)
    dim_tables: dict            # attr value -> resident dimension hash table
    gb_keys: torch.Tensor       # int64 group-by keys, uniform over [0, 10M)
    gb_vals: torch.Tensor       # int64 values
    sort_keys: torch.Tensor     # int64 sort keys


SERVE_MIX = (("fragment", 0.60), ("groupby", 0.25), ("sort", 0.15))
SERVE_GB_DISTINCT = 10_000_000


def serving_inputs(ctx, frag_rows=16_000_000, gb_rows=16_000_000, sort_rows=256_000_000, seed=SEED_BASE + 5):
    """Resident inputs of the serving mix: a config-1 fact table with its dimension hash tables
    (one per filter constant, built once and shared by every request), group-by columns and a
    sort-key column; requests are prefixes or aligned slices of these."""
    S = small_tables(ctx, frag_rows, seed=seed)
    dims = {c: small_dim_table(ctx, S.dim, attr_value=c) for c in range(7)}
    k = ctx.gen_uniform_i64(seed, 20, gb_rows, 0, SERVE_GB_DISTINCT - 1)
    v = ctx.gen_uniform_i64(seed, 21, gb_rows, -10**9, 10**9)
    sk = ctx.gen_uniform_i64(seed, 22, sort_rows, -2**62, 2**62)
    return ServingInputs(S, dims, k, v, sk)


def serving_requests(nreq, seed, I):
    """nreq requests of the mix (kind, params): fragments of 1k..16M rows (lognormal, median 256k)
    at an aligned offset with one of the 7 filter constants; group-bys over 1k..16M rows with up
    to 10M distinct keys; sorts of 1M..256M int64 keys (lognormal, median 8M)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = rng.choice(len(SERVE_MIX), size=nreq, p=[p for _, p in SERVE_MIX])
    fact_rows = I.small.fact_rows
    out = []
    for i, k in enumerate(kinds):
        kind = SERVE_MIX[k][0]
        if kind == "fragment":
            rows = int(lognormal_rows(1, seed * 7919 + i, median=256_000, hi=fact_rows)[0])
            start = int(rng.integers(0, (fact_rows - rows) // 4 + 1)) * 4
            out.append(("fragment", dict(start=start, rows=rows, attr=int(rng.integers(0, 7)))))
        elif kind == "groupby":
            rows = int(lognormal_rows(1, seed * 7919 + i, median=1_000_000, hi=I.gb_keys.numel())[0])
            out.append(("groupby", dict(rows=rows)))
        else:
            rows = int(lognormal_rows(1, seed * 7919 + i, median=8_000_000, sigma=1.2, lo=1_000_000,
                                      hi=I.sort_keys.numel())[0])
            out.append(("sort", dict(rows=rows)))
    return out


def serving_submit(srv, I, kind, p):
    """Submits one request of the mix to a Server; returns (ticket, fact rows)."""
    if kind == "fragment":
        fact = ops.Table([c[p["start"]: p["start"] + p["rows"]] for c in I.small.fact.columns])
        return srv.submit_fragment(small_fragment(fact, I.dim_tables[p["attr"]]), capacity=64), p["rows"]
    if kind == "groupby":
        n = p["rows"]
        t = ops.Table([I.gb_keys[:n], I.gb_vals[:n]])
        aggs = [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)]
        return srv.submit_groupby(t, [0], aggs, capacity=min(n, SERVE_GB_DISTINCT)), n
    n = p["rows"]
    return srv.submit_sort(ops.Table([I.sort_keys[:n]]), [0], "+"), n


def serving_plain(ctx, I, kind, p):
    """The same request through the plain entry point (for identity checks)."""
    if kind == "fragment":
        fact = ops.Table([c[p["start"]: p["start"] + p["rows"]] for c in I.small.fact.columns])
        return small_fragment(fact, I.dim_tables[p["attr"]]).execute(ctx, capacity=64)
    if kind == "groupby":
        n = p["rows"]
        t = ops.Table([I.gb_keys[:n], I.gb_vals[:n]])
        aggs = [(abi.AGG_SUM, 1), (abi.AGG_CNT, abi.COUNT_STAR)]
        return ops.GroupByPredicate([0], aggs).gpu_execute(ctx, t, capacity=min(n, SERVE_GB_DISTINCT))
    n = p["rows"]
    return [ops.sort(ctx, ops.Table([I.sort_keys[:n]]), [0], "+")]
