// Fused fragment: fact scan -> optional fact filter -> chain of dimension probes -> int64
// expression -> hash GROUP BY, in one streaming pass over the fact columns. This is the CS1
// map-task pipeline of GHive (TableScan -> Filter -> MapJoin* -> Select -> GroupBy ->
// ReduceSink), which the reference runs as separate Thrust calls with host round trips
// (FilterOperator.cu:199-225, SortMergeJoinPredicate.cu:8-461, SelectPredicate.cu:84-409,
// GroupByPredicate.cu:19-359).
//
// Each thread owns 4 consecutive rows per tile (one 16-byte load per referenced column). Join
// and aggregate counts are template parameters, so every per-row array lives in registers.
// A dimension whose keys span a compact range is probed through its dense companion: one bit
// test in an LDS copy of its existence bitmap (or in the HBM/L2 bitmap when it does not fit),
// then one payload load for the survivors. Joins run cheapest-and-most-selective first (rank =
// probe cost / (1 - selectivity)), with Hive's min/max pre-check in front of every probe. The
// surviving rows aggregate into an LDS-private group table merged once per workgroup, or
// straight into the global table when the group bound does not fit LDS. For the star shape with
// such a group bound (and at most 2^20 group-key combinations) the pipelined kernel
// (k_fragment_pipe) addresses the groups directly and keeps the next tile's loads in flight.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "agg.hpp"
#include "common.hpp"
#include "device.hpp"
#include "fragment.hpp"
#include "groupby.hpp"
#include "hash.hpp"
#include "fragment_kernel.hpp"
#include "pred_host.hpp"

namespace ghive {


int fragment_plan(ghive_ctx *ctx, const ghive_fragment *f, int late, FragPlan &P) {
  memset(&P, 0, sizeof P);
  FragDev &F = P.F;
  GroupTableDev &T = P.T;
  KeySpec &K = P.K;
  if (!f) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: NULL descriptor");
  GH_TRY(check_table(ctx, &f->fact, "fragment fact"));
  const ghive_table &fact = f->fact;
  F.n = fact.row_num;
  if (f->njoins > (uint32_t)kMaxFragJoins)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: %u joins (max %d)", f->njoins, kMaxFragJoins);
  if (f->nkeys > 2) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: %u keys (max 2)", f->nkeys);
  if (f->naggs > (uint32_t)kMaxFragAggs || f->naggs < 1)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: %u aggregates (1..%d)", f->naggs, kMaxFragAggs);
  if ((f->njoins && !f->joins) || (f->nkeys && !f->keys) || !f->aggs)
    return set_error(ctx, GHIVE_ERR_INVALID, "fragment: NULL joins/keys/aggs");
  auto int_col = [&](uint32_t c, const char *what) -> int {
    if (c >= fact.ncols) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: %s column %u of %u", what, c, fact.ncols);
    if (fact.columns[c].type != GHIVE_INT)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: %s column %u must be INT", what, c);
    if (((uintptr_t)fact.columns[c].data & 15) != 0)
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment: %s column %u is not 16-byte aligned", what, c);
    return GHIVE_OK;
  };
  std::vector<int> need_payload(f->njoins, 0);
  for (uint32_t k = 0; k < f->nkeys; ++k)
    if (f->keys[k].source == GHIVE_SRC_JOIN && f->keys[k].index < f->njoins) need_payload[f->keys[k].index] = 1;
  // probe mode of every join: LDS bitmap while the budget lasts (smallest first), else HBM
  // bitmap for dense tables, else the hash table
  const uint64_t bits_budget_words = (uint64_t)std::min<int64_t>(ctx->tune.frag_lds_budget, 48 * 1024) / 4;
  std::vector<int> mode(f->njoins, kJoinHash);
  std::vector<uint32_t> by_size;
  std::vector<const ghive_hash_table *> tables(f->njoins);
  for (uint32_t j = 0; j < f->njoins; ++j) {
    const ghive_hash_table *h = f->joins[j].table;
    if (!h) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: join %u has no table", j);
    if (!h->owner) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: join %u's table: its context was shut down", j);
    tables[j] = h;
  }
  GH_TRY(hash_tables_ready(ctx, tables.data(), (int)f->njoins));  // one wait for every pending build
  for (uint32_t j = 0; j < f->njoins; ++j) {
    const ghive_hash_table *h = f->joins[j].table;
    if (h->layout != 8 || !h->unique)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED,
                       "fragment: join %u needs a unique-key INT table with an INT payload", j);
    if (need_payload[j] && h->payload_kind == 0)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: join %u carries no attribute", j);
    GH_TRY(int_col(f->joins[j].fact_key_col, "join key"));
    if (h->bitmap && ctx->tune.frag_dense && h->entries) {  // an empty table is probed as a hash table
      mode[j] = kJoinGlobalBits;
      by_size.push_back(j);
    }
  }
  std::stable_sort(by_size.begin(), by_size.end(),
                   [&](uint32_t a, uint32_t b) { return f->joins[a].table->range < f->joins[b].table->range; });
  uint64_t used_words = 0;
  for (uint32_t j : by_size) {
    uint64_t w = (f->joins[j].table->range + 31) / 32;
    if (ctx->tune.frag_lds_bits && used_words + w <= bits_budget_words) {
      mode[j] = kJoinLdsBits;
      used_words += w;
    }
  }
  // order: rank = probe cost / (1 - selectivity), selectivity estimated as entries / rows
  std::vector<uint32_t> order(f->njoins);
  std::iota(order.begin(), order.end(), 0u);
  auto rank = [&](uint32_t j) {
    const ghive_hash_table *h = f->joins[j].table;
    double sel = h->source_rows ? (double)h->entries / (double)h->source_rows : 1.0;
    double cost = mode[j] == kJoinLdsBits ? 1.0 : (mode[j] == kJoinGlobalBits ? 4.0 : 10.0);
    return cost / std::max(1.0 - sel, 0.01);
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rank(a) < rank(b); });
  std::vector<int> pos_of(f->njoins);
  uint32_t lds_off = 0;
  F.njoins = (int)f->njoins;
  for (uint32_t i = 0; i < f->njoins; ++i) {
    const uint32_t j = order[i];
    const ghive_hash_table *h = f->joins[j].table;
    FragJoinDev &J = F.joins[i];
    J.fk = static_cast<const int32_t *>(fact.columns[f->joins[j].fact_key_col].data);
    J.slots = static_cast<const uint64_t *>(h->data);
    J.bitmap = h->bitmap;
    J.dense = h->dense;
    J.mask = h->slots - 1;
    J.log2 = h->log2_slots;
    J.mode = mode[j];
    J.key_min = h->entries ? (int32_t)h->key_min : 1;  // an empty table rejects every key
    J.key_max = h->entries ? (int32_t)h->key_max : 0;
    // NULL (INT_MAX) is never inserted, so key_max < INT_MAX and the unsigned test below is exact;
    // an empty table gets span 2^32-1 and rejects every key in its (empty) hash probe
    J.span = (uint32_t)J.key_max - (uint32_t)J.key_min;
    J.nwords = h->bitmap ? (uint32_t)((h->range + 31) / 32) : 0;
    J.need_payload = need_payload[j];
    if (J.mode == kJoinLdsBits) {
      J.lds_word_off = lds_off;
      lds_off += J.nwords;
    }
    pos_of[j] = (int)i;
  }
  F.lds_bitmap_words = (lds_off + 1) & ~1u;
  P.group_bound = 1;
  F.nkeys = (int)f->nkeys;
  K.nkeys = (int)f->nkeys;
  for (uint32_t k = 0; k < f->nkeys; ++k) {
    const ghive_fragment_key &fk = f->keys[k];
    K.type[k] = GHIVE_INT;
    K.shift[k] = 32 * (int)k;
    F.key_src[k] = (int8_t)fk.source;
    if (fk.source == GHIVE_SRC_JOIN) {
      if (fk.index >= f->njoins) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: key %u uses join %u", k, fk.index);
      F.key_idx[k] = (int8_t)pos_of[fk.index];
      const ghive_hash_table *h = f->joins[fk.index].table;
      uint64_t range = h->entries ? (uint64_t)(h->pay_max - h->pay_min + 1) : 1;
      P.group_bound = P.group_bound > (1ull << 40) / (range ? range : 1) ? (1ull << 40) : P.group_bound * range;
    } else if (fk.source == GHIVE_SRC_FACT) {
      GH_TRY(int_col(fk.index, "group key"));
      F.key_col[k] = static_cast<const int32_t *>(fact.columns[fk.index].data);
      P.group_bound = 1ull << 40;
    } else {
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment: key %u source %d", k, fk.source);
    }
  }
  F.naggs = (int)f->naggs;
  T.naggs = (int)f->naggs;
  int word = 0;
  bool any_dbl = false;
  for (uint32_t a = 0; a < f->naggs; ++a) {
    const ghive_fragment_agg &E = f->aggs[a];
    FragAggDev &D = F.aggs[a];
    AggSpec &A = T.aggs[a];
    D.expr = E.expr;
    if (E.expr != GHIVE_EXPR_COL && E.expr != GHIVE_EXPR_MUL && E.expr != GHIVE_EXPR_ONE)
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment: aggregate %u expression %d", a, E.expr);
    // DOUBLE inputs (both, for a product) go to the general kernel, which reads them per surviving
    // row; INT inputs stream with the tile
    const bool dbl = E.expr != GHIVE_EXPR_ONE && E.a < fact.ncols && fact.columns[E.a].type == GHIVE_DOUBLE;
    if (dbl) {
      for (int i = 0; i < (E.expr == GHIVE_EXPR_MUL ? 2 : 1); ++i) {
        const uint32_t c = i ? E.b : E.a;
        if (c >= fact.ncols) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: aggregate column %u of %u", c, fact.ncols);
        if (fact.columns[c].type != GHIVE_DOUBLE)
          return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: aggregate %u mixes DOUBLE and non-DOUBLE inputs", a);
        if (((uintptr_t)fact.columns[c].data & 7) != 0)
          return set_error(ctx, GHIVE_ERR_INVALID, "fragment: DOUBLE column %u is not 8-byte aligned", c);
      }
      D.dbl = 1;
      D.a = static_cast<const int32_t *>(fact.columns[E.a].data);
      D.b = static_cast<const int32_t *>(fact.columns[E.expr == GHIVE_EXPR_MUL ? E.b : E.a].data);
      any_dbl = true;
    } else {
      if (E.expr != GHIVE_EXPR_ONE) {
        GH_TRY(int_col(E.a, "aggregate"));
        D.a = static_cast<const int32_t *>(fact.columns[E.a].data);
      }
      if (E.expr == GHIVE_EXPR_MUL) {
        GH_TRY(int_col(E.b, "aggregate"));
        D.b = static_cast<const int32_t *>(fact.columns[E.b].data);
      }
    }
    switch (E.type) {
      case GHIVE_AGG_SUM: A.kind = dbl ? AK_SUMD : AK_SUMI; break;
      case GHIVE_AGG_CNT: A.kind = AK_CNT; break;
      case GHIVE_AGG_MIN: A.kind = dbl ? AK_MIND : AK_MINI; break;
      case GHIVE_AGG_MAX: A.kind = dbl ? AK_MAXD : AK_MAXI; break;
      default:
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: aggregate %u type %d (SUM/COUNT/MIN/MAX)", a, E.type);
    }
    A.in_type = dbl ? GHIVE_DOUBLE : GHIVE_LONG;
    A.out_type = (dbl && A.kind != AK_CNT) ? GHIVE_DOUBLE : GHIVE_LONG;
    A.col = -1;
    A.word0 = (int16_t)word;
    word += agg_words(A.kind);
    if (A.kind == AK_SUMD) {
      // exact fixed-point sum: the LSB from a bound of |product| (the product of the columns'
      // largest magnitudes; rounding is monotonic) and the row count
      double ma = 0, mb = 1;
      GH_TRY(column_absmax(ctx, fact.columns[E.a].data, GHIVE_DOUBLE, nullptr, F.n, &ma));
      if (E.expr == GHIVE_EXPR_MUL) GH_TRY(column_absmax(ctx, fact.columns[E.b].data, GHIVE_DOUBLE, nullptr, F.n, &mb));
      GH_TRY(sum_scale(ctx, ma * mb, F.n, &A.scale_e));
    }
  }
  T.nwords = word;
  // unused aggregate inputs read a column the tile loads anyway, so that the dense-probe kernel
  // issues every load of a tile unconditionally (the values are ignored)
  for (int a = 0; a < F.naggs && F.njoins > 0; ++a) {
    if (!F.aggs[a].a) F.aggs[a].a = F.joins[0].fk;
    if (!F.aggs[a].b) F.aggs[a].b = F.joins[0].fk;
  }
  if (f->filter_nnodes) {
    GH_TRY(flatten_pred(ctx, &fact, f->filter, f->filter_nnodes, f->filter_root, F.filter));
    F.has_filter = 1;
  }
  // LDS group table when the bound fits what the bitmaps leave, at load <= 1/2
  const size_t bitmap_bytes = (size_t)F.lds_bitmap_words * 4;
  const int64_t left = ctx->tune.frag_lds_budget - (int64_t)bitmap_bytes;
  const uint64_t want_slots = next_pow2(2 * (P.group_bound < 32 ? 32 : P.group_bound));
  P.lcap_log2 = -1;
  if (left >= 8 * 1024) {
    int lg = lds_log2_for(T.nwords, (int)left);
    if ((1ull << lg) >= want_slots) {
      while (lg > 5 && (1ull << (lg - 1)) >= want_slots) --lg;
      P.lcap_log2 = lg;
    }
  }
  P.lds_bytes = bitmap_bytes + (P.lcap_log2 >= 0 ? (size_t)((1 << P.lcap_log2) + 1) * (1 + T.nwords) * 8 : 0);
  P.late = late;
  F.late = late ? 1 : 0;
  const int rows = ctx->tune.frag_rows == 8 ? 8 : 4;
  // the lean dense-probe kernel when the fragment has its shape (no fact filter, early reads, every
  // join through a bitmap, group keys from join attributes)
  bool dense_form = !ctx->tune.frag_generic && !F.has_filter && !late && F.njoins >= 1 && !any_dbl;
  for (int j = 0; j < F.njoins; ++j) dense_form = dense_form && F.joins[j].mode != kJoinHash;
  for (int k = 0; k < F.nkeys; ++k) dense_form = dense_form && F.key_src[k] == GHIVE_SRC_JOIN;
  P.dense_form = dense_form;
  // direct group slots (pipelined kernel): every key is a join attribute, the product of the
  // attributes' ranges is small, and no key packs to the all-ones pattern that marks an empty slot
  bool pipe_form = dense_form && ctx->tune.frag_pipe && rows == 4;
  if (pipe_form) {
    uint64_t size = 1;
    bool both_hold_minus1 = F.nkeys == 2;
    for (int k = F.nkeys - 1; k >= 0; --k) {
      const ghive_hash_table *h = f->joins[f->keys[k].index].table;
      const uint64_t range = (uint64_t)(h->pay_max - h->pay_min) + 1;
      F.dmin[k] = (int32_t)h->pay_min;
      F.dmul[k] = (uint32_t)size;
      both_hold_minus1 = both_hold_minus1 && h->pay_min <= -1 && h->pay_max >= -1;
      if (range > kMaxDirectGroups || size * range > kMaxDirectGroups) {
        pipe_form = false;
        break;
      }
      size *= range;
    }
    if (both_hold_minus1) pipe_form = false;
    if (P.lcap_log2 >= 0) pipe_form = false;  // group domains that fit LDS aggregate there
    F.dsize = (uint32_t)size;
  }
  F.direct = pipe_form ? 1 : 0;
  P.pipe_form = pipe_form;
  int nh = 0;
  if (pipe_form) {  // the pipelined kernel probes the LDS bitmaps first: order the joins so
    FragJoinDev sorted[kMaxFragJoins];
    int newpos[kMaxFragJoins], c = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j < F.njoins; ++j)
        if ((F.joins[j].mode == kJoinLdsBits) == (pass == 0)) {
          newpos[j] = c;
          sorted[c++] = F.joins[j];
        }
    for (int j = 0; j < F.njoins; ++j) {
      F.joins[j] = sorted[j];
      nh += sorted[j].mode != kJoinLdsBits;
    }
    for (int k = 0; k < F.nkeys; ++k) F.key_idx[k] = (int8_t)newpos[F.key_idx[k]];
  }
  if (pipe_form) {
    // per-wave survivor queues in the LDS the bitmaps leave over at their occupancy (LDS is
    // allocated in 512-byte granules); at least 16 entries each
    const size_t gran = 512, lds_max = 160 * 1024, waves = kFragThreads / 64;
    const size_t entry = 4 * (2 + 2 * (size_t)F.naggs), base = P.lds_bytes;
    const size_t base_up = std::max<size_t>((base + gran - 1) / gran * gran, gran);
    const size_t blocks = std::min<size_t>(8, lds_max / base_up);
    const size_t budget = lds_max / blocks / gran * gran;
    size_t q = budget > base ? (budget - base) / (waves * entry) : 0;
    q = std::min<size_t>(std::max<size_t>(q, 16), 512);
    F.queue_cap = (uint32_t)q;
    P.lds_bytes = base + waves * q * entry;
  }
  auto pick_pipe = [](int nj, int nhb, int na) {
    switch (nj) {
      case 1: return pick_pipe_j1(nhb, na);
      case 2: return pick_pipe_j2(nhb, na);
      case 3: return pick_pipe_j3(nhb, na);
      default: return pick_pipe_j4(nhb, na);
    }
  };
  FragKernel kfn = pipe_form ? pick_pipe(F.njoins, nh, F.naggs)
                 : dense_form ? (rows == 8 ? pick_dense_r8(F.njoins, F.naggs) : pick_dense_r4(F.njoins, F.naggs))
                              : (rows == 8 ? pick_kernel_r8(F.njoins, F.naggs) : pick_kernel_r4(F.njoins, F.naggs));
  P.kernel = reinterpret_cast<const void *>(kfn);
  P.threads = pipe_form ? (uint32_t)kFragThreads : (uint32_t)ctx->tune.frag_block;
  P.rows = rows;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, (int)P.threads, P.lds_bytes) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  const uint64_t tiles = (F.n + (uint64_t)P.threads * rows - 1) / ((uint64_t)P.threads * rows);
  const uint64_t cap = (uint64_t)ctx->cus * per_cu;
  P.grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, cap));
  if (ctx->tune.frag_grid > 0) P.grid = (uint32_t)ctx->tune.frag_grid;
  return GHIVE_OK;
}

int fragment_launch(ghive_ctx *ctx, const FragPlan &P, unsigned long long *d_joined) {
  if (!P.F.n) return GHIVE_OK;
  FragDev F = P.F;
  GroupTableDev T = P.T;
  int lg = P.lcap_log2;
  unsigned long long *j = d_joined;
  void *args[] = {&F, &T, &lg, &j};
  GH_HIP(ctx, hipLaunchKernel(P.kernel, dim3(P.grid), dim3(P.threads), args, P.lds_bytes, ctx->stream));
  return GHIVE_OK;
}

namespace {

// Device memory of one fragment run: from the context's memory manager, or carved in order out of
// a caller-provided region (the library's static scratch array, for the entry point that
// allocates nothing).
struct FragMem {
  ghive_ctx *ctx;
  char *base = nullptr;  // non-null: carve from [base, base + cap)
  uint64_t cap = 0, used = 0;
  std::vector<DevBuf> bufs;
  int get(uint64_t bytes, void **p) {
    if (base) {
      const uint64_t off = (used + 255) & ~255ull;
      if (off + bytes > cap)
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED,
                         "fragment (no allocation): needs more than the %llu bytes of its scratch array "
                         "(group table of %llu slots); use ghive_fragment_execute",
                         (unsigned long long)cap, (unsigned long long)bytes);
      *p = base + off;
      used = off + bytes;
      return GHIVE_OK;
    }
    bufs.emplace_back();
    GH_TRY(bufs.back().alloc(ctx, bytes));
    *p = bufs.back().get();
    return GHIVE_OK;
  }
};

}  // namespace

int fragment_run(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols, uint64_t capacity,
                 uint64_t *out_groups, int late, FragMem &mem) {
  if (!out_groups) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: out_groups is NULL");
  FragPlan P;
  GH_TRY(fragment_plan(ctx, frag, late, P));
  if (capacity && !out_cols) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: out_cols is NULL");
  for (uint32_t c = 0; c < frag->nkeys + frag->naggs && capacity; ++c)
    if (!out_cols[c].data) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: output column %u has no buffer", c);
  uint64_t table_groups = capacity < P.group_bound ? capacity : P.group_bound;
  if (P.pipe_form) table_groups = P.F.dsize;  // the slots are addressed by key, whatever the capacity
  uint64_t key_bytes = 0, acc_bytes = 0;
  group_table_size(P.T, table_groups ? table_groups : 1, &key_bytes, &acc_bytes);
  void *dkeys = nullptr, *dacc = nullptr, *dcnt = nullptr, *djoined = nullptr;
  GH_TRY(mem.get(key_bytes, &dkeys));
  GH_TRY(mem.get(acc_bytes, &dacc));
  GH_TRY(mem.get(3 * sizeof(unsigned long long), &dcnt));  // entries, overflow, rows joined
  djoined = static_cast<unsigned long long *>(dcnt) + 2;
  GH_TRY(group_table_init(ctx, P.T, static_cast<uint64_t *>(dkeys), static_cast<uint64_t *>(dacc),
                          static_cast<unsigned long long *>(dcnt), 3));  // + rows joined
  const uint64_t tile_rows = (uint64_t)P.threads * (uint64_t)P.rows;
  if (P.F.n % tile_rows) {
    void *stage = nullptr;
    GH_TRY(mem.get((uint64_t)(P.F.njoins + 2 * P.F.naggs) * tile_rows * sizeof(int32_t), &stage));
    P.F.stage = static_cast<int32_t *>(stage);
  }
  P.T.capacity = capacity;
  if (frag->nkeys == 0) {  // one group, even over no rows
    if (P.pipe_form) GH_TRY(fill_u64(ctx, P.T.keys, 0ull, 1));  // direct slot 0 holds key 0
    else GH_TRY(group_table_preinsert(ctx, P.T, 0ull));
  }
  GH_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  GH_TRY(fragment_launch(ctx, P, static_cast<unsigned long long *>(djoined)));
  GH_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  OutSpec O;
  memset(&O, 0, sizeof O);
  for (uint32_t c = 0; c < frag->nkeys + frag->naggs; ++c) {
    O.col[c] = out_cols ? out_cols[c].data : nullptr;
    O.type[c] = c < frag->nkeys ? GHIVE_INT : P.T.aggs[c - frag->nkeys].out_type;
    if (out_cols) out_cols[c].type = O.type[c];
  }
  uint64_t joined = 0;
  int st = group_table_emit(ctx, P.K, P.T, O, capacity, out_groups, ctx->ev[2], static_cast<uint64_t *>(djoined),
                            &joined);
  if (st != GHIVE_OK && st != GHIVE_ERR_CAPACITY) return st;
  float ms_main = 0, ms_fin = 0;
  (void)hipEventElapsedTime(&ms_main, ctx->ev[0], ctx->ev[1]);
  (void)hipEventElapsedTime(&ms_fin, ctx->ev[1], ctx->ev[2]);
  ctx->times.main_kernel_ms = ms_main;
  ctx->times.finalize_ms = ms_fin;
  ctx->times.total_ms = ms_main + ms_fin;
  ctx->times.rows_in = P.F.n;
  ctx->times.rows_joined = joined;
  for (uint32_t c = 0; c < frag->nkeys + frag->naggs && out_cols; ++c) out_cols[c].row_num = st == GHIVE_OK ? *out_groups : 0;
  return st;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_fragment_execute(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                           uint64_t capacity, uint64_t *out_groups) {
  GH_CHECK_CTX(ctx);
  FragMem mem{ctx};
  return fragment_run(ctx, frag, out_cols, capacity, out_groups, (int)ctx->tune.frag_late, mem);
}

int ghive_fragment_execute_late(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                                uint64_t capacity, uint64_t *out_groups) {
  GH_CHECK_CTX(ctx);
  FragMem mem{ctx};
  return fragment_run(ctx, frag, out_cols, capacity, out_groups, /*late=*/1, mem);
}

int ghive_fragment_execute_noalloc(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                                   uint64_t capacity, uint64_t *out_groups) {
  GH_CHECK_CTX(ctx);
  if (!ctx->scratch) return set_error(ctx, GHIVE_ERR_HIP, "fragment (no allocation): no scratch array on this device");
  if (!frag) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: NULL descriptor");
  // finishing a deferred build allocates its dense companion: the tables must be finished already
  for (uint32_t j = 0; j < frag->njoins && frag->joins; ++j) {
    const ghive_hash_table *h = frag->joins[j].table;
    if (h && h->pending.load())
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment (no allocation): join %u's table is still being built "
                       "(build it with ghive_hash_build, or use it once first)", j);
  }
  FragMem mem{ctx};
  mem.base = static_cast<char *>(ctx->scratch);
  mem.cap = kNoAllocScratchBytes;
  return fragment_run(ctx, frag, out_cols, capacity, out_groups, (int)ctx->tune.frag_late, mem);
}

}  // extern "C"
