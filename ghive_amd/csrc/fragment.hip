// Fused fragment: fact scan -> optional fact filter -> chain of dimension hash probes ->
// int64 expression -> hash GROUP BY, in one streaming pass over the fact columns. This is the
// CS1 map-task pipeline of GHive (TableScan -> Filter -> MapJoin* -> Select -> GroupBy ->
// ReduceSink), which the reference runs as separate Thrust calls with host round trips
// (FilterOperator.cu:199-225, SortMergeJoinPredicate.cu:8-461, SelectPredicate.cu:84-409,
// GroupByPredicate.cu:19-359).
//
// Each thread owns 4 consecutive rows per tile (one 16-byte load per referenced column).
// Dimension tables are probed with the Hive-style min/max pre-check first, most selective join
// first, and the surviving rows aggregate into an LDS-private group table that is merged into
// the global table once per workgroup.
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
#include "pred_host.hpp"

namespace ghive {

constexpr int kFragThreads = 256;
constexpr int kRowsPerThread = 4;
constexpr uint64_t kFragTile = (uint64_t)kFragThreads * kRowsPerThread;

__device__ __forceinline__ void load4_i32(const int32_t *p, uint64_t r0, uint64_t n, int32_t v[4]) {
  if (r0 + 3 < n) {
    int4 x = *reinterpret_cast<const int4 *>(p + r0);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = r0 + q < n ? p[r0 + q] : 0x7fffffff;
  }
}

template <bool kLate>
__global__ __launch_bounds__(kFragThreads) void k_fragment(FragDev F, GroupTableDev T, int lcap_log2,
                                                           unsigned long long *joined) {
  extern __shared__ uint64_t lds[];
  const bool use_lds = lcap_log2 >= 0;  // uniform: LDS pre-aggregation or global table only
  const int lstride = use_lds ? (1 << lcap_log2) + 1 : 0;
  if (use_lds) lds_table_init(lds, T, lcap_log2);
  __syncthreads();
  const uint64_t n = F.n;
  const uint64_t ntiles = (n + kFragTile - 1) / kFragTile;
  uint32_t njoined = 0;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t r0 = tile * kFragTile + (uint64_t)threadIdx.x * kRowsPerThread;
    uint32_t alive = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) alive |= (r0 + q < n ? 1u : 0u) << q;
    int32_t fk[kMaxFragJoins][4];
    int32_t av[kMaxFragAggs][4], bv[kMaxFragAggs][4];
    if (!kLate) {
      // issue every column load of the tile up front: many bytes in flight per wave
      for (int j = 0; j < F.njoins; ++j) load4_i32(F.joins[j].fk, r0, n, fk[j]);
      for (int a = 0; a < F.naggs; ++a) {
        if (F.aggs[a].expr != GHIVE_EXPR_ONE) load4_i32(F.aggs[a].a, r0, n, av[a]);
        if (F.aggs[a].expr == GHIVE_EXPR_MUL) load4_i32(F.aggs[a].b, r0, n, bv[a]);
      }
    }
    if (F.has_filter && alive) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (((alive >> q) & 1u) && !eval_pred(F.filter, r0 + q)) alive &= ~(1u << q);
    }
    int32_t pay[kMaxFragJoins][4];
    for (int j = 0; j < F.njoins; ++j) {
      const FragJoinDev &J = F.joins[j];
      if (kLate && alive) load4_i32(J.fk, r0, n, fk[j]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!((alive >> q) & 1u)) continue;
        const int32_t k = fk[j][q];
        bool hit = k != 0x7fffffff && k >= J.key_min && k <= J.key_max &&
                   probe_unique8(J.slots, J.mask, J.log2, k, pay[j][q]);
        if (!hit) alive &= ~(1u << q);
      }
    }
    if (!alive) continue;
    if (kLate) {
      for (int a = 0; a < F.naggs; ++a) {
        if (F.aggs[a].expr != GHIVE_EXPR_ONE) load4_i32(F.aggs[a].a, r0, n, av[a]);
        if (F.aggs[a].expr == GHIVE_EXPR_MUL) load4_i32(F.aggs[a].b, r0, n, bv[a]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (!((alive >> q) & 1u)) continue;
      ++njoined;
      uint64_t key = 0;
      for (int k = 0; k < F.nkeys; ++k) {
        int32_t kv = F.key_src[k] == GHIVE_SRC_JOIN ? pay[F.key_idx[k]][q] : F.key_col[k][r0 + q];
        key |= (uint64_t)(uint32_t)kv << (32 * k);
      }
      const int ls = use_lds ? lds_find_insert(lds, key, lcap_log2) : -1;
      int64_t g = -1;
      if (ls < 0) {
        g = gt_find_insert(T, key);
        if (g < 0) continue;
      }
      for (int a = 0; a < F.naggs; ++a) {
        const FragAggDev &E = F.aggs[a];
        AggVal v;
        v.neg = false;
        v.b = 0;
        if (E.expr == GHIVE_EXPR_ONE) {
          v.valid = true;
          v.a = 1;
        } else {
          int32_t x = av[a][q];
          v.valid = x != 0x7fffffff;
          int64_t val = x;
          if (E.expr == GHIVE_EXPR_MUL) {
            int32_t y = bv[a][q];
            v.valid = v.valid && y != 0x7fffffff;
            val = (int64_t)((uint64_t)(int64_t)x * (uint64_t)(int64_t)y);
          }
          v.a = (uint64_t)val;
        }
        const AggSpec &A = T.aggs[a];
        if (ls >= 0)
          acc_apply(A, [&](int w) { return lds + (size_t)(1 + A.word0 + w) * lstride + ls; }, v);
        else
          acc_apply(A, [&](int w) { return T.acc + (size_t)(A.word0 + w) * T.stride + g; }, v);
      }
    }
  }
  __syncthreads();
  if (use_lds) lds_table_flush(lds, T, lcap_log2);
  // rows that survived every join (for reporting)
  unsigned long long s = njoined;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
  if (lane_id() == 0 && s) atomicAdd(joined, s);
}

int fragment_prepare(ghive_ctx *ctx, const ghive_fragment *f, FragDev &F, GroupTableDev &T, KeySpec &K,
                     uint64_t &group_bound) {
  memset(&F, 0, sizeof F);
  memset(&T, 0, sizeof T);
  memset(&K, 0, sizeof K);
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
  // most selective join first (hit-rate estimate: entries / rows of the dimension)
  std::vector<uint32_t> order(f->njoins);
  std::iota(order.begin(), order.end(), 0u);
  for (uint32_t j = 0; j < f->njoins; ++j) {
    const ghive_hash_table *h = f->joins[j].table;
    if (!h) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: join %u has no table", j);
    if (h->layout != 8 || !h->unique)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED,
                       "fragment: join %u needs a unique-key INT table with an INT payload", j);
    GH_TRY(int_col(f->joins[j].fact_key_col, "join key"));
  }
  auto rate = [&](uint32_t j) {
    const ghive_hash_table *h = f->joins[j].table;
    return h->source_rows ? (double)h->entries / (double)h->source_rows : 1.0;
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rate(a) < rate(b); });
  std::vector<int> pos_of(f->njoins);
  F.njoins = (int)f->njoins;
  for (uint32_t i = 0; i < f->njoins; ++i) {
    const ghive_hash_table *h = f->joins[order[i]].table;
    FragJoinDev &J = F.joins[i];
    J.fk = static_cast<const int32_t *>(fact.columns[f->joins[order[i]].fact_key_col].data);
    J.slots = static_cast<const uint64_t *>(h->data);
    J.mask = h->slots - 1;
    J.log2 = h->log2_slots;
    // an empty table rejects every key through an empty range
    J.key_min = h->entries ? (int32_t)h->key_min : 1;
    J.key_max = h->entries ? (int32_t)h->key_max : 0;
    pos_of[order[i]] = (int)i;
  }
  group_bound = 1;
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
      if (h->payload_kind == 0)
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: key %u: join %u carries no attribute", k, fk.index);
      uint64_t range = h->entries ? (uint64_t)(h->pay_max - h->pay_min + 1) : 1;
      group_bound = group_bound > (1ull << 40) / (range ? range : 1) ? (1ull << 40) : group_bound * range;
    } else if (fk.source == GHIVE_SRC_FACT) {
      GH_TRY(int_col(fk.index, "group key"));
      F.key_col[k] = static_cast<const int32_t *>(fact.columns[fk.index].data);
      group_bound = 1ull << 40;
    } else {
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment: key %u source %d", k, fk.source);
    }
  }
  F.naggs = (int)f->naggs;
  T.naggs = (int)f->naggs;
  int word = 0;
  for (uint32_t a = 0; a < f->naggs; ++a) {
    const ghive_fragment_agg &E = f->aggs[a];
    FragAggDev &D = F.aggs[a];
    AggSpec &A = T.aggs[a];
    D.expr = E.expr;
    if (E.expr != GHIVE_EXPR_COL && E.expr != GHIVE_EXPR_MUL && E.expr != GHIVE_EXPR_ONE)
      return set_error(ctx, GHIVE_ERR_INVALID, "fragment: aggregate %u expression %d", a, E.expr);
    if (E.expr != GHIVE_EXPR_ONE) {
      GH_TRY(int_col(E.a, "aggregate"));
      D.a = static_cast<const int32_t *>(fact.columns[E.a].data);
    }
    if (E.expr == GHIVE_EXPR_MUL) {
      GH_TRY(int_col(E.b, "aggregate"));
      D.b = static_cast<const int32_t *>(fact.columns[E.b].data);
    }
    switch (E.type) {
      case GHIVE_AGG_SUM: A.kind = AK_SUMI; break;
      case GHIVE_AGG_CNT: A.kind = AK_CNT; break;
      case GHIVE_AGG_MIN: A.kind = AK_MINI; break;
      case GHIVE_AGG_MAX: A.kind = AK_MAXI; break;
      default:
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "fragment: aggregate %u type %d (SUM/COUNT/MIN/MAX)", a, E.type);
    }
    A.in_type = GHIVE_LONG;
    A.out_type = GHIVE_LONG;
    A.col = -1;
    A.word0 = (int16_t)word;
    word += agg_words(A.kind);
  }
  T.nwords = word;
  if (f->filter_nnodes) {
    GH_TRY(flatten_pred(ctx, &fact, f->filter, f->filter_nnodes, f->filter_root, F.filter));
    F.has_filter = 1;
  }
  return GHIVE_OK;
}

int fragment_launch(ghive_ctx *ctx, const FragDev &F, const GroupTableDev &T, uint64_t group_bound,
                    unsigned long long *d_joined, int late, uint32_t grid_override) {
  // LDS pre-aggregation only when the group bound fits the budget at load <= 1/2; otherwise
  // every surviving row goes to the global table (few rows survive a selective star join)
  uint64_t want_slots = next_pow2(2 * (group_bound < 32 ? 32 : group_bound));
  int lg = lds_log2_for(T.nwords, (int)ctx->tune.frag_lds_budget);
  if ((1ull << lg) >= want_slots) {
    while (lg > 5 && (1ull << (lg - 1)) >= want_slots) --lg;
  } else {
    lg = -1;
  }
  size_t lds_bytes = lg >= 0 ? (size_t)((1 << lg) + 1) * (1 + T.nwords) * 8 : 0;
  uint64_t tiles = (F.n + kFragTile - 1) / kFragTile;
  if (!tiles) return GHIVE_OK;
  int per_cu = 0;
  auto kern = late ? k_fragment<true> : k_fragment<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kFragThreads, lds_bytes) != hipSuccess || per_cu < 1)
    per_cu = 1;
  uint64_t cap = (uint64_t)ctx->cus * per_cu;
  uint32_t grid = (uint32_t)(tiles < cap ? tiles : cap);
  if (grid_override) grid = grid_override;
  else if (ctx->tune.frag_grid > 0) grid = (uint32_t)ctx->tune.frag_grid;
  if (late)
    k_fragment<true><<<grid, kFragThreads, lds_bytes, ctx->stream>>>(F, T, lg, d_joined);
  else
    k_fragment<false><<<grid, kFragThreads, lds_bytes, ctx->stream>>>(F, T, lg, d_joined);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int fragment_run(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols, uint64_t capacity,
                 uint64_t *out_groups, int late) {
  if (!out_groups) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: out_groups is NULL");
  FragDev F;
  GroupTableDev T;
  KeySpec K;
  uint64_t bound = 0;
  GH_TRY(fragment_prepare(ctx, frag, F, T, K, bound));
  if (capacity && !out_cols) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: out_cols is NULL");
  for (uint32_t c = 0; c < frag->nkeys + frag->naggs && capacity; ++c)
    if (!out_cols[c].data) return set_error(ctx, GHIVE_ERR_INVALID, "fragment: output column %u has no buffer", c);
  uint64_t table_groups = capacity < bound ? capacity : bound;
  DevBuf dkeys, dacc, dcnt, djoined;
  GH_TRY(group_table_alloc(ctx, T, table_groups ? table_groups : 1, dkeys, dacc, dcnt));
  T.capacity = capacity;
  GH_TRY(djoined.alloc(ctx, sizeof(unsigned long long)));
  GH_TRY(fill_u64(ctx, djoined.as<uint64_t>(), 0, 1));
  if (frag->nkeys == 0) GH_TRY(group_table_preinsert(ctx, T, 0ull));  // one group, even over no rows
  GH_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  GH_TRY(fragment_launch(ctx, F, T, bound, djoined.as<unsigned long long>(), late, 0));
  GH_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  OutSpec O;
  memset(&O, 0, sizeof O);
  for (uint32_t c = 0; c < frag->nkeys + frag->naggs; ++c) {
    O.col[c] = out_cols ? out_cols[c].data : nullptr;
    O.type[c] = c < frag->nkeys ? GHIVE_INT : GHIVE_LONG;
    if (out_cols) out_cols[c].type = O.type[c];
  }
  int st = group_table_emit(ctx, K, T, O, capacity, out_groups);
  GH_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  GH_HIP(ctx, hipEventSynchronize(ctx->ev[2]));
  float ms_main = 0, ms_fin = 0;
  (void)hipEventElapsedTime(&ms_main, ctx->ev[0], ctx->ev[1]);
  (void)hipEventElapsedTime(&ms_fin, ctx->ev[1], ctx->ev[2]);
  uint64_t joined = 0;
  GH_TRY(d2h(ctx, &joined, djoined.get(), sizeof joined));
  ctx->times.main_kernel_ms = ms_main;
  ctx->times.finalize_ms = ms_fin;
  ctx->times.total_ms = ms_main + ms_fin;
  ctx->times.rows_in = F.n;
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
  return fragment_run(ctx, frag, out_cols, capacity, out_groups, (int)ctx->tune.frag_late);
}

// Same operator with late materialisation: a column is read only for lanes that still hold a
// live row (selection before projection). Exposed for measurement beside the default form.
int ghive_fragment_execute_late(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                                uint64_t capacity, uint64_t *out_groups) {
  GH_CHECK_CTX(ctx);
  return fragment_run(ctx, frag, out_cols, capacity, out_groups, /*late=*/1);
}

}  // extern "C"
