// The output table of a join from its result_offsets: BaseJoinPredicate::generate_result
// (gpu_impl/src/Operator/BaseJoinPredicate.cu:4-97) as one gather kernel per output column.
#include "common.hpp"

namespace ghive {

namespace {

constexpr int kThreads = 256;

// out[i] = in[offsets[i * width + w]], or `null_value` where the offset is negative (the NULL
// side of a LEFT OUTER join).
template <class T>
__global__ __launch_bounds__(kThreads) void k_gather_offsets(const T *in, const int32_t *offsets, uint32_t width,
                                                             uint32_t w, uint64_t n, T null_value, T *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int32_t r = offsets[i * width + w];
    out[i] = r < 0 ? null_value : in[r];
  }
}

// STRING: the (start,end) pair of each row; (-1,-1) for NULL.
struct StrPair {  // 4-byte aligned, like the int32 pairs it copies
  int32_t start, end;
};

__global__ __launch_bounds__(kThreads) void k_gather_pairs(const StrPair *in, const int32_t *offsets, uint32_t width,
                                                           uint32_t w, uint64_t n, StrPair *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int32_t r = offsets[i * width + w];
    out[i] = r < 0 ? StrPair{-1, -1} : in[r];
  }
}

uint32_t grid_of(ghive_ctx *ctx, uint64_t n) {
  const uint64_t g = (n + kThreads - 1) / kThreads, cap = (uint64_t)ctx->cus * 16;
  return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_join_result(ghive_ctx *ctx, const ghive_table *tables, uint32_t ntables, const ghive_join_cond *conds,
                      uint32_t nconds, const uint32_t *const *maintain_cols, const uint32_t *maintain_counts,
                      const int32_t *offsets, uint64_t cardinality, ghive_column *out_cols,
                      uint32_t out_capacity_cols, uint32_t *out_ncols) {
  GH_CHECK_CTX(ctx);
  if (!tables || !ntables || (nconds && !conds) || !maintain_counts || !out_ncols)
    return set_error(ctx, GHIVE_ERR_INVALID, "join_result: NULL argument");
  if (nconds != ntables - 1)
    return set_error(ctx, GHIVE_ERR_INVALID, "join_result: %u tables need %u conditions", ntables, ntables - 1);
  // result tables: the non-semi inputs in table order, as ghive_join lays out result_offsets
  std::vector<int> semi(ntables, 0);
  for (uint32_t c = 0; c < nconds; ++c) {
    if (conds[c].right >= ntables) return set_error(ctx, GHIVE_ERR_INVALID, "join_result: condition %u", c);
    semi[conds[c].right] = conds[c].type == GHIVE_LEFT_SEMI_JOIN;
  }
  std::vector<uint32_t> result_tables;
  for (uint32_t t = 0; t < ntables; ++t)
    if (!semi[t]) result_tables.push_back(t);
  const uint32_t width = (uint32_t)result_tables.size();
  // the columns to gather, in output order
  struct Out {
    uint32_t table, col, w;
  };
  std::vector<Out> plan;
  for (uint32_t w = 0; w < width; ++w) {
    const uint32_t t = result_tables[w];
    const uint32_t cnt = maintain_counts[t];
    if (cnt && (!maintain_cols || !maintain_cols[t]))
      return set_error(ctx, GHIVE_ERR_INVALID, "join_result: maintain_cols[%u] is NULL", t);
    if (cnt == 0 && w == 0) plan.push_back({t, 0, w});  // the reference keeps column 0 of the first
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t c = maintain_cols[t][k];
      if (c >= tables[t].ncols) return set_error(ctx, GHIVE_ERR_INVALID, "join_result: table %u has no column %u", t, c);
      plan.push_back({t, c, w});
    }
  }
  *out_ncols = (uint32_t)plan.size();
  if (plan.size() > out_capacity_cols)
    return set_error(ctx, GHIVE_ERR_CAPACITY, "join_result: %zu output columns, room for %u", plan.size(),
                     out_capacity_cols);
  if (cardinality && !offsets) return set_error(ctx, GHIVE_ERR_INVALID, "join_result: offsets is NULL");
  if (cardinality > kMaxRows) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "join_result: too many rows");
  for (size_t o = 0; o < plan.size(); ++o) {
    const ghive_column &in = tables[plan[o].table].columns[plan[o].col];
    ghive_column &out = out_cols[o];
    if (cardinality && (in.type == GHIVE_STRING ? !out.aux : !out.data))
      return set_error(ctx, GHIVE_ERR_INVALID, "join_result: output column %zu has no buffer", o);
    if (cardinality && tables[plan[o].table].row_num && !in.data)
      return set_error(ctx, GHIVE_ERR_INVALID, "join_result: table %u column %u has no data", plan[o].table,
                       plan[o].col);
  }
  for (size_t o = 0; o < plan.size(); ++o) {
    const ghive_column &in = tables[plan[o].table].columns[plan[o].col];
    ghive_column &out = out_cols[o];
    out.type = in.type;
    out.row_num = cardinality;
    out.char_size = in.type == GHIVE_STRING ? in.char_size : 0;
    if (!cardinality) continue;
    const uint32_t g = grid_of(ctx, cardinality), w = plan[o].w;
    switch (in.type) {
      case GHIVE_LONG:
        k_gather_offsets<int64_t><<<g, kThreads, 0, ctx->stream>>>(
            static_cast<const int64_t *>(in.data), offsets, width, w, cardinality, INT64_MAX,
            static_cast<int64_t *>(out.data));
        break;
      case GHIVE_DOUBLE:  // NULL = NaN; moved as its bit pattern
        k_gather_offsets<uint64_t><<<g, kThreads, 0, ctx->stream>>>(
            static_cast<const uint64_t *>(in.data), offsets, width, w, cardinality, 0x7ff8000000000000ull,
            static_cast<uint64_t *>(out.data));
        break;
      case GHIVE_INT:
        k_gather_offsets<int32_t><<<g, kThreads, 0, ctx->stream>>>(
            static_cast<const int32_t *>(in.data), offsets, width, w, cardinality, INT32_MAX,
            static_cast<int32_t *>(out.data));
        break;
      case GHIVE_FLOAT:
        k_gather_offsets<uint32_t><<<g, kThreads, 0, ctx->stream>>>(
            static_cast<const uint32_t *>(in.data), offsets, width, w, cardinality, 0x7fc00000u,
            static_cast<uint32_t *>(out.data));
        break;
      case GHIVE_STRING:  // the blob is shared; only the (start,end) pairs are gathered
        out.data = in.data;
        k_gather_pairs<<<g, kThreads, 0, ctx->stream>>>(static_cast<const StrPair *>(in.aux), offsets, width, w,
                                                        cardinality, static_cast<StrPair *>(out.aux));
        break;
      default:
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "join_result: column type %d", in.type);
    }
    GH_LAUNCH(ctx);
  }
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

}  // extern "C"
