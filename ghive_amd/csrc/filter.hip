// Filter with order-preserving stream compaction (FilterPredicate::process + clean_data,
// gpu_impl/src/Operator/FilterPredicate.cu:434-571, FilterOperator.cu:36-93).
//
// Pass 1 evaluates the whole predicate tree per row in registers (one coalesced read of each
// referenced column) and stores one 64-bit wave ballot per 64 rows plus a popcount per 4096-row
// tile. Pass 2 scans the tile counts. Pass 3 re-reads only the ballots (n/8 bytes) and writes the
// selected row ids in ascending order. No decoupled look-back: nothing waits on another
// workgroup.
#include "common.hpp"
#include "device.hpp"
#include "filter.hpp"
#include "pred_host.hpp"
#include "scan.hpp"

namespace ghive {

typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int kFilterThreads = 256;
constexpr int kFilterRowsPerThread = 16;
constexpr uint64_t kFilterTile = (uint64_t)kFilterThreads * kFilterRowsPerThread;  // 4096
constexpr uint64_t kWordsPerTile = kFilterTile / 64;                               // 64

// Tiles first_tile + blockIdx.x: one bit per row in masks (row order), passing rows per tile.
__global__ __launch_bounds__(kFilterThreads) void k_filter_mask(DevPred P, uint64_t n, uint64_t first_tile,
                                                                uint64_t *masks, uint32_t *tile_counts) {
  __shared__ uint32_t lds[kFilterThreads / kWave];
  const uint64_t tile = first_tile + blockIdx.x;
  const uint64_t tile0 = tile * kFilterTile;
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t cnt = 0;
#pragma unroll 4
  for (int j = 0; j < kFilterRowsPerThread; ++j) {
    uint64_t r = tile0 + (uint64_t)j * kFilterThreads + threadIdx.x;
    bool pass = r < n && eval_pred(P, r);
    uint64_t m = ballot(pass);
    uint64_t word = (tile0 >> 6) + (uint64_t)j * (kFilterThreads / kWave) + wave;
    if (lane_id() == 0 && (word << 6) < n) masks[word] = m;
    cnt += pass;
  }
  uint32_t total;
  (void)block_exclusive_scan_256(cnt, lds, total);
  if (threadIdx.x == 0) tile_counts[tile] = total;
}

// The same masks when every leaf compares a 4-byte INT column with literals (AND / OR inside):
// each lane reads 4 consecutive rows with one 16-byte load per leaf column and the 4 row groups of
// its tile together, then evaluates the nodes on the 16 values in registers. A lane's 4 result
// bits are a nibble of the row-order mask word that 16 consecutive lanes share. The table's last
// tile may be partial: there a lane whose 4 rows run past n loads the rows that exist one by one
// (the rest read as NULL, which no comparison passes), and mask words past the table stay unwritten.
constexpr int kFilterGroups = kFilterTile / (kFilterThreads * 4);  // 4 groups of 1024 rows
__global__ __launch_bounds__(kFilterThreads) void k_filter_mask_i32(DevPred P, uint64_t n, uint64_t *masks,
                                                                    uint32_t *tile_counts) {
  __shared__ uint32_t lds[kFilterThreads / kWave];
  const uint64_t tile0 = (uint64_t)blockIdx.x * kFilterTile;
  uint32_t bits[kFilterGroups][4];
#pragma unroll
  for (int g = 0; g < kFilterGroups; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) bits[g][q] = 0;
  for (int i = 0; i < P.nnodes; ++i) {
    const DevPredNode &nd = P.nodes[i];
    if (nd.op == kLeafInt) {
      const int32_t *col = static_cast<const int32_t *>(P.col_data[nd.a]) + tile0 + 4 * threadIdx.x;
      v4i x[kFilterGroups];
#pragma unroll
      for (int g = 0; g < kFilterGroups; ++g) {
        const uint64_t r = tile0 + (uint64_t)g * kFilterThreads * 4 + 4 * threadIdx.x;
        if (r + 4 <= n) {
          x[g] = *reinterpret_cast<const v4i *>(col + (uint64_t)g * kFilterThreads * 4);
        } else {  // the table's partial last tile
          const int32_t *c = col + (uint64_t)g * kFilterThreads * 4;
          x[g].x = r < n ? c[0] : 0x7fffffff;
          x[g].y = r + 1 < n ? c[1] : 0x7fffffff;
          x[g].z = r + 2 < n ? c[2] : 0x7fffffff;
          x[g].w = 0x7fffffff;
        }
      }
#pragma unroll
      for (int g = 0; g < kFilterGroups; ++g) {
        const int32_t v[4] = {x[g].x, x[g].y, x[g].z, x[g].w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          bits[g][q] |= (uint32_t)(!is_null_i32(v[q]) && cmp_mode<int64_t>(nd.mode, v[q], nd.ia, nd.ib)) << i;
      }
    } else {  // AND / OR
      const bool is_and = nd.op == kAnd;
#pragma unroll
      for (int g = 0; g < kFilterGroups; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t a = (bits[g][q] >> nd.a) & 1u, b = (bits[g][q] >> nd.b) & 1u;
          bits[g][q] |= (is_and ? (a & b) : (a | b)) << i;
        }
    }
  }
  const int top = P.nnodes - 1;
  uint32_t cnt = 0;
#pragma unroll
  for (int g = 0; g < kFilterGroups; ++g) {
    uint32_t nib = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) nib |= ((bits[g][q] >> top) & 1u) << q;
    cnt += (uint32_t)__builtin_popcount(nib);
    uint64_t word = (uint64_t)nib << (4 * (threadIdx.x & 15));
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) word |= __shfl_xor(word, d, kWave);
    const uint64_t wi = (tile0 >> 6) + (uint64_t)g * (kFilterThreads / 16) + threadIdx.x / 16;
    if ((threadIdx.x & 15) == 0 && (wi << 6) < n) masks[wi] = word;
  }
  uint32_t total;
  (void)block_exclusive_scan_256(cnt, lds, total);
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = total;
}

// Tile offsets: tile_offs (the scanned counts), or -- self_scan, for tables of at most
// kSelfScanTiles tiles -- each workgroup sums the counts of the tiles before it itself (a few KiB
// from L2), and the last one writes the total to *total: no scan launches.
constexpr uint32_t kSelfScanTiles = 4096;
__global__ __launch_bounds__(kFilterThreads) void k_filter_write(const uint64_t *masks, uint64_t nwords,
                                                                 const uint32_t *tile_offs, int self_scan,
                                                                 uint64_t *total, uint32_t *out) {
  __shared__ uint32_t word_pre[kWordsPerTile];
  __shared__ uint32_t red[kFilterThreads / kWave];
  const uint64_t w0 = (uint64_t)blockIdx.x * kWordsPerTile;
  if (threadIdx.x < kWordsPerTile) {
    uint64_t w = w0 + threadIdx.x;
    uint32_t c = w < nwords ? (uint32_t)__popcll(masks[w]) : 0u;
    uint32_t inc = wave_inclusive_scan(c);  // threads 0..63 are exactly wave 0
    word_pre[threadIdx.x] = inc - c;
  }
  uint32_t base;
  if (self_scan) {
    uint32_t part = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kFilterThreads) part += tile_offs[t];
    uint32_t sum;
    (void)block_exclusive_scan_256(part, red, sum);  // synchronises the workgroup
    base = sum;
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = (uint64_t)sum + tile_offs[blockIdx.x];
  } else {
    __syncthreads();
    base = tile_offs[blockIdx.x];
  }
  const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
  for (uint32_t q = wave; q < kWordsPerTile; q += kFilterThreads / kWave) {
    uint64_t w = w0 + q;
    if (w >= nwords) break;
    uint64_t m = masks[w];
    if ((m >> lane) & 1ull) {
      uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      out[base + word_pre[q] + below] = (uint32_t)((w << 6) + lane);
    }
  }
}

struct StrPair {  // a STRING row's (start,end); 4-byte aligned like the int32 pairs it reads
  int32_t start, end;
};

template <class T>
__global__ void k_gather(const T *in, const uint32_t *sel, uint64_t n, T *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = in[sel[i]];
}

// out_count (host, after a wait) or, when count_dev is set, the count left on the device with no
// wait (an asynchronous consumer reads it there).
static int filter_run(ghive_ctx *ctx, const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                      uint32_t root, uint32_t *out_sel, uint64_t *out_count, uint64_t *count_dev) {
  DevPred P;
  GH_TRY(flatten_pred(ctx, t, nodes, nnodes, root, P));
  const uint64_t n = t->row_num;
  if (out_count) *out_count = 0;
  if (n == 0) {
    if (count_dev) GH_TRY(fill_u64(ctx, count_dev, 0, 1));
    return GHIVE_OK;
  }
  const uint32_t ntiles = ceil_div(n, kFilterTile);
  const uint64_t nwords = (n + 63) / 64;
  DevBuf masks, counts;
  GH_TRY(masks.alloc(ctx, sizeof(uint64_t) * nwords));
  GH_TRY(counts.alloc(ctx, sizeof(uint32_t) * ntiles));
  // every tile through the INT form when every leaf compares an aligned INT column, every other
  // predicate through the general one
  bool int_form = P.nnodes > 0;
  for (int i = 0; i < P.nnodes; ++i) {
    const DevPredNode &nd = P.nodes[i];
    if (nd.op == kAnd || nd.op == kOr) continue;
    int_form = int_form && nd.op == kLeafInt && P.col_type[nd.a] == GHIVE_INT &&
               ((uintptr_t)P.col_data[nd.a] & 15) == 0;
  }
  const uint64_t full = int_form ? ntiles : 0;  // the INT form takes the partial last tile too
  if (full) {
    k_filter_mask_i32<<<(uint32_t)full, kFilterThreads, 0, ctx->stream>>>(P, n, masks.as<uint64_t>(),
                                                                         counts.as<uint32_t>());
    GH_LAUNCH(ctx);
  }
  if (full < ntiles) {
    k_filter_mask<<<(uint32_t)(ntiles - full), kFilterThreads, 0, ctx->stream>>>(P, n, full, masks.as<uint64_t>(),
                                                                              counts.as<uint32_t>());
    GH_LAUNCH(ctx);
  }
  // the asynchronous form takes the count where the scan leaves its total, with no copy after
  uint64_t *total = count_dev ? count_dev : ctx->d_counters;
  const int self_scan = ntiles <= kSelfScanTiles;
  if (!self_scan) GH_TRY(exclusive_scan_u32(ctx, counts.as<uint32_t>(), counts.as<uint32_t>(), ntiles, total));
  k_filter_write<<<ntiles, kFilterThreads, 0, ctx->stream>>>(masks.as<uint64_t>(), nwords, counts.as<uint32_t>(),
                                                             self_scan, total, out_sel);
  GH_LAUNCH(ctx);
  if (count_dev) return GHIVE_OK;
  GH_TRY(d2h(ctx, out_count, ctx->d_counters, sizeof(uint64_t)));
  return GHIVE_OK;
}

int filter_select(ghive_ctx *ctx, const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                  uint32_t root, uint32_t *out_sel, uint64_t *out_count) {
  return filter_run(ctx, t, nodes, nnodes, root, out_sel, out_count, nullptr);
}

int gather_columns(ghive_ctx *ctx, const ghive_table *t, const uint32_t *sel, uint64_t n,
                   ghive_column *out_cols) {
  for (uint32_t c = 0; c < t->ncols; ++c) {
    const ghive_column &in = t->columns[c];
    ghive_column &o = out_cols[c];
    o.type = in.type;
    o.row_num = n;
    uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 16;
    uint32_t grid = (uint32_t)(g > cap ? cap : g);
    if (in.type == GHIVE_STRING) {  // (start,end) pairs into the caller's aux; the blob is shared
      o.data = in.data;
      o.char_size = in.char_size;
      if (n == 0) continue;
      if (!o.aux) return set_error(ctx, GHIVE_ERR_INVALID, "gather: STRING output column %u has no aux buffer", c);
      k_gather<StrPair><<<grid, 256, 0, ctx->stream>>>((const StrPair *)in.aux, sel, n, (StrPair *)o.aux);
      GH_LAUNCH(ctx);
      continue;
    }
    if (n == 0) continue;
    if (!o.data) return set_error(ctx, GHIVE_ERR_INVALID, "gather: output column %u has no buffer", c);
    if (type_width(in.type) == 8)
      k_gather<uint64_t><<<grid, 256, 0, ctx->stream>>>((const uint64_t *)in.data, sel, n, (uint64_t *)o.data);
    else
      k_gather<uint32_t><<<grid, 256, 0, ctx->stream>>>((const uint32_t *)in.data, sel, n, (uint32_t *)o.data);
    GH_LAUNCH(ctx);
  }
  return GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_filter(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes, uint32_t nnodes,
                 uint32_t root, uint32_t *out_sel, uint64_t *out_count) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "filter"));
  if (!out_count) return set_error(ctx, GHIVE_ERR_INVALID, "filter: out_count is NULL");
  if (table->row_num && !out_sel) return set_error(ctx, GHIVE_ERR_INVALID, "filter: out_sel is NULL");
  return filter_select(ctx, table, nodes, nnodes, root, out_sel, out_count);
}

int ghive_filter_async(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes, uint32_t nnodes,
                       uint32_t root, uint32_t *out_sel, uint64_t *count_dev) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "filter_async"));
  if (!count_dev) return set_error(ctx, GHIVE_ERR_INVALID, "filter_async: count_dev is NULL");
  if (table->row_num && !out_sel) return set_error(ctx, GHIVE_ERR_INVALID, "filter_async: out_sel is NULL");
  return filter_run(ctx, table, nodes, nnodes, root, out_sel, nullptr, count_dev);
}

int ghive_filter_compact(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes,
                         uint32_t nnodes, uint32_t root, ghive_column *out_cols, uint64_t *out_count) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "filter_compact"));
  if (!out_count || (table->ncols && !out_cols))
    return set_error(ctx, GHIVE_ERR_INVALID, "filter_compact: NULL output");
  DevBuf sel;
  GH_TRY(sel.alloc(ctx, sizeof(uint32_t) * (table->row_num ? table->row_num : 1)));
  GH_TRY(filter_select(ctx, table, nodes, nnodes, root, sel.as<uint32_t>(), out_count));
  GH_TRY(gather_columns(ctx, table, sel.as<uint32_t>(), *out_count, out_cols));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

int ghive_gather(ghive_ctx *ctx, const ghive_table *table, const uint32_t *sel, uint64_t n,
                 ghive_column *out_cols) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "gather"));
  if (n && (!sel || !out_cols)) return set_error(ctx, GHIVE_ERR_INVALID, "gather: NULL argument");
  GH_TRY(gather_columns(ctx, table, sel, n, out_cols));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

}  // extern "C"
