// Batched form: thousands of map-task fragments of one query shape (filter on a dimension
// attribute -> join -> GROUP BY a dimension attribute, SUM(a*b)) in ONE persistent kernel,
// captured in a HIP graph. The dimension table is resident and shared; every fragment brings
// its own fact columns, row count and filter constants.
//
// Work is the concatenation of all fragments' tiles. Workgroup b owns the contiguous tile range
// [b*W/G, (b+1)*W/G), so consecutive tiles mostly belong to one fragment: partial sums stay in
// LDS (one slot per group-attribute value) and are flushed with one atomic per group only when
// the fragment changes. A fragment boundary is uniform across the workgroup (tile granularity),
// so the flush is a plain block-wide step with no waiting on other workgroups.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "hash.hpp"

struct ghive_batch {
  ghive_ctx *owner;
  uint32_t nitems;
  uint32_t groups;       // slots per item (2^group_bits)
  uint64_t total_rows;
  uint64_t total_tiles;
  ghive_batch_tuning tune;
  int mode;              // 1 LDS bitmap, 2 HBM bitmap, 3 hash
  void *d_items;         // BatchItemDev[nitems + 1]
  uint32_t *d_fbits;     // filtered bitmaps, one per distinct [lo, hi]
  uint32_t npairs;
  uint32_t nwords;
  unsigned long long *d_acc;  // sums then counts, nitems * groups each
  hipStream_t stream;
  hipEvent_t ev_in, ev_out, t0, t1;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int use_graph;
  size_t lds_bytes;
  uint32_t grid;
  float last_ms;
};

namespace ghive {

struct BatchItemDev {
  const int32_t *key, *a, *b;
  uint64_t rows;
  uint64_t tile_begin;  // first global tile of this item (item nitems: total tiles)
  int32_t lo, hi;
  uint32_t fidx;        // which filtered bitmap (distinct [lo, hi] pairs share one)
  uint32_t pad;
};

struct BatchDev {
  const BatchItemDev *items;
  uint32_t nitems;
  uint32_t tile_rows;   // multiple of kBatchThreads * 4
  uint64_t total_tiles;
  int32_t mode;
  int32_t log2;
  const uint64_t *slots;
  uint64_t mask;
  const uint32_t *bitmap;
  const int32_t *dense;
  int32_t key_min, key_max;
  uint32_t nwords;
  uint32_t fshift, fmask, gshift, gmask;
  uint32_t groups;
  unsigned long long *sums;
  unsigned long long *counts;
  const uint32_t *fbits;  // modes 1/2: per distinct filter, bitmap of dimension keys that pass it
};

// Filtered bitmaps: bit u of filter p is set when key (key_min + u) is in the dimension and its
// filter field lies in [lo_p, hi_p]. One thread per 32-key word.
__global__ void k_batch_filter_bits(const uint32_t *bitmap, const int32_t *dense, uint64_t range, uint32_t nwords,
                                    const int32_t *lohi, uint32_t npairs, uint32_t fshift, uint32_t fmask,
                                    uint32_t *fbits) {
  const uint64_t total = (uint64_t)nwords * npairs;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    const uint32_t p = (uint32_t)(i / nwords), w = (uint32_t)(i % nwords);
    const int32_t lo = lohi[2 * p], hi = lohi[2 * p + 1];
    const uint32_t present = bitmap[w];
    uint32_t out = 0;
    for (int b = 0; b < 32; ++b) {
      const uint64_t u = (uint64_t)w * 32 + b;
      if (u >= range || !((present >> b) & 1u)) continue;
      const int32_t f = (int32_t)(((uint32_t)dense[u] >> fshift) & fmask);
      if (f >= lo && f <= hi) out |= 1u << b;
    }
    fbits[(uint64_t)p * nwords + w] = out;
  }
}

typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int kBatchThreads = 256;
constexpr int kNull = 0x7fffffff;

// Fact columns are read once: non-temporal loads keep the dimension bitmap and attributes cached.
__device__ __forceinline__ void bload4(const int32_t *p, uint64_t r0, uint64_t n, int32_t v[4]) {
  if (r0 + 3 < n) {
    const v4i x = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(p + r0));
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = r0 + q < n ? p[r0 + q] : kNull;
  }
}

__device__ __forceinline__ void batch_flush(const BatchDev &B, unsigned long long *lsum, unsigned long long *lcnt,
                                            uint32_t item) {
  __syncthreads();
  for (uint32_t g = threadIdx.x; g < B.groups; g += blockDim.x) {
    unsigned long long c = lcnt[g];
    if (c) {
      atomicAdd(&B.sums[(uint64_t)item * B.groups + g], lsum[g]);
      atomicAdd(&B.counts[(uint64_t)item * B.groups + g], c);
      lsum[g] = 0;
      lcnt[g] = 0;
    }
  }
  __syncthreads();
}

// s_waitcnt immediate of "vmcnt(0)" on gfx9-family chips (expcnt and lgkmcnt at their maxima);
// OR-ing a count below 16 into the low bits gives "vmcnt(count)"
constexpr int kWaitVmcnt0 = 0x0f70;

struct BatchRows {
  int32_t k[4], a[4], b[4];
};

__device__ __forceinline__ void batch_load(const int32_t *kp, const int32_t *ap, const int32_t *bp, uint64_t r0,
                                           BatchRows &x) {
  const v4i k = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(kp + r0));
  const v4i a = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(ap + r0));
  const v4i b = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(bp + r0));
  x.k[0] = k.x; x.k[1] = k.y; x.k[2] = k.z; x.k[3] = k.w;
  x.a[0] = a.x; x.a[1] = a.y; x.a[2] = a.z; x.a[3] = a.w;
  x.b[0] = b.x; x.b[1] = b.y; x.b[2] = b.z; x.b[3] = b.w;
}

// One whole sub-tile (kBatchThreads * 4 rows) of the LDS-bitmap mode, with the next sub-tile's
// three column loads in flight: LDS bitmap test, then the attribute loads of the passing rows,
// then the next loads, then a wait for everything older than those (vector memory completes in
// order) and the LDS accumulation. No global write happens here, so the waits count loads only.
__device__ __forceinline__ void batch_sub(const BatchDev &B, const BatchItemDev &it, const uint32_t *lbits,
                                          unsigned long long *lsum, unsigned long long *lcnt, const int32_t *bp,
                                          const BatchRows &cur, BatchRows &nxt, uint64_t next_r0) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  uint32_t alive = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int32_t k = cur.k[q];
    bool ok = k != kNull && k >= B.key_min && k <= B.key_max;
    if (ok) {
      const uint32_t u = (uint32_t)((int64_t)k - B.key_min);
      ok = (lbits[u >> 5] >> (u & 31)) & 1u;
    }
    alive |= (uint32_t)ok << q;
  }
  int32_t pay[4] = {0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if ((alive >> q) & 1u) pay[q] = B.dense[(uint32_t)((int64_t)cur.k[q] - B.key_min)];
  batch_load(it.key, it.a, bp, next_r0, nxt);
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0 | 3);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!((alive >> q) & 1u)) continue;
    const int32_t f = (int32_t)(((uint32_t)pay[q] >> B.fshift) & B.fmask);
    if (f < it.lo || f > it.hi) continue;
    if (cur.a[q] == kNull || (it.b && cur.b[q] == kNull)) continue;
    const uint32_t g = ((uint32_t)pay[q] >> B.gshift) & B.gmask;
    const int64_t v = it.b ? (int64_t)((uint64_t)(int64_t)cur.a[q] * (uint64_t)(int64_t)cur.b[q]) : (int64_t)cur.a[q];
    atomicAdd(&lsum[g], (unsigned long long)v);
    atomicAdd(&lcnt[g], 1ull);
  }
}

// Sub-tiles 0 .. nfull-1 of the tile at row `base` of item `it`, all whole.
__device__ __forceinline__ void batch_tile_pipelined(const BatchDev &B, const BatchItemDev &it, const uint32_t *lbits,
                                                     unsigned long long *lsum, unsigned long long *lcnt, uint64_t base,
                                                     uint32_t nfull) {
  constexpr uint32_t kSub = kBatchThreads * 4;
  const int32_t *bp = it.b ? it.b : it.a;  // always three loads per sub-tile (b unused without it)
  const uint64_t lane_r = base + (uint64_t)threadIdx.x * 4;
  BatchRows x0, x1;
  batch_load(it.key, it.a, bp, lane_r, x0);
  for (uint32_t s = 0;; s += 2) {
    batch_sub(B, it, lbits, lsum, lcnt, bp, x0, x1, lane_r + (uint64_t)(s + 1 < nfull ? s + 1 : s) * kSub);
    if (s + 1 >= nfull) break;
    batch_sub(B, it, lbits, lsum, lcnt, bp, x1, x0, lane_r + (uint64_t)(s + 2 < nfull ? s + 2 : s + 1) * kSub);
    if (s + 2 >= nfull) break;
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // the last (repeated) loads land before the registers are reused
}

__global__ __launch_bounds__(kBatchThreads) void k_batch(BatchDev B) {
  extern __shared__ uint64_t lds64[];
  unsigned long long *lsum = reinterpret_cast<unsigned long long *>(lds64);
  unsigned long long *lcnt = lsum + B.groups;
  uint32_t *lbits = reinterpret_cast<uint32_t *>(lcnt + B.groups);
  for (uint32_t g = threadIdx.x; g < B.groups; g += blockDim.x) lsum[g] = lcnt[g] = 0;
  __syncthreads();
  const uint64_t t_begin = B.total_tiles * blockIdx.x / gridDim.x;
  const uint64_t t_end = B.total_tiles * (blockIdx.x + 1) / gridDim.x;
  if (t_begin >= t_end) return;  // uniform: the whole workgroup leaves together
  // item holding t_begin: last item with tile_begin <= t_begin (binary search, uniform)
  uint32_t lo = 0, hi = B.nitems;  // invariant: items[lo].tile_begin <= t_begin < items[hi].tile_begin
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (B.items[mid].tile_begin <= t_begin) lo = mid;
    else hi = mid;
  }
  uint32_t item = lo;
  BatchItemDev it = B.items[item];
  uint64_t next_begin = B.items[item + 1].tile_begin;
  uint32_t loaded = 0xffffffffu;  // filtered bitmap currently in LDS (mode 1)
  for (uint64_t t = t_begin; t < t_end; ++t) {
    while (t >= next_begin) {  // crossed into the next item(s)
      batch_flush(B, lsum, lcnt, item);
      ++item;
      it = B.items[item];
      next_begin = B.items[item + 1].tile_begin;
    }
    if (B.mode == 1 && it.fidx != loaded) {  // uniform: every thread sees the same item
      __syncthreads();
      const uint32_t *src = B.fbits + (uint64_t)it.fidx * B.nwords;
      for (uint32_t w = threadIdx.x; w < B.nwords; w += blockDim.x) lbits[w] = src[w];
      loaded = it.fidx;
      __syncthreads();
    }
    const uint32_t *gbits = B.mode == 2 ? B.fbits + (uint64_t)it.fidx * B.nwords : nullptr;
    const uint64_t base = (t - it.tile_begin) * B.tile_rows;
    uint32_t sub0 = 0;  // first sub-tile left to the row-by-row loop below
    if (B.mode == 1) {  // whole sub-tiles of the LDS-bitmap mode: software-pipelined
      const uint64_t lim = it.rows < base + B.tile_rows ? it.rows : base + B.tile_rows;
      const uint32_t nfull = lim > base ? (uint32_t)((lim - base) / (kBatchThreads * 4)) : 0u;
      if (nfull) batch_tile_pipelined(B, it, lbits, lsum, lcnt, base, nfull);
      sub0 = nfull * kBatchThreads * 4;
    }
    for (uint32_t sub = sub0; sub < B.tile_rows; sub += kBatchThreads * 4) {
      const uint64_t r0 = base + sub + (uint64_t)threadIdx.x * 4;
      if (r0 >= it.rows) break;
      int32_t k[4], a[4], b[4];
      bload4(it.key, r0, it.rows, k);
      bload4(it.a, r0, it.rows, a);
      if (it.b) bload4(it.b, r0, it.rows, b);
      uint32_t alive = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        bool ok = r0 + q < it.rows && k[q] != kNull && k[q] >= B.key_min && k[q] <= B.key_max;
        alive |= (ok ? 1u : 0u) << q;
      }
      int32_t pay[4] = {0, 0, 0, 0};
      if (B.mode == 1 || B.mode == 2) {
        // the filtered bitmap already holds "key present and its filter field in [lo, hi]", so
        // only the rows that pass read the payload
        if (B.mode == 1) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (!((alive >> q) & 1u)) continue;
            const uint32_t u = (uint32_t)((int64_t)k[q] - B.key_min);
            if (!((lbits[u >> 5] >> (u & 31)) & 1u)) alive &= ~(1u << q);
          }
        } else {
          uint32_t w[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            w[q] = 0;
            if ((alive >> q) & 1u) {
              const uint32_t u = (uint32_t)((int64_t)k[q] - B.key_min);
              w[q] = gbits[u >> 5] >> (u & 31);
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (!(w[q] & 1u)) alive &= ~(1u << q);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if ((alive >> q) & 1u) pay[q] = B.dense[(uint32_t)((int64_t)k[q] - B.key_min)];
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!((alive >> q) & 1u)) continue;
          int32_t p = 0;
          if (!probe_unique8(B.slots, B.mask, B.log2, k[q], p)) alive &= ~(1u << q);
          pay[q] = p;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!((alive >> q) & 1u)) continue;
        const int32_t f = (int32_t)(((uint32_t)pay[q] >> B.fshift) & B.fmask);
        if (f < it.lo || f > it.hi) continue;
        if (a[q] == kNull || (it.b && b[q] == kNull)) continue;
        const uint32_t g = ((uint32_t)pay[q] >> B.gshift) & B.gmask;
        const int64_t v = it.b ? (int64_t)((uint64_t)(int64_t)a[q] * (uint64_t)(int64_t)b[q]) : (int64_t)a[q];
        atomicAdd(&lsum[g], (unsigned long long)v);
        atomicAdd(&lcnt[g], 1ull);
      }
    }
  }
  batch_flush(B, lsum, lcnt, item);
}

__global__ void k_zero_u64(unsigned long long *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0;
}

static int batch_enqueue(ghive_ctx *ctx, ghive_batch *bt, const BatchDev &B) {
  const uint64_t acc_words = 2ull * bt->nitems * bt->groups;
  uint64_t g = (acc_words + 255) / 256, cap = (uint64_t)ctx->cus * 4;
  k_zero_u64<<<(uint32_t)std::max<uint64_t>(1, std::min(g, cap)), 256, 0, bt->stream>>>(bt->d_acc, acc_words);
  GH_LAUNCH(ctx);
  if (bt->total_tiles) {
    k_batch<<<bt->grid, kBatchThreads, bt->lds_bytes, bt->stream>>>(B);
    GH_LAUNCH(ctx);
  }
  return GHIVE_OK;
}

// HIP objects and device memory of a batch (its context still alive).
void batch_release(ghive_batch *bt) {
  ghive_ctx *c = bt->owner;
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  if (bt->d_items) c->arena->free(bt->d_items);
  if (bt->d_acc) c->arena->free(bt->d_acc);
  if (bt->d_fbits) c->arena->free(bt->d_fbits);
  for (hipEvent_t e : {bt->ev_in, bt->ev_out, bt->t0, bt->t1})
    if (e) (void)hipEventDestroy(e);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  bt->stream = nullptr;
  bt->exec = nullptr;
  bt->graph = nullptr;
  bt->d_items = nullptr;
  bt->d_acc = nullptr;
  bt->d_fbits = nullptr;
  bt->ev_in = bt->ev_out = bt->t0 = bt->t1 = nullptr;
}

void orphan_batches(ghive_ctx *ctx) {
  for (ghive_batch *bt : ctx->batches) {
    batch_release(bt);
    bt->owner = nullptr;
  }
  ctx->batches.clear();
}

}  // namespace ghive

using namespace ghive;

extern "C" {

void ghive_batch_free(ghive_ctx *ctx, ghive_batch *bt) {
  (void)ctx;
  if (!bt) return;
  ghive_ctx *c = bt->owner;  // NULL: its context was shut down and released everything already
  if (c) {
    (void)hipSetDevice(c->device);
    ghive::batch_release(bt);
    c->batches.erase(bt);
  }
  delete bt;
}

int ghive_batch_create(ghive_ctx *ctx, const ghive_batch_spec *spec, const ghive_batch_item *items,
                       uint32_t nitems, const ghive_batch_tuning *tuning, int use_graph, ghive_batch **out) {
  GH_CHECK_CTX(ctx);
  if (!spec || !out || (nitems && !items)) return set_error(ctx, GHIVE_ERR_INVALID, "batch: NULL argument");
  *out = nullptr;
  const ghive_hash_table *h = spec->dim;
  if (h && !h->owner) return set_error(ctx, GHIVE_ERR_INVALID, "batch: the dimension table's context was shut down");
  if (h) GH_TRY(hash_table_ready(ctx, h));
  if (!h || h->layout != 8 || !h->unique || h->payload_kind != 1)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: the dimension must be a unique-key INT table with an INT payload");
  if (spec->group_bits < 1 || spec->group_bits > 12 || spec->filter_bits < 1 || spec->filter_bits > 31 ||
      spec->group_shift + spec->group_bits > 32 || spec->filter_shift + spec->filter_bits > 32)
    return set_error(ctx, GHIVE_ERR_INVALID, "batch: attribute fields out of range");
  ghive_batch_tuning tn = tuning ? *tuning : ghive_batch_tuning{0, 0, 0, 0};
  if (tn.tile_rows == 0) tn.tile_rows = 8192;
  if (tn.tile_rows % (kBatchThreads * 4) != 0 || tn.tile_rows > (1u << 24))
    return set_error(ctx, GHIVE_ERR_INVALID, "batch: tile_rows %u (a multiple of %d)", tn.tile_rows, kBatchThreads * 4);
  if (tn.blocks_per_cu == 0) tn.blocks_per_cu = 4;
  if (tn.blocks_per_cu > 32) return set_error(ctx, GHIVE_ERR_INVALID, "batch: blocks_per_cu %u", tn.blocks_per_cu);
  auto *bt = new ghive_batch();
  bt->owner = ctx;
  bt->nitems = nitems;
  bt->groups = 1u << spec->group_bits;
  bt->tune = tn;
  bt->use_graph = use_graph;
  int mode = tn.probe_mode;
  const uint32_t nwords = h->bitmap ? (uint32_t)((h->range + 31) / 32) : 0;
  const size_t acc_lds = (size_t)bt->groups * 16;
  if (mode == 0) mode = h->bitmap ? ((acc_lds + (size_t)nwords * 4 <= 64 * 1024) ? 1 : 2) : 3;
  if ((mode == 1 || mode == 2) && !h->bitmap) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: probe mode %d needs a dense dimension", mode);
  }
  if (mode == 1 && acc_lds + (size_t)nwords * 4 > 160 * 1024) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: dimension bitmap does not fit LDS");
  }
  if (mode < 1 || mode > 3) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_INVALID, "batch: probe mode %d", mode);
  }
  bt->mode = mode;
  bt->lds_bytes = acc_lds + (mode == 1 ? (size_t)nwords * 4 : 0);
  std::vector<BatchItemDev> hv(nitems + 1);
  std::vector<std::pair<int32_t, int32_t>> pairs;
  uint64_t tiles = 0, rows = 0;
  for (uint32_t i = 0; i < nitems; ++i) {
    const ghive_batch_item &s = items[i];
    if (s.rows && (!s.key || !s.a))
      return delete bt, set_error(ctx, GHIVE_ERR_INVALID, "batch: item %u has no key/a column", i);
    for (const int32_t *p : {s.key, s.a, s.b})
      if (p && ((uintptr_t)p & 15))
        return delete bt, set_error(ctx, GHIVE_ERR_INVALID, "batch: item %u column is not 16-byte aligned", i);
    if (s.rows > kMaxRows) return delete bt, set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: item %u too large", i);
    uint32_t fidx = 0;
    for (; fidx < pairs.size(); ++fidx)
      if (pairs[fidx].first == s.filter_lo && pairs[fidx].second == s.filter_hi) break;
    if (fidx == pairs.size()) pairs.push_back({s.filter_lo, s.filter_hi});
    hv[i] = BatchItemDev{s.key, s.a, s.b, s.rows, tiles, s.filter_lo, s.filter_hi, fidx, 0};
    tiles += (s.rows + tn.tile_rows - 1) / tn.tile_rows;
    rows += s.rows;
  }
  hv[nitems] = BatchItemDev{nullptr, nullptr, nullptr, 0, tiles, 0, 0, 0, 0};
  bt->total_tiles = tiles;
  bt->total_rows = rows;
  bt->npairs = (uint32_t)pairs.size();
  bt->nwords = nwords;
  int st = GHIVE_OK;
  bt->d_items = ctx->arena->alloc(sizeof(BatchItemDev) * (nitems + 1));
  bt->d_acc = static_cast<unsigned long long *>(ctx->arena->alloc(16ull * nitems * bt->groups + 16));
  if (!bt->d_items || !bt->d_acc) st = set_error(ctx, GHIVE_ERR_OOM, "batch: descriptors do not fit");
  if (st == GHIVE_OK && hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess)
    st = set_error(ctx, GHIVE_ERR_HIP, "batch: stream");
  for (hipEvent_t *e : {&bt->ev_in, &bt->ev_out})
    if (st == GHIVE_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      st = set_error(ctx, GHIVE_ERR_HIP, "batch: event");
  for (hipEvent_t *e : {&bt->t0, &bt->t1})
    if (st == GHIVE_OK && hipEventCreate(e) != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "batch: event");
  if (st == GHIVE_OK && hipMemcpyAsync(bt->d_items, hv.data(), sizeof(BatchItemDev) * hv.size(),
                                       hipMemcpyHostToDevice, bt->stream) != hipSuccess)
    st = set_error(ctx, GHIVE_ERR_HIP, "batch: upload");
  if (st == GHIVE_OK && (mode == 1 || mode == 2) && bt->npairs) {
    // one filtered bitmap per distinct [lo, hi]: bit set = key present and filter field passes
    bt->d_fbits = static_cast<uint32_t *>(ctx->arena->alloc(4ull * nwords * bt->npairs));
    void *d_lohi = ctx->arena->alloc(8ull * bt->npairs);
    std::vector<int32_t> lohi;
    for (auto &pr : pairs) {
      lohi.push_back(pr.first);
      lohi.push_back(pr.second);
    }
    if (!bt->d_fbits || !d_lohi) {
      st = set_error(ctx, GHIVE_ERR_OOM, "batch: filtered bitmaps do not fit");
    } else if (hipMemcpyAsync(d_lohi, lohi.data(), 8ull * bt->npairs, hipMemcpyHostToDevice, bt->stream) != hipSuccess) {
      st = set_error(ctx, GHIVE_ERR_HIP, "batch: upload");
    } else {
      const uint64_t total = (uint64_t)nwords * bt->npairs;
      const uint32_t fmask = spec->filter_bits >= 32 ? ~0u : ((1u << spec->filter_bits) - 1);
      k_batch_filter_bits<<<(uint32_t)std::min<uint64_t>((total + 255) / 256, (uint64_t)ctx->cus * 8), 256, 0,
                            bt->stream>>>(h->bitmap, h->dense, h->range, nwords, static_cast<const int32_t *>(d_lohi),
                                          bt->npairs, spec->filter_shift, fmask, bt->d_fbits);
      if (hipGetLastError() != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "batch: filter bitmaps launch");
    }
    if (hipStreamSynchronize(bt->stream) != hipSuccess && st == GHIVE_OK) st = set_error(ctx, GHIVE_ERR_HIP, "batch: sync");
    if (d_lohi) ctx->arena->free(d_lohi);
  }
  if (st == GHIVE_OK && hipStreamSynchronize(bt->stream) != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "batch: sync");
  if (st != GHIVE_OK) {
    ghive_batch_free(ctx, bt);
    return st;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, kBatchThreads, bt->lds_bytes) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  uint32_t bpc = std::min<uint32_t>(tn.blocks_per_cu, (uint32_t)per_cu);
  bt->grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles ? tiles : 1, (uint64_t)ctx->cus * bpc));
  if (use_graph) {
    BatchDev B{};
    B.items = static_cast<const BatchItemDev *>(bt->d_items);
    B.nitems = nitems;
    B.tile_rows = tn.tile_rows;
    B.total_tiles = tiles;
    B.mode = mode;
    B.log2 = h->log2_slots;
    B.slots = static_cast<const uint64_t *>(h->data);
    B.mask = h->slots - 1;
    B.bitmap = h->bitmap;
    B.dense = h->dense;
    B.key_min = h->entries ? (int32_t)h->key_min : 1;
    B.key_max = h->entries ? (int32_t)h->key_max : 0;
    B.nwords = nwords;
    B.fshift = spec->filter_shift;
    B.fmask = spec->filter_bits >= 32 ? ~0u : ((1u << spec->filter_bits) - 1);
    B.gshift = spec->group_shift;
    B.gmask = bt->groups - 1;
    B.groups = bt->groups;
    B.sums = bt->d_acc;
    B.counts = bt->d_acc + (uint64_t)nitems * bt->groups;
    B.fbits = bt->d_fbits;
    if (hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      st = set_error(ctx, GHIVE_ERR_HIP, "batch: begin capture");
    if (st == GHIVE_OK) {
      int s2 = batch_enqueue(ctx, bt, B);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(bt->stream, &g);
      if (s2 != GHIVE_OK) st = s2;
      else if (e != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "batch: end capture");
      bt->graph = g;
    }
    if (st == GHIVE_OK && hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0) != hipSuccess)
      st = set_error(ctx, GHIVE_ERR_HIP, "batch: graph instantiate");
    if (st != GHIVE_OK) {
      ghive_batch_free(ctx, bt);
      return st;
    }
  }
  // keep the spec for direct launches
  bt->tune.probe_mode = mode;
  ctx->batches.insert(bt);
  *out = bt;
  // stash what a direct launch needs in the graph-less case
  if (!use_graph) {
    // the descriptor is rebuilt per run from the table, which must outlive the batch
  }
  return GHIVE_OK;
}

int ghive_batch_run(ghive_ctx *ctx, ghive_batch *bt, const ghive_batch_spec *spec) {
  GH_CHECK_CTX(ctx);
  if (!bt) return set_error(ctx, GHIVE_ERR_INVALID, "batch_run: NULL batch");
  if (!bt->owner) return set_error(ctx, GHIVE_ERR_INVALID, "batch_run: the batch's context was shut down");
  GH_HIP(ctx, hipEventRecord(bt->ev_in, ctx->stream));
  GH_HIP(ctx, hipStreamWaitEvent(bt->stream, bt->ev_in, 0));
  GH_HIP(ctx, hipEventRecord(bt->t0, bt->stream));
  if (bt->use_graph) {
    GH_HIP(ctx, hipGraphLaunch(bt->exec, bt->stream));
  } else {
    if (!spec || !spec->dim) return set_error(ctx, GHIVE_ERR_INVALID, "batch_run: a graph-less batch needs its spec");
    const ghive_hash_table *h = spec->dim;
    BatchDev B{};
    B.items = static_cast<const BatchItemDev *>(bt->d_items);
    B.nitems = bt->nitems;
    B.tile_rows = bt->tune.tile_rows;
    B.total_tiles = bt->total_tiles;
    B.mode = bt->mode;
    B.log2 = h->log2_slots;
    B.slots = static_cast<const uint64_t *>(h->data);
    B.mask = h->slots - 1;
    B.bitmap = h->bitmap;
    B.dense = h->dense;
    B.key_min = h->entries ? (int32_t)h->key_min : 1;
    B.key_max = h->entries ? (int32_t)h->key_max : 0;
    B.nwords = h->bitmap ? (uint32_t)((h->range + 31) / 32) : 0;
    B.fshift = spec->filter_shift;
    B.fmask = spec->filter_bits >= 32 ? ~0u : ((1u << spec->filter_bits) - 1);
    B.gshift = spec->group_shift;
    B.gmask = bt->groups - 1;
    B.groups = bt->groups;
    B.sums = bt->d_acc;
    B.counts = bt->d_acc + (uint64_t)bt->nitems * bt->groups;
    B.fbits = bt->d_fbits;
    GH_TRY(batch_enqueue(ctx, bt, B));
  }
  GH_HIP(ctx, hipEventRecord(bt->t1, bt->stream));
  GH_HIP(ctx, hipEventRecord(bt->ev_out, bt->stream));
  GH_HIP(ctx, hipStreamWaitEvent(ctx->stream, bt->ev_out, 0));
  GH_HIP(ctx, hipEventSynchronize(bt->t1));
  (void)hipEventElapsedTime(&bt->last_ms, bt->t0, bt->t1);
  return GHIVE_OK;
}

int ghive_batch_results(ghive_ctx *ctx, ghive_batch *bt, int64_t *sums, int64_t *counts) {
  GH_CHECK_CTX(ctx);
  if (!bt || !sums || !counts) return set_error(ctx, GHIVE_ERR_INVALID, "batch_results: NULL argument");
  if (!bt->owner) return set_error(ctx, GHIVE_ERR_INVALID, "batch_results: the batch's context was shut down");
  const uint64_t n = (uint64_t)bt->nitems * bt->groups;
  GH_HIP(ctx, hipMemcpyAsync(sums, bt->d_acc, n * 8, hipMemcpyDeviceToHost, bt->stream));
  GH_HIP(ctx, hipMemcpyAsync(counts, bt->d_acc + n, n * 8, hipMemcpyDeviceToHost, bt->stream));
  GH_HIP(ctx, hipStreamSynchronize(bt->stream));
  return GHIVE_OK;
}

int ghive_batch_info(const ghive_batch *bt, uint32_t *groups_per_item, uint64_t *total_rows, uint64_t *total_tiles,
                     int32_t *probe_mode, uint32_t *grid, float *last_ms) {
  if (!bt) return GHIVE_ERR_INVALID;
  if (groups_per_item) *groups_per_item = bt->groups;
  if (total_rows) *total_rows = bt->total_rows;
  if (total_tiles) *total_tiles = bt->total_tiles;
  if (probe_mode) *probe_mode = bt->mode;
  if (grid) *grid = bt->grid;
  if (last_ms) *last_ms = bt->last_ms;
  return GHIVE_OK;
}

}  // extern "C"
