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
#include "scan.hpp"

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
  uint8_t *d_franks;     // mode 4: per distinct [lo, hi], ranked words and group codes
  uint32_t fstride, fcopy, nwe;
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
  // mode 4: per distinct filter, fstride bytes: nwe uint2 words {bitmap word, passing keys in the
  // words before it}, then one uint8 group code per passing key in key order
  const uint8_t *franks;
  uint32_t fstride;  // bytes per filter in franks
  uint32_t fcopy;    // bytes of it copied to LDS: the words and the codes of the most passing keys
  uint32_t nwe;      // nwords rounded up to even (16-byte aligned codes)
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

// Mode 4 tables, one workgroup per distinct filter: each 32-key word of the filtered bitmap with
// the count of passing keys in the words before it (a block-wide scan, chunk by chunk), and the
// group field of every passing key, stored at its rank among the passing keys.
__global__ __launch_bounds__(1024) void k_batch_rank_codes(const uint32_t *fbits, const int32_t *dense, uint32_t nwords,
                                                          uint32_t nwe, uint32_t fstride, uint32_t gshift, uint32_t gmask,
                                                          uint8_t *franks, uint32_t *npass) {
  __shared__ uint32_t sc[16];
  const uint32_t p = blockIdx.x;
  const uint32_t *bits = fbits + (uint64_t)p * nwords;
  uint8_t *base = franks + (uint64_t)p * fstride;
  uint2 *words = reinterpret_cast<uint2 *>(base);
  uint8_t *codes = base + (uint64_t)nwe * 8;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nwe; c0 += blockDim.x) {
    const uint32_t w = c0 + threadIdx.x;
    const uint32_t b = w < nwords ? bits[w] : 0u;
    uint32_t total;
    const uint32_t r0 = carry + block_exclusive_scan_256((uint32_t)__popc(b), sc, total);
    if (w < nwe) words[w] = make_uint2(b, r0);
    uint32_t r = r0, m = b;
    while (m) {
      const uint32_t j = (uint32_t)__ffs(m) - 1;
      m &= m - 1;
      codes[r++] = (uint8_t)(((uint32_t)dense[(uint64_t)w * 32 + j] >> gshift) & gmask);
    }
    carry += total;
  }
  if (threadIdx.x == 0) npass[p] = carry;
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

// The pipelined sub-tile: kBatchVec 16-byte loads per column per lane, kBatchVec * kBatchThreads * 4
// rows; lane l holds rows l*4 .. l*4+3 of each kBatchThreads*4-row slice (coalesced loads).
constexpr int kBatchVec = 2;
constexpr int kBatchLaneRows = 4 * kBatchVec;

struct BatchRows {
  int32_t k[kBatchLaneRows], a[kBatchLaneRows], b[kBatchLaneRows];
};

__device__ __forceinline__ void batch_load(const int32_t *kp, const int32_t *ap, const int32_t *bp, uint64_t r0,
                                           BatchRows &x) {
#pragma unroll
  for (int v = 0; v < kBatchVec; ++v) {
    // global (not flat) loads: a flat load also counts against the LDS wait counter
    typedef const __attribute__((address_space(1))) v4i gv4i;
    const uint64_t r = r0 + (uint64_t)v * kBatchThreads * 4;
    const v4i k = __builtin_nontemporal_load((gv4i *)(kp + r));
    const v4i a = __builtin_nontemporal_load((gv4i *)(ap + r));
    const v4i b = __builtin_nontemporal_load((gv4i *)(bp + r));
    x.k[4 * v] = k.x; x.k[4 * v + 1] = k.y; x.k[4 * v + 2] = k.z; x.k[4 * v + 3] = k.w;
    x.a[4 * v] = a.x; x.a[4 * v + 1] = a.y; x.a[4 * v + 2] = a.z; x.a[4 * v + 3] = a.w;
    x.b[4 * v] = b.x; x.b[4 * v + 1] = b.y; x.b[4 * v + 2] = b.z; x.b[4 * v + 3] = b.w;
  }
}

// Rows of the previous sub-tile that passed the bitmap, and their dimension payloads (loads in
// flight until the next step's wait).
struct BatchPend {
  int32_t pay[kBatchLaneRows];
  uint32_t alive;
};

// Adds the rows of `pend` (payloads landed) to the LDS group slots; x holds their fact values.
__device__ __forceinline__ void batch_accumulate(const BatchDev &B, const BatchItemDev &it, unsigned long long *lsum,
                                                 unsigned long long *lcnt, const BatchRows &x, const BatchPend &pend) {
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) {
    if (!((pend.alive >> q) & 1u)) continue;
    const int32_t f = (int32_t)(((uint32_t)pend.pay[q] >> B.fshift) & B.fmask);
    if (f < it.lo || f > it.hi) continue;
    if (x.a[q] == kNull || (it.b && x.b[q] == kNull)) continue;
    const uint32_t g = ((uint32_t)pend.pay[q] >> B.gshift) & B.gmask;
    const int64_t v = it.b ? (int64_t)((uint64_t)(int64_t)x.a[q] * (uint64_t)(int64_t)x.b[q]) : (int64_t)x.a[q];
    atomicAdd(&lsum[g], (unsigned long long)v);
    atomicAdd(&lcnt[g], 1ull);
  }
}

// Step N of the three-stage pipeline of the LDS-bitmap mode. On entry the loads in flight are, in
// issue order: sub-tile N (into cur), the payloads of sub-tile N-1 (into pend), sub-tile N+1. The
// step waits for all but the last 3 * kBatchVec of them (vector memory completes in order), adds
// sub-tile N-1 (its fact values in prev), tests sub-tile N against the LDS bitmap, issues its
// payload loads, and reloads prev with sub-tile N+2. So two sub-tiles of fact columns are in
// flight while the LDS work runs, and no global write happens in the loop.
__device__ __forceinline__ void batch_step(const BatchDev &B, const BatchItemDev &it, const uint32_t *lbits,
                                           unsigned long long *lsum, unsigned long long *lcnt, const int32_t *bp,
                                           const BatchRows &cur, BatchRows &prev, BatchPend &pend, uint64_t r_next2) {
  static_assert(3 * kBatchVec < 16, "vmcnt immediate holds counts below 16 in its low bits");
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0 | (3 * kBatchVec));
  batch_accumulate(B, it, lsum, lcnt, prev, pend);
  // every row reads its bitmap word (word 0 when out of range), so the LDS reads issue together
  // and complete under one wait
  uint32_t alive = 0;
  uint32_t w[kBatchLaneRows];
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) {
    const int32_t k = cur.k[q];
    const bool in = k != kNull && k >= B.key_min && k <= B.key_max;
    const uint32_t u = in ? (uint32_t)((int64_t)k - B.key_min) : 0u;
    w[q] = lbits[u >> 5] >> (u & 31);
    alive |= (uint32_t)in << q;
  }
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) alive &= ~((~w[q] & 1u) << q);
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) {
    pend.pay[q] = 0;
    if ((alive >> q) & 1u) pend.pay[q] = B.dense[(uint32_t)((int64_t)cur.k[q] - B.key_min)];
  }
  pend.alive = alive;
  batch_load(it.key, it.a, bp, r_next2, prev);
}

// Pipelined sub-tiles 0 .. nfull-1 of the run of rows starting at `base` of item `it`, all whole.
// Loads past the last sub-tile repeat it (in bounds, never used).
__device__ __forceinline__ void batch_run_pipelined(const BatchDev &B, const BatchItemDev &it, const uint32_t *lbits,
                                                    unsigned long long *lsum, unsigned long long *lcnt, uint64_t base,
                                                    uint32_t nfull) {
  constexpr uint32_t kSubP = kBatchThreads * kBatchLaneRows;
  const int32_t *bp = it.b ? it.b : it.a;  // always three loads per slice (b unused without it)
  const uint64_t lane_r = base + (uint64_t)threadIdx.x * 4;
  const uint32_t last = nfull - 1;
  auto row = [&](uint32_t s) { return lane_r + (uint64_t)(s < last ? s : last) * kSubP; };
  BatchRows x0, x1, x2;
  BatchPend pend;
  pend.alive = 0;
  batch_load(it.key, it.a, bp, row(0), x0);
  batch_load(it.key, it.a, bp, row(1), x1);
  uint32_t s = 0;
  for (;; s += 3) {
    batch_step(B, it, lbits, lsum, lcnt, bp, x0, x2, pend, row(s + 2));
    if (s + 1 > last) break;
    batch_step(B, it, lbits, lsum, lcnt, bp, x1, x0, pend, row(s + 3));
    if (s + 2 > last) break;
    batch_step(B, it, lbits, lsum, lcnt, bp, x2, x1, pend, row(s + 4));
    if (s + 3 > last) break;
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // payloads of the last sub-tile, and the repeated loads
  // the last step consumed set (last % 3); its fact values are still there
  const uint32_t m = last % 3;
  batch_accumulate(B, it, lsum, lcnt, m == 0 ? x0 : (m == 1 ? x1 : x2), pend);
}

// Mode 4: group of fact key k from the LDS tables of the item's filter, or -1 when the key is
// outside the dimension or its row fails the filter. Both LDS reads happen for every row (word 0 /
// code 0 otherwise), so a sub-tile's reads issue together.
__device__ __forceinline__ int32_t batch_code(const BatchDev &B, const uint2 *lw, const uint8_t *lcodes, int32_t k) {
  const bool in = k != kNull && k >= B.key_min && k <= B.key_max;
  const uint32_t u = in ? (uint32_t)((int64_t)k - B.key_min) : 0u;
  const uint2 w = lw[u >> 5];
  const uint32_t sh = u & 31;
  const bool pass = in && ((w.x >> sh) & 1u);
  const uint32_t rank = pass ? w.y + (uint32_t)__popc(w.x & ((1u << sh) - 1u)) : 0u;
  const int32_t g = lcodes[rank];
  return pass ? g : -1;
}

__device__ __forceinline__ void batch_add_coded(const BatchItemDev &it, unsigned long long *lsum,
                                                unsigned long long *lcnt, int32_t g, int32_t a, int32_t b) {
  if (g < 0 || a == kNull || (it.b && b == kNull)) return;
  const int64_t v = it.b ? (int64_t)((uint64_t)(int64_t)a * (uint64_t)(int64_t)b) : (int64_t)a;
  atomicAdd(&lsum[g], (unsigned long long)v);
  atomicAdd(&lcnt[g], 1ull);
}

// Step N of the mode-4 pipeline: on entry sub-tiles N and N+1 are in flight (in that order); the
// step waits for N, issues N+2 into `spare` (sub-tile N-1's registers, consumed), and resolves
// and adds sub-tile N from LDS alone.
__device__ __forceinline__ void batch_step_coded(const BatchDev &B, const BatchItemDev &it, const uint2 *lw,
                                                 const uint8_t *lcodes, unsigned long long *lsum,
                                                 unsigned long long *lcnt, const int32_t *bp, const BatchRows &cur,
                                                 BatchRows &spare, uint64_t r_next2) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0 | (3 * kBatchVec));
  batch_load(it.key, it.a, bp, r_next2, spare);
  int32_t g[kBatchLaneRows];
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) g[q] = batch_code(B, lw, lcodes, cur.k[q]);
#pragma unroll
  for (int q = 0; q < kBatchLaneRows; ++q) batch_add_coded(it, lsum, lcnt, g[q], cur.a[q], cur.b[q]);
}

__device__ __forceinline__ void batch_run_coded(const BatchDev &B, const BatchItemDev &it, const uint2 *lw,
                                                const uint8_t *lcodes, unsigned long long *lsum,
                                                unsigned long long *lcnt, uint64_t base, uint32_t nfull) {
  constexpr uint32_t kSubP = kBatchThreads * kBatchLaneRows;
  const int32_t *bp = it.b ? it.b : it.a;
  const uint64_t lane_r = base + (uint64_t)threadIdx.x * 4;
  const uint32_t last = nfull - 1;
  auto row = [&](uint32_t s) { return lane_r + (uint64_t)(s < last ? s : last) * kSubP; };
  BatchRows x0, x1, x2;
  batch_load(it.key, it.a, bp, row(0), x0);
  batch_load(it.key, it.a, bp, row(1), x1);
  for (uint32_t s = 0;; s += 3) {
    batch_step_coded(B, it, lw, lcodes, lsum, lcnt, bp, x0, x2, row(s + 2));
    if (s + 1 > last) break;
    batch_step_coded(B, it, lw, lcodes, lsum, lcnt, bp, x1, x0, row(s + 3));
    if (s + 2 > last) break;
    batch_step_coded(B, it, lw, lcodes, lsum, lcnt, bp, x2, x1, row(s + 4));
    if (s + 3 > last) break;
  }
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);  // the repeated loads land before the registers are reused
}

// One sub-tile (kBatchThreads * 4 rows from r_sub; rows at or past it.rows are skipped) in any
// probe mode, without pipelining: the HBM-bitmap and hash modes, and the last partial sub-tile
// of an item in the LDS-bitmap mode.
__device__ __forceinline__ void batch_sub_plain(const BatchDev &B, const BatchItemDev &it, const uint32_t *lbits,
                                                const uint32_t *gbits, unsigned long long *lsum,
                                                unsigned long long *lcnt, uint64_t r_sub) {
  const uint64_t r0 = r_sub + (uint64_t)threadIdx.x * 4;
  if (r0 >= it.rows) return;
  int32_t k[4], a[4], b[4];
  bload4(it.key, r0, it.rows, k);
  bload4(it.a, r0, it.rows, a);
  if (it.b) bload4(it.b, r0, it.rows, b);
  if (B.mode == 4) {
    const uint2 *lw = reinterpret_cast<const uint2 *>(lbits);
    const uint8_t *lcodes = reinterpret_cast<const uint8_t *>(lw + B.nwe);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (r0 + q < it.rows) batch_add_coded(it, lsum, lcnt, batch_code(B, lw, lcodes, k[q]), a[q], b[q]);
    return;
  }
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
    bool live[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) live[q] = (alive >> q) & 1u;
    probe_unique8_rows<4>(B.slots, B.mask, B.log2, k, live, pay);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (!live[q]) alive &= ~(1u << q);
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

__global__ __launch_bounds__(kBatchThreads) void k_batch(BatchDev B) {
  constexpr uint32_t kSub = kBatchThreads * 4;
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
  // The tiles [t_begin, t_end) cut into runs, one per item they touch; each run is streamed as
  // one sequence of sub-tiles, so the load pipeline drains only where the item changes.
  for (uint64_t t = t_begin; t < t_end;) {
    while (t >= next_begin) {  // crossed into the next item(s)
      batch_flush(B, lsum, lcnt, item);
      ++item;
      it = B.items[item];
      next_begin = B.items[item + 1].tile_begin;
    }
    const uint64_t t_stop = t_end < next_begin ? t_end : next_begin;
    if ((B.mode == 1 || B.mode == 4) && it.fidx != loaded) {  // uniform: every thread sees the same item
      __syncthreads();
      if (B.mode == 1) {
        const uint32_t *src = B.fbits + (uint64_t)it.fidx * B.nwords;
        for (uint32_t w = threadIdx.x; w < B.nwords; w += blockDim.x) lbits[w] = src[w];
      } else {  // fstride is a multiple of 16 bytes, and so is the LDS offset of lbits
        const v4i *src = reinterpret_cast<const v4i *>(B.franks + (uint64_t)it.fidx * B.fstride);
        v4i *dst = reinterpret_cast<v4i *>(lbits);
        for (uint32_t w = threadIdx.x; w < B.fcopy / 16; w += blockDim.x) dst[w] = src[w];
      }
      loaded = it.fidx;
      __syncthreads();
    }
    const uint32_t *gbits = B.mode == 2 ? B.fbits + (uint64_t)it.fidx * B.nwords : nullptr;
    // rows [base, lim) of the item: whole tiles, the item's last one possibly partial
    const uint64_t base = (t - it.tile_begin) * B.tile_rows;
    const uint64_t run_end = (t_stop - it.tile_begin) * B.tile_rows;
    const uint64_t lim = it.rows < run_end ? it.rows : run_end;
    uint64_t r = base;
    if ((B.mode == 1 || B.mode == 4) && lim > base) {  // whole sub-tiles of the LDS modes: software-pipelined
      constexpr uint32_t kSubP = kBatchThreads * kBatchLaneRows;
      const uint32_t nfull = (uint32_t)((lim - base) / kSubP);
      if (nfull && B.mode == 1) batch_run_pipelined(B, it, lbits, lsum, lcnt, base, nfull);
      if (nfull && B.mode == 4) {
        const uint2 *lw = reinterpret_cast<const uint2 *>(lbits);
        batch_run_coded(B, it, lw, reinterpret_cast<const uint8_t *>(lw + B.nwe), lsum, lcnt, base, nfull);
      }
      r = base + (uint64_t)nfull * kSubP;
    }
    // base and r are multiples of kSub and lim is either one or the item's row count, so every
    // sub-tile below either ends at or before lim or is the item's last
    for (; r < lim; r += kSub) batch_sub_plain(B, it, lbits, gbits, lsum, lcnt, r);
    t = t_stop;
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
  if (bt->d_franks) c->arena->free(bt->d_franks);
  for (hipEvent_t e : {bt->ev_in, bt->ev_out, bt->t0, bt->t1})
    if (e) (void)hipEventDestroy(e);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  bt->stream = nullptr;
  bt->exec = nullptr;
  bt->graph = nullptr;
  bt->d_items = nullptr;
  bt->d_acc = nullptr;
  bt->d_fbits = nullptr;
  bt->d_franks = nullptr;
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
  // automatic choice: mode 4 when its tables leave LDS for two workgroups per CU (decided once
  // the passing keys are counted), else the LDS bitmap when it fits 64 KB, else the HBM bitmap
  const bool auto_mode = mode == 0;
  const bool coded_ok = h->bitmap && spec->group_bits <= 8;
  if (auto_mode) mode = coded_ok ? 4 : h->bitmap ? ((acc_lds + (size_t)nwords * 4 <= 64 * 1024) ? 1 : 2) : 3;
  if ((mode == 1 || mode == 2 || mode == 4) && !h->bitmap) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: probe mode %d needs a dense dimension", mode);
  }
  if (mode == 4 && spec->group_bits > 8) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: probe mode 4 needs group_bits <= 8");
  }
  if (mode == 1 && acc_lds + (size_t)nwords * 4 > 160 * 1024) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: dimension bitmap does not fit LDS");
  }
  if (mode < 1 || mode > 4) {
    delete bt;
    return set_error(ctx, GHIVE_ERR_INVALID, "batch: probe mode %d", mode);
  }
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
  if (st == GHIVE_OK && (mode == 1 || mode == 2 || mode == 4) && bt->npairs) {
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
  if (st == GHIVE_OK && mode == 4) {
    // ranked words and group codes per filter; the codes area has room for every key to pass,
    // and the LDS copy covers the largest passing count
    bt->nwe = (nwords + 1) & ~1u;
    bt->fstride = bt->nwe * 8 + ((nwords * 32 + 15) & ~15u);
    const uint32_t np = std::max<uint32_t>(bt->npairs, 1);
    bt->d_franks = static_cast<uint8_t *>(ctx->arena->alloc((uint64_t)bt->fstride * np));
    uint32_t *d_npass = static_cast<uint32_t *>(ctx->arena->alloc(4ull * np));
    std::vector<uint32_t> npass(np, 0);
    if (!bt->d_franks || !d_npass) {
      st = set_error(ctx, GHIVE_ERR_OOM, "batch: ranked filter tables do not fit");
    } else if (bt->npairs) {
      k_batch_rank_codes<<<bt->npairs, 1024, 0, bt->stream>>>(bt->d_fbits, h->dense, nwords, bt->nwe, bt->fstride,
                                                               spec->group_shift, bt->groups - 1, bt->d_franks, d_npass);
      if (hipGetLastError() != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "batch: ranked tables launch");
      else if (hipMemcpyAsync(npass.data(), d_npass, 4ull * bt->npairs, hipMemcpyDeviceToHost, bt->stream) != hipSuccess ||
               hipStreamSynchronize(bt->stream) != hipSuccess)
        st = set_error(ctx, GHIVE_ERR_HIP, "batch: ranked tables readback");
    }
    if (d_npass) ctx->arena->free(d_npass);
    const uint32_t maxpass = *std::max_element(npass.begin(), npass.end());
    bt->fcopy = bt->nwe * 8 + ((std::max<uint32_t>(maxpass, 1) + 15) & ~15u);
    if (st == GHIVE_OK && auto_mode && acc_lds + bt->fcopy > 80 * 1024) {
      mode = (acc_lds + (size_t)nwords * 4 <= 64 * 1024) ? 1 : 2;  // keeps the filtered bitmaps built above
    } else if (st == GHIVE_OK && acc_lds + bt->fcopy > 160 * 1024) {
      st = set_error(ctx, GHIVE_ERR_UNSUPPORTED, "batch: ranked filter tables do not fit LDS");
    }
    if (mode == 4) {
      bt->lds_bytes = acc_lds + bt->fcopy;
    } else {
      bt->lds_bytes = acc_lds + (mode == 1 ? (size_t)nwords * 4 : 0);
      // modes 1 and 2 never read the ranked tables: give their memory back now
      if (hipStreamSynchronize(bt->stream) != hipSuccess && st == GHIVE_OK)
        st = set_error(ctx, GHIVE_ERR_HIP, "batch: sync");
      if (bt->d_franks) ctx->arena->free(bt->d_franks);
      bt->d_franks = nullptr;
      bt->fstride = bt->fcopy = bt->nwe = 0;
    }
  }
  bt->mode = mode;
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
    B.franks = bt->d_franks;
    B.fstride = bt->fstride;
    B.fcopy = bt->fcopy;
    B.nwe = bt->nwe;
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
    B.franks = bt->d_franks;
    B.fstride = bt->fstride;
    B.fcopy = bt->fcopy;
    B.nwe = bt->nwe;
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
