// Hash-join build and probe, and the N-way join of BaseJoinPredicate::process
// (gpu_impl/src/Operator/SortMergeJoinPredicate.cu:8-461 produces the same result_offsets
// with a GPU sort plus a single-threaded CPU merge; here both sides stay on the device).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "groupby.hpp"
#include "hash.hpp"
#include "pred_host.hpp"
#include "scan.hpp"

namespace ghive {

constexpr int kHashThreads = 256;

__device__ __forceinline__ long long wave_min_i64(long long v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    long long o = __shfl_xor(v, d, kWave);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ long long wave_max_i64(long long v) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    long long o = __shfl_xor(v, d, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// Build statistics per workgroup, part[6 * blockIdx.x + i]: [0] entries, [1] duplicate-key flag,
// [2] key min, [3] key max, [4] payload min, [5] payload max (k_build_stats combines them; no
// atomics on shared counters, which serialise across the XCDs).
constexpr int kStatWords = 6;

// One workgroup's statistics from its lanes' values into part[blockIdx.x * kStatWords ..]; every
// thread of the workgroup calls it (it holds a barrier).
template <int TH>
__device__ void stat_partial(unsigned long long *part, uint32_t inserted, uint32_t dup, long long kmin, long long kmax,
                             long long pmin, long long pmax) {
  kmin = wave_min_i64(kmin);
  kmax = wave_max_i64(kmax);
  pmin = wave_min_i64(pmin);
  pmax = wave_max_i64(pmax);
  const unsigned long long ins = wave_sum_u64(inserted);
  const bool anydup = ballot(dup != 0) != 0;
  __shared__ long long red[TH / kWave][kStatWords];
  const int wv = threadIdx.x / kWave;
  if (lane_id() == 0) {
    red[wv][0] = (long long)ins;
    red[wv][1] = anydup;
    red[wv][2] = kmin;
    red[wv][3] = kmax;
    red[wv][4] = pmin;
    red[wv][5] = pmax;
  }
  __syncthreads();
  if (threadIdx.x < kStatWords) {
    const int w = threadIdx.x;
    long long v = red[0][w];
    for (int j = 1; j < TH / kWave; ++j) {
      const long long o = red[j][w];
      v = w < 2 ? v + o : ((w & 1) == 0 ? (o < v ? o : v) : (o > v ? o : v));  // sums, mins, maxes
    }
    part[(uint64_t)blockIdx.x * kStatWords + w] = (unsigned long long)v;
  }
}
// n_dev (optional): the row count on the device (a filter's count), at most n.
__global__ __launch_bounds__(kHashThreads) void k_hash_build(
    const void *keys, int key_type, const uint32_t *sel, uint64_t n, const uint64_t *n_dev, const void *pay,
    int pay_type, int layout, uint64_t *slots, uint64_t mask, int log2, unsigned long long *part) {
  long long kmin = LLONG_MAX, kmax = LLONG_MIN, pmin = LLONG_MAX, pmax = LLONG_MIN;
  uint32_t inserted = 0, dup = 0;
  if (n_dev && *n_dev < n) n = *n_dev;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = sel ? sel[i] : i;
    int64_t k;
    if (!col_load_i64(keys, key_type, r, k)) continue;  // NULL keys never join
    int64_t p;
    if (pay) {
      if (!col_load_i64(pay, pay_type, r, p)) p = pay_type == GHIVE_INT ? 0x7fffffffll : 0x7fffffffffffffffll;
    } else {
      p = (int64_t)r;
    }
    if (layout == 8) {
      uint64_t word = (uint64_t)(uint32_t)(int32_t)k | ((uint64_t)(uint32_t)p << 32);
      uint64_t s = hash_u32((uint32_t)k, log2);
      for (uint64_t step = 0; step <= mask; ++step) {
        unsigned long long old = atomicCAS((unsigned long long *)&slots[s], kEmpty8, word);
        if (old == kEmpty8) break;
        if ((int32_t)(uint32_t)old == (int32_t)k) dup = 1;
        s = (s + 1) & mask;
      }
    } else {
      uint64_t s = hash_u64((uint64_t)k, log2);
      for (uint64_t step = 0; step <= mask; ++step) {
        unsigned long long old = atomicCAS((unsigned long long *)&slots[2 * s], kEmpty16, (unsigned long long)k);
        if (old == kEmpty16) {
          slots[2 * s + 1] = (uint64_t)p;
          break;
        }
        if ((int64_t)old == k) dup = 1;
        s = (s + 1) & mask;
      }
    }
    ++inserted;
    kmin = k < kmin ? k : kmin;
    kmax = k > kmax ? k : kmax;
    pmin = p < pmin ? p : pmin;
    pmax = p > pmax ? p : pmax;
  }
  stat_partial<kHashThreads>(part, inserted, dup, kmin, kmax, pmin, pmax);
}

// Combines the per-workgroup build statistics into stats[0..5] (one workgroup, kHashThreads).
__device__ void build_stats_combine(const unsigned long long *part, uint32_t nparts, unsigned long long *stats) {
  long long v[kStatWords] = {0, 0, LLONG_MAX, LLONG_MIN, LLONG_MAX, LLONG_MIN};
  for (uint32_t b = threadIdx.x; b < nparts; b += blockDim.x) {
    const unsigned long long *p = part + (uint64_t)b * kStatWords;
    v[0] += (long long)p[0];
    v[1] |= (long long)p[1];
    v[2] = (long long)p[2] < v[2] ? (long long)p[2] : v[2];
    v[3] = (long long)p[3] > v[3] ? (long long)p[3] : v[3];
    v[4] = (long long)p[4] < v[4] ? (long long)p[4] : v[4];
    v[5] = (long long)p[5] > v[5] ? (long long)p[5] : v[5];
  }
  v[0] = (long long)wave_sum_u64((unsigned long long)v[0]);
  v[1] = ballot(v[1] != 0) != 0;
  v[2] = wave_min_i64(v[2]);
  v[3] = wave_max_i64(v[3]);
  v[4] = wave_min_i64(v[4]);
  v[5] = wave_max_i64(v[5]);
  __shared__ long long red[kHashThreads / kWave][kStatWords];
  const int wv = threadIdx.x / kWave;
  if (lane_id() == 0)
    for (int w = 0; w < kStatWords; ++w) red[wv][w] = v[w];
  __syncthreads();
  if (threadIdx.x < kStatWords) {
    const int w = threadIdx.x;
    long long r = red[0][w];
    for (int j = 1; j < (int)(blockDim.x / kWave); ++j) {
      const long long o = red[j][w];
      r = w < 2 ? (w == 0 ? r + o : (r | o)) : ((w & 1) == 0 ? (o < r ? o : r) : (o > r ? o : r));
    }
    stats[w] = (unsigned long long)r;
  }
}
__global__ __launch_bounds__(kHashThreads) void k_build_stats(const unsigned long long *part, uint32_t nparts,
                                                             unsigned long long *stats) {
  build_stats_combine(part, nparts, stats);
}

// ---- filtered builds: a table's filter and its hash build in one pass -------------------------
// Several tables' fills and statistics each run as one launch for all of them (blockIdx.y or
// blockIdx.x = table), their filter-and-insert passes one launch per table (the predicate travels
// by value).
struct TableFill {
  uint64_t *p;
  uint64_t words;
  uint64_t value;
};
struct TableFills {
  TableFill t[GHIVE_MAX_FILTERED_BUILDS];
};
__global__ void k_fill_tables(TableFills F) {
  const TableFill &f = F.t[blockIdx.y];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.words; i += stride) f.p[i] = f.value;
}

struct TableStats {
  const unsigned long long *part;
  uint32_t nparts;
  unsigned long long *stats;
};
struct TableStatsAll {
  TableStats t[GHIVE_MAX_FILTERED_BUILDS];
};
__global__ __launch_bounds__(kHashThreads) void k_build_stats_tables(TableStatsAll S) {
  build_stats_combine(S.t[blockIdx.x].part, S.t[blockIdx.x].nparts, S.t[blockIdx.x].stats);
}

// Rows r < rows that pass P (nnodes 0: every row) and whose key is not NULL: inserted (the loop of
// k_hash_build), with the statistics partials of k_hash_build. No list of the rows is kept: a
// counter shared by every wave would serialise their appends at the memory side.
__global__ __launch_bounds__(kHashThreads) void k_filtered_build(DevPred P, const void *keys, int key_type,
                                                                const void *pay, int pay_type, uint64_t rows,
                                                                int layout, uint64_t *slots, uint64_t mask, int log2,
                                                                unsigned long long *part) {
  long long kmin = LLONG_MAX, kmax = LLONG_MIN, pmin = LLONG_MAX, pmax = LLONG_MIN;
  uint32_t inserted = 0, dup = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += stride) {
    int64_t k = 0;
    if (P.nnodes && !eval_pred(P, r)) continue;
    if (!col_load_i64(keys, key_type, r, k)) continue;  // NULL keys never join
    int64_t p;
    if (pay) {
      if (!col_load_i64(pay, pay_type, r, p)) p = pay_type == GHIVE_INT ? 0x7fffffffll : 0x7fffffffffffffffll;
    } else {
      p = (int64_t)r;
    }
    if (layout == 8) {
      const uint64_t word = (uint64_t)(uint32_t)(int32_t)k | ((uint64_t)(uint32_t)p << 32);
      uint64_t s = hash_u32((uint32_t)k, log2);
      for (uint64_t step = 0; step <= mask; ++step) {
        const unsigned long long old = atomicCAS((unsigned long long *)&slots[s], kEmpty8, word);
        if (old == kEmpty8) break;
        if ((int32_t)(uint32_t)old == (int32_t)k) dup = 1;
        s = (s + 1) & mask;
      }
    } else {
      uint64_t s = hash_u64((uint64_t)k, log2);
      for (uint64_t step = 0; step <= mask; ++step) {
        const unsigned long long old = atomicCAS((unsigned long long *)&slots[2 * s], kEmpty16, (unsigned long long)k);
        if (old == kEmpty16) {
          slots[2 * s + 1] = (uint64_t)p;
          break;
        }
        if ((int64_t)old == k) dup = 1;
        s = (s + 1) & mask;
      }
    }
    ++inserted;
    kmin = k < kmin ? k : kmin;
    kmax = k > kmax ? k : kmax;
    pmin = p < pmin ? p : pmin;
    pmax = p > pmax ? p : pmax;
  }
  stat_partial<kHashThreads>(part, inserted, dup, kmin, kmax, pmin, pmax);
}

// ---- regional build: large tables without device-scope atomics ---------------------------------
// A key's home slot is the top log2 bits of its Fibonacci product w (k * kFibMul64, or for L8 the
// 32-bit product in w's high half), and the product determines the key. Rows sorted by the top
// bits of w therefore arrive grouped by table region (2^kRegionLog2 consecutive slots). One
// workgroup builds each region in LDS (linear probing inside the region with LDS compare-and-swap)
// and writes it out whole, so the table needs no fill pass and no insert reaches past the XCD's L2;
// device-scope atomics on this GPU execute at the memory side, and a per-key global CAS loop moved
// a 128-byte line per insert. A key whose probe runs past its region's last slot goes to an
// overflow list and is inserted afterwards by the ordinary global loop from its home slot: every
// slot from there to the region's end is occupied, so its probe sequence is the one a sequential
// build would give it.
constexpr int kRegionLog2 = 13;
constexpr int kRegionThreads = 512;

// W[i] = row i's payload (the build row id without a payload column), W[n + i] = its product
// (NULL keys: the product of the empty marker, which the region build skips); statistics partials
// as k_hash_build writes them (no duplicate flag: the region build finds duplicates).
__global__ __launch_bounds__(kHashThreads) void k_build_words(const void *keys, int key_type, const uint32_t *sel,
                                                             uint64_t n, const void *pay, int pay_type, int layout,
                                                             uint64_t *W, unsigned long long *part) {
  long long kmin = LLONG_MAX, kmax = LLONG_MIN, pmin = LLONG_MAX, pmax = LLONG_MIN;
  uint32_t inserted = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = sel ? sel[i] : i;
    int64_t k;
    const bool valid = col_load_i64(keys, key_type, r, k);
    int64_t p;
    if (pay) {
      if (!col_load_i64(pay, pay_type, r, p)) p = pay_type == GHIVE_INT ? 0x7fffffffll : 0x7fffffffffffffffll;
    } else {
      p = (int64_t)r;
    }
    if (!valid) k = layout == 8 ? (int64_t)0x7fffffff : (int64_t)kEmpty16;
    W[i] = (uint64_t)p;
    W[n + i] = layout == 8 ? (uint64_t)((uint32_t)k * kFibMul32) << 32 : (uint64_t)k * kFibMul64;
    if (!valid) continue;
    ++inserted;
    kmin = k < kmin ? k : kmin;
    kmax = k > kmax ? k : kmax;
    pmin = p < pmin ? p : pmin;
    pmax = p > pmax ? p : pmax;
  }
  stat_partial<kHashThreads>(part, inserted, 0, kmin, kmax, pmin, pmax);
}

// starts[r] = the first row whose product's top rbits are >= r, for r in [0, 2^rbits] (one binary
// search per region over the sorted products).
__global__ void k_region_starts(const uint64_t *w, uint64_t n, int rbits, uint64_t *starts) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nr = 1ull << rbits;
  if (r > nr) return;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if ((w[mid] >> (64 - rbits)) < r) lo = mid + 1;
    else hi = mid;
  }
  starts[r] = lo;
}

// The key of product w (decoded by the multiplier's inverse).
template <int L>
__device__ __forceinline__ int64_t region_key(uint64_t w) {
  if constexpr (L == 16) return (int64_t)(w * kFibInv64);
  else return (int64_t)(int32_t)((uint32_t)(w >> 32) * kFibInv32);
}

// Regions blockIdx.x, + gridDim.x, ... of a table of 2^log2 slots from the sorted rows (W: payloads,
// W + n: products; starts from k_region_starts). ov (2 x ov_cap words: products, then payloads)
// takes the rows that run past their region, *ov_count their number.
template <int L>
__global__ __launch_bounds__(kRegionThreads) void k_region_build(const uint64_t *W, uint64_t n, const uint64_t *starts,
                                                                 uint32_t nregions, int log2, uint64_t *table,
                                                                 uint64_t *ov, unsigned long long *ov_count,
                                                                 uint64_t ov_cap, unsigned long long *part) {
  constexpr uint32_t kSlots = 1u << kRegionLog2;
  __shared__ __attribute__((aligned(16))) uint64_t tk[kSlots];  // L16 keys; L8 slot words
  __shared__ __attribute__((aligned(16))) uint64_t tp[L == 16 ? kSlots : 1];  // L16 payloads
  constexpr uint64_t kEmpty = L == 16 ? kEmpty16 : kEmpty8;
  uint32_t dup = 0;
  for (uint32_t reg = blockIdx.x; reg < nregions; reg += gridDim.x) {
    for (uint32_t s = threadIdx.x; s < kSlots; s += kRegionThreads) {
      tk[s] = kEmpty;
      if constexpr (L == 16) tp[s] = kEmpty16;  // an empty slot's payload word, as the fill pass leaves it
    }
    __syncthreads();
    const uint64_t b = starts[reg], e = starts[reg + 1];
    const uint64_t base = (uint64_t)reg << kRegionLog2;
    for (uint64_t i = b + threadIdx.x; i < e; i += kRegionThreads) {
      const uint64_t w = W[n + i], p = W[i];
      const int64_t k = region_key<L>(w);
      if (L == 16 ? (uint64_t)k == kEmpty16 : (int32_t)k == 0x7fffffff) continue;  // a NULL key: never joins
      uint32_t s = (uint32_t)((w >> (64 - log2)) - base);  // home slot inside the region
      bool placed = false;
      for (; s < kSlots; ++s) {
        if constexpr (L == 16) {
          const unsigned long long old = atomicCAS((unsigned long long *)&tk[s], kEmpty16, (unsigned long long)k);
          if (old == kEmpty16) {
            tp[s] = p;
            placed = true;
            break;
          }
          if ((int64_t)old == k) dup = 1;
        } else {
          const uint64_t word = (uint64_t)(uint32_t)(int32_t)k | ((uint64_t)(uint32_t)p << 32);
          const unsigned long long old = atomicCAS((unsigned long long *)&tk[s], kEmpty8, word);
          if (old == kEmpty8) {
            placed = true;
            break;
          }
          if ((int32_t)(uint32_t)old == (int32_t)k) dup = 1;
        }
      }
      if (!placed) {
        const unsigned long long o = atomicAdd(ov_count, 1ull);
        if (o < ov_cap) {
          ov[o] = w;
          ov[ov_cap + o] = p;
        }
      }
    }
    __syncthreads();
    if constexpr (L == 16) {  // {key, payload} pairs as 16-byte stores
      typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
      u64x2 *out = reinterpret_cast<u64x2 *>(table) + base;
      for (uint32_t s = threadIdx.x; s < kSlots; s += kRegionThreads) out[s] = u64x2{tk[s], tp[s]};
    } else {
      for (uint32_t s = threadIdx.x; s < kSlots; s += kRegionThreads) table[base + s] = tk[s];
    }
    __syncthreads();  // the region is written out before the next one reuses the LDS table
  }
  stat_partial<kRegionThreads>(part, 0, dup, LLONG_MAX, LLONG_MIN, LLONG_MAX, LLONG_MIN);
}

// The overflow rows of the region build, inserted by the global loop from their home slots.
__global__ __launch_bounds__(kHashThreads) void k_overflow_insert(const uint64_t *ov, uint64_t ov_cap,
                                                                 const unsigned long long *ov_count, int layout,
                                                                 uint64_t *slots, uint64_t mask, int log2,
                                                                 unsigned long long *part) {
  const uint64_t cnt = *ov_count < ov_cap ? *ov_count : ov_cap;
  uint32_t dup = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += stride) {
    const uint64_t w = ov[i], p = ov[ov_cap + i];
    uint64_t s = w >> (64 - log2);
    if (layout == 8) {
      const int32_t k = (int32_t)region_key<8>(w);
      const uint64_t word = (uint64_t)(uint32_t)k | ((uint64_t)(uint32_t)p << 32);
      for (uint64_t step = 0; step <= mask; ++step) {
        const unsigned long long old = atomicCAS((unsigned long long *)&slots[s], kEmpty8, word);
        if (old == kEmpty8) break;
        if ((int32_t)(uint32_t)old == k) dup = 1;
        s = (s + 1) & mask;
      }
    } else {
      const int64_t k = region_key<16>(w);
      for (uint64_t step = 0; step <= mask; ++step) {
        const unsigned long long old = atomicCAS((unsigned long long *)&slots[2 * s], kEmpty16, (unsigned long long)k);
        if (old == kEmpty16) {
          slots[2 * s + 1] = p;
          break;
        }
        if ((int64_t)old == k) dup = 1;
        s = (s + 1) & mask;
      }
    }
  }
  stat_partial<kHashThreads>(part, 0, dup, LLONG_MAX, LLONG_MIN, LLONG_MAX, LLONG_MIN);
}

// mode 0 inner: every match; 1 left semi: the first match only; 2 left outer: every match, or
// one (row, -1) when there is none. Probe row i reads key row sel[i] (sel == nullptr: i) or
// tup[i] (-1 = NULL side of an earlier outer join). Emits (emit_index ? i : row, payload).
__global__ __launch_bounds__(kHashThreads) void k_hash_probe(
    DevHash h, const void *keys, int key_type, const uint32_t *sel, const int32_t *tup, uint64_t n,
    int mode, int emit_index, int unique, uint32_t *out_a, int64_t *out_b, uint64_t cap,
    unsigned long long *counter) {
  uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of outer iterations
  const uint64_t iters = (n + stride - 1) / stride;
  for (uint64_t it = 0; it < iters; ++it) {
    uint64_t i = i0 + it * stride;
    bool active = i < n;
    uint64_t r = 0;
    int64_t k = 0;
    if (active) {
      if (tup) {
        int32_t t = tup[i];
        active = t >= 0;
        r = (uint64_t)(t < 0 ? 0 : t);
      } else {
        r = sel ? sel[i] : i;
      }
    }
    bool valid = active && col_load_i64(keys, key_type, r, k) && k >= h.key_min && k <= h.key_max;
    if (valid && h.layout == 8 && (k < INT32_MIN || k > INT32_MAX)) valid = false;
    bool searching = valid;
    bool any = false;
    uint64_t s = valid ? home_slot(h, k) : 0;
    uint64_t steps = 0;  // slots visited; a scan never wraps the whole table
    bool pending_outer = (mode == 2) && (i < n);
    for (;;) {
      bool match = false;
      int64_t payload = -1;
      if (searching) {
        for (;;) {
          if (steps++ > h.mask) { searching = false; break; }
          uint64_t kk, pp;
          if (h.layout == 8) {
            uint64_t w = h.slots[s];
            kk = (uint64_t)(int64_t)(int32_t)(uint32_t)w;
            pp = (uint64_t)(int64_t)(int32_t)(uint32_t)(w >> 32);
            if ((uint32_t)w == 0x7fffffffu) { searching = false; break; }
          } else {
            kk = h.slots[2 * s];
            if (kk == kEmpty16) { searching = false; break; }
            pp = h.slots[2 * s + 1];
          }
          s = (s + 1) & h.mask;
          if ((int64_t)kk == k) {
            match = true;
            payload = (int64_t)pp;
            if (unique || mode == 1) searching = false;
            break;
          }
        }
      }
      any |= match;
      bool emit_outer = pending_outer && !searching && !any;
      bool emit = match || emit_outer;
      if (emit_outer) pending_outer = false;
      if (match) pending_outer = false;
      uint64_t pos = wave_append(counter, emit);
      if (emit && pos < cap) {
        out_a[pos] = emit_index ? (uint32_t)i : (uint32_t)r;
        out_b[pos] = match ? payload : -1;
      }
      if (!ballot(searching)) break;
    }
  }
}

// Probe of a unique-key table (or the first match only, LEFT SEMI): at most one output per probe
// row. It takes the place of the reference's join front end, which sorts every input's row ids
// with thrust::sort(sort_comparator) (gpu_impl/src/Operator/SortMergeJoinPredicate.cu:103-115)
// before a CPU merge: here the build side is hashed once and each probe row is one table lookup.
// A workgroup probes a tile of blockDim.x * R rows (row = tile base + q * blockDim.x + lane,
// coalesced key loads), counts its matches with wave ballots and one LDS scan, and reserves its
// output range with ONE atomic per tile -- a per-wave append would put ~n/64 device-scope
// atomics on a single counter, which serialise across the XCDs.
//
// The table is probed at random, so the kernel is bound by how many slot loads are in flight:
// each lane first loads the keys of its R rows, then issues the R home-slot loads (one 8- or
// 16-byte load per slot) before it examines any of them, and only the rows whose home slot holds
// another key go on to further rounds, all R of them again issued together. KT is the key column
// type (GHIVE_INT or GHIVE_LONG), L the slot layout (8 or 16 bytes).
constexpr int kProbeThreads = 512;
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
template <int R, int KT, int L>
__global__ __launch_bounds__(kProbeThreads) void k_probe_unique(
    DevHash h, const void *keys, const uint32_t *sel, const int32_t *tup, uint64_t n,
    int emit_index, uint32_t *out_a, int64_t *out_b, uint64_t cap, unsigned long long *counter) {
  __shared__ uint32_t wave_total[kProbeThreads / kWave];
  __shared__ unsigned long long tile_base;
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave, nwaves = blockDim.x / kWave;
  const uint64_t tile_rows = (uint64_t)blockDim.x * R;
  const uint64_t ntiles = (n + tile_rows - 1) / tile_rows;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint64_t row[R], s[R];
    int64_t k[R], p[R];
    bool act[R], pend[R], match[R];
    // key rows (each branch issues its R index loads together)
    if (tup) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const uint64_t i = tile * tile_rows + (uint64_t)q * blockDim.x + threadIdx.x;
        const int32_t t = i < n ? tup[i] : -1;
        act[q] = t >= 0;
        row[q] = (uint64_t)(t < 0 ? 0 : t);
      }
    } else if (sel) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const uint64_t i = tile * tile_rows + (uint64_t)q * blockDim.x + threadIdx.x;
        act[q] = i < n;
        row[q] = sel[act[q] ? i : n - 1];  // n >= 1: the kernel is launched only for rows
      }
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const uint64_t i = tile * tile_rows + (uint64_t)q * blockDim.x + threadIdx.x;
        act[q] = i < n;
        row[q] = i;
      }
    }
    // the keys of all R rows in flight together
#pragma unroll
    for (int q = 0; q < R; ++q) {
      k[q] = 0;
      if (act[q])
        k[q] = KT == GHIVE_INT ? (int64_t)static_cast<const int32_t *>(keys)[row[q]]
                               : static_cast<const int64_t *>(keys)[row[q]];
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      match[q] = false;
      p[q] = -1;
      const bool notnull = KT == GHIVE_INT ? !is_null_i32((int32_t)k[q]) : !is_null_i64(k[q]);
      pend[q] = act[q] && notnull && k[q] >= h.key_min && k[q] <= h.key_max &&
                (L == 16 || (k[q] >= INT32_MIN && k[q] <= INT32_MAX));
      s[q] = pend[q] ? home_slot(h, k[q]) : 0;
    }
    // rounds of slot loads: every pending row's next slot, issued before any is examined; a
    // table always keeps empty slots, and a row is never followed past a full cycle of the table
    for (uint64_t round = 0; round <= h.mask; ++round) {
      uint64_t w0[R], w1[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (L == 16) {
          const u64x2 v = pend[q] ? *reinterpret_cast<const u64x2 *>(h.slots + 2 * s[q]) : u64x2{0, 0};
          w0[q] = v.x;
          w1[q] = v.y;
        } else {
          w0[q] = pend[q] ? h.slots[s[q]] : 0;
          w1[q] = 0;
        }
      }
      bool more = false;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        if (!pend[q]) continue;
        int64_t kk, pp;
        bool empty;
        if (L == 16) {
          kk = (int64_t)w0[q];
          pp = (int64_t)w1[q];
          empty = w0[q] == kEmpty16;
        } else {
          kk = (int64_t)(int32_t)(uint32_t)w0[q];
          pp = (int64_t)(int32_t)(uint32_t)(w0[q] >> 32);
          empty = (uint32_t)w0[q] == 0x7fffffffu;
        }
        if (!empty && kk == k[q]) {
          match[q] = true;
          p[q] = pp;
          pend[q] = false;
        } else if (empty) {
          pend[q] = false;
        } else {
          s[q] = (s[q] + 1) & h.mask;
          more = true;
        }
      }
      if (!ballot(more)) break;
    }
    // offsets: per q a ballot; the tile's matches are laid out q-major, wave-major, lane-minor
    uint32_t below[R], wave_count = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint64_t m = ballot(match[q]);
      below[q] = wave_count + lanes_below(m);
      wave_count += (uint32_t)__popcll(m);
    }
    if (lane == 0) wave_total[wave] = wave_count;
    __syncthreads();
    uint32_t wave_off = 0, tile_count = 0;
    for (uint32_t w = 0; w < nwaves; ++w) {
      const uint32_t c = wave_total[w];
      wave_off += w < wave ? c : 0u;
      tile_count += c;
    }
    if (threadIdx.x == 0) tile_base = tile_count ? atomicAdd(counter, (unsigned long long)tile_count) : 0ull;
    __syncthreads();
    const uint64_t base = tile_base + wave_off;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint64_t pos = base + below[q];
      if (match[q] && pos < cap) {
        const uint64_t i = tile * tile_rows + (uint64_t)q * blockDim.x + threadIdx.x;
        out_a[pos] = (uint32_t)(emit_index ? i : row[q]);
        out_b[pos] = p[q];
      }
    }
    __syncthreads();  // wave_total / tile_base are rewritten by the next tile
  }
}

// Dense companions of up to kDenseJobs tables per launch (blockIdx.y = table): one launch clears
// their bitmaps, the next sets the bit and dense payload of every inserted key (unique keys, so no
// write races on dense).
constexpr int kDenseJobs = 8;
struct DenseJob {
  const void *key;
  const uint32_t *sel;
  const uint64_t *n_dev;  // the row count on the device (at most n), optional
  const void *pay;
  uint64_t n;             // rows, or slots of slots8
  const uint64_t *slots8; // set: the entries are the L8 table's n slots, not build rows
  int64_t key_min;
  uint64_t range;
  uint32_t *bitmap;
  int32_t *dense;
  int key_type, pay_type;
};
struct DenseJobs {
  DenseJob j[kDenseJobs];
};

__global__ __launch_bounds__(kHashThreads) void k_dense_clear(DenseJobs J) {
  const DenseJob &d = J.j[blockIdx.y];
  const uint64_t words = (d.range + 31) / 32;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) d.bitmap[i] = 0u;
}

__global__ __launch_bounds__(kHashThreads) void k_dense_fill(DenseJobs J) {
  const DenseJob &d = J.j[blockIdx.y];
  uint64_t n = d.n;
  if (d.n_dev && *d.n_dev < n) n = *d.n_dev;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (d.slots8) {  // every occupied slot: (int32 key, int32 payload)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      const uint64_t w = d.slots8[i];
      if (w == kEmpty8) continue;
      const uint64_t idx = (uint64_t)((int64_t)(int32_t)(uint32_t)w - d.key_min);
      if (idx >= d.range) continue;
      atomicOr(&d.bitmap[idx >> 5], 1u << (idx & 31));
      d.dense[idx] = (int32_t)(uint32_t)(w >> 32);
    }
    return;
  }
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = d.sel ? d.sel[i] : i;
    int64_t k;
    if (!col_load_i64(d.key, d.key_type, r, k)) continue;
    const uint64_t idx = (uint64_t)(k - d.key_min);
    if (idx >= d.range) continue;
    int64_t p;
    if (d.pay) {
      if (!col_load_i64(d.pay, d.pay_type, r, p)) p = 0x7fffffffll;
    } else {
      p = (int64_t)r;
    }
    atomicOr(&d.bitmap[idx >> 5], 1u << (idx & 31));
    d.dense[idx] = (int32_t)p;
  }
}

__global__ void k_seq_i32(int32_t *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (int32_t)i;
}

// new[t][j] = old[t][pair_i[j]]; the joined table gets the build row (or -1).
__global__ void k_expand_tuples(const int32_t *old_col, const uint32_t *pair_i, uint64_t n, int32_t *new_col) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) new_col[j] = old_col[pair_i[j]];
}
__global__ void k_payload_to_i32(const int64_t *pay, uint64_t n, int32_t *out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) out[j] = (int32_t)pay[j];
}
__global__ void k_interleave(const int32_t *const *cols, uint32_t ncols, uint64_t n, int32_t *out) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n * ncols; j += stride) out[j] = cols[j % ncols][j / ncols];
}

static uint32_t grid_for(ghive_ctx *ctx, uint64_t n, uint32_t threads = kHashThreads) {
  uint64_t g = (n + threads - 1) / threads, cap = (uint64_t)ctx->cus * 16;
  return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

// Host-side statistics of a build (host[kStatWords] as k_build_stats writes them), then the dense
// companion of a unique INT table whose keys span a compact range (surrogate keys of a star
// schema): an existence bitmap over [key_min, key_max] and a payload array indexed by
// key - key_min. A probe is then one bit test (LDS-resident when small) plus, for survivors, one
// load -- no probe chain. Rows: sel[0 .. min(n, *n_dev)) (n_dev optional). When the table gets a
// companion, *job describes its fill (dense_launch) and *has_job is set.
static void finish_stats(ghive_ctx *ctx, ghive_hash_table *h, const uint64_t (&host)[kStatWords], const void *key,
                         const uint32_t *sel, uint64_t n, const uint64_t *n_dev, const void *pay, int pay_type,
                         DenseJob *job, bool *has_job) {
  *has_job = false;
  h->entries = host[0];
  h->unique = host[1] == 0;
  h->key_min = (int64_t)host[2];
  h->key_max = (int64_t)host[3];
  h->pay_min = (int64_t)host[4];
  h->pay_max = (int64_t)host[5];
  const uint64_t range = h->entries ? (uint64_t)(h->key_max - h->key_min) + 1 : 0;
  if (!(h->layout == 8 && h->unique && h->entries && range <= (1ull << 27) && range <= 64 * h->entries + 65536))
    return;
  const uint64_t nwords = (range + 31) / 32;
  h->bitmap = static_cast<uint32_t *>(ctx->arena->alloc(nwords * 4));
  h->dense = static_cast<int32_t *>(ctx->arena->alloc(range * 4));
  if (!h->bitmap || !h->dense) {  // no room: the table is probed through its slots
    if (h->bitmap) ctx->arena->free(h->bitmap);
    if (h->dense) ctx->arena->free(h->dense);
    h->bitmap = nullptr;
    h->dense = nullptr;
    return;
  }
  h->range = range;
  if (h->dense_from_slots)
    *job = DenseJob{key, nullptr, nullptr, pay, h->slots, static_cast<const uint64_t *>(h->data), h->key_min, range,
                    h->bitmap, h->dense, h->key_type, pay_type};
  else
    *job = DenseJob{key, sel, n_dev, pay, n, nullptr, h->key_min, range, h->bitmap, h->dense, h->key_type, pay_type};
  *has_job = true;
}

// The fills of njobs dense companions: one clearing launch and one filling launch per
// kDenseJobs tables.
static int dense_launch(ghive_ctx *ctx, const DenseJob *jobs, int njobs) {
  for (int b = 0; b < njobs; b += kDenseJobs) {
    const int nb = std::min(njobs - b, kDenseJobs);
    DenseJobs J;
    memset(&J, 0, sizeof J);
    uint64_t words = 1, rows = 1;
    for (int i = 0; i < nb; ++i) {
      J.j[i] = jobs[b + i];
      words = std::max(words, (jobs[b + i].range + 31) / 32);
      rows = std::max(rows, jobs[b + i].n);
    }
    k_dense_clear<<<dim3(grid_for(ctx, words), nb), kHashThreads, 0, ctx->stream>>>(J);
    k_dense_fill<<<dim3(grid_for(ctx, rows), nb), kHashThreads, 0, ctx->stream>>>(J);
    if (hipGetLastError() != hipSuccess) return set_error(ctx, GHIVE_ERR_HIP, "hash_build: dense launch failed");
  }
  return GHIVE_OK;
}

static int finish_build(ghive_ctx *ctx, ghive_hash_table *h, const uint64_t (&host)[kStatWords], const void *key,
                        const uint32_t *sel, uint64_t n, const uint64_t *n_dev, const void *pay, int pay_type) {
  DenseJob job;
  bool has_job = false;
  finish_stats(ctx, h, host, key, sel, n, n_dev, pay, pay_type, &job, &has_job);
  return has_job ? dense_launch(ctx, &job, 1) : GHIVE_OK;
}

// Rows sel[0 .. n) (n_dev: the count on the device, at most n). Synchronous unless `async`:
// then the statistics stay on the device and the table is finished on its first use.
// The regional build's launches (scratch allocated by the caller: W0 and W1 2n words each, starts
// 2^rbits + 1 words, one counter word, part (g1 + g2 + g3 + 1) x kStatWords words): products and
// payloads, the stable passes over the products' top bytes, region starts, regions, overflow rows,
// then the combined statistics at stats.
static int launch_regional(ghive_ctx *ctx, ghive_hash_table *h, const ghive_column *key, const uint32_t *sel,
                           uint64_t n, const void *pay, int pay_type, uint64_t *W0, uint64_t *W1, uint64_t *starts,
                           unsigned long long *ov_count, unsigned long long *part, uint32_t g1, uint32_t g2,
                           uint32_t g3, unsigned long long *stats) {
  const int rbits = h->log2_slots - kRegionLog2;
  const uint32_t nregions = 1u << rbits;
  k_build_words<<<g1, kHashThreads, 0, ctx->stream>>>(key->data, key->type, sel, n, pay, pay_type, h->layout, W0, part);
  GH_LAUNCH(ctx);
  std::vector<int> digits;  // the bytes of word 1 (the products) that hold their top rbits
  for (int b = (64 - rbits) / 8; b < 8; ++b) digits.push_back(8 + b);
  uint64_t *words = nullptr;
  uint32_t *unused = nullptr;
  const ghive_sort_info last_sort = ctx->sort_info;  // the last ORDER BY's record stays as it was
  const int sst = radix_sort_u64x2_ids(ctx, n, W0, W1, nullptr, nullptr, digits, &words, &unused, false);
  ctx->sort_info = last_sort;
  GH_TRY(sst);
  uint64_t *ov = words == W0 ? W1 : W0;  // the other buffer: 2n words for overflow rows
  GH_TRY(fill_u64(ctx, reinterpret_cast<uint64_t *>(ov_count), 0, 1));
  k_region_starts<<<(nregions + 1 + 255) / 256, 256, 0, ctx->stream>>>(words + n, n, rbits, starts);
  GH_LAUNCH(ctx);
  uint64_t *table = static_cast<uint64_t *>(h->data);
  if (h->layout == 16)
    k_region_build<16><<<g2, kRegionThreads, 0, ctx->stream>>>(words, n, starts, nregions, h->log2_slots, table, ov,
                                                                ov_count, n, part + (uint64_t)g1 * kStatWords);
  else
    k_region_build<8><<<g2, kRegionThreads, 0, ctx->stream>>>(words, n, starts, nregions, h->log2_slots, table, ov,
                                                               ov_count, n, part + (uint64_t)g1 * kStatWords);
  GH_LAUNCH(ctx);
  k_overflow_insert<<<g3, kHashThreads, 0, ctx->stream>>>(ov, n, ov_count, h->layout, table, h->slots - 1,
                                                         h->log2_slots, part + (uint64_t)(g1 + g2) * kStatWords);
  GH_LAUNCH(ctx);
  k_build_stats<<<1, kHashThreads, 0, ctx->stream>>>(part, g1 + g2 + g3, stats);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

static int hash_build_core(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
                           const uint64_t *n_dev, const ghive_column *payload, float load_factor, bool async,
                           ghive_hash_table **out) {
  if (!key || !out) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build: NULL argument");
  *out = nullptr;
  if (!is_int_type(key->type))
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "hash_build: key type %d (INT/LONG only)", key->type);
  if (payload && !is_int_type(payload->type))
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "hash_build: payload type %d (INT/LONG only)", payload->type);
  if (payload && payload->row_num != key->row_num)
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_build: payload and key row counts differ");
  if (key->row_num > kMaxRows) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "hash_build: too many rows");
  if (!sel && n && n != key->row_num)  // n == 0: an empty build (a selection that kept nothing)
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_build: n (%llu) != rows (%llu) without a selection",
                     (unsigned long long)n, (unsigned long long)key->row_num);
  if (n && !key->data) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build: key has no data");
  if (load_factor <= 0.f) load_factor = 0.5f;
  if (load_factor > 0.95f) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build: load factor %.3f > 0.95", load_factor);
  auto *h = new ghive_hash_table();
  h->owner = ctx;
  h->key_type = key->type;
  h->source_rows = key->row_num;
  h->payload_kind = payload ? (payload->type == GHIVE_INT ? 1 : 2) : 0;
  h->layout = (key->type == GHIVE_INT && h->payload_kind != 2) ? 8 : 16;
  uint64_t want = (uint64_t)((double)(n ? n : 1) / load_factor) + 1;
  h->slots = next_pow2(want < 64 ? 64 : want);
  h->log2_slots = __builtin_ctzll(h->slots);
  uint64_t words = h->slots * (h->layout == 8 ? 1 : 2);
  h->data = ctx->arena->alloc(words * 8);
  if (!h->data) {
    delete h;
    return set_error(ctx, GHIVE_ERR_OOM, "hash_build: cannot allocate %llu slots", (unsigned long long)h->slots);
  }
  int st = GHIVE_OK;
  // A large synchronous build runs in the regional form (every slot written by its region, no fill
  // pass); its scratch lives until the statistics are read back. Too little memory for the scratch:
  // the global form.
  DevBuf rw0, rw1, rstarts, rcount;
  bool regional = !async && !n_dev && ctx->tune.hash_regional_rows > 0 && n >= (uint64_t)ctx->tune.hash_regional_rows &&
                  h->log2_slots >= kRegionLog2 && h->log2_slots <= (h->layout == 8 ? 32 : 48);
  if (regional) {
    const uint64_t nregions = 1ull << (h->log2_slots - kRegionLog2);
    void *a = ctx->arena->alloc(16 * n), *b = ctx->arena->alloc(16 * n), *c = ctx->arena->alloc(8 * (nregions + 1)),
         *d = ctx->arena->alloc(8);
    regional = a && b && c && d;
    for (void *p : {a, b, c, d})
      if (p && !regional) ctx->arena->free(p);
    if (regional) {
      rw0.adopt(ctx, a);
      rw1.adopt(ctx, b);
      rstarts.adopt(ctx, c);
      rcount.adopt(ctx, d);
    }
  }
  if (regional) {
  } else if (h->layout == 8) {
    st = fill_u64(ctx, static_cast<uint64_t *>(h->data), kEmpty8, h->slots);
  } else {
    // {key = LONG_MAX, payload = LONG_MAX}: keys and payloads filled as one u64 array
    st = fill_u64(ctx, static_cast<uint64_t *>(h->data), kEmpty16, words);
  }
  const void *pay = payload ? payload->data : nullptr;
  const int pay_type = payload ? payload->type : 0;
  // statistics: per-workgroup partials combined by one small kernel, then one copy to the host
  uint64_t host[kStatWords] = {0, 0, (uint64_t)LLONG_MAX, (uint64_t)LLONG_MIN, (uint64_t)LLONG_MAX,
                               (uint64_t)LLONG_MIN};
  if (st == GHIVE_OK && n) {
    const uint32_t grid = grid_for(ctx, n);
    // regional form: partials of the product pass, the region workgroups and the overflow pass
    uint32_t g2 = 0, g3 = 0;
    if (regional) {
      int per_cu = 0;
      const void *kfn = h->layout == 16 ? reinterpret_cast<const void *>(k_region_build<16>)
                                        : reinterpret_cast<const void *>(k_region_build<8>);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kRegionThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
      const uint64_t nregions = 1ull << (h->log2_slots - kRegionLog2), slots = (uint64_t)ctx->cus * (uint32_t)per_cu;
      g2 = (uint32_t)(nregions < slots ? nregions : slots);
      g3 = (uint32_t)ctx->cus * 2;
    }
    const uint32_t nparts = grid + g2 + g3;
    DevBuf part;
    st = part.alloc(ctx, ((uint64_t)nparts + 1) * kStatWords * sizeof(uint64_t));
    unsigned long long *stats = st == GHIVE_OK ? part.as<unsigned long long>() + (uint64_t)nparts * kStatWords : nullptr;
    if (st == GHIVE_OK && regional) {
      st = launch_regional(ctx, h, key, sel, n, pay, pay_type, rw0.as<uint64_t>(), rw1.as<uint64_t>(),
                           rstarts.as<uint64_t>(), rcount.as<unsigned long long>(), part.as<unsigned long long>(),
                           grid, g2, g3, stats);
    } else if (st == GHIVE_OK) {
      k_hash_build<<<grid, kHashThreads, 0, ctx->stream>>>(key->data, key->type, sel, n, n_dev, pay, pay_type, h->layout,
                                                           static_cast<uint64_t *>(h->data), h->slots - 1, h->log2_slots,
                                                           part.as<unsigned long long>());
      k_build_stats<<<1, kHashThreads, 0, ctx->stream>>>(part.as<unsigned long long>(), grid, stats);
      if (hipGetLastError() != hipSuccess) st = set_error(ctx, GHIVE_ERR_HIP, "hash_build: launch failed");
    }
    if (st == GHIVE_OK && async) {
      h->pending = true;
      h->stats_dev = stats;
      h->stats_buf = part.release();
      h->in_key = key->data;
      h->in_pay = pay;
      h->in_pay_type = pay_type;
      h->in_sel = sel;
      h->in_n_dev = n_dev;
      h->in_n_max = n;
      ctx->tables.insert(h);
      *out = h;
      return GHIVE_OK;
    }
    if (st == GHIVE_OK) st = d2h(ctx, host, stats, sizeof host);
  }
  if (st != GHIVE_OK) {
    (void)hipStreamSynchronize(ctx->stream);  // no kernel may still write the memory given back
    ctx->arena->free(h->data);
    delete h;
    return st;
  }
  st = finish_build(ctx, h, host, key->data, sel, n, n_dev, pay, pay_type);
  if (st != GHIVE_OK) {
    ghive_hash_table_free(ctx, h);
    return st;
  }
  ctx->tables.insert(h);
  *out = h;
  return GHIVE_OK;
}

int hash_build(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
               const ghive_column *payload, float load_factor, ghive_hash_table **out) {
  return hash_build_core(ctx, key, sel, n, nullptr, payload, load_factor, false, out);
}

int hash_build_deferred(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
                        const ghive_column *payload, float load_factor, ghive_hash_table **out) {
  return hash_build_core(ctx, key, sel, n, nullptr, payload, load_factor, true, out);
}

// A table whose deferred finish failed keeps the status and message; consumers get them back.
static int fail_table(ghive_ctx *owner, ghive_hash_table *h, int st) {
  h->failed = st;
  h->fail_msg = owner->last_error;
  if (h->stats_buf) {
    (void)hipStreamSynchronize(owner->stream);  // no kernel may still write the memory given back
    owner->arena->free(h->stats_buf);
  }
  h->stats_buf = nullptr;
  h->stats_dev = nullptr;
  h->pending.store(false);
  return st;
}

static int table_failed(ghive_ctx *caller, const ghive_hash_table *h) {
  return set_error(caller, h->failed, "%s", h->fail_msg.c_str());
}

// A table built on another context: its owner's stream is waited for, so that the statistics and
// then the dense companion are complete before the caller's stream reads them.
static int ready_foreign(ghive_ctx *caller, ghive_hash_table *h) {
  std::lock_guard<std::mutex> lock(h->ready_mu);
  if (!h->pending.load()) return h->failed ? table_failed(caller, h) : GHIVE_OK;
  ghive_ctx *ctx = h->owner;
  if (!ctx) return set_error(caller, GHIVE_ERR_INVALID, "hash table: the context that built it was shut down");
  uint64_t host[kStatWords];
  int st = d2h(ctx, host, h->stats_dev, sizeof host);
  if (st == GHIVE_OK) {
    ctx->arena->free(h->stats_buf);
    h->stats_buf = nullptr;
    h->stats_dev = nullptr;
    st = finish_build(ctx, h, host, h->in_key, h->in_sel, h->in_n_max, h->in_n_dev, h->in_pay, h->in_pay_type);
  }
  if (st == GHIVE_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
    st = set_error(ctx, GHIVE_ERR_HIP, "hash table: waiting for the building context failed");
  if (st != GHIVE_OK) {
    fail_table(ctx, h, st);
    return table_failed(caller, h);
  }
  h->pending.store(false);
  return GHIVE_OK;
}

int hash_tables_ready(ghive_ctx *caller, const ghive_hash_table *const *hs, int n) {
  std::vector<ghive_hash_table *> mine;  // pending tables the caller built
  for (int i = 0; i < n; ++i) {
    auto *h = const_cast<ghive_hash_table *>(hs[i]);  // finishing is a deferred part of the build
    if (!h) continue;
    if (!h->pending.load()) {
      if (h->failed) return table_failed(caller, h);
      continue;
    }
    if (h->owner != caller) {
      GH_TRY(ready_foreign(caller, h));
      continue;
    }
    if (std::find(mine.begin(), mine.end(), h) == mine.end()) mine.push_back(h);
  }
  std::sort(mine.begin(), mine.end());  // one locking order for every caller
  constexpr size_t kBatch = kMaxReadWords / kStatWords;
  for (size_t b = 0; b < mine.size(); b += kBatch) {
    const size_t e = std::min(mine.size(), b + kBatch);
    std::vector<std::unique_lock<std::mutex>> locks;
    std::vector<ghive_hash_table *> todo;
    for (size_t i = b; i < e; ++i) {
      locks.emplace_back(mine[i]->ready_mu);
      if (mine[i]->pending.load()) todo.push_back(mine[i]);
    }
    int st = GHIVE_OK;
    for (size_t i = 0; i < todo.size() && st == GHIVE_OK; ++i)
      if (hipMemcpyAsync(caller->h_words + kStatWords * i, todo[i]->stats_dev, kStatWords * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, caller->stream) != hipSuccess)
        st = set_error(caller, GHIVE_ERR_HIP, "hash table: statistics read-back failed");
    if (st == GHIVE_OK && hipStreamSynchronize(caller->stream) != hipSuccess)
      st = set_error(caller, GHIVE_ERR_HIP, "hash table: statistics read-back failed");
    std::vector<DenseJob> jobs;  // the dense companions of these tables, filled together
    std::vector<ghive_hash_table *> done;
    for (size_t i = 0; i < todo.size(); ++i) {
      ghive_hash_table *h = todo[i];
      if (st != GHIVE_OK) {
        fail_table(caller, h, st);
        continue;
      }
      uint64_t host[kStatWords];
      for (int w = 0; w < kStatWords; ++w) host[w] = caller->h_words[kStatWords * i + w];
      caller->arena->free(h->stats_buf);  // the read-back above has completed
      h->stats_buf = nullptr;
      h->stats_dev = nullptr;
      DenseJob job;
      bool has_job = false;
      finish_stats(caller, h, host, h->in_key, h->in_sel, h->in_n_max, h->in_n_dev, h->in_pay, h->in_pay_type, &job,
                   &has_job);
      if (has_job) jobs.push_back(job);
      done.push_back(h);
    }
    if (st == GHIVE_OK && !jobs.empty()) {
      const int s = dense_launch(caller, jobs.data(), (int)jobs.size());
      if (s != GHIVE_OK) {
        for (ghive_hash_table *h : done) fail_table(caller, h, s);
        return s;
      }
    }
    for (ghive_hash_table *h : done) h->pending.store(false);
    if (st != GHIVE_OK) return st;
  }
  return GHIVE_OK;
}

int hash_table_ready(ghive_ctx *caller, const ghive_hash_table *h) { return hash_tables_ready(caller, &h, 1); }

// Probe into caller buffers; on overflow returns CAPACITY with the needed count in *count.
static int probe_run(ghive_ctx *ctx, const ghive_hash_table *h, const ghive_column *key,
                     const uint32_t *sel, const int32_t *tup, uint64_t n, int mode, int emit_index,
                     uint32_t *out_a, int64_t *out_b, uint64_t cap, uint64_t *count) {
  GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 1));
  if (n && mode != 2 && (h->unique || mode == 1)) {  // at most one output per probe row
    constexpr int R = 8;
    const uint64_t tiles = (n + (uint64_t)kProbeThreads * R - 1) / ((uint64_t)kProbeThreads * R);
    const uint64_t cap_grid = (uint64_t)ctx->cus * 8;
    const uint32_t grid = (uint32_t)(tiles < cap_grid ? tiles : cap_grid);
    auto *cnt = reinterpret_cast<unsigned long long *>(ctx->d_counters);
    const DevHash d = dev_hash(h);
    if (key->type == GHIVE_INT && h->layout == 8)
      k_probe_unique<R, GHIVE_INT, 8><<<grid, kProbeThreads, 0, ctx->stream>>>(d, key->data, sel, tup, n, emit_index,
                                                                              out_a, out_b, cap, cnt);
    else if (key->type == GHIVE_INT)
      k_probe_unique<R, GHIVE_INT, 16><<<grid, kProbeThreads, 0, ctx->stream>>>(d, key->data, sel, tup, n,
                                                                               emit_index, out_a, out_b, cap, cnt);
    else if (h->layout == 8)
      k_probe_unique<R, GHIVE_LONG, 8><<<grid, kProbeThreads, 0, ctx->stream>>>(d, key->data, sel, tup, n,
                                                                               emit_index, out_a, out_b, cap, cnt);
    else
      k_probe_unique<R, GHIVE_LONG, 16><<<grid, kProbeThreads, 0, ctx->stream>>>(d, key->data, sel, tup, n,
                                                                                emit_index, out_a, out_b, cap, cnt);
    GH_LAUNCH(ctx);
  } else if (n) {
    k_hash_probe<<<grid_for(ctx, n), kHashThreads, 0, ctx->stream>>>(
        dev_hash(h), key->data, key->type, sel, tup, n, mode, emit_index, h->unique, out_a, out_b,
        cap, reinterpret_cast<unsigned long long *>(ctx->d_counters));
    GH_LAUNCH(ctx);
  }
  GH_TRY(d2h(ctx, count, ctx->d_counters, sizeof(uint64_t)));
  return *count > cap ? GHIVE_ERR_CAPACITY : GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

namespace ghive {

void orphan_tables(ghive_ctx *ctx) {
  for (ghive_hash_table *t : ctx->tables) {
    t->owner = nullptr;
    t->data = nullptr;
    t->bitmap = nullptr;
    t->dense = nullptr;
  }
  ctx->tables.clear();
}

}  // namespace ghive

extern "C" {

int ghive_hash_build(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
                     const ghive_column *payload, float load_factor, ghive_hash_table **out) {
  GH_CHECK_CTX(ctx);
  return hash_build(ctx, key, sel, n, payload, load_factor, out);
}

int ghive_hash_build_async(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, const uint64_t *n_dev,
                           uint64_t n_max, const ghive_column *payload, float load_factor, ghive_hash_table **out) {
  GH_CHECK_CTX(ctx);
  if (!key || !out) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_async: NULL argument");
  if (!n_dev) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_async: n_dev is NULL");
  if (!sel && n_max != key->row_num)
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_async: n_max != rows without a selection");
  return hash_build_core(ctx, key, sel, n_max, n_dev, payload, load_factor, true, out);
}

int ghive_hash_build_filtered(ghive_ctx *ctx, const ghive_filtered_build *jobs, uint32_t n, ghive_hash_table **out) {
  GH_CHECK_CTX(ctx);
  if (!jobs || !out) return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_filtered: NULL argument");
  if (n == 0 || n > GHIVE_MAX_FILTERED_BUILDS)
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_filtered: %u builds (1..%d)", n, GHIVE_MAX_FILTERED_BUILDS);
  for (uint32_t i = 0; i < n; ++i) out[i] = nullptr;
  // every build is checked before anything is allocated or launched
  std::vector<DevPred> preds(n);
  for (uint32_t i = 0; i < n; ++i) {
    const ghive_filtered_build &J = jobs[i];
    GH_TRY(check_table(ctx, J.table, "hash_build_filtered"));
    const ghive_table &T = *J.table;
    if (J.key_col >= T.ncols || !is_int_type(T.columns[J.key_col].type))
      return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_filtered: build %u: key column %u (an INT/LONG column)", i,
                       J.key_col);
    if (J.payload_col >= 0 && ((uint32_t)J.payload_col >= T.ncols || !is_int_type(T.columns[J.payload_col].type)))
      return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_filtered: build %u: payload column %d (an INT/LONG column or -1)",
                       i, J.payload_col);
    if (J.load_factor > 0.95f)
      return set_error(ctx, GHIVE_ERR_INVALID, "hash_build_filtered: build %u: load factor %.3f > 0.95", i, J.load_factor);
    memset(&preds[i], 0, sizeof(DevPred));
    if (J.nnodes) GH_TRY(flatten_pred(ctx, J.table, J.nodes, J.nnodes, J.root, preds[i]));
  }
  // tables sized for every row (as an asynchronous build over a selection of unknown count); per
  // build one block of statistics partials; the dense companion comes from the slots
  std::vector<ghive_hash_table *> hs;
  std::vector<uint32_t> grids(n);
  auto undo = [&](int st) {
    for (ghive_hash_table *h : hs) {
      if (h->data) ctx->arena->free(h->data);
      if (h->stats_buf) ctx->arena->free(h->stats_buf);
      delete h;
    }
    return st;
  };
  TableFills F;
  TableStatsAll S;
  memset(&F, 0, sizeof F);
  memset(&S, 0, sizeof S);
  uint64_t max_words = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const ghive_filtered_build &J = jobs[i];
    const ghive_column &key = J.table->columns[J.key_col];
    const ghive_column *pay = J.payload_col >= 0 ? &J.table->columns[J.payload_col] : nullptr;
    const uint64_t rows = J.table->row_num;
    auto *h = new ghive_hash_table();
    hs.push_back(h);
    h->owner = ctx;
    h->key_type = key.type;
    h->source_rows = rows;
    h->payload_kind = pay ? (pay->type == GHIVE_INT ? 1 : 2) : 0;
    h->layout = (key.type == GHIVE_INT && h->payload_kind != 2) ? 8 : 16;
    const float lf = J.load_factor > 0.f ? J.load_factor : 0.5f;
    const uint64_t want = (uint64_t)((double)(rows ? rows : 1) / lf) + 1;
    h->slots = next_pow2(want < 64 ? 64 : want);
    h->log2_slots = __builtin_ctzll(h->slots);
    const uint64_t words = h->slots * (h->layout == 8 ? 1 : 2);
    grids[i] = grid_for(ctx, rows ? rows : 1);
    h->data = ctx->arena->alloc(words * 8);
    h->stats_buf = ctx->arena->alloc(((uint64_t)grids[i] + 1) * kStatWords * sizeof(uint64_t));
    if (!h->data || !h->stats_buf)
      return undo(set_error(ctx, GHIVE_ERR_OOM, "hash_build_filtered: cannot allocate build %u (%llu slots)", i,
                            (unsigned long long)h->slots));
    auto *part = static_cast<unsigned long long *>(h->stats_buf);
    F.t[i] = TableFill{static_cast<uint64_t *>(h->data), words, h->layout == 8 ? kEmpty8 : kEmpty16};
    S.t[i] = TableStats{part, grids[i], part + (uint64_t)grids[i] * kStatWords};
    max_words = words > max_words ? words : max_words;
    h->stats_dev = S.t[i].stats;
    h->in_key = key.data;
    h->in_pay = pay ? pay->data : nullptr;
    h->in_pay_type = pay ? pay->type : 0;
    h->in_sel = nullptr;
    h->in_n_dev = nullptr;
    h->in_n_max = rows;
    h->dense_from_slots = true;
  }
  const uint64_t fg = (max_words + 1023) / 1024, fcap = (uint64_t)ctx->cus * 8;
  k_fill_tables<<<dim3((uint32_t)(fg < 1 ? 1 : (fg > fcap ? fcap : fg)), n), 256, 0, ctx->stream>>>(F);
  if (hipGetLastError() != hipSuccess) return undo(set_error(ctx, GHIVE_ERR_HIP, "hash_build_filtered: launch failed"));
  for (uint32_t i = 0; i < n; ++i) {
    const ghive_filtered_build &J = jobs[i];
    const ghive_column &key = J.table->columns[J.key_col];
    const ghive_column *pay = J.payload_col >= 0 ? &J.table->columns[J.payload_col] : nullptr;
    ghive_hash_table *h = hs[i];
    k_filtered_build<<<grids[i], kHashThreads, 0, ctx->stream>>>(
        preds[i], key.data, key.type, pay ? pay->data : nullptr, pay ? pay->type : 0, J.table->row_num, h->layout,
        static_cast<uint64_t *>(h->data), h->slots - 1, h->log2_slots, static_cast<unsigned long long *>(h->stats_buf));
  }
  k_build_stats_tables<<<n, kHashThreads, 0, ctx->stream>>>(S);
  if (hipGetLastError() != hipSuccess) {
    (void)hipStreamSynchronize(ctx->stream);  // no kernel may still write the memory given back
    return undo(set_error(ctx, GHIVE_ERR_HIP, "hash_build_filtered: launch failed"));
  }
  for (uint32_t i = 0; i < n; ++i) {
    hs[i]->pending = true;
    ctx->tables.insert(hs[i]);
    out[i] = hs[i];
  }
  return GHIVE_OK;
}

int ghive_hash_table_info(const ghive_hash_table *ht, uint64_t *entries, uint64_t *slots,
                          uint64_t *bytes, int32_t *unique_keys) {
  if (!ht) return GHIVE_ERR_INVALID;
  if (ht->owner) GH_TRY(hash_table_ready(ht->owner, ht));
  if (entries) *entries = ht->entries;
  if (slots) *slots = ht->slots;
  if (bytes) *bytes = ht->slots * (ht->layout == 8 ? 8 : 16);
  if (unique_keys) *unique_keys = ht->unique;
  return GHIVE_OK;
}

void ghive_hash_table_free(ghive_ctx *ctx, ghive_hash_table *ht) {
  if (!ht) return;
  (void)ctx;
  ghive_ctx *c = ht->owner;  // the table's memory belongs to the context that built it (NULL:
  if (c) {                   // that context was shut down and released the memory already)
    // Stream-ordered release, no wait: the memory goes back to the owner's memory manager, whose
    // next user is work on the same stream, ordered after every kernel that read the table
    // (ghive_set_stream waits for the old stream before switching). Work of other contexts that
    // reads the table must have finished, as the API requires.
    (void)hipSetDevice(c->device);
    if (ht->data) c->arena->free(ht->data);
    if (ht->bitmap) c->arena->free(ht->bitmap);
    if (ht->dense) c->arena->free(ht->dense);
    if (ht->stats_buf) c->arena->free(ht->stats_buf);
    c->tables.erase(ht);
  }
  delete ht;
}

int ghive_hash_probe(ghive_ctx *ctx, const ghive_hash_table *ht, const ghive_column *key,
                     const uint32_t *sel, uint64_t n, uint32_t *out_probe_rows, int64_t *out_payload,
                     uint64_t capacity, uint64_t *out_count) {
  GH_CHECK_CTX(ctx);
  if (!ht || !key || !out_count) return set_error(ctx, GHIVE_ERR_INVALID, "hash_probe: NULL argument");
  if (!ht->owner) return set_error(ctx, GHIVE_ERR_INVALID, "hash_probe: the table's context was shut down");
  GH_TRY(hash_table_ready(ctx, ht));
  if (!is_int_type(key->type))
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "hash_probe: key type %d (INT/LONG only)", key->type);
  if (!sel && n != key->row_num)
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_probe: n != rows without a selection");
  if (key->row_num > kMaxRows) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "hash_probe: too many rows");
  if (capacity && (!out_probe_rows || !out_payload))
    return set_error(ctx, GHIVE_ERR_INVALID, "hash_probe: NULL output");
  int s = probe_run(ctx, ht, key, sel, nullptr, n, 0, 0, out_probe_rows, out_payload, capacity, out_count);
  if (s == GHIVE_ERR_CAPACITY)
    set_error(ctx, s, "hash_probe: %llu matches exceed capacity %llu", (unsigned long long)*out_count,
              (unsigned long long)capacity);
  return s;
}

int ghive_join(ghive_ctx *ctx, const ghive_table *tables, uint32_t ntables, const uint32_t *join_cols,
               const ghive_join_cond *conds, uint32_t nconds, int32_t *out_offsets, uint64_t capacity,
               uint64_t *out_cardinality) {
  GH_CHECK_CTX(ctx);
  if (!tables || !join_cols || !conds || !out_cardinality)
    return set_error(ctx, GHIVE_ERR_INVALID, "join: NULL argument");
  if (ntables < 2 || nconds != ntables - 1)
    return set_error(ctx, GHIVE_ERR_INVALID, "join: %u tables need %u conditions, got %u", ntables,
                     ntables ? ntables - 1 : 0, nconds);
  for (uint32_t t = 0; t < ntables; ++t) {
    GH_TRY(check_table(ctx, &tables[t], "join"));
    if (join_cols[t] >= tables[t].ncols)
      return set_error(ctx, GHIVE_ERR_INVALID, "join: table %u has no column %u", t, join_cols[t]);
    if (!is_int_type(tables[t].columns[join_cols[t]].type))
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "join: table %u key type %d (INT/LONG only)", t,
                       tables[t].columns[join_cols[t]].type);
  }
  std::vector<int> joined(ntables, 0), semi(ntables, 0);
  joined[0] = 1;
  for (uint32_t c = 0; c < nconds; ++c) {
    const ghive_join_cond &jc = conds[c];
    if (jc.type != GHIVE_INNER_JOIN && jc.type != GHIVE_LEFT_OUTER_JOIN && jc.type != GHIVE_LEFT_SEMI_JOIN)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "join: join type %d (INNER, LEFT OUTER, LEFT SEMI only)", jc.type);
    if (jc.left >= ntables || jc.right >= ntables || !joined[jc.left] || joined[jc.right])
      return set_error(ctx, GHIVE_ERR_INVALID, "join: condition %u (%u -> %u) is not left-deep", c, jc.left, jc.right);
    joined[jc.right] = 1;
    semi[jc.right] = jc.type == GHIVE_LEFT_SEMI_JOIN;
  }
  // tuples, column-major: tup[t][i] = row of table t in result row i (-1 = NULL)
  const uint64_t n0 = tables[0].row_num;
  std::vector<DevBuf> tup(ntables);
  uint64_t cur = n0;
  for (uint32_t t = 0; t < ntables; ++t) GH_TRY(tup[t].alloc(ctx, sizeof(int32_t) * (cur ? cur : 1)));
  if (cur) {
    k_seq_i32<<<grid_for(ctx, cur), 256, 0, ctx->stream>>>(tup[0].as<int32_t>(), cur);
    GH_LAUNCH(ctx);
  }
  for (uint32_t c = 0; c < nconds; ++c) {
    const ghive_join_cond &jc = conds[c];
    const uint32_t L = jc.left, R = jc.right;
    ghive_hash_table *ht = nullptr;
    GH_TRY(hash_build(ctx, &tables[R].columns[join_cols[R]], nullptr, tables[R].row_num, nullptr, 0.5f, &ht));
    const int mode = jc.type == GHIVE_LEFT_SEMI_JOIN ? 1 : (jc.type == GHIVE_LEFT_OUTER_JOIN ? 2 : 0);
    const ghive_column *lk = &tables[L].columns[join_cols[L]];
    uint64_t need = 0;
    // count pass, then the emit pass into exactly sized buffers
    int s = probe_run(ctx, ht, lk, nullptr, tup[L].as<int32_t>(), cur, mode, 1, nullptr, nullptr, 0, &need);
    if (s != GHIVE_OK && s != GHIVE_ERR_CAPACITY) {
      ghive_hash_table_free(ctx, ht);
      return s;
    }
    DevBuf pa, pb;
    s = pa.alloc(ctx, sizeof(uint32_t) * (need ? need : 1));
    if (s == GHIVE_OK) s = pb.alloc(ctx, sizeof(int64_t) * (need ? need : 1));
    uint64_t got = 0;
    if (s == GHIVE_OK)
      s = probe_run(ctx, ht, lk, nullptr, tup[L].as<int32_t>(), cur, mode, 1, pa.as<uint32_t>(),
                    pb.as<int64_t>(), need, &got);
    ghive_hash_table_free(ctx, ht);
    if (s != GHIVE_OK) return s;
    std::vector<DevBuf> nt(ntables);
    for (uint32_t t = 0; t < ntables; ++t) {
      GH_TRY(nt[t].alloc(ctx, sizeof(int32_t) * (got ? got : 1)));
      if (!got) continue;
      if (t == R) {
        k_payload_to_i32<<<grid_for(ctx, got), 256, 0, ctx->stream>>>(pb.as<int64_t>(), got, nt[t].as<int32_t>());
      } else {
        k_expand_tuples<<<grid_for(ctx, got), 256, 0, ctx->stream>>>(tup[t].as<int32_t>(), pa.as<uint32_t>(), got,
                                                                     nt[t].as<int32_t>());
      }
      GH_LAUNCH(ctx);
    }
    for (uint32_t t = 0; t < ntables; ++t) tup[t] = std::move(nt[t]);
    cur = got;
  }
  uint32_t nout = 0;
  std::vector<const int32_t *> cols;
  for (uint32_t t = 0; t < ntables; ++t)
    if (!semi[t]) {
      cols.push_back(tup[t].as<int32_t>());
      ++nout;
    }
  *out_cardinality = cur;
  if (cur > capacity)
    return set_error(ctx, GHIVE_ERR_CAPACITY, "join: cardinality %llu exceeds capacity %llu",
                     (unsigned long long)cur, (unsigned long long)capacity);
  if (cur) {
    if (!out_offsets) return set_error(ctx, GHIVE_ERR_INVALID, "join: out_offsets is NULL");
    DevBuf dcols;
    GH_TRY(dcols.alloc(ctx, sizeof(void *) * nout));
    GH_TRY(h2d(ctx, dcols.get(), cols.data(), sizeof(void *) * nout));
    k_interleave<<<grid_for(ctx, cur * nout), 256, 0, ctx->stream>>>(dcols.as<const int32_t *>(), nout, cur,
                                                                     out_offsets);
    GH_LAUNCH(ctx);
    GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GHIVE_OK;
}

}  // extern "C"
