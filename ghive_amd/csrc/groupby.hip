// Hash group-by (GroupByPredicate::gpu_execute, gpu_impl/src/Operator/GroupByPredicate.cu:19-359,
// which sorts row ids with a comparator and runs reduce_by_key; here one pass aggregates into
// LDS-private tables per workgroup that spill to and merge into one global table).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "agg.hpp"
#include "common.hpp"
#include "device.hpp"
#include "groupby.hpp"
#include "scan.hpp"

namespace ghive {

constexpr int kGbThreads = 256;

__device__ __forceinline__ uint64_t pack_key(const KeySpec &K, uint64_t r) {
  uint64_t key = 0;
  for (int k = 0; k < K.nkeys; ++k) {
    uint64_t bits;
    switch (K.type[k]) {
      case GHIVE_INT: bits = (uint32_t) static_cast<const int32_t *>(K.data[k])[r]; break;
      case GHIVE_LONG: bits = (uint64_t) static_cast<const int64_t *>(K.data[k])[r]; break;
      case GHIVE_DOUBLE: {
        double d = static_cast<const double *>(K.data[k])[r];
        bits = __builtin_isnan(d) ? 0x7ff8000000000000ull : (uint64_t)__double_as_longlong(d);
        break;
      }
      default: {
        float f = static_cast<const float *>(K.data[k])[r];
        bits = __builtin_isnan(f) ? 0x7fc00000u : (uint32_t)__float_as_uint(f);
        break;
      }
    }
    key |= bits << K.shift[k];
  }
  return key;
}

// Raw bits of row r of a column (4-byte types zero-extended).
__device__ __forceinline__ uint64_t col_bits(const void *p, int type, uint64_t r) {
  return (type == GHIVE_LONG || type == GHIVE_DOUBLE) ? static_cast<const uint64_t *>(p)[r]
                                                      : (uint64_t) static_cast<const uint32_t *>(p)[r];
}

// A value from its raw bits, as col_load_i64 / col_load_f64 read it; false when NULL.
__device__ __forceinline__ bool bits_i64(uint64_t bits, int type, int64_t &v) {
  if (type == GHIVE_INT) {
    const int32_t x = (int32_t)(uint32_t)bits;
    v = x;
    return x != 0x7fffffff;
  }
  v = (int64_t)bits;
  return v != 0x7fffffffffffffffll;
}
__device__ __forceinline__ bool bits_f64(uint64_t bits, int type, double &v) {
  switch (type) {
    case GHIVE_DOUBLE: v = __longlong_as_double((long long)bits); return !__builtin_isnan(v);
    case GHIVE_FLOAT: v = (double)__uint_as_float((uint32_t)bits); return !__builtin_isnan(v);
    case GHIVE_INT: {
      const int32_t x = (int32_t)(uint32_t)bits;
      v = (double)x;
      return x != 0x7fffffff;
    }
    default: {
      const int64_t x = (int64_t)bits;
      v = (double)x;
      return x != 0x7fffffffffffffffll;
    }
  }
}

// One row's contribution to aggregate A, from the raw bits of its input column.
__device__ __forceinline__ AggVal agg_value_bits(const AggSpec &A, uint64_t bits) {
  AggVal v;
  v.valid = false;
  v.neg = false;
  v.a = v.b = 0;
  if (A.col < 0) {
    v.valid = true;
    return v;
  }
  switch (A.kind) {
    case AK_CNT: {
      double d;
      v.valid = bits_f64(bits, A.in_type, d);
      break;
    }
    case AK_SUMI:
    case AK_MINI:
    case AK_MAXI: {
      int64_t x;
      v.valid = bits_i64(bits, A.in_type, x);
      v.a = (uint64_t)x;
      break;
    }
    case AK_MIND:
    case AK_MAXD: {
      double d;
      v.valid = bits_f64(bits, A.in_type, d);
      v.a = dbl_order(d);
      break;
    }
    default: {
      double d;
      v.valid = bits_f64(bits, A.in_type, d);
      if (v.valid) v.neg = to_fixed(d, A.scale_e, v.a, v.b);
      break;
    }
  }
  return v;
}

__device__ __forceinline__ AggVal agg_value(const AggSpec &A, uint64_t r) {
  return agg_value_bits(A, A.col < 0 ? 0 : col_bits(A.data, A.in_type, r));
}

// A workgroup aggregates into its LDS table first and merges it into the global table at the end.
// When its rows carry more distinct keys than the LDS table holds (a quarter of the tries miss
// after 8192 tries), it stops trying the LDS table and updates the global table directly.
__global__ __launch_bounds__(kGbThreads) void k_groupby(KeySpec K, GroupTableDev T, const uint32_t *sel,
                                                        uint64_t n, int lcap_log2) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t tries, misses, lds_off;
  const int lstride = (1 << lcap_log2) + 1;
  lds_table_init(lds, T, lcap_log2);
  if (threadIdx.x == 0) {
    tries = 0;
    misses = 0;
    lds_off = 0;
  }
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t r = sel ? sel[i] : i;
    const uint64_t key = pack_key(K, r);
    int ls = -1;
    if (!lds_off) {
      ls = lds_find_insert(lds, key, lcap_log2);
      const uint64_t miss = ballot(ls < 0), all = ballot(true);
      if (lane_id() == 0) {
        const uint32_t t = atomicAdd(&tries, (uint32_t)__popcll(all)) + (uint32_t)__popcll(all);
        const uint32_t m = atomicAdd(&misses, (uint32_t)__popcll(miss)) + (uint32_t)__popcll(miss);
        if (t >= 8192 && 4u * m > t) lds_off = 1;
      }
    }
    if (ls >= 0) {
      for (int a = 0; a < T.naggs; ++a) {
        const AggSpec &A = T.aggs[a];
        AggVal v = agg_value(A, r);
        acc_apply(A, [&](int j) { return lds + (size_t)(1 + A.word0 + j) * lstride + ls; }, v);
      }
    } else {
      int64_t g = gt_find_insert(T, key);
      if (g < 0) continue;
      for (int a = 0; a < T.naggs; ++a) {
        const AggSpec &A = T.aggs[a];
        AggVal v = agg_value(A, r);
        acc_apply(A, [&](int j) { return acc_word(T, A.word0 + j, (uint64_t)g); }, v);
      }
    }
  }
  __syncthreads();
  lds_table_flush(lds, T, lcap_log2);
}

// ---- high-cardinality form: hash partitions aggregated in LDS --------------------------------
// When the groups cannot repeat inside a workgroup's share of the rows, every row of the direct
// form costs device-scope atomics on a table far larger than L2. This form instead (1) writes
// (mix(key), row id) per row, (2) radix-sorts those pairs by the top bits of mix(key) so that each
// hash partition is a contiguous range, and (3) aggregates every partition in one workgroup's
// LDS table, then stores that table with plain stores into the partition's own region of the
// global table (regions are disjoint, so no atomics), where the usual finalize emits it.
constexpr int kGbPartThreads = 1024;  // one partition per workgroup; its LDS table allows one per CU
constexpr int kEmitRows = 8;          // rows per thread per step of the partition emit

// splitmix64 finalizer (Steele, Lea, Flood, OOPSLA 2014): every input bit reaches the top bits.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Inverse of mix64: the key comes back from its hash, so the passes need not carry it.
__device__ __forceinline__ uint64_t unmix64(uint64_t x) {
  x ^= (x >> 31) ^ (x >> 62);
  x *= 0x319642b2d24d8ec3ull;  // inverse of 0x94d049bb133111eb mod 2^64
  x ^= (x >> 27) ^ (x >> 54);
  x *= 0x96de1b173f119089ull;  // inverse of 0xbf58476d1ce4e5b9 mod 2^64
  return x ^ (x >> 30) ^ (x >> 60);
}

// h[i] = mix64(key of row r); v[i] = the raw bits of the carried value column (when vdata).
__global__ __launch_bounds__(256) void k_gb_part_keys(KeySpec K, const uint32_t *sel, uint64_t n, uint64_t *h,
                                                      uint32_t *ids, const void *vdata, int vtype, uint64_t *v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t r = sel ? sel[i] : i;
    h[i] = mix64(pack_key(K, r));
    if (ids) ids[i] = (uint32_t)r;
    if (vdata) v[i] = col_bits(vdata, vtype, r);
  }
}

// starts[p] = first position of partition p (top pbits of h) in h, which the partition passes
// leave ordered by those bits; starts[P] = n. One thread per partition runs a binary search
// (~log2 n dependent loads, the upper levels shared through L2) instead of a pass over every h.
__global__ void k_gb_part_bounds(const uint64_t *h, uint64_t n, int pbits, uint64_t *starts) {
  const uint64_t P = 1ull << pbits;
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P) return;
  uint64_t lo = 0, hi = n;  // the first i whose partition is >= q
  if (q == P) {
    lo = n;
  } else if (pbits) {
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((h[mid] >> (64 - pbits)) < q) lo = mid + 1;
      else hi = mid;
    }
  }
  starts[q] = lo;
}

// ---- wide keys: composite keys wider than 64 bits, more than kMaxKeys columns, or STRING -------
// A dictionary gives every row a dense group code; the 64-bit group-by then aggregates on the
// code, and each group's key values are gathered from its representative row. A dictionary slot
// holds (32 hash bits << 32 | representative row id): the first row of a key to claim a slot
// represents it, and a row matches a slot when the hash bits agree and its key columns equal the
// representative's. Slots never change once claimed and the input columns are immutable, so a
// reader never waits for a writer. Row ids stay below 2^32 - 1, so no slot word equals EMPTY.
struct WideOut {
  void *col[kMaxWideKeys];
  int32_t *aux[kMaxWideKeys];
};

// Key bits as the 64-bit form compares them: every NaN is the one NULL.
__device__ __forceinline__ uint64_t wide_bits(int type, const void *p, uint64_t r) {
  switch (type) {
    case GHIVE_INT: return (uint32_t) static_cast<const int32_t *>(p)[r];
    case GHIVE_LONG: return (uint64_t) static_cast<const int64_t *>(p)[r];
    case GHIVE_DOUBLE: {
      const double d = static_cast<const double *>(p)[r];
      return __builtin_isnan(d) ? 0x7ff8000000000000ull : (uint64_t)__double_as_longlong(d);
    }
    default: {
      const float f = static_cast<const float *>(p)[r];
      return __builtin_isnan(f) ? 0x7fc00000u : (uint32_t)__float_as_uint(f);
    }
  }
}

__device__ uint64_t wide_hash(const WideKeySpec &K, uint64_t r) {
  uint64_t h = 0x243f6a8885a308d3ull;
  for (int k = 0; k < K.nkeys; ++k) {
    if (K.type[k] == GHIVE_STRING) {
      const int32_t s = K.aux[k][2 * r], e = K.aux[k][2 * r + 1];
      uint64_t x = 0x9e3779b97f4a7c15ull;  // NULL
      if (s >= 0) {
        const uint8_t *b = static_cast<const uint8_t *>(K.data[k]);
        x = 0xcbf29ce484222325ull;  // FNV-1a over the bytes, then the length
        for (int32_t i = s; i < e; ++i) x = (x ^ b[i]) * 0x100000001b3ull;
        x ^= (uint64_t)(uint32_t)(e - s);
      }
      h = mix64(h ^ x);
    } else {
      h = mix64(h ^ wide_bits(K.type[k], K.data[k], r));
    }
  }
  return h;
}

__device__ bool wide_equal(const WideKeySpec &K, uint64_t r1, uint64_t r2) {
  for (int k = 0; k < K.nkeys; ++k) {
    if (K.type[k] == GHIVE_STRING) {
      const int32_t s1 = K.aux[k][2 * r1], e1 = K.aux[k][2 * r1 + 1];
      const int32_t s2 = K.aux[k][2 * r2], e2 = K.aux[k][2 * r2 + 1];
      if (s1 < 0 || s2 < 0) {
        if ((s1 < 0) != (s2 < 0)) return false;
        continue;
      }
      if (e1 - s1 != e2 - s2) return false;
      const uint8_t *b = static_cast<const uint8_t *>(K.data[k]);
      for (int32_t i = 0; i < e1 - s1; ++i)
        if (b[s1 + i] != b[s2 + i]) return false;
    } else if (wide_bits(K.type[k], K.data[k], r1) != wide_bits(K.type[k], K.data[k], r2)) {
      return false;
    }
  }
  return true;
}

// codes[r] = dictionary slot of row r's key. cnt[0] counts keys, cnt[1] is set on overflow.
__global__ __launch_bounds__(256) void k_gbw_codes(WideKeySpec K, uint64_t *dict, uint64_t mask, uint64_t cap,
                                                   unsigned long long *cnt, const uint32_t *sel, uint64_t n,
                                                   uint32_t *codes) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t r = sel ? sel[i] : i;
    const uint64_t h = wide_hash(K, r);
    const uint64_t tag = h >> 32, mine = (tag << 32) | r;
    uint64_t s = h & mask;
    bool found = false;
    for (uint64_t step = 0; step <= mask; ++step) {
      uint64_t w = __atomic_load_n(dict + s, __ATOMIC_RELAXED);
      if (w == kEmptyKey) {
        const unsigned long long old = atomicCAS((unsigned long long *)(dict + s), kEmptyKey, mine);
        if (old == kEmptyKey) {
          if (atomicAdd(cnt, 1ull) >= cap) atomicOr(cnt + 1, 1ull);
          found = true;
          break;
        }
        w = old;
      }
      if ((w >> 32) == tag && wide_equal(K, w & 0xffffffffull, r)) {
        found = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!found) atomicOr(cnt + 1, 1ull);
    codes[r] = (uint32_t)s;
  }
}

// Key columns of group g from the representative row of its code.
__global__ __launch_bounds__(256) void k_gbw_emit_keys(WideKeySpec K, const uint64_t *dict, const int32_t *gcode,
                                                       uint64_t groups, WideOut O) {
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; g < groups; g += stride) {
    const uint64_t r = dict[(uint32_t)gcode[g]] & 0xffffffffull;
    for (int k = 0; k < K.nkeys; ++k) {
      switch (K.type[k]) {
        case GHIVE_LONG:
        case GHIVE_DOUBLE:
          static_cast<uint64_t *>(O.col[k])[g] = static_cast<const uint64_t *>(K.data[k])[r];
          break;
        case GHIVE_STRING:
          O.aux[k][2 * g] = K.aux[k][2 * r];
          O.aux[k][2 * g + 1] = K.aux[k][2 * r + 1];
          break;
        default: static_cast<uint32_t *>(O.col[k])[g] = static_cast<const uint32_t *>(K.data[k])[r]; break;
      }
    }
  }
}

int key_dictionary(ghive_ctx *ctx, const WideKeySpec &K, uint64_t rows, const uint32_t *sel, uint64_t n,
                   uint64_t capacity, DevBuf &dict, DevBuf &codes, uint64_t *slots, uint64_t *count) {
  if (capacity > (1ull << 31))
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "key dictionary: more than 2^31 distinct keys");
  *slots = next_pow2(capacity * 2 < 64 ? 64 : capacity * 2);
  *count = 0;
  GH_TRY(dict.alloc(ctx, *slots * 8));
  GH_TRY(codes.alloc(ctx, (rows ? rows : 1) * 4));
  GH_TRY(fill_u64(ctx, dict.as<uint64_t>(), kEmptyKey, *slots));
  GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 2));
  if (n) {
    const uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 16;
    k_gbw_codes<<<(uint32_t)(g > cap ? cap : g), 256, 0, ctx->stream>>>(
        K, dict.as<uint64_t>(), *slots - 1, capacity, reinterpret_cast<unsigned long long *>(ctx->d_counters), sel,
        n, codes.as<uint32_t>());
    GH_LAUNCH(ctx);
  }
  uint64_t c[2] = {0, 0};
  GH_TRY(d2h(ctx, c, ctx->d_counters, sizeof c));
  *count = c[0];
  if (c[1] || c[0] > capacity)
    return set_error(ctx, GHIVE_ERR_CAPACITY, "key dictionary: more than %llu distinct keys",
                     (unsigned long long)capacity);
  return GHIVE_OK;
}

__global__ void k_gt_preinsert(GroupTableDev T, uint64_t key) {
  if (threadIdx.x == 0 && blockIdx.x == 0) (void)gt_find_insert(T, key);
}

// max |x| over the selected rows of a floating column, as ordered bits of a non-negative double
__global__ void k_absmax(const void *data, int type, const uint32_t *sel, uint64_t n, unsigned long long *out) {
  uint64_t m = 0;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double d;
    if (col_load_f64(data, type, sel ? sel[i] : i, d)) {
      uint64_t b = (uint64_t)__double_as_longlong(fabs(d));
      m = b > m ? b : m;
    }
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    uint64_t o = __shfl_xor((unsigned long long)m, d, kWave);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m) atomicMax(out, (unsigned long long)m);
}

int column_absmax(ghive_ctx *ctx, const void *data, int type, const uint32_t *sel, uint64_t n, double *out) {
  uint64_t m = 0;
  if (n) {
    GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 1));
    const uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 8;
    k_absmax<<<(uint32_t)(g > cap ? cap : g), 256, 0, ctx->stream>>>(data, type, sel, n,
                                                                     reinterpret_cast<unsigned long long *>(ctx->d_counters));
    GH_LAUNCH(ctx);
    GH_TRY(d2h(ctx, &m, ctx->d_counters, sizeof m));
  }
  memcpy(out, &m, sizeof *out);
  return GHIVE_OK;
}

int sum_scale(ghive_ctx *ctx, double mx, uint64_t n, int32_t *scale_e) {
  if (std::isinf(mx)) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "SUM over a double whose magnitude overflows");
  if (!(mx > 0.0)) {
    *scale_e = -1074;
    return GHIVE_OK;
  }
  int lg_n = 0;
  while ((1ull << lg_n) < n + 1 && lg_n < 63) ++lg_n;
  // 2^(E + 126) > n * max |x|: the positive and the negative 128-bit halves never overflow
  const int e = std::ilogb(mx) + 1 + lg_n - 126;
  *scale_e = e < -1074 ? -1074 : e;
  return GHIVE_OK;
}

__device__ __forceinline__ void write_typed(void *col, int type, uint64_t idx, int64_t iv, double dv, bool is_int) {
  switch (type) {
    case GHIVE_LONG: static_cast<int64_t *>(col)[idx] = is_int ? iv : (int64_t)dv; break;
    case GHIVE_INT: static_cast<int32_t *>(col)[idx] = (int32_t)iv; break;
    case GHIVE_DOUBLE: static_cast<double *>(col)[idx] = is_int ? (double)iv : dv; break;
    default: static_cast<float *>(col)[idx] = (float)(is_int ? (double)iv : dv); break;
  }
}

__device__ __forceinline__ void write_null(void *col, int type, uint64_t idx) {
  switch (type) {
    case GHIVE_LONG: static_cast<int64_t *>(col)[idx] = 0x7fffffffffffffffll; break;
    case GHIVE_INT: static_cast<int32_t *>(col)[idx] = 0x7fffffff; break;
    case GHIVE_DOUBLE: static_cast<double *>(col)[idx] = __longlong_as_double(0x7ff8000000000000ll); break;
    default: static_cast<float *>(col)[idx] = __uint_as_float(0x7fc00000u); break;
  }
}

// Output value of aggregate A from its accumulator words W(j), at row idx of col.
template <class Words>
__device__ __forceinline__ void finalize_words(const AggSpec &A, Words W, void *col, uint64_t idx) {
  switch (A.kind) {
    case AK_CNT: static_cast<int64_t *>(col)[idx] = (int64_t)W(0); return;
    case AK_SUMI: {
      uint64_t c = W(1);
      if (!c) return write_null(col, A.out_type, idx);
      if (A.is_avg) static_cast<double *>(col)[idx] = (double)(int64_t)W(0) / (double)c;
      else static_cast<int64_t *>(col)[idx] = (int64_t)W(0);
      return;
    }
    case AK_MINI:
    case AK_MAXI: {
      if (!W(1)) return write_null(col, A.out_type, idx);
      write_typed(col, A.out_type, idx, (int64_t)W(0), 0.0, true);
      return;
    }
    case AK_MIND:
    case AK_MAXD: {
      if (!W(1)) return write_null(col, A.out_type, idx);
      write_typed(col, A.out_type, idx, 0, dbl_unorder(W(0)), false);
      return;
    }
    default: {
      uint64_t c = W(2);
      if (!c) return write_null(col, A.out_type, idx);
      uint64_t lo = W(0), hi = W(1);  // two's complement
      const bool neg = hi >> 63;
      if (neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0 ? 1ull : 0ull);
      }
      double d = fixed_to_double(lo, hi, A.scale_e);
      if (neg) d = -d;
      if (A.is_avg) d /= (double)c;
      static_cast<double *>(col)[idx] = d;
      return;
    }
  }
}

__device__ __forceinline__ void finalize_agg(const AggSpec &A, const GroupTableDev &T, uint64_t s, void *col,
                                             uint64_t idx) {
  finalize_words(A, [&](int j) { return *acc_word(T, A.word0 + j, s); }, col, idx);
}

// The key columns of a packed key at output row idx.
__device__ __forceinline__ void emit_key(const KeySpec &K, const OutSpec &O, uint64_t key, uint64_t idx) {
  for (int k = 0; k < K.nkeys; ++k) {
    const uint64_t bits = key >> K.shift[k];
    switch (K.type[k]) {
      case GHIVE_INT: static_cast<int32_t *>(O.col[k])[idx] = (int32_t)(uint32_t)bits; break;
      case GHIVE_LONG: static_cast<int64_t *>(O.col[k])[idx] = (int64_t)bits; break;
      case GHIVE_DOUBLE: static_cast<uint64_t *>(O.col[k])[idx] = bits; break;
      default: static_cast<uint32_t *>(O.col[k])[idx] = (uint32_t)bits; break;
    }
  }
}

__global__ void k_groupby_finalize(KeySpec K, GroupTableDev T, OutSpec O, unsigned long long *counter, uint64_t cap) {
  const uint64_t total = T.stride;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t iters = (total + stride - 1) / stride;
  for (uint64_t it = 0; it < iters; ++it) {
    uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + it * stride;
    bool used = false;
    uint64_t key = 0;
    if (s < total) {
      key = T.keys[s];
      used = (s == total - 1) ? key == 1ull : key != kEmptyKey;
      if (s == total - 1) key = kEmptyKey;
    }
    uint64_t idx = wave_append(counter, used);
    if (!used || idx >= cap) continue;
    emit_key(K, O, key, idx);
    for (int a = 0; a < T.naggs; ++a) finalize_agg(T.aggs[a], T, s, O.col[K.nkeys + a], idx);
  }
}

// ---- high-cardinality form, aggregation: one workgroup per hash partition --------------------
// The rows of partition p are [starts[p], starts[p+1]) of the sorted pairs: the key is unmix64 of
// the carried hash, the aggregate input the carried value bits (VAL 1), the input columns at the
// row id (VAL 2), or nothing (VAL 0: COUNT(*) only). The partition's groups are aggregated in an
// LDS table of lcap slots (unbounded linear probing); more than 7/8 of that sets *fail and the
// caller uses the direct form. Then the workgroup reserves its range of output rows with one
// atomic and writes its groups there: no global table is built. The one key equal to the EMPTY
// pattern is aggregated in the small global table T instead, emitted afterwards.
// sum_nocount: the carried value column holds no NULL and no exact double SUM is an AVG: every
// group then has a non-NULL value, and a SUM's non-NULL count word is neither updated nor read
// (one LDS atomic a row less).
// NA (1..4): the number of aggregates, fixed when the kernel is compiled, so that every loop over
// them unrolls and their descriptors stay in scalar registers (the per-row descriptor reloads and
// switches were most of the kernel's instructions); 0: T.naggs at run time.
template <int VAL, int NA>
__global__ __launch_bounds__(kGbPartThreads) void k_gb_part_emit(KeySpec K, GroupTableDev T, const uint64_t *starts,
                                                                 const uint64_t *hs, const uint64_t *vs,
                                                                 const uint32_t *ids, uint32_t lcap, uint32_t nparts,
                                                                 unsigned long long *fail,
                                                                 unsigned long long *out_count, OutSpec O,
                                                                 uint64_t cap, int sum_nocount) {
  // One resident workgroup per CU takes partitions blockIdx.x, blockIdx.x + gridDim.x, ... in turn
  // (its table fills most of the CU's LDS; a new workgroup per partition would leave the CU idle
  // between them). The partition's table: keys[lcap] (EMPTY = all ones), then the accumulator
  // words, lcap + 1 apart (slot lcap is unused: the one key equal to EMPTY goes to the small
  // global table).
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t used, full, wpos;
  __shared__ unsigned long long base;
  const uint32_t lstride = lcap + 1;
  const uint32_t limit = lcap - lcap / 8;
  auto for_aggs = [&](auto &&f) {  // f(a) for every aggregate a
    if constexpr (NA > 0) {
#pragma unroll
      for (int a = 0; a < NA; ++a) f(a);
    } else {
      for (int a = 0; a < T.naggs; ++a) f(a);
    }
  };
  for (uint32_t part = blockIdx.x; part < nparts; part += gridDim.x) {
    for (uint32_t s = threadIdx.x; s < lstride; s += blockDim.x) lds[s] = kEmptyKey;
    for_aggs([&](int a) {
      const AggSpec &A = T.aggs[a];
      for (int w = 0; w < agg_words(A.kind); ++w) {
        const uint64_t v = init_word(A.kind, w);
        uint64_t *p = lds + (size_t)(1 + A.word0 + w) * lstride;
        for (uint32_t s = threadIdx.x; s < lstride; s += blockDim.x) p[s] = v;
      }
    });
    if (threadIdx.x == 0) {
      used = 0;
      full = 0;
      wpos = 0;
    }
    __syncthreads();
    const uint64_t b = starts[part], e = starts[part + 1];
    // kEmitRows rows a thread per step, all their loads issued first: the loads in flight then
    // cover HBM latency (one row a thread leaves a CU a few KiB in flight)
    for (uint64_t i0 = b; i0 < e && !full; i0 += (uint64_t)blockDim.x * kEmitRows) {
      uint64_t hv[kEmitRows], vv[kEmitRows], rv[kEmitRows];
#pragma unroll
      for (int q = 0; q < kEmitRows; ++q) {
        const uint64_t i = i0 + (uint64_t)q * blockDim.x + threadIdx.x;
        const uint64_t c = i < e ? i : e - 1;  // past the end: the last row again (not used)
        hv[q] = hs[c];
        vv[q] = VAL == 1 ? vs[c] : 0;
        rv[q] = VAL == 2 ? ids[c] : 0;
      }
#pragma unroll
      for (int q = 0; q < kEmitRows; ++q) {
        if (i0 + (uint64_t)q * blockDim.x + threadIdx.x >= e || full) break;
        // the table holds h = mix64(key) (a bijection): no unmix per row, and h's low bits give the
        // slot (its top bits are the partition's, equal for every row here)
        const uint64_t h = hv[q];
        const uint64_t vbits = vv[q];
        const uint64_t r = rv[q];
        if (h == kEmptyKey) {  // the one key whose hash is the EMPTY pattern: the small global table
          const int64_t g = gt_find_insert(T, unmix64(h));
          if (g < 0) continue;
          for_aggs([&](int a) {
            const AggSpec &A = T.aggs[a];
            const AggVal v = VAL == 2 ? agg_value(A, r) : agg_value_bits(A, vbits);
            acc_apply(A, [&](int j) { return acc_word(T, A.word0 + j, (uint64_t)g); }, v);
          });
          continue;
        }
        // slot: the low 32 bits of h scaled to [0, lcap), then linear probing
        uint32_t s = (uint32_t)(((h & 0xffffffffull) * (uint64_t)lcap) >> 32);
        int slot = -1;
        for (uint32_t step = 0; step < lcap; ++step) {
          const uint64_t k = __atomic_load_n(lds + s, __ATOMIC_RELAXED);
          if (k == h) {
            slot = (int)s;
            break;
          }
          if (k == kEmptyKey) {
            const unsigned long long old = atomicCAS((unsigned long long *)(lds + s), kEmptyKey, h);
            if (old == kEmptyKey) {
              if (atomicAdd(&used, 1u) + 1u > limit) full = 1;
              slot = (int)s;
              break;
            }
            if (old == h) {
              slot = (int)s;
              break;
            }
          }
          s = s + 1 == lcap ? 0 : s + 1;
        }
        if (slot < 0) {
          full = 1;
          break;
        }
        for_aggs([&](int a) {
          const AggSpec &A = T.aggs[a];
          const AggVal v = VAL == 2 ? agg_value(A, r) : agg_value_bits(A, vbits);
          auto word = [&](int j) { return lds + (size_t)(1 + A.word0 + j) * lstride + slot; };
          if (sum_nocount && A.kind == AK_SUMD) {
            if (!v.valid) return;
            uint64_t lo = v.a, hi = v.b;
            if (v.neg) {
              lo = ~lo + 1ull;
              hi = ~hi + (lo == 0 ? 1ull : 0ull);
            }
            add128(word(0), word(1), lo, hi);
            return;
          }
          acc_apply(A, word, v);
        });
      }
    }
    __syncthreads();
    if (full) {  // uniform: the whole result is discarded (the direct form runs instead)
      if (threadIdx.x == 0) atomicOr(fail, 1ull);
      return;
    }
    if (threadIdx.x == 0) base = atomicAdd(out_count, (unsigned long long)used);
    __syncthreads();
    for (uint32_t s0 = 0; s0 < lcap; s0 += blockDim.x) {  // the same trip count in every lane
      const uint32_t s = s0 + threadIdx.x;
      const uint64_t h = s < lcap ? lds[s] : kEmptyKey;
      // output rows: one LDS atomic per wave, ranks from the ballot
      const uint64_t m = ballot(h != kEmptyKey);
      uint32_t wbase = 0;
      if (lane_id() == 0 && m) wbase = atomicAdd(&wpos, (uint32_t)__popcll(m));
      wbase = __shfl(wbase, 0, kWave);
      if (h == kEmptyKey) continue;
      const uint64_t idx = base + wbase + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
      if (idx >= cap) continue;
      emit_key(K, O, unmix64(h), idx);
      for_aggs([&](int a) {
        const AggSpec &A = T.aggs[a];
        const bool one = sum_nocount && A.kind == AK_SUMD;  // the count word reads as non-zero
        finalize_words(A, [&](int j) { return one && j == 2 ? 1ull : lds[(size_t)(1 + A.word0 + j) * lstride + s]; },
                       O.col[K.nkeys + a], idx);
      });
    }
    __syncthreads();  // the table is reset for the next partition only after every slot was read
  }
}

int lds_log2_for(int nwords, int budget_bytes) {
  int lg = 12;
  while (lg > 4 && (((1 << lg) + 1) * (1 + nwords) * 8) > budget_bytes) --lg;
  return lg;
}

void group_table_size(GroupTableDev &T, uint64_t capacity, uint64_t *key_bytes, uint64_t *acc_bytes) {
  const uint64_t slots = next_pow2(capacity * 2 < 64 ? 64 : capacity * 2);
  T.stride = slots + 1;
  T.mask = slots - 1;
  T.log2 = __builtin_ctzll(slots);
  T.capacity = capacity;
  T.rec = record_words(T.nwords);
  *key_bytes = T.stride * 8;
  *acc_bytes = T.stride * 8 * (uint64_t)T.rec;
}

int group_table_alloc(ghive_ctx *ctx, GroupTableDev &T, uint64_t capacity, DevBuf &keys, DevBuf &acc,
                      DevBuf &counters) {
  uint64_t kb = 0, ab = 0;
  group_table_size(T, capacity, &kb, &ab);
  GH_TRY(keys.alloc(ctx, kb));
  GH_TRY(acc.alloc(ctx, ab));
  GH_TRY(counters.alloc(ctx, 2 * sizeof(unsigned long long)));
  return group_table_init(ctx, T, keys.as<uint64_t>(), acc.as<uint64_t>(), counters.as<unsigned long long>());
}

// The whole initial state of a group table in one launch: keys EMPTY (the extra slot 0), every
// record's accumulator words at their identities (init[w]: MIN / MAX start from the extreme
// value, the rest from 0), and the table's counters 0.
struct TableInit {
  uint64_t init[32];  // per record word (rec <= 32)
};
__global__ void k_group_table_init(uint64_t *keys, uint64_t *acc, uint64_t stride, uint32_t rec, TableInit I,
                                   unsigned long long *counters, uint32_t ncounters) {
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = i0; i < stride; i += step) keys[i] = i + 1 < stride ? kEmptyKey : 0ull;
  for (uint64_t i = i0; i < stride * rec; i += step) acc[i] = I.init[i % rec];
  if (i0 < ncounters) counters[i0] = 0ull;
}

int group_table_init(ghive_ctx *ctx, GroupTableDev &T, uint64_t *keys, uint64_t *acc, unsigned long long *counters,
                     uint32_t ncounters) {
  T.keys = keys;
  T.acc = acc;
  T.nent = counters;
  T.overflow = T.nent + 1;
  TableInit I;
  memset(&I, 0, sizeof I);
  if (T.rec > 32) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "group table: %d words per record", T.rec);
  for (int a = 0; a < T.naggs; ++a) {  // MIN / MAX start from their identities
    const AggSpec &A = T.aggs[a];
    if (A.kind == AK_MINI) I.init[A.word0] = 0x7fffffffffffffffull;
    if (A.kind == AK_MAXI) I.init[A.word0] = 0x8000000000000000ull;
    if (A.kind == AK_MIND) I.init[A.word0] = ~0ull;
  }
  const uint64_t words = T.stride * (uint64_t)T.rec;
  const uint64_t g = (words + 255) / 256, cap = (uint64_t)ctx->cus * 8;
  k_group_table_init<<<(uint32_t)(g < 1 ? 1 : (g > cap ? cap : g)), 256, 0, ctx->stream>>>(
      T.keys, T.acc, T.stride, (uint32_t)T.rec, I, counters, ncounters < 2 ? 2 : ncounters);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

// ---- partitioning of the high-cardinality form ---------------------------------------------------
// Stable LSD passes over the partition bits of h = mix64(key) (its top pbits bits), DB = 8 or 9
// bits a pass, carrying the aggregated value's bits beside h (VAL). The first pass reads the key
// and value columns and computes h itself, so no separate hashing pass writes (h, value) first.
// Per pass: a per-tile digit histogram, one scan of the digit x tile matrix, then the scatter: every
// element ranked among the equal digits of its wave with ballots (stable), the tile staged through
// LDS in digit order and each digit's run written to consecutive addresses. A 9-bit pass uses
// 4096-element tiles, an 8-bit pass 8192.
// TH threads a workgroup (256, 512 or 1024, option groupby.part_threads; 512 by default, 4 ms
// faster than 256 at 1e9 rows): the same tile in every case.
template <int DB, int TH>
struct Gbp {
  static constexpr int kBins = 1 << DB;
  static constexpr uint32_t kTile = DB == 9 ? 4096u : 8192u;
  static constexpr int kPer = (int)(kTile / TH);
  static constexpr int kBpt = kBins >= TH ? kBins / TH : 1;  // bins per thread (threads >= kBins: one)
  typedef typename std::conditional<DB == 9, uint16_t, uint8_t>::type Digit;
};

struct GbpSrc {
  KeySpec K;              // the first pass: key columns,
  const uint32_t *sel;    // the rows (NULL: all),
  const void *vdata;      // the value column
  int vtype;
  const uint64_t *h;      // later passes: the previous pass's output
  const uint64_t *v;
};

// Where a pass reads: kSrcPass the previous pass's (h, v); kSrcCols the key and value columns at the
// selected rows (h computed); kSrcCol64 one 8-byte key column and an 8-byte value column over all
// rows, read as plain words so that every load of a tile issues before any hash is computed.
enum GbpSrcKind { kSrcPass = 0, kSrcCols = 1, kSrcCol64 = 2 };

template <int SRC>
__device__ __forceinline__ uint64_t gbp_hash(const GbpSrc &S, uint64_t e) {
  if (SRC == kSrcPass) return S.h[e];
  if (SRC == kSrcCol64) return mix64(static_cast<const uint64_t *>(S.K.data[0])[e]);
  return mix64(pack_key(S.K, S.sel ? S.sel[e] : e));
}

template <int DB, int SRC, int TH>
__global__ __launch_bounds__(TH) void k_gbp_hist(GbpSrc S, uint64_t n, int shift, uint32_t *hist, uint32_t *counts,
                                                 uint32_t nblocks) {
  constexpr int B = Gbp<DB, TH>::kBins;
  __shared__ uint32_t hc[B];
  for (int b = threadIdx.x; b < B; b += TH) hc[b] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * Gbp<DB, TH>::kTile;
#pragma unroll 4
  for (int j = 0; j < Gbp<DB, TH>::kPer; ++j) {
    const uint64_t e = t0 + (uint64_t)j * TH + threadIdx.x;
    if (e < n) atomicAdd(&hc[(gbp_hash<SRC>(S, e) >> shift) & (B - 1)], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < B; b += TH) {
    hist[(uint64_t)b * nblocks + blockIdx.x] = hc[b];
    counts[(uint64_t)blockIdx.x * B + b] = hc[b];
  }
}

// An 8-bit pass's counts from the digit plane the previous pass wrote (one byte per row instead of
// the 8-byte hash). The plane is padded to whole 8192-row tiles; bytes past n are not counted.
template <int TH>
__global__ __launch_bounds__(TH) void k_gbp_hist_plane(const uint8_t *plane, uint64_t n, uint32_t *hist,
                                                       uint32_t *counts, uint32_t nblocks) {
  constexpr uint32_t kTile = Gbp<8, TH>::kTile;
  __shared__ uint32_t hc[256];
  for (int b = threadIdx.x; b < 256; b += TH) hc[b] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
  for (uint32_t c = threadIdx.x; c < kTile / 16; c += TH) {
    const uint4 v = reinterpret_cast<const uint4 *>(plane + t0)[c];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    const uint64_t e0 = t0 + 16ull * c;
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (e0 + b < n) atomicAdd(&hc[(w[b >> 2] >> (8 * (b & 3))) & 255u], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 256; b += TH) {
    hist[(uint64_t)b * nblocks + blockIdx.x] = hc[b];
    counts[(uint64_t)blockIdx.x * 256 + b] = hc[b];
  }
}

// plane_out (optional): the next pass's 8-bit digit, (h >> next_shift) & 255, as a byte per row.
// part_max (optional, a first pass that reads the value column): the workgroup's max |x| over its
// rows' non-NULL values, as the bits of a non-negative double, at part_max[blockIdx.x], with bit 63
// (kNullSeen, never set in such bits) set when one of its rows is NULL -- the exact SUM's scale,
// and whether a group can lack a non-NULL value, come from these without a pass over the column.
// OCC: workgroups per CU the kernel is compiled for (the launch bound's second figure is waves per
// SIMD, TH / 256 per workgroup; 2 holds a 512-thread pass to 128 VGPRs, and its 8-bit tile's LDS
// is under 80 KiB, so two share a CU; option groupby.part_occupancy).
constexpr unsigned long long kNullSeen = 1ull << 63;
template <int DB, int SRC, bool VAL, int TH, int OCC = 1>
__global__ __launch_bounds__(TH, TH / 256 * OCC) void k_gbp_scatter(GbpSrc S, uint64_t n, int shift, const uint32_t *offs,
                                                    const uint32_t *counts, uint32_t nblocks, uint64_t *h_out,
                                                    uint64_t *v_out, uint8_t *plane_out, int next_shift,
                                                    unsigned long long *part_max) {
  constexpr int B = Gbp<DB, TH>::kBins, kPer = Gbp<DB, TH>::kPer, kWaves = TH / kWave, kBpt = Gbp<DB, TH>::kBpt;
  constexpr uint32_t kTile = Gbp<DB, TH>::kTile, kSeg = (uint32_t)kPer * kWave;
  typedef typename Gbp<DB, TH>::Digit Digit;
  __shared__ uint64_t stage[kTile];
  __shared__ Digit sdigit[kTile];
  __shared__ uint32_t lbase[B], goff[B];
  __shared__ uint16_t wcnt[kWaves][B];  // per-wave digit counts, then their prefix over waves (< 2^16)
  __shared__ uint32_t scan_tmp[kWaves];
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
  const uint32_t tn = (uint32_t)(n - t0 < kTile ? n - t0 : kTile);
  const bool has_bins = B >= TH || tid < (uint32_t)B;  // threads past the last bin hold none
  {  // this tile's first position of every digit, and the digit's global offset
    uint32_t c[kBpt], s = 0, total;
#pragma unroll
    for (int q = 0; q < kBpt; ++q) {
      c[q] = has_bins ? counts[(uint64_t)blockIdx.x * B + tid * kBpt + q] : 0u;
      s += c[q];
    }
    uint32_t lb = block_exclusive_scan_256(s, scan_tmp, total);
    if (has_bins) {
#pragma unroll
      for (int q = 0; q < kBpt; ++q) {
        lbase[tid * kBpt + q] = lb;
        lb += c[q];
        goff[tid * kBpt + q] = offs[(uint64_t)(tid * kBpt + q) * nblocks + blockIdx.x];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < B / kWave; ++q) wcnt[wave][lane + q * kWave] = 0;
  // every load of the tile first: element wave * kSeg + j * 64 + lane; past the end a lane
  // re-reads the last element (ignored)
  uint64_t hv[kPer], vv[kPer];
  // lpos: a row's rank among its wave's equal digits, with the digit in bits 16.., then its position
  // in the tile (one register a row: registers decide whether two workgroups share a CU)
  uint32_t lpos[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t el = wave * kSeg + (uint32_t)j * kWave + lane;
    const uint64_t e = t0 + (el < tn ? el : tn - 1);
    if (SRC == kSrcCols) {
      const uint64_t r = S.sel ? S.sel[e] : e;
      hv[j] = mix64(pack_key(S.K, r));
      vv[j] = VAL ? col_bits(S.vdata, S.vtype, r) : 0;
    } else if (SRC == kSrcCol64) {  // raw key words now, hashed below once every load is in flight
      hv[j] = static_cast<const uint64_t *>(S.K.data[0])[e];
      vv[j] = VAL ? static_cast<const uint64_t *>(S.vdata)[e] : 0;
    } else {
      hv[j] = S.h[e];
      vv[j] = VAL ? S.v[e] : 0;
    }
  }
  if (SRC == kSrcCol64) {
#pragma unroll
    for (int j = 0; j < kPer; ++j) hv[j] = mix64(hv[j]);
  }
  if (VAL && SRC != kSrcPass && part_max) {  // uniform: a kernel argument
    __shared__ unsigned long long wmax[kWaves];
    unsigned long long m = 0;
    bool null_row = false;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      double d;
      if (wave * kSeg + (uint32_t)j * kWave + lane < tn) {
        if (bits_f64(vv[j], S.vtype, d)) {
          const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(d));
          m = b > m ? b : m;
        } else {
          null_row = true;
        }
      }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      const unsigned long long o = __shfl_xor(m, d, kWave);
      m = o > m ? o : m;
    }
    if (lane == 0) wmax[wave] = m | (ballot(null_row) ? kNullSeen : 0ull);
    __syncthreads();
    if (tid == 0) {
      unsigned long long nf = 0;
      m = 0;
      for (int w = 0; w < kWaves; ++w) {
        const unsigned long long x = wmax[w] & ~kNullSeen;
        m = x > m ? x : m;
        nf |= wmax[w] & kNullSeen;
      }
      part_max[blockIdx.x] = m | nf;
    }
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const bool valid = wave * kSeg + (uint32_t)j * kWave + lane < tn;
    const uint32_t digit = (uint32_t)(hv[j] >> shift) & (uint32_t)(B - 1);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < DB; ++b) {
      const uint64_t bb = ballot((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t before = wcnt[wave][digit];
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wcnt[wave][digit] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
    lpos[j] = digit << 16 | (before + rank);  // before + rank < kSeg <= 2^16
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kBpt; ++q) {  // exclusive prefix of each digit's counts over the waves
    if (!has_bins) break;
    const uint32_t d = tid * kBpt + q;
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = wcnt[w][d];
      wcnt[w][d] = (uint16_t)acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (wave * kSeg + (uint32_t)j * kWave + lane < tn) {
      const uint32_t digit = lpos[j] >> 16;
      const uint32_t pos = lbase[digit] + wcnt[wave][digit] + (lpos[j] & 0xFFFFu);
      lpos[j] = pos;
      sdigit[pos] = (Digit)digit;
    }
  }
#pragma unroll
  for (int q = 0; q < (VAL ? 2 : 1); ++q) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j)
      if (wave * kSeg + (uint32_t)j * kWave + lane < tn) stage[lpos[j]] = q ? vv[j] : hv[j];
    __syncthreads();
    uint64_t *out = q ? v_out : h_out;
    for (uint32_t p = tid; p < tn; p += TH) {
      const uint32_t d = sdigit[p];
      const uint64_t g = (uint64_t)goff[d] + (p - lbase[d]);
      out[g] = stage[p];
      if (plane_out && q == 0) plane_out[g] = (uint8_t)(stage[p] >> next_shift);
    }
  }
}

// The stable passes over the top pbits bits of h, low digits first (at most 9 bits a pass, the
// first pass as wide as the rest allow). h/v hold n words each; the partitioned (h, v) end in the
// buffers *h_res / *v_res.
template <int TH>
static int gbp_partition_th(ghive_ctx *ctx, const GbpSrc &src, uint64_t n, int pbits, bool val, uint64_t *h0,
                         uint64_t *v0, uint64_t *h1, uint64_t *v1, uint64_t **h_res, uint64_t **v_res,
                         unsigned long long *part_max) {
  std::vector<int> widths;  // digit widths, low digit first (one 8-bit pass at least)
  for (int left = pbits < 8 ? 8 : pbits; left > 0;) {
    const int passes = (left + 8) / 9;
    const int w = (left + passes - 1) / passes;
    widths.push_back(w > 9 ? 9 : w);
    left -= widths.back();
  }
  const uint32_t nb_max = ceil_div(n, Gbp<9, TH>::kTile);  // the smaller tile
  DevBuf hist, counts;
  GH_TRY(hist.alloc(ctx, sizeof(uint32_t) * 512ull * nb_max));
  GH_TRY(counts.alloc(ctx, sizeof(uint32_t) * 512ull * nb_max));
  GbpSrc S = src;
  // one LONG key column (a DOUBLE key's NaNs need canonical bits), an 8-byte value column (or none),
  // every row: the plain-word first pass
  const bool fast = !src.sel && src.K.nkeys == 1 && src.K.shift[0] == 0 && src.K.type[0] == GHIVE_LONG &&
                    (!val || src.vtype == GHIVE_LONG || src.vtype == GHIVE_DOUBLE);
  // each pass's digit width and shift: a narrower digit runs as an 8-bit one over its bits; a digit
  // narrower than DB bits must not reach past bit 63, so it is shifted to end at 63 when it is the
  // last one, else it takes DB bits starting at its low bit (the extra high bits are the next
  // digit's, which the later, stable passes order anyway)
  std::vector<int> dbs, shs;
  int lo = 64 - (pbits < 8 ? 8 : pbits);
  for (int w : widths) {
    const int db = w > 8 ? 9 : 8;
    dbs.push_back(db);
    shs.push_back(lo + db > 64 ? 64 - db : lo);
    lo += w;
  }
  // an 8-bit pass after the first reads its counts from a byte plane the pass before it wrote
  DevBuf plane_buf;
  for (size_t p = 1; p < dbs.size() && !plane_buf.get(); ++p)
    if (dbs[p] == 8) GH_TRY(plane_buf.alloc(ctx, (n + 8191) / 8192 * 8192));
  uint64_t *hin = nullptr, *vin = nullptr, *hout = h0, *vout = v0;
  bool have_plane = false;
  for (size_t p = 0; p < widths.size(); ++p) {
    const bool first = p == 0;
    const int db = dbs[p], sh = shs[p];
    const bool write_plane = p + 1 < dbs.size() && dbs[p + 1] == 8;
    uint8_t *pout = write_plane ? plane_buf.as<uint8_t>() : nullptr;
    const int nsh = write_plane ? shs[p + 1] : 0;
    unsigned long long *pmax = first ? part_max : nullptr;
    const uint32_t tile = db == 9 ? Gbp<9, TH>::kTile : Gbp<8, TH>::kTile;
    const uint32_t nb = ceil_div(n, tile);
    const uint32_t bins = db == 9 ? 512 : 256;
    if (!first) {
      S.h = hin;
      S.v = vin;
    }
    const int src = !first ? kSrcPass : (fast ? kSrcCol64 : kSrcCols);
#define GBP_HIST(D, F) k_gbp_hist<D, F, TH><<<nb, TH, 0, ctx->stream>>>(S, n, sh, hist.as<uint32_t>(), counts.as<uint32_t>(), nb)
    if (have_plane && db == 8) {
      k_gbp_hist_plane<TH><<<nb, TH, 0, ctx->stream>>>(plane_buf.as<uint8_t>(), n, hist.as<uint32_t>(),
                                                       counts.as<uint32_t>(), nb);
    } else if (db == 9) {
      if (src == kSrcCols) GBP_HIST(9, kSrcCols); else if (src == kSrcCol64) GBP_HIST(9, kSrcCol64); else GBP_HIST(9, kSrcPass);
    } else {
      if (src == kSrcCols) GBP_HIST(8, kSrcCols); else if (src == kSrcCol64) GBP_HIST(8, kSrcCol64); else GBP_HIST(8, kSrcPass);
    }
#undef GBP_HIST
    GH_LAUNCH(ctx);
    GH_TRY(exclusive_scan_u32(ctx, hist.as<uint32_t>(), hist.as<uint32_t>(), (uint64_t)bins * nb, nullptr));
#define GBP_SCATTER(D, F, V)                                                                                     \
  do {                                                                                                          \
    bool two_ = false;                                                                                          \
    if constexpr (D == 8 && TH == 512) {                                                                         \
      if (ctx->tune.gb_part_occupancy == 2) {                                                                   \
        k_gbp_scatter<D, F, V, TH, 2><<<nb, TH, 0, ctx->stream>>>(S, n, sh, hist.as<uint32_t>(),                 \
                                                                  counts.as<uint32_t>(), nb, hout, vout, pout, nsh, pmax); \
        two_ = true;                                                                                            \
      }                                                                                                         \
    }                                                                                                           \
    if (!two_)                                                                                                  \
      k_gbp_scatter<D, F, V, TH><<<nb, TH, 0, ctx->stream>>>(S, n, sh, hist.as<uint32_t>(), counts.as<uint32_t>(),   \
                                                             nb, hout, vout, pout, nsh, pmax);                  \
  } while (0)
    if (db == 9) {
      if (src == kSrcCols) { if (val) GBP_SCATTER(9, kSrcCols, true); else GBP_SCATTER(9, kSrcCols, false); }
      else if (src == kSrcCol64) { if (val) GBP_SCATTER(9, kSrcCol64, true); else GBP_SCATTER(9, kSrcCol64, false); }
      else { if (val) GBP_SCATTER(9, kSrcPass, true); else GBP_SCATTER(9, kSrcPass, false); }
    } else {
      if (src == kSrcCols) { if (val) GBP_SCATTER(8, kSrcCols, true); else GBP_SCATTER(8, kSrcCols, false); }
      else if (src == kSrcCol64) { if (val) GBP_SCATTER(8, kSrcCol64, true); else GBP_SCATTER(8, kSrcCol64, false); }
      else { if (val) GBP_SCATTER(8, kSrcPass, true); else GBP_SCATTER(8, kSrcPass, false); }
    }
#undef GBP_SCATTER
    GH_LAUNCH(ctx);
    have_plane = write_plane;
    hin = hout;
    vin = vout;
    hout = hin == h0 ? h1 : h0;
    vout = vin == v0 ? v1 : v0;
  }
  *h_res = hin;
  *v_res = vin;
  return GHIVE_OK;
}

// part_max (optional): filled by the first pass, one entry per workgroup of it (at most
// ceil(n / 4096) entries), see k_gbp_scatter.
static int gbp_partition(ghive_ctx *ctx, const GbpSrc &src, uint64_t n, int pbits, bool val, uint64_t *h0,
                         uint64_t *v0, uint64_t *h1, uint64_t *v1, uint64_t **h_res, uint64_t **v_res,
                         unsigned long long *part_max) {
  switch (ctx->tune.gb_part_threads) {
    case 256: return gbp_partition_th<256>(ctx, src, n, pbits, val, h0, v0, h1, v1, h_res, v_res, part_max);
    case 1024: return gbp_partition_th<1024>(ctx, src, n, pbits, val, h0, v0, h1, v1, h_res, v_res, part_max);
    default: return gbp_partition_th<512>(ctx, src, n, pbits, val, h0, v0, h1, v1, h_res, v_res, part_max);
  }
}

// Runs the high-cardinality form: 2^pbits hash partitions, each aggregated in an LDS table of 2^lg
// slots and written straight to the output columns O (rows in unspecified order). T carries the
// aggregate specs. *failed: a partition held more groups than its LDS table (nothing usable was
// written; use the direct form). *groups = groups found (CAPACITY when above capacity).
// The largest of n partial maxima (bits of non-negative doubles) into *out (one workgroup).
// *out = the max of the partials' low 63 bits, with kNullSeen set when any partial has it.
__global__ __launch_bounds__(256) void k_max_u64(const unsigned long long *part, uint64_t n, unsigned long long *out) {
  __shared__ unsigned long long wm[256 / kWave];
  unsigned long long m = 0, nf = 0;
  // one workgroup: 16 independent loads in flight per lane, not one (the reduction is latency-bound)
  constexpr int U = 16;
  for (uint64_t i0 = 0; i0 < n; i0 += 256 * U) {
    unsigned long long x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + (uint64_t)u * 256 + threadIdx.x;
      x[u] = i < n ? part[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nf |= x[u] & kNullSeen;
      const unsigned long long v = x[u] & ~kNullSeen;
      m = v > m ? v : m;
    }
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(m, d, kWave);
    m = o > m ? o : m;
    nf |= __shfl_xor(nf, d, kWave);
  }
  if (lane_id() == 0) wm[threadIdx.x / kWave] = m | nf;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = 0;
    for (int w = 0; w < 256 / kWave; ++w) {
      const unsigned long long v = wm[w] & ~kNullSeen;
      m = v > m ? v : m;
      nf |= wm[w] & kNullSeen;
    }
    *out = m | nf;
  }
}

int groupby_partitioned(ghive_ctx *ctx, const KeySpec &K, GroupTableDev &Tspec, const uint32_t *sel,
                        uint64_t n, uint32_t lcap, int pbits, const OutSpec &O, uint64_t capacity, bool *failed,
                        uint64_t *groups, bool scale_pending) {
  *failed = false;
  *groups = 0;
  // the aggregate inputs: none (COUNT(*) only), one column carried through the passes, or several
  // (read at the row id)
  int vcol = -1, val = 0;
  for (int a = 0; a < Tspec.naggs; ++a) {
    const int c = Tspec.aggs[a].col;
    if (c < 0) continue;
    if (vcol < 0) {
      vcol = c;
      val = 1;
    } else if (c != vcol) {
      val = 2;
    }
  }
  const void *vdata = nullptr;
  int vtype = 0;
  for (int a = 0; a < Tspec.naggs && val == 1; ++a)
    if (Tspec.aggs[a].col == vcol) {
      vdata = Tspec.aggs[a].data;
      vtype = Tspec.aggs[a].in_type;
    }
  // exact double sums not yet scaled: over the carried column, max |x| comes out of the first
  // partition pass; otherwise it is read here
  bool fused_absmax = false;
  for (int a = 0; a < Tspec.naggs && scale_pending; ++a) {
    AggSpec &A = Tspec.aggs[a];
    if (A.kind != AK_SUMD) continue;
    if (val == 1 && A.col == vcol) {
      fused_absmax = true;
      continue;
    }
    double mx = 0;
    GH_TRY(column_absmax(ctx, A.data, A.in_type, sel, n, &mx));
    GH_TRY(sum_scale(ctx, mx, n, &A.scale_e));
  }
  const int KW = val == 1 ? 2 : 1;
  bool null_seen = true;  // until the first pass shows that the carried column has no NULL
  DevBuf k0, k1, i0, i1, starts, sk, sa, sc, pmax;
  const uint64_t npmax = (n + 4095) / 4096;  // the first pass's workgroups, at most
  if (fused_absmax) {
    GH_TRY(pmax.alloc(ctx, 8 * npmax));
    GH_TRY(fill_u64(ctx, pmax.as<uint64_t>(), 0, npmax));
  }
  GH_TRY(k0.alloc(ctx, 8ull * KW * n));
  GH_TRY(k1.alloc(ctx, 8ull * KW * n));
  if (val == 2) {
    GH_TRY(i0.alloc(ctx, 4 * n));
    GH_TRY(i1.alloc(ctx, 4 * n));
  }
  GH_TRY(starts.alloc(ctx, 8 * ((1ull << pbits) + 1)));
  uint64_t *ks = nullptr, *vs_part = nullptr;
  uint32_t *ids = i0.as<uint32_t>();
  if (val != 2) {  // hashed inside the first partition pass; the value word travels beside the hash
    GbpSrc S;
    memset(&S, 0, sizeof S);
    S.K = K;
    S.sel = sel;
    S.vdata = vdata;
    S.vtype = vtype;
    GH_TRY(gbp_partition(ctx, S, n, pbits, val == 1, k0.as<uint64_t>(), val == 1 ? k0.as<uint64_t>() + n : nullptr,
                         k1.as<uint64_t>(), val == 1 ? k1.as<uint64_t>() + n : nullptr, &ks, &vs_part,
                         fused_absmax ? reinterpret_cast<unsigned long long *>(pmax.get()) : nullptr));
    if (fused_absmax) {
      k_max_u64<<<1, 256, 0, ctx->stream>>>(reinterpret_cast<const unsigned long long *>(pmax.get()), npmax,
                                            reinterpret_cast<unsigned long long *>(ctx->d_counters));
      GH_LAUNCH(ctx);
      uint64_t mb = 0;
      GH_TRY(d2h(ctx, &mb, ctx->d_counters, sizeof mb));
      null_seen = (mb & kNullSeen) != 0;
      mb &= ~kNullSeen;
      double mx = 0;
      memcpy(&mx, &mb, sizeof mx);
      int32_t e = 0;
      GH_TRY(sum_scale(ctx, mx, n, &e));
      for (int a = 0; a < Tspec.naggs; ++a)
        if (Tspec.aggs[a].kind == AK_SUMD) Tspec.aggs[a].scale_e = e;
    }
  } else {  // several value columns: (hash, row id) pairs, the values read at the row id
    const uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 16;
    k_gb_part_keys<<<(uint32_t)(g > cap ? cap : g), 256, 0, ctx->stream>>>(K, sel, n, k0.as<uint64_t>(),
                                                                           i0.as<uint32_t>(), nullptr, 0, nullptr);
    GH_LAUNCH(ctx);
    std::vector<int> digits;  // the 8-bit digits of the hash that cover its top pbits bits
    for (int d = pbits ? (64 - pbits) / 8 : 8; d < 8; ++d) digits.push_back(d);
    ks = k0.as<uint64_t>();
    const ghive_sort_info last_sort = ctx->sort_info;  // the last ORDER BY's record stays as it was
    const int sst = radix_sort_u64_ids(ctx, n, k0.as<uint64_t>(), k1.as<uint64_t>(), i0.as<uint32_t>(),
                                       i1.as<uint32_t>(), digits, &ks, &ids, true);
    ctx->sort_info = last_sort;
    GH_TRY(sst);
  }
  {
    const uint64_t parts = (1ull << pbits) + 1;
    k_gb_part_bounds<<<(uint32_t)((parts + 255) / 256), 256, 0, ctx->stream>>>(ks, n, pbits, starts.as<uint64_t>());
    GH_LAUNCH(ctx);
  }
  // the small global table that takes the one key equal to the EMPTY pattern
  GroupTableDev T = Tspec;
  GH_TRY(group_table_alloc(ctx, T, 1, sk, sa, sc));
  const size_t lds_bytes = (size_t)(lcap + 1) * (1 + T.nwords) * 8;
  unsigned long long *fail = reinterpret_cast<unsigned long long *>(ctx->d_counters);
  unsigned long long *count = fail + 1;
  GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 2));
  const uint32_t P = (uint32_t)(1ull << pbits);
  const uint32_t egrid = std::min<uint32_t>(P, (uint32_t)ctx->cus);  // one resident workgroup per CU
  int sum_nocount = !null_seen;
  for (int a = 0; a < T.naggs; ++a)
    if (T.aggs[a].kind == AK_SUMD && T.aggs[a].is_avg) sum_nocount = 0;
  const uint64_t *vs = vs_part;
  // VAL x the aggregate count (1..4 compiled, 0 = any)
  const int na = T.naggs >= 1 && T.naggs <= 4 ? T.naggs : 0;
#define GB_EMIT(V, N)                                                                                           \
  do {                                                                                                          \
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_gb_part_emit<V, N>),                             \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);                       \
    k_gb_part_emit<V, N><<<egrid, kGbPartThreads, lds_bytes, ctx->stream>>>(K, T, starts.as<uint64_t>(), ks, vs, ids, \
                                                                            lcap, P, fail, count, O, capacity,    \
                                                                            sum_nocount);                         \
  } while (0)
#define GB_EMIT_N(V)        \
  switch (na) {             \
    case 1: GB_EMIT(V, 1); break; \
    case 2: GB_EMIT(V, 2); break; \
    case 3: GB_EMIT(V, 3); break; \
    case 4: GB_EMIT(V, 4); break; \
    default: GB_EMIT(V, 0); break; \
  }
  switch (val) {
    case 0: GB_EMIT_N(0); break;
    case 1: GB_EMIT_N(1); break;
    default: GB_EMIT_N(2); break;
  }
#undef GB_EMIT_N
#undef GB_EMIT
  GH_LAUNCH(ctx);
  uint64_t f = 0;
  GH_TRY(d2h(ctx, &f, fail, sizeof f));
  if (f) {
    *failed = true;
    return GHIVE_OK;
  }
  // the EMPTY-pattern key's group, if any, after the partitions' groups (same counter)
  k_groupby_finalize<<<1, 256, 0, ctx->stream>>>(K, T, O, count, capacity);
  GH_LAUNCH(ctx);
  GH_TRY(d2h(ctx, groups, count, sizeof(uint64_t)));
  if (*groups > capacity)
    return set_error(ctx, GHIVE_ERR_CAPACITY, "group-by: more than %llu groups (capacity)",
                     (unsigned long long)capacity);
  return GHIVE_OK;
}

int group_table_preinsert(ghive_ctx *ctx, const GroupTableDev &T, uint64_t key) {
  k_gt_preinsert<<<1, 64, 0, ctx->stream>>>(T, key);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int group_table_emit(ghive_ctx *ctx, const KeySpec &K, const GroupTableDev &T, const OutSpec &O,
                     uint64_t capacity, uint64_t *out_groups, hipEvent_t after, const uint64_t *extra_src,
                     uint64_t *extra_dst) {
  GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 1));
  uint64_t g = (T.stride + 255) / 256, cap = (uint64_t)ctx->cus * 16;
  k_groupby_finalize<<<(uint32_t)(g > cap ? cap : g), 256, 0, ctx->stream>>>(
      K, T, O, reinterpret_cast<unsigned long long *>(ctx->d_counters), capacity);
  GH_LAUNCH(ctx);
  if (after) GH_HIP(ctx, hipEventRecord(after, ctx->stream));
  uint64_t host[3] = {0, 0, 0};
  uint64_t *dst[3] = {&host[0], &host[1], &host[2]};
  const uint64_t *src[3] = {ctx->d_counters, reinterpret_cast<const uint64_t *>(T.overflow), extra_src};
  GH_TRY(d2h_words(ctx, extra_src ? 3 : 2, dst, src));
  if (extra_src && extra_dst) *extra_dst = host[2];
  *out_groups = host[0];
  if (host[1] || host[0] > capacity) {
    if (*out_groups <= capacity) *out_groups = capacity + 1;
    return set_error(ctx, GHIVE_ERR_CAPACITY, "group-by: more than %llu groups (capacity)",
                     (unsigned long long)capacity);
  }
  return GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_groupby(ghive_ctx *ctx, const ghive_table *table, const uint32_t *sel, uint64_t n,
                  const uint32_t *keys, uint32_t nkeys, const ghive_agg *aggs, uint32_t naggs,
                  ghive_column *out_cols, uint64_t capacity, uint64_t *out_groups) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "groupby"));
  if (!out_groups) return set_error(ctx, GHIVE_ERR_INVALID, "groupby: out_groups is NULL");
  if (!sel && n != table->row_num)
    return set_error(ctx, GHIVE_ERR_INVALID, "groupby: n != rows without a selection");
  if (nkeys > (uint32_t)kMaxWideKeys)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "groupby: %u keys (max %d)", nkeys, kMaxWideKeys);
  if (naggs > (uint32_t)kMaxAggs) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "groupby: %u aggregates (max %d)", naggs, kMaxAggs);
  if ((nkeys && !keys) || (naggs && !aggs)) return set_error(ctx, GHIVE_ERR_INVALID, "groupby: NULL keys/aggs");
  if (nkeys + naggs && !out_cols) return set_error(ctx, GHIVE_ERR_INVALID, "groupby: out_cols is NULL");
  // keys that pack into 64 bits aggregate on the packed key; wider ones through the dictionary
  KeySpec K;
  memset(&K, 0, sizeof K);
  WideKeySpec WK;
  memset(&WK, 0, sizeof WK);
  WK.nkeys = (int)nkeys;
  bool wide = nkeys > (uint32_t)kMaxKeys;
  int bits = 0;
  for (uint32_t k = 0; k < nkeys; ++k) {
    if (keys[k] >= table->ncols) return set_error(ctx, GHIVE_ERR_INVALID, "groupby: key column %u", keys[k]);
    const ghive_column &c = table->columns[keys[k]];
    WK.type[k] = c.type;
    WK.data[k] = c.data;
    WK.aux[k] = static_cast<const int32_t *>(c.aux);
    if (c.type == GHIVE_STRING) wide = true;
    else bits += 8 * type_width(c.type);
  }
  if (bits > 64) wide = true;
  if (!wide) {
    K.nkeys = (int)nkeys;
    int shift = 0;
    for (uint32_t k = 0; k < nkeys; ++k) {
      K.type[k] = WK.type[k];
      K.data[k] = WK.data[k];
      K.shift[k] = shift;
      shift += 8 * type_width(WK.type[k]);
    }
  }
  GroupTableDev T;
  memset(&T, 0, sizeof T);
  T.naggs = (int)naggs;
  int word = 0;
  for (uint32_t a = 0; a < naggs; ++a) {
    AggSpec &A = T.aggs[a];
    int ty = aggs[a].type;
    bool star = ty == GHIVE_AGG_CNT && aggs[a].index == GHIVE_COUNT_STAR;
    if (ty < GHIVE_AGG_SUM || ty > GHIVE_AGG_CNT)
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "groupby: aggregation type %d", ty);
    int in_type = GHIVE_LONG;
    if (!star) {
      if (aggs[a].index >= table->ncols)
        return set_error(ctx, GHIVE_ERR_INVALID, "groupby: aggregate column %u", aggs[a].index);
      in_type = table->columns[aggs[a].index].type;
      if (!is_numeric_type(in_type))
        return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "groupby: aggregates over STRING are not supported");
      A.col = (int)aggs[a].index;
      A.data = table->columns[aggs[a].index].data;
    } else {
      A.col = -1;
    }
    bool fp = in_type == GHIVE_DOUBLE || in_type == GHIVE_FLOAT;
    A.in_type = (int8_t)in_type;
    A.out_type = (int8_t)ghive_agg_output_type(ty, in_type);
    A.is_avg = ty == GHIVE_AGG_AVG;
    switch (ty) {
      case GHIVE_AGG_CNT: A.kind = AK_CNT; break;
      case GHIVE_AGG_SUM:
      case GHIVE_AGG_AVG: A.kind = fp ? AK_SUMD : AK_SUMI; break;
      case GHIVE_AGG_MIN: A.kind = fp ? AK_MIND : AK_MINI; break;
      default: A.kind = fp ? AK_MAXD : AK_MAXI; break;
    }
    A.word0 = (int16_t)word;
    word += agg_words(A.kind);
    ghive_column &o = out_cols[nkeys + a];
    if (capacity && !o.data) return set_error(ctx, GHIVE_ERR_INVALID, "groupby: output column %u has no buffer", nkeys + a);
    o.type = A.out_type;
  }
  for (uint32_t k = 0; k < nkeys; ++k) {
    ghive_column &o = out_cols[k];
    if (WK.type[k] == GHIVE_STRING) {  // (start,end) pairs into the caller's aux; the blob is the input's
      if (capacity && !o.aux)
        return set_error(ctx, GHIVE_ERR_INVALID, "groupby: STRING output column %u has no aux buffer", k);
      o.data = const_cast<void *>(WK.data[k]);
      o.char_size = table->columns[keys[k]].char_size;
    } else if (capacity && !o.data) {
      return set_error(ctx, GHIVE_ERR_INVALID, "groupby: output column %u has no buffer", k);
    }
    o.type = WK.type[k];
  }
  T.nwords = word;
  // the high-cardinality form: a large group bound over many rows (or forced by the option)
  const int64_t part_opt = ctx->tune.gb_partition;
  const bool want_part = nkeys > 0 && n && part_opt != 0 &&
                         (part_opt == 2 || (n >= (1ull << 22) && capacity >= (1ull << 20)));
  // exact double sums: pick the fixed-point LSB from max |x| and the row count (the partitioned
  // form does this itself, from its first pass where it can)
  for (uint32_t a = 0; a < naggs && !want_part; ++a) {
    AggSpec &A = T.aggs[a];
    if (A.kind != AK_SUMD) continue;
    double mx = 0;
    GH_TRY(column_absmax(ctx, A.data, A.in_type, sel, n, &mx));
    GH_TRY(sum_scale(ctx, mx, n, &A.scale_e));
  }
  DevBuf wdict, wcodes, wgcode;
  uint64_t *dict = nullptr;
  if (wide) {
    if (capacity > (1ull << 31))
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "groupby: wide keys with a group bound above 2^31");
    uint64_t slots = 0, count = 0;
    GH_TRY(wgcode.alloc(ctx, (capacity ? capacity : 1) * 4));
    const int st = key_dictionary(ctx, WK, table->row_num, sel, n, capacity, wdict, wcodes, &slots, &count);
    if (st == GHIVE_ERR_CAPACITY) {
      *out_groups = count > capacity ? count : capacity + 1;
      for (uint32_t k = 0; k < nkeys + naggs; ++k) out_cols[k].row_num = 0;
      return set_error(ctx, GHIVE_ERR_CAPACITY, "group-by: more than %llu groups (capacity)",
                       (unsigned long long)capacity);
    }
    GH_TRY(st);
    dict = wdict.as<uint64_t>();
    K.nkeys = 1;  // aggregate on the dictionary code
    K.type[0] = GHIVE_INT;
    K.shift[0] = 0;
    K.data[0] = wcodes.as<uint32_t>();
  }
  OutSpec O;
  memset(&O, 0, sizeof O);
  if (wide) {
    O.col[0] = wgcode.as<int32_t>();
    O.type[0] = GHIVE_INT;
    for (uint32_t a = 0; a < naggs; ++a) {
      O.col[1 + a] = out_cols[nkeys + a].data;
      O.type[1 + a] = out_cols[nkeys + a].type;
    }
  } else {
    for (uint32_t c = 0; c < nkeys + naggs; ++c) {
      O.col[c] = out_cols[c].data;
      O.type[c] = out_cols[c].type;
    }
  }
  DevBuf dkeys, dacc, dcnt;
  bool partitioned = false;
  int st = GHIVE_OK;
  if (want_part) {
    // LDS table slots per partition: a multiple of 512 that fits 150 KiB; partitions so that the
    // group bound fills each table to at most 60% on average (the emit falls back to the direct form
    // if a partition ever exceeds 7/8)
    uint32_t lcap = 4096;
    while (lcap > 512 && (size_t)(lcap + 1) * (1 + T.nwords) * 8 > 150 * 1024) lcap -= 512;
    int pbits = 0;
    while (pbits < 24 && (double)capacity > 0.6 * (double)lcap * (double)(1ull << pbits)) ++pbits;
    bool failed = false;
    st = groupby_partitioned(ctx, K, T, sel, n, lcap, pbits, O, capacity, &failed, out_groups, /*scale_pending=*/true);
    if (st != GHIVE_OK && st != GHIVE_ERR_CAPACITY) return st;
    partitioned = !failed;
    if (failed) ++ctx->gb_part_fallbacks;
  }
  if (!partitioned) GH_TRY(group_table_alloc(ctx, T, capacity, dkeys, dacc, dcnt));
  if (nkeys == 0) GH_TRY(group_table_preinsert(ctx, T, 0ull));  // a global aggregate: one group, even over no rows
  if (n && !partitioned) {
    int lg = lds_log2_for(T.nwords, (int)ctx->tune.gb_lds_budget);
    size_t lds_bytes = (size_t)((1 << lg) + 1) * (1 + T.nwords) * 8;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_groupby, kGbThreads, lds_bytes) != hipSuccess || per_cu < 1)
      per_cu = 1;
    uint64_t want = (n + kGbThreads - 1) / kGbThreads, cap = (uint64_t)ctx->cus * per_cu;
    k_groupby<<<(uint32_t)(want > cap ? cap : want), kGbThreads, lds_bytes, ctx->stream>>>(K, T, sel, n, lg);
    GH_LAUNCH(ctx);
  }
  if (!partitioned) st = group_table_emit(ctx, K, T, O, capacity, out_groups);
  if (st == GHIVE_OK && wide && *out_groups) {
    WideOut WO;
    memset(&WO, 0, sizeof WO);
    for (uint32_t k = 0; k < nkeys; ++k) {
      WO.col[k] = out_cols[k].data;
      WO.aux[k] = static_cast<int32_t *>(out_cols[k].aux);
    }
    const uint64_t g = (*out_groups + 255) / 256, cap = (uint64_t)ctx->cus * 16;
    k_gbw_emit_keys<<<(uint32_t)(g > cap ? cap : g), 256, 0, ctx->stream>>>(WK, dict, wgcode.as<int32_t>(),
                                                                             *out_groups, WO);
    GH_LAUNCH(ctx);
  }
  for (uint32_t c = 0; c < nkeys + naggs; ++c) out_cols[c].row_num = st == GHIVE_OK ? *out_groups : 0;
  return st;
}

}  // extern "C"
