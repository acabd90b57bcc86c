// Hash aggregation state shared by the group-by operator and the fused fragment kernel.
//
// A group table is an open-addressing array of packed uint64 keys (EMPTY = all ones; a real
// key with that bit pattern lives in the extra slot at index `slots`) plus structure-of-arrays
// accumulator words: word j of slot s is acc[j * stride + s]. Each workgroup first aggregates
// into a private LDS table of the same shape and merges it into the global table once, at the
// end, with one atomic per word -- so low-cardinality group-bys cost LDS atomics only.
//
// Accumulators per aggregate kind:
//   CNT  : count                               SUMI : int64 sum (wraps like a Java long), count
//   MINI/MAXI : int64 value, count             MIND/MAXD : order-preserving double bits, count
//   SUMD : exact fixed-point sum of doubles, LSB 2^E, as one two's-complement 128-bit integer
//          (E is chosen so that |sum| < 2^126: no overflow) -- integer addition, so the result
//          does not depend on the order rows arrive in -- plus count.
#pragma once

#include "device.hpp"

namespace ghive {

enum AggKind : int8_t { AK_CNT = 0, AK_SUMI = 1, AK_MINI = 2, AK_MAXI = 3, AK_MIND = 4, AK_MAXD = 5, AK_SUMD = 6 };
constexpr int kMaxAggs = 8;
constexpr int kMaxKeys = 4;
constexpr uint64_t kEmptyKey = ~0ull;

__host__ __device__ inline int agg_words(int kind) { return kind == AK_CNT ? 1 : (kind == AK_SUMD ? 3 : 2); }

struct AggSpec {
  int8_t kind;
  int8_t in_type;   // input column type (GHIVE_*)
  int8_t out_type;  // output column type
  int8_t is_avg;
  int16_t word0;    // first accumulator word
  int16_t pad;
  int32_t scale_e;  // SUMD: LSB exponent
  int32_t col;      // input column index (-1: COUNT(*) or an expression)
  const void *data; // input column data
};

struct GroupTableDev {
  uint64_t *keys;    // slots + 1
  uint64_t *acc;     // stride records of `rec` words: all accumulator words of a group together
  uint64_t stride;   // slots + 1
  uint64_t mask;     // slots - 1
  int32_t log2;
  int32_t nwords;
  int32_t naggs;
  int32_t rec;       // words per record: nwords rounded up to 1, 2, 4 or a multiple of 8, so that a
                     // record never straddles a 64-byte line and one row's updates share a line
  unsigned long long *nent;      // entries inserted (extra slot included)
  unsigned long long *overflow;  // set when an insert could not be placed
  uint64_t capacity;             // max entries accepted
  AggSpec aggs[kMaxAggs];
};

// Accumulator word w of group slot g.
__host__ __device__ __forceinline__ uint64_t *acc_word(const GroupTableDev &T, int w, uint64_t g) {
  return T.acc + g * (uint64_t)T.rec + (uint64_t)w;
}
__host__ __device__ inline int record_words(int nwords) {
  if (nwords <= 1) return 1;
  if (nwords <= 2) return 2;
  if (nwords <= 4) return 4;
  return (nwords + 7) & ~7;
}

__device__ __forceinline__ uint64_t init_word(int kind, int w) {
  if (w == 0) {
    if (kind == AK_MINI) return 0x7fffffffffffffffull;
    if (kind == AK_MAXI) return 0x8000000000000000ull;
    if (kind == AK_MIND) return ~0ull;
  }
  return 0ull;
}

// order-preserving map of a double to uint64 (larger double -> larger integer)
__device__ __forceinline__ uint64_t dbl_order(double d) {
  if (d == 0.0) d = 0.0;  // -0 and +0 compare equal
  uint64_t b = (uint64_t)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dbl_unorder(uint64_t u) {
  uint64_t b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
  return __longlong_as_double((long long)b);
}

// |x| / 2^E rounded to nearest as an unsigned 128-bit integer (lo, hi); returns x < 0.
__device__ __forceinline__ bool to_fixed(double x, int E, uint64_t &lo, uint64_t &hi) {
  uint64_t bits = (uint64_t)__double_as_longlong(x);
  bool neg = bits >> 63;
  int ex = (int)((bits >> 52) & 0x7ff);
  uint64_t mant = bits & ((1ull << 52) - 1);
  if (ex == 0) ex = 1;
  else mant |= 1ull << 52;
  lo = hi = 0;
  if (!mant) return false;
  int sh = ex - 1075 - E;
  if (sh >= 64) {
    hi = mant << (sh - 64);
  } else if (sh > 0) {
    lo = mant << sh;
    hi = mant >> (64 - sh);
  } else if (sh == 0) {
    lo = mant;
  } else if (sh > -64) {
    int rs = -sh;
    lo = (mant + (1ull << (rs - 1))) >> rs;
  }
  return neg;
}

// nearest double to (hi:lo) * 2^E, ties to even
__device__ __forceinline__ double fixed_to_double(uint64_t lo, uint64_t hi, int E) {
  if (!hi && !lo) return 0.0;
  int p = hi ? 127 - __clzll((long long)hi) : 63 - __clzll((long long)lo);  // top bit index
  if (p <= 52) return ldexp((double)lo, E);
  int sh = p - 52;
  // mantissa = bits [sh, p]; remainder = bits below sh
  uint64_t mant, rem_hi_bit, rem_rest;
  if (sh >= 64) {
    mant = hi >> (sh - 64);
    uint64_t below_hi = (sh - 64) ? (hi & ((1ull << (sh - 64)) - 1)) : 0;
    rem_hi_bit = (sh - 65) >= 0 ? (hi >> (sh - 65)) & 1 : (lo >> 63) & 1;
    uint64_t rest_hi = (sh - 65) > 0 ? (hi & ((1ull << (sh - 65)) - 1)) : 0;
    rem_rest = rest_hi | ((sh - 65) >= 0 ? lo : (lo & 0x7fffffffffffffffull));
    (void)below_hi;
  } else {
    mant = (lo >> sh) | (sh ? (hi << (64 - sh)) : 0);
    rem_hi_bit = (lo >> (sh - 1)) & 1;
    rem_rest = (sh - 1) ? (lo & ((1ull << (sh - 1)) - 1)) : 0;
  }
  mant &= (1ull << 53) - 1;
  if (rem_hi_bit && (rem_rest || (mant & 1))) {
    ++mant;
    if (mant >> 53) {
      mant >>= 1;
      ++sh;
    }
  }
  return ldexp((double)mant, E + sh);
}

__device__ __forceinline__ void add128(uint64_t *lo_w, uint64_t *hi_w, uint64_t lo, uint64_t hi) {
  if (lo) {
    unsigned long long old = atomicAdd((unsigned long long *)lo_w, (unsigned long long)lo);
    if (old + lo < old) ++hi;  // carry out of the low word
  }
  if (hi) atomicAdd((unsigned long long *)hi_w, (unsigned long long)hi);
}

// One row's contribution of one aggregate, already converted.
struct AggVal {
  bool valid;
  bool neg;
  uint64_t a;  // value / ordered bits / fixed-point low word
  uint64_t b;  // fixed-point high word
};

template <class Word>
__device__ __forceinline__ void acc_apply(const AggSpec &A, Word word, const AggVal &v) {
  if (!v.valid) return;
  switch (A.kind) {
    case AK_CNT: atomicAdd((unsigned long long *)word(0), 1ull); break;
    case AK_SUMI:
      atomicAdd((unsigned long long *)word(0), (unsigned long long)v.a);
      atomicAdd((unsigned long long *)word(1), 1ull);
      break;
    case AK_MINI:
      atomicMin((long long *)word(0), (long long)v.a);
      atomicAdd((unsigned long long *)word(1), 1ull);
      break;
    case AK_MAXI:
      atomicMax((long long *)word(0), (long long)v.a);
      atomicAdd((unsigned long long *)word(1), 1ull);
      break;
    case AK_MIND:
      atomicMin((unsigned long long *)word(0), (unsigned long long)v.a);
      atomicAdd((unsigned long long *)word(1), 1ull);
      break;
    case AK_MAXD:
      atomicMax((unsigned long long *)word(0), (unsigned long long)v.a);
      atomicAdd((unsigned long long *)word(1), 1ull);
      break;
    default: {  // AK_SUMD: the signed value in two's complement, added modulo 2^128
      uint64_t lo = v.a, hi = v.b;
      if (v.neg) {
        lo = ~lo + 1ull;
        hi = ~hi + (lo == 0 ? 1ull : 0ull);
      }
      add128(word(0), word(1), lo, hi);
      atomicAdd((unsigned long long *)word(2), 1ull);
      break;
    }
  }
}

// Merge accumulated words src(j) into dst(j).
template <class Dst, class Src>
__device__ __forceinline__ void acc_merge(const AggSpec &A, Dst dst, Src src) {
  switch (A.kind) {
    case AK_CNT: {
      uint64_t c = *src(0);
      if (c) atomicAdd((unsigned long long *)dst(0), (unsigned long long)c);
      break;
    }
    case AK_SUMI: {
      uint64_t c = *src(1);
      if (!c) break;
      atomicAdd((unsigned long long *)dst(0), (unsigned long long)*src(0));
      atomicAdd((unsigned long long *)dst(1), (unsigned long long)c);
      break;
    }
    case AK_MINI:
    case AK_MAXI:
    case AK_MIND:
    case AK_MAXD: {
      uint64_t c = *src(1);
      if (!c) break;
      uint64_t v = *src(0);
      if (A.kind == AK_MINI) atomicMin((long long *)dst(0), (long long)v);
      else if (A.kind == AK_MAXI) atomicMax((long long *)dst(0), (long long)v);
      else if (A.kind == AK_MIND) atomicMin((unsigned long long *)dst(0), (unsigned long long)v);
      else atomicMax((unsigned long long *)dst(0), (unsigned long long)v);
      atomicAdd((unsigned long long *)dst(1), (unsigned long long)c);
      break;
    }
    default: {
      uint64_t c = *src(2);
      if (!c) break;
      add128(dst(0), dst(1), *src(0), *src(1));
      atomicAdd((unsigned long long *)dst(2), (unsigned long long)c);
      break;
    }
  }
}

// Global table: slot of `key`, inserting it if new; -1 when the table is full.
__device__ __forceinline__ int64_t gt_find_insert(const GroupTableDev &T, uint64_t key) {
  if (key == kEmptyKey) {
    uint64_t *flag = T.keys + (T.stride - 1);
    if (atomicCAS((unsigned long long *)flag, 0ull, 1ull) == 0ull) atomicAdd(T.nent, 1ull);
    return (int64_t)(T.stride - 1);
  }
  uint64_t s = hash_u64(key, T.log2);
  for (uint64_t step = 0; step <= T.mask; ++step) {
    uint64_t k = __atomic_load_n(T.keys + s, __ATOMIC_RELAXED);
    if (k == key) return (int64_t)s;
    if (k == kEmptyKey) {
      // the entry count is claimed first, so at most `capacity` keys are ever placed
      unsigned long long old = atomicCAS((unsigned long long *)(T.keys + s), kEmptyKey, key);
      if (old == kEmptyKey) {
        if (atomicAdd(T.nent, 1ull) >= T.capacity) {
          atomicOr(T.overflow, 1ull);
        }
        return (int64_t)s;
      }
      if (old == key) return (int64_t)s;
    }
    s = (s + 1) & T.mask;
  }
  atomicOr(T.overflow, 1ull);
  return -1;
}

constexpr int kLdsProbe = 16;

// ---- LDS-private group tables: keys[lcap + 1] then nwords x (lcap + 1) words ----
__device__ __forceinline__ int lds_find_insert(uint64_t *lkeys, uint64_t key, int lcap_log2) {
  const int lcap = 1 << lcap_log2;
  if (key == kEmptyKey) {
    lkeys[lcap] = 1ull;
    return lcap;
  }
  uint32_t s = (uint32_t)hash_u64(key, lcap_log2);
  for (int step = 0; step < kLdsProbe; ++step) {
    uint64_t k = __atomic_load_n(lkeys + s, __ATOMIC_RELAXED);
    if (k == key) return (int)s;
    if (k == kEmptyKey) {
      unsigned long long old = atomicCAS((unsigned long long *)(lkeys + s), kEmptyKey, key);
      if (old == kEmptyKey || old == key) return (int)s;
    }
    s = (s + 1) & (uint32_t)(lcap - 1);
  }
  return -1;
}

__device__ __forceinline__ void lds_table_init(uint64_t *lds, const GroupTableDev &T, int lcap_log2) {
  const int lstride = (1 << lcap_log2) + 1;
  for (int s = threadIdx.x; s < lstride; s += blockDim.x) lds[s] = s == lstride - 1 ? 0ull : kEmptyKey;
  for (int a = 0; a < T.naggs; ++a) {
    const AggSpec &A = T.aggs[a];
    for (int w = 0; w < agg_words(A.kind); ++w) {
      uint64_t v = init_word(A.kind, w);
      uint64_t *p = lds + (size_t)(1 + A.word0 + w) * lstride;
      for (int s = threadIdx.x; s < lstride; s += blockDim.x) p[s] = v;
    }
  }
}

__device__ __forceinline__ void lds_table_flush(uint64_t *lds, const GroupTableDev &T, int lcap_log2) {
  const int lstride = (1 << lcap_log2) + 1;
  for (int s = threadIdx.x; s < lstride; s += blockDim.x) {
    uint64_t k = lds[s];
    bool used = s == lstride - 1 ? k == 1ull : k != kEmptyKey;
    if (!used) continue;
    int64_t g = gt_find_insert(T, s == lstride - 1 ? kEmptyKey : k);
    if (g < 0) continue;
    for (int a = 0; a < T.naggs; ++a) {
      const AggSpec &A = T.aggs[a];
      acc_merge(A, [&](int j) { return acc_word(T, A.word0 + j, (uint64_t)g); },
                [&](int j) { return lds + (size_t)(1 + A.word0 + j) * lstride + s; });
    }
  }
}

}  // namespace ghive
