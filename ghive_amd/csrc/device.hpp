// Device-side building blocks shared by the kernels: wave64 helpers, typed column access with the
// reference's NULL sentinels, the flattened predicate program, and Philox4x32-10.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/ghive/ghive.h"

namespace ghive {

constexpr int kWave = 64;

// ---- wave64 helpers ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of `mask` in lanes below this one
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ int first_lane(uint64_t mask) { return __ffsll((unsigned long long)mask) - 1; }
__device__ __forceinline__ uint64_t read_lane_u64(uint64_t v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t read_lane_u32(uint32_t v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

// Wave-aggregated append: every lane with `want` gets a distinct slot of *counter (one atomic
// per wave). Returns the slot (valid only when want).
__device__ __forceinline__ uint64_t wave_append(unsigned long long *counter, bool want) {
  uint64_t m = ballot(want);
  if (!m) return 0;
  int leader = first_lane(m);
  uint64_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = read_lane_u64(base, leader);
  return base + lanes_below(m);
}

// ---- columns and NULL sentinels (gpu_impl/inc/DataFlow/Table.hpp:29-52) -------------------
__device__ __forceinline__ bool is_null_i32(int32_t v) { return v == 0x7fffffff; }
__device__ __forceinline__ bool is_null_i64(int64_t v) { return v == 0x7fffffffffffffffll; }

__device__ __forceinline__ bool col_load_i64(const void *p, int type, uint64_t r, int64_t &v) {
  if (type == GHIVE_INT) {
    int32_t x = static_cast<const int32_t *>(p)[r];
    v = x;
    return !is_null_i32(x);
  }
  int64_t x = static_cast<const int64_t *>(p)[r];
  v = x;
  return !is_null_i64(x);
}

__device__ __forceinline__ bool col_load_f64(const void *p, int type, uint64_t r, double &v) {
  switch (type) {
    case GHIVE_DOUBLE: v = static_cast<const double *>(p)[r]; return !__builtin_isnan(v);
    case GHIVE_FLOAT: v = (double)static_cast<const float *>(p)[r]; return !__builtin_isnan(v);
    case GHIVE_INT: {
      int32_t x = static_cast<const int32_t *>(p)[r];
      v = (double)x;
      return !is_null_i32(x);
    }
    default: {
      int64_t x = static_cast<const int64_t *>(p)[r];
      v = (double)x;
      return !is_null_i64(x);
    }
  }
}

// ---- flattened predicate program ----------------------------------------------------------
constexpr int kMaxPredNodes = 32;
constexpr int kMaxPredCols = 8;
constexpr int kPredLitBytes = 256;  // string literals of one predicate

enum DevPredOp : int8_t {
  kLeafInt = 0,   // column (as int64) OP literal(s)
  kLeafDbl = 1,   // column (as double) OP literal(s)
  kColColInt = 2,
  kColColDbl = 3,
  kNotNull = 4,
  kAnd = 5,
  kOr = 6,
  kLeafStr = 7,     // STRING column OP string literal(s)
  kColColStr = 8,   // STRING column OP STRING column
  kNotNullStr = 9,  // STRING column is not NULL
};

struct DevPredNode {
  int8_t op;
  int8_t mode;
  int8_t a;  // column slot, or child node index for AND/OR
  int8_t b;
  int32_t pad;     // kLeafStr: number of literals
  int64_t ia, ib;  // kLeafStr: literal k at lit[ia/ib & 0xffffffff], length ia/ib >> 32
  double da, db;
};

// Nodes in evaluation order (children first); the value of the last node is the predicate.
struct DevPred {
  int32_t nnodes;
  int32_t ncols;
  const void *col_data[kMaxPredCols];
  const int32_t *col_aux[kMaxPredCols];  // STRING: (start,end) pairs
  int32_t col_type[kMaxPredCols];
  DevPredNode nodes[kMaxPredNodes];
  uint8_t lit[kPredLitBytes];
};

// Hive's string order (StringExpr.compare, storage-api .../vector/expressions/StringExpr.java:37-47):
// unsigned bytes, then length. Returns <0, 0, >0.
__device__ __forceinline__ int str_compare(const uint8_t *a, int la, const uint8_t *b, int lb) {
  const int m = la < lb ? la : lb;
  for (int i = 0; i < m; ++i)
    if (a[i] != b[i]) return (int)a[i] - (int)b[i];
  return la - lb;
}

template <class T>
__device__ __forceinline__ bool cmp_mode(int mode, T x, T a, T b) {
  switch (mode) {
    case GHIVE_FILTER_EQ: return x == a;
    case GHIVE_FILTER_LE: return x <= a;
    case GHIVE_FILTER_GE: return x >= a;
    case GHIVE_FILTER_LT: return x < a;
    case GHIVE_FILTER_GT: return x > a;
    case GHIVE_FILTER_RANGE: return x > a && x < b;
    case GHIVE_FILTER_EQ_RANGE: return x >= a && x <= b;
    case GHIVE_FILTER_NOT_EQ: return x != a;
    case GHIVE_FILTER_NOT_RANGE: return x < a || x > b;
    case GHIVE_FILTER_NOT_EQ_RANGE: return x <= a || x >= b;
    default: return true;
  }
}

// Value of node n for row r, given the values of the nodes before it (bit i of bits).
__device__ __forceinline__ bool eval_pred_node(const DevPred &P, const DevPredNode &n, uint32_t bits, uint64_t r) {
  bool v;
  switch (n.op) {
    case kLeafInt: {
      int64_t x;
      v = col_load_i64(P.col_data[n.a], P.col_type[n.a], r, x) && cmp_mode<int64_t>(n.mode, x, n.ia, n.ib);
      break;
    }
    case kLeafDbl: {
      double x;
      v = col_load_f64(P.col_data[n.a], P.col_type[n.a], r, x) && cmp_mode<double>(n.mode, x, n.da, n.db);
      break;
    }
    case kColColInt: {
      int64_t x, y;
      bool ok = col_load_i64(P.col_data[n.a], P.col_type[n.a], r, x);
      ok = col_load_i64(P.col_data[n.b], P.col_type[n.b], r, y) && ok;
      v = ok && cmp_mode<int64_t>(n.mode, x, y, 0);
      break;
    }
    case kColColDbl: {
      double x, y;
      bool ok = col_load_f64(P.col_data[n.a], P.col_type[n.a], r, x);
      ok = col_load_f64(P.col_data[n.b], P.col_type[n.b], r, y) && ok;
      v = ok && cmp_mode<double>(n.mode, x, y, 0.0);
      break;
    }
    case kNotNull: {
      double x;
      v = col_load_f64(P.col_data[n.a], P.col_type[n.a], r, x);
      break;
    }
    case kLeafStr: {
      const int32_t s = P.col_aux[n.a][2 * r], e = P.col_aux[n.a][2 * r + 1];
      if (s < 0) {
        v = false;
        break;
      }
      const uint8_t *x = static_cast<const uint8_t *>(P.col_data[n.a]) + s;
      const int c0 = str_compare(x, e - s, P.lit + (uint32_t)n.ia, (int)(n.ia >> 32));
      const int c1 = n.pad == 2 ? str_compare(x, e - s, P.lit + (uint32_t)n.ib, (int)(n.ib >> 32)) : 0;
      v = cmp_mode<int>(n.mode, 0, -c0, -c1);  // 0 OP -c  <=>  string OP literal
      break;
    }
    case kColColStr: {
      const int32_t s0 = P.col_aux[n.a][2 * r], e0 = P.col_aux[n.a][2 * r + 1];
      const int32_t s1 = P.col_aux[n.b][2 * r], e1 = P.col_aux[n.b][2 * r + 1];
      if (s0 < 0 || s1 < 0) {
        v = false;
        break;
      }
      const int c = str_compare(static_cast<const uint8_t *>(P.col_data[n.a]) + s0, e0 - s0,
                                static_cast<const uint8_t *>(P.col_data[n.b]) + s1, e1 - s1);
      v = cmp_mode<int>(n.mode, 0, -c, 0);
      break;
    }
    case kNotNullStr: v = P.col_aux[n.a][2 * r] >= 0; break;
    case kAnd: v = ((bits >> n.a) & 1u) && ((bits >> n.b) & 1u); break;
    default: v = ((bits >> n.a) & 1u) || ((bits >> n.b) & 1u); break;
  }
  return v;
}

__device__ __forceinline__ bool eval_pred(const DevPred &P, uint64_t r) {
  uint32_t bits = 0;
  for (int i = 0; i < P.nnodes; ++i) bits |= (uint32_t)eval_pred_node(P, P.nodes[i], bits, r) << i;
  return (bits >> (P.nnodes - 1)) & 1u;
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ---- keys and skew for the join/sort workloads ------------------------------------------------
// A bijection of [0, 2^62): add a seed-derived offset, then xorshift / odd-multiply rounds, all
// mod 2^62. Distinct ranks give distinct keys; keys stay below LONG_MAX (the NULL sentinel).
__host__ __device__ __forceinline__ uint64_t key62(uint64_t seed, uint64_t rank) {
  const uint64_t M = (1ull << 62) - 1;
  uint64_t x = (rank + seed * 0x9E3779B97F4A7C15ull) & M;
  x ^= x >> 31;
  x = (x * 0x7fb5d329728ea185ull) & M;
  x ^= x >> 27;
  x = (x * 0x81dadef4bc2dd44dull) & M;
  x ^= x >> 33;
  return x;
}

// Rejection-inversion sampling of a Zipf(s) rank in [1, n] (W. Hoermann, G. Derflinger,
// "Rejection-inversion to generate variates from monotone discrete distributions", ACM TOMACS
// 6(3), 1996): invert the integral H of h(x) = x^-s, then accept or reject the rounded value.
struct ZipfParams {
  double s, n, h_x1, h_n, sx;
};
__host__ __device__ __forceinline__ double zipf_helper1(double x) {  // log1p(x)/x
  return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
}
__host__ __device__ __forceinline__ double zipf_helper2(double x) {  // expm1(x)/x
  return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
}
__host__ __device__ __forceinline__ double zipf_H(double s, double x) {
  double lx = log(x);
  return zipf_helper2((1.0 - s) * lx) * lx;
}
__host__ __device__ __forceinline__ double zipf_h(double s, double x) { return exp(-s * log(x)); }
__host__ __device__ __forceinline__ double zipf_Hinv(double s, double y) {
  double t = y * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zipf_helper1(t) * y);
}
inline ZipfParams zipf_params(double s, uint64_t n) {
  ZipfParams p;
  p.s = s;
  p.n = (double)n;
  p.h_x1 = zipf_H(s, 1.5) - 1.0;
  p.h_n = zipf_H(s, p.n + 0.5);
  p.sx = 2.0 - zipf_Hinv(s, zipf_H(s, 2.5) - zipf_h(s, 2.0));
  return p;
}
// One attempt with uniform u in [0,1); returns the rank, or 0 when rejected.
__host__ __device__ __forceinline__ uint64_t zipf_try(const ZipfParams &p, double u) {
  double uu = p.h_n + u * (p.h_x1 - p.h_n);
  double x = zipf_Hinv(p.s, uu);
  double kf = floor(x + 0.5);
  if (kf < 1.0) kf = 1.0;
  if (kf > p.n) kf = p.n;
  if (kf - x <= p.sx || uu >= zipf_H(p.s, kf + 0.5) - zipf_h(p.s, kf)) return (uint64_t)kf;
  return 0;
}

// ---- hashing -------------------------------------------------------------------------------
// Fibonacci hashing: multiply by 2^64/phi and keep the top bits (Knuth, TAOCP 6.4). The
// multipliers are odd, so the full product determines the key (kFibInv*: their inverses modulo
// 2^64 / 2^32, by Newton's iteration x <- x (2 - a x), which doubles the correct low bits).
constexpr uint64_t kFibMul64 = 0x9E3779B97F4A7C15ull;
constexpr uint32_t kFibMul32 = 0x9E3779B1u;
constexpr uint64_t inverse_odd64(uint64_t a) {
  uint64_t x = a;  // correct to 3 bits for any odd a
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}
constexpr uint64_t kFibInv64 = inverse_odd64(kFibMul64);
constexpr uint32_t kFibInv32 = (uint32_t)inverse_odd64(kFibMul32);
static_assert(kFibMul64 * kFibInv64 == 1ull, "inverse modulo 2^64");
static_assert((uint32_t)(kFibMul32 * kFibInv32) == 1u, "inverse modulo 2^32");
__device__ __forceinline__ uint64_t hash_u64(uint64_t k, int log2_slots) {
  return (k * kFibMul64) >> (64 - log2_slots);
}
__device__ __forceinline__ uint32_t hash_u32(uint32_t k, int log2_slots) {
  return (uint32_t)(((uint64_t)(k * kFibMul32)) >> (32 - log2_slots));
}

}  // namespace ghive
