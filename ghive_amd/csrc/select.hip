// Projection arithmetic: CalculationSelectPredicate::process (gpu_impl/src/Operator/
// SelectPredicate.cu:84-409) with its functors (gpu_impl/inc/Operator/SelectPredicate.hpp:193-234)
// as one elementwise kernel per expression node.
#include "common.hpp"
#include "device.hpp"

namespace ghive {

namespace {

constexpr int kThreads = 256;

enum ArithKind { AR_I32 = 0, AR_I64 = 1, AR_F64 = 2 };

// Operand value of row i as the arithmetic type (a 1-row column broadcasts); false when NULL.
__device__ __forceinline__ bool operand_i64(const void *p, int type, uint64_t rows, uint64_t i, int64_t &v) {
  const uint64_t r = rows == 1 ? 0 : i;
  if (type == GHIVE_INT) {
    const int32_t x = static_cast<const int32_t *>(p)[r];
    v = x;
    return x != 0x7fffffff;
  }
  const int64_t x = static_cast<const int64_t *>(p)[r];
  v = x;
  return x != 0x7fffffffffffffffll;
}

__device__ __forceinline__ bool operand_f64(const void *p, int type, uint64_t rows, uint64_t i, double &v) {
  const uint64_t r = rows == 1 ? 0 : i;
  switch (type) {
    case GHIVE_INT: {
      const int32_t x = static_cast<const int32_t *>(p)[r];
      v = (double)x;
      return x != 0x7fffffff;
    }
    case GHIVE_LONG: {
      const int64_t x = static_cast<const int64_t *>(p)[r];
      v = (double)x;
      return x != 0x7fffffffffffffffll;
    }
    case GHIVE_FLOAT: {
      const float x = static_cast<const float *>(p)[r];
      v = (double)x;
      return !__builtin_isnan(x);
    }
    default: {
      const double x = static_cast<const double *>(p)[r];
      v = x;
      return !__builtin_isnan(x);
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_arith(const void *a, int ta, uint64_t ra, const void *b, int tb,
                                                    uint64_t rb, char op, int kind, uint64_t n, void *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    if (kind == AR_F64) {
      double x, y;
      const bool ok = (int)operand_f64(a, ta, ra, i, x) & (int)operand_f64(b, tb, rb, i, y);
      double r = __builtin_nan("");
      if (ok) {
        switch (op) {
          case '+': r = x + y; break;
          case '-': r = x - y; break;
          case '*': r = x * y; break;
          default: r = y == 0.0 ? __builtin_nan("") : x / y; break;  // Hive: x / 0 is NULL
        }
      }
      static_cast<double *>(out)[i] = r;
    } else {
      int64_t x, y;
      const bool ok = (int)operand_i64(a, ta, ra, i, x) & (int)operand_i64(b, tb, rb, i, y);
      // integer arithmetic wraps like Java's int / long (two's complement, through unsigned)
      uint64_t r = 0;
      switch (op) {
        case '+': r = (uint64_t)x + (uint64_t)y; break;
        case '-': r = (uint64_t)x - (uint64_t)y; break;
        default: r = (uint64_t)x * (uint64_t)y; break;
      }
      if (kind == AR_I32)
        static_cast<int32_t *>(out)[i] = ok ? (int32_t)(uint32_t)r : 0x7fffffff;
      else
        static_cast<int64_t *>(out)[i] = ok ? (int64_t)r : 0x7fffffffffffffffll;
    }
  }
}

}  // namespace

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_select_arith(ghive_ctx *ctx, const ghive_column *a, const ghive_column *b, char op, uint64_t n,
                       ghive_column *out) {
  GH_CHECK_CTX(ctx);
  if (!a || !b || !out) return set_error(ctx, GHIVE_ERR_INVALID, "select_arith: NULL argument");
  if (op != '+' && op != '-' && op != '*' && op != '/')
    return set_error(ctx, GHIVE_ERR_INVALID, "select_arith: operator '%c'", op);
  for (const ghive_column *c : {a, b}) {
    if (!is_numeric_type(c->type)) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "select_arith: STRING operand");
    if (c->row_num != n && c->row_num != 1)
      return set_error(ctx, GHIVE_ERR_INVALID, "select_arith: operand of %llu rows for %llu",
                       (unsigned long long)c->row_num, (unsigned long long)n);
    if (n && !c->data) return set_error(ctx, GHIVE_ERR_INVALID, "select_arith: operand has no data");
  }
  // result type: '/' is DOUBLE; otherwise INT op INT is INT, integer op LONG is LONG (Hive's
  // widening; the reference accepts only equal types), anything with a floating operand DOUBLE
  const bool fp = a->type == GHIVE_DOUBLE || a->type == GHIVE_FLOAT || b->type == GHIVE_DOUBLE ||
                  b->type == GHIVE_FLOAT;
  int kind, otype;
  if (op == '/' || fp) {
    kind = AR_F64;
    otype = GHIVE_DOUBLE;
  } else if (a->type == GHIVE_INT && b->type == GHIVE_INT) {
    kind = AR_I32;
    otype = GHIVE_INT;
  } else {
    kind = AR_I64;
    otype = GHIVE_LONG;
  }
  if (n && !out->data) return set_error(ctx, GHIVE_ERR_INVALID, "select_arith: output has no buffer");
  out->type = otype;
  out->row_num = n;
  if (!n) return GHIVE_OK;
  const uint64_t g = (n + kThreads - 1) / kThreads, cap = (uint64_t)ctx->cus * 16;
  k_arith<<<(uint32_t)(g > cap ? cap : g), kThreads, 0, ctx->stream>>>(a->data, a->type, a->row_num, b->data, b->type,
                                                                        b->row_num, op, kind, n, out->data);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

}  // extern "C"
