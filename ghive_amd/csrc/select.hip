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

// ---- STRING functions ------------------------------------------------------------------------
// Character.isWhitespace for an ASCII byte (HiveStringUtils.findLeadingSpaces / findTrailingSpaces).
__device__ __forceinline__ bool java_ascii_space(uint8_t c) {
  return c == ' ' || (c >= 0x09 && c <= 0x0d) || (c >= 0x1c && c <= 0x1f);
}

// CAST(string AS INT / BIGINT): UDFToInteger.evaluate(Text) / UDFToLong.evaluate(Text)
// (ql/.../udf/UDFToInteger.java:169-186) = LazyUtils.isNumberMaybe on the raw bytes, then
// LazyInteger.parseInt / LazyLong.parseLong with trim (serde2/lazy/LazyInteger.java): ASCII
// whitespace trimmed, an optional sign, decimal digits accumulated negatively with overflow
// checks, a '.' ends the integer part (the rest must be digits: truncation). Anything else is NULL.
template <class T>
__device__ bool parse_integer(const uint8_t *b, int32_t len, T &out) {
  if (len == 0) return false;  // isNumberMaybe
  auto digit = [](uint8_t c) { return c >= '0' && c <= '9'; };
  if (len == 1 && !digit(b[0])) return false;
  if (len == 2 && !digit(b[0]) && !digit(b[1])) return false;
  if (len == 4 && (b[0] == 'N' || b[0] == 'n')) return false;
  int32_t s = 0, e = len;
  while (s < e && java_ascii_space(b[s])) ++s;
  while (e > s && java_ascii_space(b[e - 1])) --e;
  if (s == e) return false;
  const bool negative = b[s] == '-';
  int32_t i = s;
  if (negative || b[s] == '+') {
    ++i;
    if (e - s == 1) return false;
  }
  const T lim = (T)(((T)1) << (sizeof(T) * 8 - 1));  // MIN_VALUE
  const T max = lim / 10;
  T result = 0;
  for (; i < e; ++i) {
    const uint8_t c = b[i];
    if (!digit(c)) {
      if (c == '.') break;
      return false;
    }
    if (max > result) return false;  // result * 10 would pass MIN_VALUE
    const T r10 = (T)(result * 10), d = (T)(c - '0');
    if (r10 < lim + d) return false;  // r10 - d would pass it (Java: the wrapped next > result)
    result = (T)(r10 - d);
  }
  if (i < e) {  // the fraction after '.': digits only
    for (++i; i < e; ++i)
      if (!digit(b[i])) return false;
  }
  if (!negative) {  // -MIN_VALUE does not fit: Java's negation wraps and its check rejects it
    if (result == lim) return false;
    result = (T)(-result);
  }
  out = result;
  return true;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_cast_str_int(const uint8_t *blob, const int32_t *pairs, uint64_t n,
                                                           T null_value, T *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int32_t s = pairs[2 * i], e = pairs[2 * i + 1];
    T v = null_value;
    if (s >= 0 && e >= s && !parse_integer<T>(blob + s, e - s, v)) v = null_value;
    out[i] = v;
  }
}

// Numeric CAST as Java's conversions: integer widening keeps the value, narrowing truncates to the
// low bits, floating to integer rounds toward zero and saturates (NaN is NULL here, its sentinel).
__global__ __launch_bounds__(kThreads) void k_cast_num(const void *in, int tin, uint64_t n, int tout, void *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const bool fin = tin == GHIVE_DOUBLE || tin == GHIVE_FLOAT;
    int64_t x = 0;
    double d = 0;
    const bool ok = fin ? operand_f64(in, tin, n, i, d) : operand_i64(in, tin, n, i, x);
    switch (tout) {
      case GHIVE_INT: {
        int32_t v = 0x7fffffff;
        if (ok) {
          if (!fin) v = (int32_t)(uint32_t)(uint64_t)x;
          else v = d >= 2147483647.0 ? 2147483647 : (d <= -2147483648.0 ? (int32_t)0x80000000 : (int32_t)d);
        }
        static_cast<int32_t *>(out)[i] = v;
        break;
      }
      case GHIVE_LONG: {
        int64_t v = 0x7fffffffffffffffll;
        if (ok) {
          if (!fin) v = x;
          else v = d >= 9223372036854775807.0 ? 0x7fffffffffffffffll
                                              : (d <= -9223372036854775808.0 ? (int64_t)0x8000000000000000ull : (int64_t)d);
        }
        static_cast<int64_t *>(out)[i] = v;
        break;
      }
      case GHIVE_FLOAT:
        static_cast<float *>(out)[i] = ok ? (fin ? (float)d : (float)x) : __builtin_nanf("");
        break;
      default:
        static_cast<double *>(out)[i] = ok ? (fin ? d : (double)x) : __builtin_nan("");
        break;
    }
  }
}

// UTF-16 code units of the UTF-8 sequence starting with byte c (Text.toString then String.length):
// 4-byte sequences are surrogate pairs; a stray continuation byte decodes to one replacement unit.
__device__ __forceinline__ int utf16_units(uint8_t c) { return (c & 0xf8) == 0xf0 ? 2 : 1; }
__device__ __forceinline__ int utf8_len(uint8_t c) {
  return c < 0x80 ? 1 : ((c & 0xe0) == 0xc0 ? 2 : ((c & 0xf0) == 0xe0 ? 3 : ((c & 0xf8) == 0xf0 ? 4 : 1)));
}

// SUBSTR(str, pos, len): UDFSubstr.evaluate(Text, IntWritable, IntWritable) with makeIndex
// (ql/.../udf/UDFSubstr.java:61-105): positions count UTF-16 units, pos is 1-based (negative:
// from the end, 0: the start), len <= 0 or |pos| > length gives the empty string. The result is
// a (start,end) byte range of the same blob.
__global__ __launch_bounds__(kThreads) void k_substr(const uint8_t *blob, const int32_t *pairs, uint64_t n, int32_t pos,
                                                     int32_t len, int32_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const int32_t s = pairs[2 * i], e = pairs[2 * i + 1];
    int32_t os = -1, oe = -1;
    if (s >= 0 && e >= s) {
      os = oe = s;  // the empty string
      int64_t L = 0;
      for (int32_t p = s; p < e;) {
        L += utf16_units(blob[p]);
        const int32_t step = utf8_len(blob[p]);
        p = p + step > e ? e : p + step;
      }
      const int64_t apos = pos < 0 ? -(int64_t)pos : pos;
      if (len > 0 && apos <= L) {
        const int64_t start = pos > 0 ? pos - 1 : (pos < 0 ? L + pos : 0);
        const int64_t end = (L - start < len) ? L : start + len;
        // byte offsets of units start and end (a boundary inside a surrogate pair moves past the pair)
        int64_t u = 0;
        int32_t p = s, bs = e, be = e;
        while (p < e) {
          if (u >= start && bs == e) bs = p;
          if (u >= end) {
            be = p;
            break;
          }
          u += utf16_units(blob[p]);
          const int32_t step = utf8_len(blob[p]);
          p = p + step > e ? e : p + step;
        }
        if (bs == e && u >= start) bs = p;
        os = bs;
        oe = be < bs ? bs : be;
      }
    }
    out[2 * i] = os;
    out[2 * i + 1] = oe;
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

int ghive_select_cast(ghive_ctx *ctx, const ghive_column *in, int32_t to_type, uint64_t n, ghive_column *out) {
  GH_CHECK_CTX(ctx);
  if (!in || !out) return set_error(ctx, GHIVE_ERR_INVALID, "select_cast: NULL argument");
  if (!is_numeric_type(to_type)) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "select_cast: to type %d", to_type);
  if (in->row_num != n) return set_error(ctx, GHIVE_ERR_INVALID, "select_cast: input of %llu rows for %llu",
                                         (unsigned long long)in->row_num, (unsigned long long)n);
  if (n && (!out->data || !in->data || (in->type == GHIVE_STRING && !in->aux)))
    return set_error(ctx, GHIVE_ERR_INVALID, "select_cast: a buffer is NULL");
  out->type = to_type;
  out->row_num = n;
  if (!n) return GHIVE_OK;
  const uint64_t g = (n + kThreads - 1) / kThreads, cap = (uint64_t)ctx->cus * 16;
  const uint32_t grid = (uint32_t)(g > cap ? cap : g);
  if (in->type == GHIVE_STRING) {
    const uint8_t *blob = static_cast<const uint8_t *>(in->data);
    const int32_t *pairs = static_cast<const int32_t *>(in->aux);
    if (to_type == GHIVE_INT)
      k_cast_str_int<int32_t><<<grid, kThreads, 0, ctx->stream>>>(blob, pairs, n, 0x7fffffff,
                                                                  static_cast<int32_t *>(out->data));
    else if (to_type == GHIVE_LONG)
      k_cast_str_int<int64_t><<<grid, kThreads, 0, ctx->stream>>>(blob, pairs, n, 0x7fffffffffffffffll,
                                                                  static_cast<int64_t *>(out->data));
    else
      return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "select_cast: STRING to a floating type");
  } else {
    k_cast_num<<<grid, kThreads, 0, ctx->stream>>>(in->data, in->type, n, to_type, out->data);
  }
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_select_substr(ghive_ctx *ctx, const ghive_column *str, int32_t pos, int32_t len, uint64_t n,
                        ghive_column *out) {
  GH_CHECK_CTX(ctx);
  if (!str || !out) return set_error(ctx, GHIVE_ERR_INVALID, "select_substr: NULL argument");
  if (str->type != GHIVE_STRING) return set_error(ctx, GHIVE_ERR_INVALID, "select_substr: not a STRING column");
  if (str->row_num != n) return set_error(ctx, GHIVE_ERR_INVALID, "select_substr: input of %llu rows for %llu",
                                          (unsigned long long)str->row_num, (unsigned long long)n);
  if (n && (!str->aux || !out->aux)) return set_error(ctx, GHIVE_ERR_INVALID, "select_substr: a buffer is NULL");
  out->type = GHIVE_STRING;
  out->row_num = n;
  out->data = str->data;
  out->char_size = str->char_size;
  if (!n) return GHIVE_OK;
  const uint64_t g = (n + kThreads - 1) / kThreads, cap = (uint64_t)ctx->cus * 16;
  k_substr<<<(uint32_t)(g > cap ? cap : g), kThreads, 0, ctx->stream>>>(
      static_cast<const uint8_t *>(str->data), static_cast<const int32_t *>(str->aux), n, pos, len,
      static_cast<int32_t *>(out->aux));
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

}  // extern "C"
