// Synthetic column generation on the device: counter-based Philox4x32-10, counter =
// (row_lo, row_hi, column_id, 0), key = (seed_lo, seed_hi). Row r of column c is a pure function
// of (seed, c, r), so the host can regenerate any slice bit-for-bit for integer columns.
#include "common.hpp"
#include "device.hpp"

namespace ghive {

__global__ void k_gen_i32(uint64_t seed, uint32_t col, uint64_t row0, uint64_t n, int32_t lo,
                          uint64_t range, int32_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = row0 + i;
    uint32_t w[4];
    philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    out[i] = (int32_t)((int64_t)lo + (int64_t)(((uint64_t)w[0] * range) >> 32));
  }
}

__global__ void k_gen_i64(uint64_t seed, uint32_t col, uint64_t row0, uint64_t n, int64_t lo,
                          uint64_t range, int64_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = row0 + i;
    uint32_t w[4];
    philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    uint64_t x = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    uint64_t off = range ? __umul64hi(x, range) : x;
    out[i] = (int64_t)((uint64_t)lo + off);
  }
}

__global__ void k_gen_f32(uint64_t seed, uint32_t col, uint64_t row0, uint64_t n, float *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = row0 + i;
    uint32_t w[4];
    philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    out[i] = (float)(w[0] >> 8) * (1.0f / 16777216.0f);
  }
}

__global__ void k_gen_normal(uint64_t seed, uint32_t col, uint64_t row0, uint64_t n, double *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const double two_pi = 6.283185307179586476925286766559;
  for (; i < n; i += stride) {
    uint64_t r = row0 + i;
    uint32_t w[4];
    philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    uint64_t a = ((uint64_t)w[0] << 32) | w[1];
    uint64_t b = ((uint64_t)w[2] << 32) | w[3];
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
}

__global__ void k_gen_rank_keys(uint64_t seed, uint64_t rank0, uint64_t n, int64_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (int64_t)key62(seed, rank0 + i);
}

// Probe keys: with probability miss_frac a key of a rank above nranks (never a build key), else
// the key of a Zipf(s) rank in [1, nranks]. Attempt a of row r draws Philox counter
// (r_lo, r_hi, column, a); rejection-inversion accepts in ~1.1 attempts.
__global__ void k_gen_zipf_keys(uint64_t seed, uint32_t col, uint64_t row0, uint64_t n, ZipfParams zp,
                                uint64_t nranks, uint32_t miss_threshold, int64_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t r = row0 + i;
    uint32_t w[4];
    philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    uint64_t rank;
    if (w[0] < miss_threshold) {
      rank = nranks + 1 + (((uint64_t)w[1] << 8) | (w[2] & 0xffu));  // 40-bit offset past the build ranks
    } else {
      rank = 0;
      for (uint32_t a = 0; a < 64 && rank == 0; ++a) {
        if (a) philox4x32_10((uint32_t)r, (uint32_t)(r >> 32), col, a, (uint32_t)seed, (uint32_t)(seed >> 32), w);
        const uint64_t bits = ((uint64_t)w[2] << 32) | w[3];
        rank = zipf_try(zp, (double)(bits >> 11) * (1.0 / 9007199254740992.0));
      }
      if (rank == 0) rank = 1;  // 64 rejections in a row (probability ~1e-60)
    }
    out[i] = (int64_t)key62(seed, rank);
  }
}

template <class T>
__global__ void k_seq(uint64_t row0, uint64_t n, T start, T *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (T)(start + (T)(row0 + i));
}

static uint32_t gen_grid(ghive_ctx *ctx, uint64_t n) {
  uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 16;
  return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_gen_uniform_i32(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, int32_t lo, int32_t hi, int32_t *out) {
  GH_CHECK_CTX(ctx);
  if (hi < lo) return set_error(ctx, GHIVE_ERR_INVALID, "gen: hi < lo");
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1u;
  k_gen_i32<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, column_id, row_begin, n, lo, range, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_uniform_i64(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, int64_t lo, int64_t hi, int64_t *out) {
  GH_CHECK_CTX(ctx);
  if (hi < lo) return set_error(ctx, GHIVE_ERR_INVALID, "gen: hi < lo");
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  uint64_t range = (uint64_t)hi - (uint64_t)lo + 1u;
  k_gen_i64<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, column_id, row_begin, n, lo, range, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_sequence_i32(ghive_ctx *ctx, uint64_t row_begin, uint64_t n, int32_t start, int32_t *out) {
  GH_CHECK_CTX(ctx);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  k_seq<int32_t><<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(row_begin, n, start, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_sequence_i64(ghive_ctx *ctx, uint64_t row_begin, uint64_t n, int64_t start, int64_t *out) {
  GH_CHECK_CTX(ctx);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  k_seq<int64_t><<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(row_begin, n, start, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_uniform_f32(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, float *out) {
  GH_CHECK_CTX(ctx);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  k_gen_f32<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, column_id, row_begin, n, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_rank_keys(ghive_ctx *ctx, uint64_t seed, uint64_t rank_begin, uint64_t n, int64_t *out) {
  GH_CHECK_CTX(ctx);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  k_gen_rank_keys<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, rank_begin, n, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_zipf_keys(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                        uint64_t nranks, double theta, double miss_frac, int64_t *out) {
  GH_CHECK_CTX(ctx);
  if (nranks < 1 || nranks > (1ull << 40) || !(theta > 0.0) || theta > 10.0 || !(miss_frac >= 0.0) ||
      miss_frac > 1.0)
    return set_error(ctx, GHIVE_ERR_INVALID, "gen_zipf: nranks %llu theta %g miss %g", (unsigned long long)nranks,
                     theta, miss_frac);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  ZipfParams zp = zipf_params(theta, nranks);
  const uint32_t thr = miss_frac >= 1.0 ? 0xffffffffu : (uint32_t)(miss_frac * 4294967296.0);
  k_gen_zipf_keys<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, column_id, row_begin, n, zp, nranks, thr, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int ghive_gen_normal_f64(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                         uint64_t n, double *out) {
  GH_CHECK_CTX(ctx);
  if (!n) return GHIVE_OK;
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "gen: out is NULL");
  k_gen_normal<<<gen_grid(ctx, n), 256, 0, ctx->stream>>>(seed, column_id, row_begin, n, out);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

}  // extern "C"
