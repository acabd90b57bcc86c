// Small device utilities: fill kernels (used instead of hipMemsetAsync so that the whole
// operator sequence stays capturable in a HIP graph) and a reduce-then-scan exclusive prefix
// sum over uint32 counts (filter tile counts, radix-sort digit histograms).
#include "common.hpp"
#include "device.hpp"
#include "scan.hpp"

namespace ghive {

__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u64_strided(uint64_t *p, uint64_t v, uint64_t n, uint64_t step) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i * step] = v;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

static uint32_t fill_grid(ghive_ctx *ctx, uint64_t n) {
  uint64_t g = (n + 255) / 256;
  uint64_t cap = (uint64_t)ctx->cus * 8;
  return (uint32_t)(g < 1 ? 1 : (g > cap ? cap : g));
}

int fill_u64(ghive_ctx *ctx, uint64_t *p, uint64_t value, uint64_t n) {
  if (!n) return GHIVE_OK;
  k_fill_u64<<<fill_grid(ctx, n), 256, 0, ctx->stream>>>(p, value, n);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int fill_u64_strided(ghive_ctx *ctx, uint64_t *p, uint64_t value, uint64_t n, uint64_t step) {
  if (!n) return GHIVE_OK;
  k_fill_u64_strided<<<fill_grid(ctx, n), 256, 0, ctx->stream>>>(p, value, n, step);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

int fill_u32(ghive_ctx *ctx, uint32_t *p, uint32_t value, uint64_t n) {
  if (!n) return GHIVE_OK;
  k_fill_u32<<<fill_grid(ctx, n), 256, 0, ctx->stream>>>(p, value, n);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

// LONG -> 32-bit copy for kernels that stream int32 columns: out[i] = in[i], LONG NULL -> INT NULL.
// *bad becomes nonzero when a non-NULL value lies outside [INT_MIN, INT_MAX - 1] (INT_MAX is the
// INT NULL sentinel), i.e. when the copy would not hold the same values.
__global__ void k_narrow_i64(const int64_t *in, uint64_t n, int32_t *out, unsigned long long *bad) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  bool b = false;
  for (; i < n; i += stride) {
    const int64_t v = in[i];
    const bool null = v == 0x7fffffffffffffffll;
    b |= !null && (v < -2147483648ll || v > 2147483646ll);
    out[i] = null ? 0x7fffffff : (int32_t)v;
  }
  if (__ballot(b) && lane_id() == 0) atomicOr(bad, 1ull);
}

int narrow_long(ghive_ctx *ctx, const int64_t *in, uint64_t n, int32_t *out, uint64_t *bad_dev) {
  if (!n) return GHIVE_OK;
  k_narrow_i64<<<fill_grid(ctx, n), 256, 0, ctx->stream>>>(in, n, out, reinterpret_cast<unsigned long long *>(bad_dev));
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr uint64_t kScanChunk = (uint64_t)kScanThreads * kScanPerThread;

// per-chunk totals
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *in, uint64_t n,
                                                              uint32_t *sums) {
  __shared__ uint32_t lds[kScanThreads / kWave];
  uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
  uint32_t s = 0;
  for (int j = 0; j < kScanPerThread; ++j) {
    uint64_t i = base + (uint64_t)j * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t total;
  (void)block_exclusive_scan_256(s, lds, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of one chunk, offset by offs[block] (offs == nullptr: 0)
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t *in, uint32_t *out,
                                                             uint64_t n, const uint32_t *offs,
                                                             uint64_t *d_total, uint32_t nblocks) {
  __shared__ uint32_t lds[kScanThreads / kWave];
  uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanPerThread;
  uint32_t v[kScanPerThread];
  uint32_t s = 0;
  for (int j = 0; j < kScanPerThread; ++j) {
    uint64_t i = base + j;
    v[j] = i < n ? in[i] : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan_256(s, lds, total);
  uint32_t off = offs ? offs[blockIdx.x] : 0u;
  uint32_t run = off + pre;
  for (int j = 0; j < kScanPerThread; ++j) {
    uint64_t i = base + j;
    if (i < n) out[i] = run;
    run += v[j];
  }
  if (d_total && blockIdx.x == nblocks - 1 && threadIdx.x == 0) *d_total = (uint64_t)off + total;
}

int exclusive_scan_u32(ghive_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n,
                       uint64_t *d_total) {
  if (n == 0) {
    if (d_total) return fill_u64(ctx, d_total, 0, 1);
    return GHIVE_OK;
  }
  uint32_t nb = ceil_div(n, kScanChunk);
  if (nb == 1) {
    k_scan_apply<<<1, kScanThreads, 0, ctx->stream>>>(in, out, n, nullptr, d_total, 1);
    GH_LAUNCH(ctx);
    return GHIVE_OK;
  }
  DevBuf sums;
  GH_TRY(sums.alloc(ctx, sizeof(uint32_t) * nb));
  k_scan_reduce<<<nb, kScanThreads, 0, ctx->stream>>>(in, n, sums.as<uint32_t>());
  GH_LAUNCH(ctx);
  GH_TRY(exclusive_scan_u32(ctx, sums.as<uint32_t>(), sums.as<uint32_t>(), nb, nullptr));
  k_scan_apply<<<nb, kScanThreads, 0, ctx->stream>>>(in, out, n, sums.as<uint32_t>(), d_total, nb);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

// ---- STREAM-style bandwidth probes (achievable HBM read / copy rate of the device) --------
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Each workgroup streams one contiguous range (consecutive DRAM pages), four 16-byte
// non-temporal loads per thread in flight -- the access pattern of the fused kernels.
__global__ __launch_bounds__(256) void k_stream_read(const v4u *p, uint64_t n16, unsigned long long *sink) {
  const uint64_t b0 = n16 * blockIdx.x / gridDim.x, b1 = n16 * (blockIdx.x + 1) / gridDim.x;
  v4u acc = {0u, 0u, 0u, 0u};
  uint64_t i = b0 + threadIdx.x;
  for (; i + 3 * blockDim.x < b1; i += 4 * blockDim.x) {
    const v4u a = __builtin_nontemporal_load(p + i), b = __builtin_nontemporal_load(p + i + blockDim.x);
    const v4u c = __builtin_nontemporal_load(p + i + 2 * blockDim.x), d = __builtin_nontemporal_load(p + i + 3 * blockDim.x);
    acc ^= a ^ b ^ c ^ d;
  }
  for (; i < b1; i += blockDim.x) acc ^= __builtin_nontemporal_load(p + i);
  uint32_t x = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (x == 0x9e3779b9u) atomicAdd(sink, 1ull);  // keeps the loads alive; almost never taken
}

// The fused star-join kernel's pattern without its work: src holds kStreamCols equal columns and
// each workgroup streams the same contiguous range of every column, a 16-byte non-temporal load per
// column per lane, two row groups in flight.
constexpr int kStreamCols = 5;
__global__ __launch_bounds__(256) void k_stream_read_cols(const v4u *p, uint64_t n16, unsigned long long *sink) {
  const uint64_t b0 = n16 * blockIdx.x / gridDim.x, b1 = n16 * (blockIdx.x + 1) / gridDim.x;
  v4u acc = {0u, 0u, 0u, 0u};
  uint64_t i = b0 + threadIdx.x;
  for (; i + blockDim.x < b1; i += 2 * blockDim.x) {
    v4u x[2 * kStreamCols];
#pragma unroll
    for (int c = 0; c < kStreamCols; ++c) {
      x[2 * c] = __builtin_nontemporal_load(p + (uint64_t)c * n16 + i);
      x[2 * c + 1] = __builtin_nontemporal_load(p + (uint64_t)c * n16 + i + blockDim.x);
    }
#pragma unroll
    for (int c = 0; c < 2 * kStreamCols; ++c) acc ^= x[c];
  }
  for (; i < b1; i += blockDim.x)
#pragma unroll
    for (int c = 0; c < kStreamCols; ++c) acc ^= __builtin_nontemporal_load(p + (uint64_t)c * n16 + i);
  uint32_t x = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (x == 0x9e3779b9u) atomicAdd(sink, 1ull);  // keeps the loads alive; almost never taken
}

__global__ __launch_bounds__(256) void k_stream_copy(const v4u *src, v4u *dst, uint64_t n16) {
  const uint64_t b0 = n16 * blockIdx.x / gridDim.x, b1 = n16 * (blockIdx.x + 1) / gridDim.x;
  uint64_t i = b0 + threadIdx.x;
  for (; i + blockDim.x < b1; i += 2 * blockDim.x) {
    const v4u a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + blockDim.x);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + blockDim.x);
  }
  for (; i < b1; i += blockDim.x) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace ghive

extern "C" int ghive_bench_stream(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes, int mode,
                                  int iters, float *ms_per_iter) {
  using namespace ghive;
  GH_CHECK_CTX(ctx);
  if (!src || (mode == 1 && !dst) || mode < 0 || mode > 2 || bytes < 16 || (bytes & 15) || iters < 1 ||
      !ms_per_iter || ((uintptr_t)src & 15) || (dst && ((uintptr_t)dst & 15)) ||
      (mode == 2 && bytes % (16 * kStreamCols)))
    return set_error(ctx, GHIVE_ERR_INVALID, "bench_stream: bad arguments");
  const uint64_t n16 = bytes / 16;
  const uint32_t grid = (uint32_t)ctx->cus * 8;
  GH_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int it = 0; it < iters; ++it) {
    if (mode == 0)
      k_stream_read<<<grid, 256, 0, ctx->stream>>>(static_cast<const v4u *>(src), n16,
                                                    reinterpret_cast<unsigned long long *>(ctx->d_counters));
    else if (mode == 2)
      k_stream_read_cols<<<grid, 256, 0, ctx->stream>>>(static_cast<const v4u *>(src), n16 / kStreamCols,
                                                         reinterpret_cast<unsigned long long *>(ctx->d_counters));
    else
      k_stream_copy<<<grid, 256, 0, ctx->stream>>>(static_cast<const v4u *>(src), static_cast<v4u *>(dst), n16);
    GH_LAUNCH(ctx);
  }
  GH_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  GH_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  GH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  *ms_per_iter = ms / (float)iters;
  return GHIVE_OK;
}
