// Shared host-side infrastructure of libghive_mi355x.so: the context (stream, last error,
// device memory manager, stage timers) and the status/error helpers used by every entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <set>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ghive/ghive.h"

namespace ghive {

// Device memory manager. The reference has none (every cudaMalloc in gpu_impl leaks); here all
// intermediates come from chunks obtained once with hipMalloc and recycled through a first-fit
// free list with coalescing, so repeated fragments (batched and serving forms) do not touch the
// HIP allocator and the footprint is accounted for admission control.
class DeviceArena {
 public:
  explicit DeviceArena(uint64_t limit) : limit_(limit) {}
  ~DeviceArena();
  // Returns nullptr when the limit would be exceeded or hipMalloc fails.
  void *alloc(uint64_t bytes);
  bool free(void *p);
  void release_all();
  uint64_t in_use() const { return in_use_; }
  uint64_t peak() const { return peak_; }
  uint64_t limit() const { return limit_; }
  uint64_t reserved() const { return reserved_; }
  void set_limit(uint64_t l) { limit_ = l; }

 private:
  struct Chunk {
    char *base;
    uint64_t size;
  };
  std::mutex mu_;
  std::vector<Chunk> chunks_;
  std::map<uintptr_t, uint64_t> free_;  // address -> size (free blocks)
  std::map<uintptr_t, uint64_t> used_;  // address -> size (live blocks)
  uint64_t limit_, in_use_ = 0, peak_ = 0, reserved_ = 0;
};

}  // namespace ghive

// Tunables of one context (set through ghive_set_option; the batched form's autotuner writes
// them).
struct ghive_tuning {
  int64_t frag_lds_budget = 64 * 1024;  // bytes of LDS for the fragment's private group table
  int64_t frag_late = 0;                // 1: read a column only for lanes holding live rows
  int64_t frag_grid = 0;                // 0: one resident wave of workgroups
  int64_t frag_lds_bits = 1;            // 1: copy small dimension bitmaps into LDS
  int64_t frag_dense = 1;               // 1: probe dense dimensions through bitmap + payload array
  int64_t frag_rows = 4;                // fact rows per thread per tile (4 or 8)
  int64_t frag_block = 256;             // fused-kernel workgroup size (256, 512 or 1024)
  int64_t frag_generic = 0;             // 1: always the general fused kernel (no dense-probe form)
  int64_t frag_pipe = 1;                // 1: pipelined star kernel where group keys fit direct slots
  int64_t gb_lds_budget = 64 * 1024;    // bytes of LDS for the group-by's private table
  int64_t gb_partition = 1;             // high-cardinality form: 0 never, 1 by size, 2 always
  int64_t plan_fuse = 1;                // 1: plan vertices of the map-join shape run as one fragment
  int64_t gb_part_threads = 512;        // workgroup size of the group-by partition passes (256, 512, 1024)
  int64_t gb_part_occupancy = 2;        // workgroups per CU of its 512-thread 8-bit passes (1 or 2)
  int64_t sort_shape = 1;               // LSD tile: 0 256 x 32, 1 512 x 16, 2 512 x 32, 3 1024 x 8
  int64_t hash_regional_rows = 1 << 22;  // builds of at least this many rows take the regional form (0: never)
  int64_t sort_pack = 0;                // 1: key word + id (and the word-0-dropping pass) as 16-byte LDS records
  int64_t sort_merge = 1;               // 1: key word + id become one word once the low half's digits are done
  int64_t sort_occupancy = 2;           // workgroups per CU of the 8-byte-staging 512 x 16 passes (1 or 2)
};

struct ghive_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string last_error;
  ghive::DeviceArena *arena = nullptr;
  hipDeviceProp_t props{};
  int cus = 256;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  ghive_stage_times times{};
  ghive_plan_exec_info plan_info{};        // what the last ghive_plan_execute did
  ghive_process_times process_times{};     // the last ghive_gpu_process
  ghive_sort_info sort_info{};             // the last ghive_sort
  uint64_t gb_part_fallbacks = 0;          // partitioned group-bys whose result was discarded for the direct form
  // the device address of this device's static scratch array (ctx.hip), looked up at init
  void *scratch = nullptr;
  ghive_tuning tune{};
  // reusable small device scratch for counters (8 x uint64) and its host mirror
  uint64_t *d_counters = nullptr;
  // pinned host words that small device-to-host read-backs land in, so that several of them
  // are issued asynchronously and waited for once
  uint64_t *h_words = nullptr;
  // live objects whose memory comes from this context; ghive_shutdown orphans them so that a
  // later ghive_hash_table_free / ghive_batch_free only releases the handle
  std::set<ghive_hash_table *> tables;
  std::set<ghive_batch *> batches;
};

namespace ghive {

// Bytes of the static device array that ghive_fragment_execute_noalloc carves its group table,
// counters and staging from (one array per device, shared by every call on that device).
constexpr uint64_t kNoAllocScratchBytes = 8ull << 20;

int set_error(ghive_ctx *ctx, int code, const char *fmt, ...);
// Called by ghive_shutdown before the context's memory is released (hash.hip, batch.hip).
void orphan_tables(ghive_ctx *ctx);
void orphan_batches(ghive_ctx *ctx);

#define GH_CHECK_CTX(ctx)                         \
  do {                                            \
    if (!(ctx)) return GHIVE_ERR_INVALID;         \
    if (hipSetDevice((ctx)->device) != hipSuccess) \
      return ghive::set_error(ctx, GHIVE_ERR_HIP, "hipSetDevice(%d) failed", (ctx)->device); \
  } while (0)

#define GH_HIP(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ghive::set_error(ctx, GHIVE_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                              hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

#define GH_LAUNCH(ctx)                                                                      \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess)                                                                   \
      return ghive::set_error(ctx, GHIVE_ERR_HIP, "kernel launch failed: %s (%s:%d)",       \
                              hipGetErrorString(e_), __FILE__, __LINE__);                   \
  } while (0)

#define GH_TRY(expr)          \
  do {                        \
    int s_ = (expr);          \
    if (s_ != GHIVE_OK) return s_; \
  } while (0)

// RAII device buffer from the context arena.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ctx_(o.ctx_), p_(o.p_) { o.p_ = nullptr; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      ctx_ = o.ctx_;
      p_ = o.p_;
      o.p_ = nullptr;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  int alloc(ghive_ctx *ctx, uint64_t bytes) {
    reset();
    ctx_ = ctx;
    p_ = ctx->arena->alloc(bytes ? bytes : 1);
    if (!p_)
      return set_error(ctx, GHIVE_ERR_OOM, "device memory manager: cannot allocate %llu bytes "
                       "(in use %llu of limit %llu)", (unsigned long long)bytes,
                       (unsigned long long)ctx->arena->in_use(), (unsigned long long)ctx->arena->limit());
    return GHIVE_OK;
  }
  void reset() {
    if (p_ && ctx_) ctx_->arena->free(p_);
    p_ = nullptr;
  }
  void *release() {
    void *p = p_;
    p_ = nullptr;
    return p;
  }
  void adopt(ghive_ctx *ctx, void *p) {  // takes over a block from ctx's arena
    reset();
    ctx_ = ctx;
    p_ = p;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p_); }
  void *get() const { return p_; }

 private:
  ghive_ctx *ctx_ = nullptr;
  void *p_ = nullptr;
};

inline int type_width(int t) {
  switch (t) {
    case GHIVE_LONG: case GHIVE_DOUBLE: return 8;
    case GHIVE_INT: case GHIVE_FLOAT: return 4;
    default: return 0;
  }
}
inline bool is_int_type(int t) { return t == GHIVE_INT || t == GHIVE_LONG; }
inline bool is_numeric_type(int t) {
  return t == GHIVE_INT || t == GHIVE_LONG || t == GHIVE_DOUBLE || t == GHIVE_FLOAT;
}

constexpr uint64_t kMaxRows = 0xFFFFFFFEull;  // row ids are uint32

int check_table(ghive_ctx *ctx, const ghive_table *t, const char *what);
// Synchronous device->host copy on the context stream.
int d2h(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes);
// Reads up to kMaxReadWords device words (src[i] -> dst[i]) with one wait on the stream.
constexpr int kMaxReadWords = 64;
int d2h_words(ghive_ctx *ctx, int n, uint64_t *const *dst, const uint64_t *const *src);
int h2d(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes);
// Launch helpers (defined in kernels_util.hip).
int fill_u64(ghive_ctx *ctx, uint64_t *p, uint64_t value, uint64_t n);
int fill_u32(ghive_ctx *ctx, uint32_t *p, uint32_t value, uint64_t n);
// p[i * step] = value for i < n.
int fill_u64_strided(ghive_ctx *ctx, uint64_t *p, uint64_t value, uint64_t n, uint64_t step);
// ghive_hash_build without a host wait: the statistics are read when the table is first used
// (hash.hip).
int hash_build_deferred(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
                        const ghive_column *payload, float load_factor, ghive_hash_table **out);
// 32-bit copy of a LONG column (NULL kept); *bad_dev (device, zeroed by the caller) becomes nonzero
// when some value does not fit.
int narrow_long(ghive_ctx *ctx, const int64_t *in, uint64_t n, int32_t *out, uint64_t *bad_dev);
// Exclusive scan of n uint32 values into out (out may alias in); writes the total to
// *d_total (device) when non-null.
int exclusive_scan_u32(ghive_ctx *ctx, const uint32_t *in, uint32_t *out, uint64_t n,
                       uint64_t *d_total);

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace ghive
