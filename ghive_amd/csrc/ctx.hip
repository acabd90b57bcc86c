// Context, errors, device memory manager and device query of libghive_mi355x.so.
// Replaces the implicit process-global state of gpu_impl (static maps in Parser.cu:16-23 and
// JoinOperator.cu:11-12): everything here hangs off a ghive_ctx handle.
#include <cstring>

#include "common.hpp"

namespace ghive {

int set_error(ghive_ctx *ctx, int code, const char *fmt, ...) {
  if (!ctx) return code;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  ctx->last_error = buf;
  return code;
}

DeviceArena::~DeviceArena() { release_all(); }

void DeviceArena::release_all() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto &c : chunks_) (void)hipFree(c.base);
  chunks_.clear();
  free_.clear();
  used_.clear();
  in_use_ = reserved_ = 0;
}

void *DeviceArena::alloc(uint64_t bytes) {
  constexpr uint64_t kAlign = 256;
  bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
  std::lock_guard<std::mutex> g(mu_);
  if (in_use_ + bytes > limit_) return nullptr;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= bytes) {
      uintptr_t a = it->first;
      uint64_t s = it->second;
      free_.erase(it);
      if (s > bytes) free_[a + bytes] = s - bytes;
      used_[a] = bytes;
      in_use_ += bytes;
      if (in_use_ > peak_) peak_ = in_use_;
      return reinterpret_cast<void *>(a);
    }
  }
  // new chunk: at least 256 MiB so that small intermediates share one hipMalloc
  uint64_t csize = bytes < (256ull << 20) ? (256ull << 20) : bytes;
  if (reserved_ + csize > limit_ + (256ull << 20)) csize = bytes;
  void *p = nullptr;
  if (hipMalloc(&p, csize) != hipSuccess) {
    (void)hipGetLastError();
    if (csize == bytes) return nullptr;
    csize = bytes;
    if (hipMalloc(&p, csize) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  chunks_.push_back({static_cast<char *>(p), csize});
  reserved_ += csize;
  uintptr_t a = reinterpret_cast<uintptr_t>(p);
  if (csize > bytes) free_[a + bytes] = csize - bytes;
  used_[a] = bytes;
  in_use_ += bytes;
  if (in_use_ > peak_) peak_ = in_use_;
  return p;
}

bool DeviceArena::free(void *p) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = used_.find(reinterpret_cast<uintptr_t>(p));
  if (it == used_.end()) return false;
  uintptr_t a = it->first;
  uint64_t s = it->second;
  used_.erase(it);
  in_use_ -= s;
  // coalesce with neighbours that lie in the same chunk
  auto chunk_of = [&](uintptr_t x) -> const Chunk * {
    for (auto &c : chunks_)
      if (x >= reinterpret_cast<uintptr_t>(c.base) && x < reinterpret_cast<uintptr_t>(c.base) + c.size)
        return &c;
    return nullptr;
  };
  const Chunk *ch = chunk_of(a);
  auto nx = free_.lower_bound(a);
  if (nx != free_.end() && nx->first == a + s && chunk_of(nx->first) == ch) {
    s += nx->second;
    free_.erase(nx);
  }
  auto pv = free_.lower_bound(a);
  if (pv != free_.begin()) {
    --pv;
    if (pv->first + pv->second == a && chunk_of(pv->first) == ch) {
      pv->second += s;
      return true;
    }
  }
  free_[a] = s;
  return true;
}

int check_table(ghive_ctx *ctx, const ghive_table *t, const char *what) {
  if (!t) return set_error(ctx, GHIVE_ERR_INVALID, "%s: table is NULL", what);
  if (t->row_num > kMaxRows)
    return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "%s: %llu rows exceed the uint32 row-id range",
                     what, (unsigned long long)t->row_num);
  if (t->ncols && !t->columns) return set_error(ctx, GHIVE_ERR_INVALID, "%s: columns is NULL", what);
  for (uint32_t i = 0; i < t->ncols; ++i) {
    const ghive_column &c = t->columns[i];
    if (c.type < GHIVE_LONG || c.type > GHIVE_FLOAT)
      return set_error(ctx, GHIVE_ERR_INVALID, "%s: column %u has type %d", what, i, c.type);
    if (c.row_num != t->row_num)
      return set_error(ctx, GHIVE_ERR_INVALID, "%s: column %u has %llu rows, table %llu", what, i,
                       (unsigned long long)c.row_num, (unsigned long long)t->row_num);
    if (t->row_num && !c.data)
      return set_error(ctx, GHIVE_ERR_INVALID, "%s: column %u has no data", what, i);
    if (c.type == GHIVE_STRING && t->row_num && !c.aux)
      return set_error(ctx, GHIVE_ERR_INVALID, "%s: STRING column %u has no (start,end) pairs", what, i);
  }
  return GHIVE_OK;
}

int d2h(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!bytes) return GHIVE_OK;
  GH_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

int d2h_words(ghive_ctx *ctx, int n, uint64_t *const *dst, const uint64_t *const *src) {
  if (n < 0 || n > kMaxReadWords) return set_error(ctx, GHIVE_ERR_INVALID, "d2h_words: %d words", n);
  for (int i = 0; i < n; ++i)
    GH_HIP(ctx, hipMemcpyAsync(ctx->h_words + i, src[i], sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) *dst[i] = ctx->h_words[i];
  return GHIVE_OK;
}

int h2d(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!bytes) return GHIVE_OK;
  GH_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_abi_version(void) { return GHIVE_ABI_VERSION; }

// The scratch memory of ghive_fragment_execute_noalloc: one array per device for the life of the
// process, so that the call needs no allocation. In host code `scratch` names a host-side stand-in,
// not device memory: only the address hipGetSymbolAddress returns for the current device is used.
static __device__ __align__(256) __attribute__((used)) uint8_t scratch[ghive::kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device -> device address of its copy of `scratch`

// Looked up once per device, with that device current.
int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return GHIVE_ERR_HIP;
    }
    it = g_scratch_ptr.emplace(device, p).first;
  }
  *out = it->second;
  return GHIVE_OK;
}
}  // namespace

int ghive_init(int device, uint64_t mem_limit_bytes, ghive_ctx **out) {
  if (!out) return GHIVE_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return GHIVE_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= n) return GHIVE_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return GHIVE_ERR_HIP;
  ghive_ctx *ctx = new ghive_ctx();
  ctx->device = device;
  if (hipGetDeviceProperties(&ctx->props, device) != hipSuccess) {
    delete ctx;
    return GHIVE_ERR_HIP;
  }
  ctx->cus = ctx->props.multiProcessorCount > 0 ? ctx->props.multiProcessorCount : 256;
  if (hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return GHIVE_ERR_HIP;
  }
  ctx->stream = ctx->own_stream;
  for (auto &e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return GHIVE_ERR_HIP;
    }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 1ull << 34;
  uint64_t limit = mem_limit_bytes ? mem_limit_bytes : (uint64_t)(0.9 * (double)free_b);
  if (scratch_address(device, &ctx->scratch) != GHIVE_OK) {
    delete ctx;
    return GHIVE_ERR_HIP;
  }
  ctx->arena = new DeviceArena(limit);
  ctx->d_counters = static_cast<uint64_t *>(ctx->arena->alloc(64 * sizeof(uint64_t)));
  if (!ctx->d_counters || hipHostMalloc((void **)&ctx->h_words, kMaxReadWords * sizeof(uint64_t),
                                        hipHostMallocDefault) != hipSuccess) {
    delete ctx->arena;
    delete ctx;
    return GHIVE_ERR_OOM;
  }
  *out = ctx;
  return GHIVE_OK;
}

void ghive_shutdown(ghive_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ghive::orphan_batches(ctx);
  ghive::orphan_tables(ctx);
  delete ctx->arena;
  if (ctx->h_words) (void)hipHostFree(ctx->h_words);
  for (auto &e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

const char *ghive_last_error(const ghive_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "no context";
}

// Memory the context's manager hands out again is used in the order of the context's stream, so
// the old stream's work is waited for before the stream changes.
int ghive_set_stream(ghive_ctx *ctx, void *hip_stream) {
  GH_CHECK_CTX(ctx);
  if (ctx->stream != static_cast<hipStream_t>(hip_stream)) GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = static_cast<hipStream_t>(hip_stream);  // NULL: the HIP null stream
  return GHIVE_OK;
}

int ghive_use_own_stream(ghive_ctx *ctx) {
  GH_CHECK_CTX(ctx);
  if (ctx->stream != ctx->own_stream) GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = ctx->own_stream;
  return GHIVE_OK;
}

void *ghive_get_stream(ghive_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int ghive_sync(ghive_ctx *ctx) {
  GH_CHECK_CTX(ctx);
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

int ghive_device_query(ghive_ctx *ctx, ghive_device_info *out) {
  GH_CHECK_CTX(ctx);
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "device_query: out is NULL");
  memset(out, 0, sizeof *out);
  const hipDeviceProp_t &p = ctx->props;
  snprintf(out->name, sizeof out->name, "%s", p.name);
  snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->wavefront_size = p.warpSize;
  out->lds_per_block = p.sharedMemPerBlock;
  out->hbm_bytes = p.totalGlobalMem;
  out->l2_bytes = p.l2CacheSize;
  out->clock_khz = p.clockRate;
  out->mem_clock_khz = p.memoryClockRate;
  out->mem_bus_width = p.memoryBusWidth;
  return GHIVE_OK;
}

int ghive_malloc(ghive_ctx *ctx, uint64_t bytes, void **out) {
  GH_CHECK_CTX(ctx);
  if (!out) return set_error(ctx, GHIVE_ERR_INVALID, "malloc: out is NULL");
  *out = ctx->arena->alloc(bytes ? bytes : 1);
  if (!*out)
    return set_error(ctx, GHIVE_ERR_OOM, "malloc: cannot allocate %llu bytes",
                     (unsigned long long)bytes);
  return GHIVE_OK;
}

int ghive_free(ghive_ctx *ctx, void *ptr) {
  GH_CHECK_CTX(ctx);
  if (!ptr) return GHIVE_OK;
  if (!ctx->arena->free(ptr))
    return set_error(ctx, GHIVE_ERR_INVALID, "free: %p was not allocated by this context", ptr);
  return GHIVE_OK;
}

int ghive_memcpy(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind) {
  GH_CHECK_CTX(ctx);
  if (!bytes) return GHIVE_OK;
  if (!dst || !src) return set_error(ctx, GHIVE_ERR_INVALID, "memcpy: NULL pointer");
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                              : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  if (kind < 0 || kind > 2) return set_error(ctx, GHIVE_ERR_INVALID, "memcpy: kind %d", kind);
  GH_HIP(ctx, hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

int ghive_mem_stats(ghive_ctx *ctx, uint64_t *in_use, uint64_t *peak, uint64_t *limit,
                    uint64_t *reserved) {
  GH_CHECK_CTX(ctx);
  if (in_use) *in_use = ctx->arena->in_use();
  if (peak) *peak = ctx->arena->peak();
  if (limit) *limit = ctx->arena->limit();
  if (reserved) *reserved = ctx->arena->reserved();
  return GHIVE_OK;
}

int ghive_last_stage_times(ghive_ctx *ctx, ghive_stage_times *out) {
  if (!ctx || !out) return GHIVE_ERR_INVALID;
  *out = ctx->times;
  return GHIVE_OK;
}

int ghive_set_option(ghive_ctx *ctx, const char *name, int64_t value) {
  if (!ctx || !name) return GHIVE_ERR_INVALID;
  std::string k(name);
  if (k == "fragment.lds_budget") {
    if (value < 0 || value > 160 * 1024) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld", name, (long long)value);
    ctx->tune.frag_lds_budget = value;
  } else if (k == "fragment.late") {
    ctx->tune.frag_late = value ? 1 : 0;
  } else if (k == "fragment.grid") {
    if (value < 0 || value > (1 << 24)) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld", name, (long long)value);
    ctx->tune.frag_grid = value;
  } else if (k == "fragment.lds_bits") {
    ctx->tune.frag_lds_bits = value ? 1 : 0;
  } else if (k == "fragment.dense") {
    ctx->tune.frag_dense = value ? 1 : 0;
  } else if (k == "fragment.rows_per_thread") {
    if (value != 4 && value != 8) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (4 or 8)", name, (long long)value);
    ctx->tune.frag_rows = value;
  } else if (k == "fragment.generic") {
    ctx->tune.frag_generic = value ? 1 : 0;
  } else if (k == "fragment.pipe") {
    ctx->tune.frag_pipe = value ? 1 : 0;
  } else if (k == "fragment.block") {
    if (value != 256 && value != 512 && value != 1024)
      return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (256, 512 or 1024)", name, (long long)value);
    ctx->tune.frag_block = value;
  } else if (k == "plan.fuse") {
    ctx->tune.plan_fuse = value ? 1 : 0;
  } else if (k == "groupby.part_threads") {
    if (value != 256 && value != 512 && value != 1024)
      return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (256, 512 or 1024)", name, (long long)value);
    ctx->tune.gb_part_threads = value;
  } else if (k == "groupby.part_occupancy") {
    if (value != 1 && value != 2) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (1 or 2)", name, (long long)value);
    ctx->tune.gb_part_occupancy = value;
  } else if (k == "sort.shape") {
    if (value < 0 || value > 3) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (0..3)", name, (long long)value);
    ctx->tune.sort_shape = value;
  } else if (k == "sort.pack") {
    ctx->tune.sort_pack = value ? 1 : 0;
  } else if (k == "hash.regional_rows") {
    if (value < 0) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (>= 0)", name, (long long)value);
    ctx->tune.hash_regional_rows = value;
  } else if (k == "sort.merge") {
    ctx->tune.sort_merge = value ? 1 : 0;
  } else if (k == "sort.occupancy") {
    if (value != 1 && value != 2) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (1 or 2)", name, (long long)value);
    ctx->tune.sort_occupancy = value;
  } else if (k == "groupby.partition") {
    if (value < 0 || value > 2) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld (0, 1 or 2)", name, (long long)value);
    ctx->tune.gb_partition = value;
  } else if (k == "groupby.lds_budget") {
    if (value < 0 || value > 160 * 1024) return set_error(ctx, GHIVE_ERR_INVALID, "option %s: %lld", name, (long long)value);
    ctx->tune.gb_lds_budget = value;
  } else {
    return set_error(ctx, GHIVE_ERR_INVALID, "unknown option '%s'", name);
  }
  return GHIVE_OK;
}

int ghive_get_option(ghive_ctx *ctx, const char *name, int64_t *value) {
  if (!ctx || !name || !value) return GHIVE_ERR_INVALID;
  std::string k(name);
  if (k == "fragment.lds_budget") *value = ctx->tune.frag_lds_budget;
  else if (k == "fragment.late") *value = ctx->tune.frag_late;
  else if (k == "fragment.grid") *value = ctx->tune.frag_grid;
  else if (k == "fragment.lds_bits") *value = ctx->tune.frag_lds_bits;
  else if (k == "fragment.dense") *value = ctx->tune.frag_dense;
  else if (k == "fragment.rows_per_thread") *value = ctx->tune.frag_rows;
  else if (k == "fragment.block") *value = ctx->tune.frag_block;
  else if (k == "fragment.generic") *value = ctx->tune.frag_generic;
  else if (k == "fragment.pipe") *value = ctx->tune.frag_pipe;
  else if (k == "groupby.lds_budget") *value = ctx->tune.gb_lds_budget;
  else if (k == "groupby.partition") *value = ctx->tune.gb_partition;
  else if (k == "plan.fuse") *value = ctx->tune.plan_fuse;
  else if (k == "sort.shape") *value = ctx->tune.sort_shape;
  else if (k == "sort.pack") *value = ctx->tune.sort_pack;
  else if (k == "sort.merge") *value = ctx->tune.sort_merge;
  else if (k == "hash.regional_rows") *value = ctx->tune.hash_regional_rows;
  else if (k == "sort.occupancy") *value = ctx->tune.sort_occupancy;
  else if (k == "groupby.part_threads") *value = ctx->tune.gb_part_threads;
  else if (k == "groupby.part_occupancy") *value = ctx->tune.gb_part_occupancy;
  else if (k == "groupby.partition_fallbacks") *value = (int64_t)ctx->gb_part_fallbacks;
  else return set_error(ctx, GHIVE_ERR_INVALID, "unknown option '%s'", name);
  return GHIVE_OK;
}

int ghive_last_plan_info(ghive_ctx *ctx, ghive_plan_exec_info *out) {
  if (!ctx || !out) return GHIVE_ERR_INVALID;
  *out = ctx->plan_info;
  return GHIVE_OK;
}

int ghive_agg_output_type(int32_t agg_type, int32_t input_type) {
  switch (agg_type) {
    case GHIVE_AGG_CNT: return GHIVE_LONG;
    case GHIVE_AGG_AVG: return GHIVE_DOUBLE;
    case GHIVE_AGG_SUM:
      return (input_type == GHIVE_INT || input_type == GHIVE_LONG) ? GHIVE_LONG : GHIVE_DOUBLE;
    case GHIVE_AGG_MAX:
    case GHIVE_AGG_MIN: return input_type;
    default: return -1;
  }
}

}  // extern "C"
