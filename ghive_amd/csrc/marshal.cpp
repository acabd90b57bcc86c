// The JNI boundary's data marshalling without the JVM: GTables (the Java side's input object,
// ghive/GTables.java:13-31, read by gpu_impl/src/ghive_JNIInterface.cu:88-204) given as the plain
// addresses of its direct ByteBuffers is staged into device tables that feed ghive_plan_execute,
// and the GPUResult column numbering (ghive_JNIInterface.cu:223-342) is computed for a result. A
// JNI shim only extracts the arrays (GetDirectBufferAddress, GetIntArrayElements, GetStringUTFChars)
// and calls these.
//
// Staging goes through two pinned host buffers: while the DMA engine copies one chunk to HBM, the
// host copies the next chunk of the (pageable) Java buffer into the other, so the host copy and the
// PCIe transfer overlap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "plan_internal.hpp"

using namespace ghive;

struct ghive_staged {
  std::vector<std::vector<ghive_column>> cols;
  std::vector<ghive_table> tables;
  std::vector<std::string> names;
  std::vector<ghive_plan_input> inputs;
  std::vector<void *> owned;  // device buffers
  // table name -> per column: 1 when a LONG column was sent as 32-bit integers (its values fit)
  std::map<std::string, std::vector<char>> narrowed;
  uint64_t h2d_bytes = 0;
  uint32_t narrowed_columns = 0;
};

namespace {

constexpr uint64_t kChunk = 16ull << 20;  // bytes per pinned staging buffer

// fn(begin, end) over [0, n) on up to 8 host threads (one for small n)
template <class F>
void parallel_for(uint64_t n, F fn) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const uint64_t parts = std::min<uint64_t>(std::min(8u, hw), std::max<uint64_t>(1, n / (256 * 1024)));
  if (parts <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (uint64_t p = 1; p < parts; ++p) th.emplace_back(fn, n * p / parts, n * (p + 1) / parts);
  fn(0, n / parts);
  for (auto &t : th) t.join();
}

struct PinnedPipe {
  void *buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  int next = 0;
  ~PinnedPipe() {
    for (int b = 0; b < 2; ++b) {
      if (done[b]) {
        (void)hipEventSynchronize(done[b]);
        (void)hipEventDestroy(done[b]);
      }
      if (buf[b]) (void)hipHostFree(buf[b]);
    }
  }
  int init(ghive_ctx *ctx) {
    for (int b = 0; b < 2; ++b) {
      GH_HIP(ctx, hipHostMalloc(&buf[b], kChunk, hipHostMallocDefault));
      GH_HIP(ctx, hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
    }
    return GHIVE_OK;
  }
  // host -> device through the two pinned buffers, in chunks, on the context stream
  int copy(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    const char *s = static_cast<const char *>(src);
    char *d = static_cast<char *>(dst);
    for (uint64_t off = 0; off < bytes; off += kChunk) {
      const uint64_t n = bytes - off < kChunk ? bytes - off : kChunk;
      const int b = next;
      next ^= 1;
      if (used[b]) GH_HIP(ctx, hipEventSynchronize(done[b]));  // its previous DMA has finished
      char *o = static_cast<char *>(buf[b]);
      parallel_for(n, [&](uint64_t lo, uint64_t hi) { memcpy(o + lo, s + off + lo, hi - lo); });
      GH_HIP(ctx, hipMemcpyAsync(d + off, buf[b], n, hipMemcpyHostToDevice, ctx->stream));
      GH_HIP(ctx, hipEventRecord(done[b], ctx->stream));
      used[b] = true;
    }
    return GHIVE_OK;
  }
  // host int64 values -> int32 on the device (LONG NULL -> INT NULL), converted into the pinned
  // buffers by the host threads; fits = false as soon as a non-NULL value lies outside
  // [INT_MIN, INT_MAX - 1] (the device copy is then incomplete and must not be used)
  int copy_narrow(ghive_ctx *ctx, int32_t *dst, const int64_t *src, uint64_t rows, bool &fits) {
    fits = true;
    const uint64_t per = kChunk / sizeof(int32_t);
    for (uint64_t off = 0; off < rows; off += per) {
      const uint64_t n = rows - off < per ? rows - off : per;
      const int b = next;
      next ^= 1;
      if (used[b]) GH_HIP(ctx, hipEventSynchronize(done[b]));
      int32_t *o = static_cast<int32_t *>(buf[b]);
      const int64_t *in = src + off;
      std::atomic<bool> ok{true};
      parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        bool good = true;
        for (uint64_t i = lo; i < hi; ++i) {
          const int64_t v = in[i];
          const bool null = v == LLONG_MAX;
          good &= null || (v >= INT_MIN && v <= INT_MAX - 1);
          o[i] = null ? INT_MAX : (int32_t)v;
        }
        if (!good) ok.store(false);
      });
      if (!ok.load()) {
        fits = false;
        return GHIVE_OK;
      }
      GH_HIP(ctx, hipMemcpyAsync(dst + off, buf[b], n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
      GH_HIP(ctx, hipEventRecord(done[b], ctx->stream));
      used[b] = true;
    }
    return GHIVE_OK;
  }
};

int stage_column(ghive_ctx *ctx, PinnedPipe &pipe, ghive_staged &S, const void *src, uint64_t bytes, void **dev) {
  void *p = nullptr;
  GH_TRY(ghive_malloc(ctx, bytes ? bytes : 1, &p));
  S.owned.push_back(p);
  if (bytes && !src) return set_error(ctx, GHIVE_ERR_INVALID, "gtables: a column buffer is NULL");
  GH_TRY(pipe.copy(ctx, p, src, bytes));
  *dev = p;
  return GHIVE_OK;
}

void release(ghive_ctx *ctx, ghive_staged *S) {
  for (void *p : S->owned) ghive_free(ctx, p);
  delete S;
}

// GTables -> device tables. cand (optional): table name -> positions of LONG columns to send as
// 32-bit integers when every value fits (the columns a fused kernel streams).
int stage_tables(ghive_ctx *ctx, const ghive_gtables *g, const std::map<std::string, std::set<uint32_t>> *cand,
                 ghive_staged **out) {
  if (!g || !out) return set_error(ctx, GHIVE_ERR_INVALID, "gtables: NULL argument");
  *out = nullptr;
  if (g->ntables && (!g->types || !g->ncols || !g->row_cnt || !g->vertex_name))
    return set_error(ctx, GHIVE_ERR_INVALID, "gtables: NULL table arrays");
  auto *S = new ghive_staged();
  PinnedPipe pipe;
  int st = pipe.init(ctx);
  uint32_t li = 0, di = 0, ii = 0, si = 0;  // next buffer of each type, consumed in table order
  for (uint32_t t = 0; st == GHIVE_OK && t < g->ntables; ++t) {
    if (!g->vertex_name[t]) {  // the reference returns nullptr: an input is missing
      st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: table %u has no vertex name (missing input)", t);
      break;
    }
    if (g->row_cnt[t] < 0) {
      st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: table %u has %d rows", t, g->row_cnt[t]);
      break;
    }
    const uint64_t rows = (uint64_t)g->row_cnt[t];
    std::vector<ghive_column> cols;
    std::vector<char> narrowed;
    const std::set<uint32_t> *want = nullptr;
    if (cand) {
      const auto it = cand->find(g->vertex_name[t]);
      if (it != cand->end()) want = &it->second;
    }
    for (uint32_t c = 0; st == GHIVE_OK && c < g->ncols[t]; ++c) {
      ghive_column col;
      memset(&col, 0, sizeof col);
      col.row_num = rows;
      switch (g->types[t][c]) {
        case 0: {
          col.type = GHIVE_LONG;
          if (li >= g->nlong) {
            st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few long buffers");
            break;
          }
          const void *src = g->long_cols[li++];
          if (want && want->count((uint32_t)cols.size()) && rows && src) {
            void *p = nullptr;
            st = ghive_malloc(ctx, 4 * rows, &p);
            bool fits = false;
            if (st == GHIVE_OK) {
              S->owned.push_back(p);
              st = pipe.copy_narrow(ctx, static_cast<int32_t *>(p), static_cast<const int64_t *>(src), rows, fits);
            }
            if (st == GHIVE_OK && fits) {
              col.data = p;
              col.type = GHIVE_INT;
              S->h2d_bytes += 4 * rows;
              ++S->narrowed_columns;
              break;
            }
            if (st != GHIVE_OK) break;
          }
          st = stage_column(ctx, pipe, *S, src, 8 * rows, &col.data);
          S->h2d_bytes += 8 * rows;
          break;
        }
        case 1:
          if (di >= g->ndouble) st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few double buffers");
          else st = stage_column(ctx, pipe, *S, g->double_cols[di++], 8 * rows, &col.data);
          S->h2d_bytes += 8 * rows;
          col.type = GHIVE_DOUBLE;
          break;
        case 2:
          if (ii >= g->nint) st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few int buffers");
          else st = stage_column(ctx, pipe, *S, g->int_cols[ii++], 4 * rows, &col.data);
          S->h2d_bytes += 4 * rows;
          col.type = GHIVE_INT;
          break;
        case 3: {
          if (si >= g->nstring) {
            st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few string buffers");
            break;
          }
          // the blob's size is the largest (start,end) of the pairs, as the reference computes it
          const int32_t *pairs = static_cast<const int32_t *>(g->string_idx_cols[si]);
          int64_t size = 0;
          for (uint64_t r = 0; pairs && r < 2 * rows; ++r) size = pairs[r] > size ? pairs[r] : size;
          if (rows && !pairs) {
            st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: a string index buffer is NULL");
            break;
          }
          st = stage_column(ctx, pipe, *S, g->string_cols[si], (uint64_t)size, &col.data);
          if (st == GHIVE_OK) st = stage_column(ctx, pipe, *S, pairs, 8 * rows, &col.aux);
          S->h2d_bytes += (uint64_t)size + 8 * rows;
          ++si;
          col.type = GHIVE_STRING;
          col.char_size = (uint64_t)size;
          break;
        }
        default:
          continue;  // an unknown type code is skipped, as the reference does
      }
      if (st == GHIVE_OK) {
        narrowed.push_back(col.type == GHIVE_INT && g->types[t][c] == 0);
        cols.push_back(col);
      }
    }
    S->cols.push_back(std::move(cols));
    S->names.push_back(g->vertex_name[t]);
    S->narrowed[g->vertex_name[t]] = narrowed;
  }
  if (st == GHIVE_OK) st = ghive_sync(ctx);
  if (st != GHIVE_OK) {
    (void)hipStreamSynchronize(ctx->stream);  // no copy may still write into buffers given back
    release(ctx, S);
    return st;
  }
  for (uint32_t t = 0; t < g->ntables; ++t) {
    const uint32_t keys = g->key_cnt ? (uint32_t)(g->key_cnt[t] < 0 ? 0 : g->key_cnt[t]) : 0;
    S->tables.push_back(ghive_table{S->cols[t].data(), (uint32_t)S->cols[t].size(), keys,
                                    (uint64_t)g->row_cnt[t]});
  }
  for (uint32_t t = 0; t < g->ntables; ++t)
    S->inputs.push_back(ghive_plan_input{S->names[t].c_str(), &S->tables[t], nullptr});
  *out = S;
  return GHIVE_OK;
}

}  // namespace

extern "C" {

int ghive_gtables_stage(ghive_ctx *ctx, const ghive_gtables *g, ghive_staged **out) {
  GH_CHECK_CTX(ctx);
  return stage_tables(ctx, g, nullptr, out);
}

int ghive_staged_inputs(const ghive_staged *s, const ghive_plan_input **inputs, uint32_t *n) {
  if (!s || !inputs || !n) return GHIVE_ERR_INVALID;
  *inputs = s->inputs.data();
  *n = (uint32_t)s->inputs.size();
  return GHIVE_OK;
}

void ghive_staged_free(ghive_ctx *ctx, ghive_staged *s) {
  if (s) release(ctx, s);
}

int ghive_gpuresult_sequence(const int32_t *types, uint32_t ncols, int32_t *sequence, uint32_t counts[4]) {
  if (ncols && (!types || !sequence)) return GHIVE_ERR_INVALID;
  uint32_t n[4] = {0, 0, 0, 0};  // LONG, DOUBLE, INT, STRING
  auto group = [](int32_t t) {
    switch (t) {
      case GHIVE_LONG: return 0;
      case GHIVE_DOUBLE:
      case GHIVE_FLOAT: return 1;  // GPUResult has no float[][]: widened to double
      case GHIVE_INT: return 2;
      case GHIVE_STRING: return 3;
      default: return -1;
    }
  };
  for (uint32_t c = 0; c < ncols; ++c) {
    const int g = group(types[c]);
    if (g < 0) return GHIVE_ERR_INVALID;
    ++n[g];
  }
  const uint32_t base[4] = {0, n[0], n[0] + n[1], n[0] + n[1] + n[2]};
  uint32_t seen[4] = {0, 0, 0, 0};
  for (uint32_t c = 0; c < ncols; ++c) {
    const int g = group(types[c]);
    sequence[c] = (int32_t)(base[g] + seen[g]++);
  }
  if (counts)
    for (int g = 0; g < 4; ++g) counts[g] = n[g];
  return GHIVE_OK;
}

// ---- GPUProcess in one call -------------------------------------------------------------------
struct ghive_gpu_result {
  uint64_t rows = 0;
  std::vector<int32_t> types, sequence;
  uint32_t counts[4] = {0, 0, 0, 0};
  std::vector<std::vector<int64_t>> longs;
  std::vector<std::vector<double>> doubles;
  std::vector<std::vector<int32_t>> ints;
  std::vector<std::vector<char>> blobs;
  std::vector<std::vector<int32_t>> pairs;
};

int ghive_gpu_process(ghive_ctx *ctx, const ghive_plan *plan, const ghive_gtables *g, const char *this_vertex,
                      char *err, uint64_t errcap, ghive_gpu_result **out) {
  GH_CHECK_CTX(ctx);
  auto say = [&](const std::string &m) {
    if (err && errcap) {
      const size_t n = std::min<size_t>(m.size(), (size_t)errcap - 1);
      memcpy(err, m.data(), n);
      err[n] = 0;
    }
  };
  if (!plan || !g || !this_vertex || !out) {
    say("gpu_process: NULL argument");
    return set_error(ctx, GHIVE_ERR_INVALID, "gpu_process: NULL argument");
  }
  *out = nullptr;
  memset(&ctx->process_times, 0, sizeof ctx->process_times);
  const auto t0 = std::chrono::steady_clock::now();
  auto ms_since = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
  };
  // a table without a vertex name is a missing input: no result, as GPUProcess returns nullptr
  // (ghive_JNIInterface.cu:108-120,210-220)
  for (uint32_t t = 0; t < g->ntables; ++t)
    if (g->vertex_name && !g->vertex_name[t]) return GHIVE_OK;
  std::map<std::string, std::set<uint32_t>> cand;
  plan_narrow_candidates(plan, this_vertex, cand);
  ghive_staged *S = nullptr;
  int st = stage_tables(ctx, g, cand.empty() ? nullptr : &cand, &S);
  if (st != GHIVE_OK) {
    say(ghive_last_error(ctx));
    return st;
  }
  ghive_process_times &PT = ctx->process_times;
  PT.stage_ms = ms_since(t0);
  PT.h2d_bytes = S->h2d_bytes;
  PT.narrowed_columns = S->narrowed_columns;
  const auto t1 = std::chrono::steady_clock::now();
  ghive_result *R = nullptr;
  st = plan_execute_staged(ctx, plan, this_vertex, S->inputs.data(), (uint32_t)S->inputs.size(), &S->narrowed, err,
                           errcap, &R);
  if (st != GHIVE_OK) {
    set_error(ctx, st, "%s", err ? err : "gpu_process: the vertex failed");
    ghive_staged_free(ctx, S);
    return st;
  }
  PT.execute_ms = ms_since(t1);
  PT.fused = ctx->plan_info.fused;
  const auto t2 = std::chrono::steady_clock::now();
  // generate_tbl_result (ghive_JNIInterface.cu:223-342): columns grouped LONG, DOUBLE, INT, STRING
  auto *G = new ghive_gpu_result();
  uint64_t rows = 0;
  uint32_t ncols = 0;
  (void)ghive_result_info(R, &rows, &ncols);
  G->rows = rows;
  for (uint32_t c = 0; c < ncols && st == GHIVE_OK; ++c) {
    ghive_column col;
    (void)ghive_result_column(R, c, &col, nullptr);
    G->types.push_back(col.type);
    switch (col.type) {
      case GHIVE_LONG:
        G->longs.emplace_back(rows);
        if (rows) st = ghive_memcpy(ctx, G->longs.back().data(), col.data, 8 * rows, 1);
        break;
      case GHIVE_DOUBLE:
        G->doubles.emplace_back(rows);
        if (rows) st = ghive_memcpy(ctx, G->doubles.back().data(), col.data, 8 * rows, 1);
        break;
      case GHIVE_FLOAT: {  // GPUResult has no float[][]: widened to double
        std::vector<float> f(rows);
        if (rows) st = ghive_memcpy(ctx, f.data(), col.data, 4 * rows, 1);
        G->doubles.emplace_back(f.begin(), f.end());
        break;
      }
      case GHIVE_INT:
        G->ints.emplace_back(rows);
        if (rows) st = ghive_memcpy(ctx, G->ints.back().data(), col.data, 4 * rows, 1);
        break;
      case GHIVE_STRING: {
        G->pairs.emplace_back(2 * rows);
        std::vector<int32_t> &pr = G->pairs.back();
        if (rows) st = ghive_memcpy(ctx, pr.data(), col.aux, 8 * rows, 1);
        int64_t size = 0;
        for (int32_t x : pr) size = x > size ? x : size;
        G->blobs.emplace_back((size_t)size);
        if (st == GHIVE_OK && size) st = ghive_memcpy(ctx, G->blobs.back().data(), col.data, (uint64_t)size, 1);
        break;
      }
      default:
        st = set_error(ctx, GHIVE_ERR_UNSUPPORTED, "gpu_process: result column %u has type %d", c, col.type);
    }
  }
  if (st == GHIVE_OK) {
    G->sequence.resize(ncols);
    st = ghive_gpuresult_sequence(G->types.data(), ncols, G->sequence.data(), G->counts);
  }
  ghive_result_free(ctx, R);
  ghive_staged_free(ctx, S);
  if (st != GHIVE_OK) {
    say(ghive_last_error(ctx));
    delete G;
    return st;
  }
  PT.result_ms = ms_since(t2);
  PT.total_ms = ms_since(t0);
  *out = G;
  return GHIVE_OK;
}

int ghive_gpu_result_info(const ghive_gpu_result *r, uint64_t *rows, uint32_t *ncols, uint32_t counts[4],
                          const int32_t **sequence) {
  if (!r) return GHIVE_ERR_INVALID;
  if (rows) *rows = r->rows;
  if (ncols) *ncols = (uint32_t)r->sequence.size();
  if (counts)
    for (int i = 0; i < 4; ++i) counts[i] = r->counts[i];
  if (sequence) *sequence = r->sequence.data();
  return GHIVE_OK;
}

const int64_t *ghive_gpu_result_long(const ghive_gpu_result *r, uint32_t i) {
  return r && i < r->longs.size() ? r->longs[i].data() : nullptr;
}
const double *ghive_gpu_result_double(const ghive_gpu_result *r, uint32_t i) {
  return r && i < r->doubles.size() ? r->doubles[i].data() : nullptr;
}
const int32_t *ghive_gpu_result_int(const ghive_gpu_result *r, uint32_t i) {
  return r && i < r->ints.size() ? r->ints[i].data() : nullptr;
}
int ghive_gpu_result_string(const ghive_gpu_result *r, uint32_t i, const char **blob, uint64_t *blob_bytes,
                            const int32_t **pairs) {
  if (!r || i >= r->pairs.size()) return GHIVE_ERR_INVALID;
  if (blob) *blob = r->blobs[i].data();
  if (blob_bytes) *blob_bytes = r->blobs[i].size();
  if (pairs) *pairs = r->pairs[i].data();
  return GHIVE_OK;
}
void ghive_gpu_result_free(ghive_gpu_result *r) { delete r; }

int ghive_last_process_times(ghive_ctx *ctx, ghive_process_times *out) {
  if (!ctx || !out) return GHIVE_ERR_INVALID;
  *out = ctx->process_times;
  return GHIVE_OK;
}

}  // extern "C"
