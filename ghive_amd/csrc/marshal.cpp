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

#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace ghive;

struct ghive_staged {
  std::vector<std::vector<ghive_column>> cols;
  std::vector<ghive_table> tables;
  std::vector<std::string> names;
  std::vector<ghive_plan_input> inputs;
  std::vector<void *> owned;  // device buffers
};

namespace {

constexpr uint64_t kChunk = 8ull << 20;  // bytes per pinned staging buffer

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
      memcpy(buf[b], s + off, n);
      GH_HIP(ctx, hipMemcpyAsync(d + off, buf[b], n, hipMemcpyHostToDevice, ctx->stream));
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

}  // namespace

extern "C" {

int ghive_gtables_stage(ghive_ctx *ctx, const ghive_gtables *g, ghive_staged **out) {
  GH_CHECK_CTX(ctx);
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
    for (uint32_t c = 0; st == GHIVE_OK && c < g->ncols[t]; ++c) {
      ghive_column col;
      memset(&col, 0, sizeof col);
      col.row_num = rows;
      switch (g->types[t][c]) {
        case 0:
          if (li >= g->nlong) st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few long buffers");
          else st = stage_column(ctx, pipe, *S, g->long_cols[li++], 8 * rows, &col.data);
          col.type = GHIVE_LONG;
          break;
        case 1:
          if (di >= g->ndouble) st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few double buffers");
          else st = stage_column(ctx, pipe, *S, g->double_cols[di++], 8 * rows, &col.data);
          col.type = GHIVE_DOUBLE;
          break;
        case 2:
          if (ii >= g->nint) st = set_error(ctx, GHIVE_ERR_INVALID, "gtables: too few int buffers");
          else st = stage_column(ctx, pipe, *S, g->int_cols[ii++], 4 * rows, &col.data);
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
          ++si;
          col.type = GHIVE_STRING;
          col.char_size = (uint64_t)size;
          break;
        }
        default:
          continue;  // an unknown type code is skipped, as the reference does
      }
      if (st == GHIVE_OK) cols.push_back(col);
    }
    S->cols.push_back(std::move(cols));
    S->names.push_back(g->vertex_name[t]);
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

}  // extern "C"
