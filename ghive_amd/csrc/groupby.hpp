// Group-table helpers shared by the group-by operator and the fused fragment.
#pragma once

#include <vector>

#include "agg.hpp"
#include "common.hpp"

namespace ghive {

// Packing of up to kMaxKeys key columns into one uint64 (key k at bit K.shift[k]).
struct KeySpec {
  int32_t nkeys;
  int32_t type[kMaxKeys];
  int32_t shift[kMaxKeys];
  const void *data[kMaxKeys];
};

// Key columns of any type, STRING included, compared by value (the dictionary form of the group-by).
constexpr int kMaxWideKeys = 8;
struct WideKeySpec {
  int32_t nkeys;
  int32_t type[kMaxWideKeys];
  const void *data[kMaxWideKeys];    // values; STRING: the character blob
  const int32_t *aux[kMaxWideKeys];  // STRING: (start,end) per row, (-1,-1) = NULL
};

struct OutSpec {
  void *col[kMaxKeys + kMaxAggs];
  int32_t type[kMaxKeys + kMaxAggs];
};

// Largest power-of-two LDS table (log2) whose keys + words fit in budget_bytes.
int lds_log2_for(int nwords, int budget_bytes);
// Sizes the global table for `capacity` groups (T.aggs/naggs/nwords set): fills T's shape and the
// bytes of its key and accumulator arrays (the two counters take 16 more bytes).
void group_table_size(GroupTableDev &T, uint64_t capacity, uint64_t *key_bytes, uint64_t *acc_bytes);
// Initialises a table sized by group_table_size in caller memory (fill kernels only).
// ncounters: table counters zeroed with it (entries, overflow, then any the caller keeps after them)
int group_table_init(ghive_ctx *ctx, GroupTableDev &T, uint64_t *keys, uint64_t *acc, unsigned long long *counters,
                     uint32_t ncounters = 2);
// Allocates and initialises the global table for `capacity` groups (T.aggs/naggs/nwords set).
int group_table_alloc(ghive_ctx *ctx, GroupTableDev &T, uint64_t capacity, DevBuf &keys, DevBuf &acc,
                      DevBuf &counters);
// Inserts `key` (so that a keyless aggregate yields its one group even over no rows).
int group_table_preinsert(ghive_ctx *ctx, const GroupTableDev &T, uint64_t key);
// Stable LSD radix passes (sort.hip) over 8-bit `digits` of one uint64 word per row with a uint32
// id; *keys_out / *ids_out point at the buffers holding the result.
int radix_sort_u64_ids(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                       const std::vector<int> &digits, uint64_t **keys_out, uint32_t **ids_out, bool ids = true);
// The same over two uint64 words per row (word-major: word q of row i at q * n + i); the digits
// index word 0 (0..7) and word 1 (8..15); both words travel with the id.
int radix_sort_u64x2_ids(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                         const std::vector<int> &digits, uint64_t **keys_out, uint32_t **ids_out, bool ids = true);
// Key dictionary of rows sel[0..n) (sel NULL: rows 0..n): codes[r] (rows entries) = the slot of
// row r's key; dict[slot] = (32 hash bits << 32) | the key's representative row, or all ones when
// empty; *slots = dictionary size (a power of two >= 2 x capacity); *count = distinct keys.
// CAPACITY when there are more than `capacity` distinct keys.
int key_dictionary(ghive_ctx *ctx, const WideKeySpec &K, uint64_t rows, const uint32_t *sel, uint64_t n,
                   uint64_t capacity, DevBuf &dict, DevBuf &codes, uint64_t *slots, uint64_t *count);
// Writes every group to the output columns; CAPACITY error when there are more than capacity.
// `after` (optional) is recorded on the stream after the write-out kernel; extra_src / extra_dst
// (optional) is one more device word read back with the group count.
// Largest |x| over rows sel[0..n) (sel NULL: 0..n) of a numeric column, NULLs skipped (0 if none).
int column_absmax(ghive_ctx *ctx, const void *data, int type, const uint32_t *sel, uint64_t n, double *out);
// LSB exponent of an exact fixed-point double sum of n values of magnitude <= mx;
// GHIVE_ERR_UNSUPPORTED when mx is infinite.
int sum_scale(ghive_ctx *ctx, double mx, uint64_t n, int32_t *scale_e);
int group_table_emit(ghive_ctx *ctx, const KeySpec &K, const GroupTableDev &T, const OutSpec &O,
                     uint64_t capacity, uint64_t *out_groups, hipEvent_t after = nullptr,
                     const uint64_t *extra_src = nullptr, uint64_t *extra_dst = nullptr);

}  // namespace ghive
