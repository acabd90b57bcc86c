// Device description of a fused fragment (see fragment.hip).
#pragma once

#include "agg.hpp"
#include "common.hpp"
#include "device.hpp"
#include "groupby.hpp"

namespace ghive {

constexpr int kMaxFragJoins = 4;
constexpr int kMaxFragAggs = 4;
constexpr uint32_t kMaxDirectGroups = 1u << 20;  // largest group-key domain given direct slots

// How a join is probed inside the fused kernel.
enum FragJoinMode : int32_t {
  kJoinHash = 0,        // open-addressing L8 table in HBM/L2
  kJoinGlobalBits = 1,  // dense companion: bitmap + payload array in HBM/L2
  kJoinLdsBits = 2,     // dense companion with the bitmap copied into LDS per workgroup
};

struct FragJoinDev {
  const int32_t *fk;        // fact foreign-key column
  const uint64_t *slots;    // L8 slot array
  const uint32_t *bitmap;   // dense existence bitmap over [key_min, key_max]
  const int32_t *dense;     // dense payload array
  uint64_t mask;
  int32_t log2;
  int32_t mode;
  int32_t key_min, key_max;  // inserted key range (empty table: min > max)
  uint32_t lds_word_off;     // kJoinLdsBits: first uint32 word of this bitmap in LDS
  uint32_t nwords;           // bitmap words
  int32_t need_payload;      // payload used as a group key
  uint32_t span;             // key_max - key_min: key k is in range iff (uint32)(k - key_min) <= span
};

struct FragAggDev {
  int32_t expr;
  int32_t dbl;  // 1: a and b are DOUBLE columns (general kernel: read per surviving row)
  const int32_t *a;
  const int32_t *b;
};

struct FragDev {
  uint64_t n;
  int32_t njoins;
  int32_t nkeys;
  int32_t naggs;
  int32_t has_filter;
  int32_t late;  // read a column only for lanes that still hold a row
  int32_t pad;
  int8_t key_src[2];
  int8_t key_idx[2];
  uint32_t lds_bitmap_words;  // words of LDS bitmaps (even, so the group table is 8-aligned)
  // direct group slots (join-attribute keys of small ranges): the group of key values v[k] sits
  // at slot sum_k (v[k] - dmin[k]) * dmul[k], no hashing; dsize slots are used
  int32_t direct;
  uint32_t dsize;
  int32_t dmin[2];
  uint32_t dmul[2];
  uint32_t queue_cap;  // entries of each wave's LDS queue of survivors (global direct table)
  uint32_t pad2;
  const int32_t *key_col[2];
  // NULL-padded copy of the partial last tile: one tile (blockDim.x * R rows) per streamed column,
  // joins first (slot j), then aggregate inputs (slots NJ + 2a and NJ + 2a + 1)
  int32_t *stage;
  FragJoinDev joins[kMaxFragJoins];
  FragAggDev aggs[kMaxFragAggs];
  DevPred filter;
};

// Validated device description plus launch plan of one fragment.
struct FragPlan {
  FragDev F;
  GroupTableDev T;
  KeySpec K;
  uint64_t group_bound;  // product of the payload ranges of the join keys
  int lcap_log2;         // LDS group table (log2 slots) or -1: global table only
  size_t lds_bytes;
  uint32_t grid;
  uint32_t threads;  // workgroup size
  int rows;          // fact rows per lane per tile
  bool dense_form;   // k_fragment_dense or k_fragment_pipe (else the general k_fragment)
  bool pipe_form;    // k_fragment_pipe (direct group slots, next tile's loads in flight)
  int late;
  const void *kernel;
};

// Validates the descriptor, chooses the probe mode and order of every join, sizes LDS and grid.
int fragment_plan(ghive_ctx *ctx, const ghive_fragment *f, int late, FragPlan &P);
// Launches the fused kernel on ctx->stream (rows surviving every join counted in d_joined).
int fragment_launch(ghive_ctx *ctx, const FragPlan &P, unsigned long long *d_joined);

}  // namespace ghive
