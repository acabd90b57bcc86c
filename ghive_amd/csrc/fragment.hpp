// Device description of a fused fragment (see fragment.hip).
#pragma once

#include "agg.hpp"
#include "common.hpp"
#include "device.hpp"
#include "groupby.hpp"

namespace ghive {

constexpr int kMaxFragJoins = 4;
constexpr int kMaxFragAggs = 4;

struct FragJoinDev {
  const int32_t *fk;      // fact foreign-key column
  const uint64_t *slots;  // L8 slot array of the dimension table
  uint64_t mask;
  int32_t log2;
  int32_t key_min, key_max;  // inserted key range
  int32_t pad;
};

struct FragAggDev {
  int32_t expr;
  int32_t pad;
  const int32_t *a;
  const int32_t *b;
};

struct FragDev {
  uint64_t n;
  int32_t njoins;
  int32_t nkeys;
  int32_t naggs;
  int32_t has_filter;
  int8_t key_src[2];
  int8_t key_idx[2];
  int32_t pad;
  const int32_t *key_col[2];
  FragJoinDev joins[kMaxFragJoins];
  FragAggDev aggs[kMaxFragAggs];
  DevPred filter;
};

// Validates the descriptor and fills the device description; group_bound is an upper bound on
// the number of groups (product of the payload ranges of the join keys).
int fragment_prepare(ghive_ctx *ctx, const ghive_fragment *f, FragDev &F, GroupTableDev &T, KeySpec &K,
                     uint64_t &group_bound);
// Launches the fused kernel on ctx->stream (rows that survive the joins counted in d_joined).
int fragment_launch(ghive_ctx *ctx, const FragDev &F, const GroupTableDev &T, uint64_t group_bound,
                    unsigned long long *d_joined, int late, uint32_t grid_override);

}  // namespace ghive
