// Filter building blocks shared with the join and fragment code.
#pragma once

#include "common.hpp"

namespace ghive {

// Selected row ids of `t` in ascending order into out_sel (capacity t->row_num); synchronises.
int filter_select(ghive_ctx *ctx, const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                  uint32_t root, uint32_t *out_sel, uint64_t *out_count);
// out_cols[c].data[i] = t->columns[c].data[sel[i]] (asynchronous on the context stream).
int gather_columns(ghive_ctx *ctx, const ghive_table *t, const uint32_t *sel, uint64_t n,
                   ghive_column *out_cols);

}  // namespace ghive
