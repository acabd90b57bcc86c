// Host-side conversion of a ghive_pred tree into the flattened DevPred program.
#pragma once

#include "common.hpp"
#include "device.hpp"

namespace ghive {

// Validates the tree (modes, column indexes and types, cycles) and flattens it in post-order.
int flatten_pred(ghive_ctx *ctx, const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                 uint32_t root, DevPred &out);

}  // namespace ghive
