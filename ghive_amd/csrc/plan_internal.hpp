// Host-side entry points of the plan executor used by the one-call JNI boundary (marshal.cpp).
#pragma once

#include <map>
#include <set>
#include <string>
#include <vector>

#include "../../include/ghive/ghive.h"

namespace ghive {

// The input columns the fused form of `vertex` streams as 32-bit integers (join keys, integer
// aggregate inputs, group keys of its fact table and dimensions): input name -> column positions.
// Empty when the vertex has no fused form.
void plan_narrow_candidates(const ghive_plan *plan, const char *vertex,
                            std::map<std::string, std::set<uint32_t>> &out);

// ghive_plan_execute over staged inputs some of whose LONG columns hold 32-bit copies
// (narrowed: input name -> per column flag).
int plan_execute_staged(ghive_ctx *ctx, const ghive_plan *plan, const char *vertex, const ghive_plan_input *inputs,
                        uint32_t ninputs, const std::map<std::string, std::vector<char>> *narrowed, char *err,
                        uint64_t errcap, ghive_result **out);

}  // namespace ghive
