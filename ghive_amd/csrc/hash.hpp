// Device hash tables for the join (build side resident in HBM, probed from L2/MALL).
//
// Two slot layouts, both open addressing with linear probing over a power-of-two slot array:
//   L8  : one uint64 per slot = (uint32)key | (uint32)payload << 32, for INT keys whose payload
//         fits 32 bits (an INT attribute or a build row id). One 8-byte load answers a probe and
//         a 128-byte line holds 16 consecutive slots.
//   L16 : two uint64 per slot = {key, payload}, for LONG keys or LONG payloads.
// The empty marker is the NULL sentinel of the key (INT_MAX / LONG_MAX): NULL keys never join,
// so no real key can be confused with an empty slot.
#pragma once

#include <atomic>
#include <mutex>

#include "common.hpp"
#include "device.hpp"

struct ghive_hash_table {
  int layout;        // 8 or 16
  int key_type;      // GHIVE_INT or GHIVE_LONG
  int payload_kind;  // 0 build row id, 1 INT column, 2 LONG column
  int log2_slots;
  uint64_t slots;
  uint64_t entries;
  uint64_t source_rows;  // rows of the build-side table (hit-rate estimate for join ordering)
  int unique;
  int64_t key_min, key_max;
  int64_t pay_min, pay_max;
  void *data;  // device slot array
  // dense companion (unique INT keys over a compact range), else null
  uint32_t *bitmap;  // bit (key - key_min) set for every inserted key
  int32_t *dense;    // payload at (key - key_min)
  uint64_t range;    // key_max - key_min + 1
  ghive_ctx *owner;
  // Asynchronous build: the statistics above are still on the device (read, and the dense
  // companion built, on the table's first use); the build inputs are kept for that.
  std::atomic<bool> pending{false};
  void *stats_buf = nullptr;              // arena block: per-workgroup partials, then the statistics
  const unsigned long long *stats_dev = nullptr;
  const void *in_key = nullptr, *in_pay = nullptr;
  int in_pay_type = 0;
  const uint32_t *in_sel = nullptr;
  const uint64_t *in_n_dev = nullptr;
  uint64_t in_n_max = 0;
  std::mutex ready_mu;
  // finishing an asynchronous build failed: its status and message, returned to every consumer
  int failed = 0;
  std::string fail_msg;
};

namespace ghive {

// Reads the statistics of an asynchronously built table and builds its dense companion, once;
// every consumer of a table calls it first (caller: the consumer's context, whose stream must
// see the finished table). A no-op for tables built synchronously.
int hash_table_ready(ghive_ctx *caller, const ghive_hash_table *h);
// The same for n tables at once: the statistics of the pending tables the caller owns are read
// back with one wait on its stream.
int hash_tables_ready(ghive_ctx *caller, const ghive_hash_table *const *tables, int n);

constexpr uint64_t kEmpty8 = 0x000000007fffffffull;
constexpr uint64_t kEmpty16 = 0x7fffffffffffffffull;

struct DevHash {
  const uint64_t *slots;
  uint64_t mask;
  int32_t log2;
  int32_t layout;
  int64_t key_min, key_max;  // inserted key range (probe pre-check, as Hive's min/max)
};

inline DevHash dev_hash(const ghive_hash_table *h) {
  DevHash d;
  d.slots = static_cast<const uint64_t *>(h->data);
  d.mask = h->slots - 1;
  d.log2 = h->log2_slots;
  d.layout = h->layout;
  d.key_min = h->key_min;
  d.key_max = h->key_max;
  return d;
}

__device__ __forceinline__ uint64_t home_slot(const DevHash &h, int64_t key) {
  return h.layout == 8 ? (uint64_t)hash_u32((uint32_t)key, h.log2) : hash_u64((uint64_t)key, h.log2);
}

// First match of a unique-key table; returns false on a miss.
__device__ __forceinline__ bool probe_unique8(const uint64_t *slots, uint64_t mask, int log2,
                                              int32_t key, int32_t &payload) {
  uint64_t s = hash_u32((uint32_t)key, log2);
  for (uint64_t step = 0; step <= mask; ++step) {  // a table always has empty slots; bounded anyway
    uint64_t w = slots[s];
    int32_t k = (int32_t)(uint32_t)w;
    if (k == key) {
      payload = (int32_t)(w >> 32);
      return true;
    }
    if (k == 0x7fffffff) return false;
    s = (s + 1) & mask;
  }
  return false;
}

int hash_build(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
               const ghive_column *payload, float load_factor, ghive_hash_table **out);

}  // namespace ghive
