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
  bool dense_from_slots = false;          // the dense companion is filled from the table's slots (a
                                          // filtered build keeps no list of its rows)
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

// R lookups of one thread in a unique-key L8 table at once: every row's home slot is loaded
// before any is examined, and only the rows whose slot holds another key go on, again together,
// so a lane keeps R slot loads in flight instead of one. alive[q] in: probe row q; out: matched.
template <int R>
__device__ __forceinline__ void probe_unique8_rows(const uint64_t *slots, uint64_t mask, int log2,
                                                   const int32_t (&key)[R], bool (&alive)[R],
                                                   int32_t (&payload)[R]) {
  uint64_t s[R];
  bool pend[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    pend[q] = alive[q];
    s[q] = pend[q] ? hash_u32((uint32_t)key[q], log2) : 0;
    alive[q] = false;
    payload[q] = 0;
  }
  for (uint64_t round = 0; round <= mask; ++round) {  // a table always has empty slots; bounded anyway
    uint64_t w[R];
#pragma unroll
    for (int q = 0; q < R; ++q) w[q] = pend[q] ? slots[s[q]] : 0;
    bool more = false;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (!pend[q]) continue;
      const int32_t k = (int32_t)(uint32_t)w[q];
      if (k == 0x7fffffff) {  // empty (NULL keys are never inserted)
        pend[q] = false;
      } else if (k == key[q]) {
        alive[q] = true;
        payload[q] = (int32_t)(w[q] >> 32);
        pend[q] = false;
      } else {
        s[q] = (s[q] + 1) & mask;
        more = true;
      }
    }
    if (!more) break;
  }
}

int hash_build(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
               const ghive_column *payload, float load_factor, ghive_hash_table **out);

}  // namespace ghive
