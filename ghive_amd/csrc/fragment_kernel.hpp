// The fused fragment kernel template (see fragment.hip for the design). Instantiated in
// fragment_r4.hip and fragment_r8.hip so that the instantiations compile in parallel.
#pragma once

#include "agg.hpp"
#include "device.hpp"
#include "fragment.hpp"
#include "hash.hpp"

namespace ghive {

constexpr int kFragThreads = 256;     // default workgroup size
constexpr int kFragMaxThreads = 1024; // launch bound; the size is chosen at launch
constexpr int kNullI32 = 0x7fffffff;

typedef int v4i __attribute__((ext_vector_type(4)));

// R consecutive int32 values of row r0..r0+R-1 (NULL past the end); 16-byte loads, optionally
// non-temporal (streamed fact columns should not evict the dimension bitmaps from L2).
template <int R>
__device__ __forceinline__ void loadR(const int32_t *p, uint64_t r0, uint64_t n, int32_t *v, bool nt) {
  if (r0 + R <= n) {
#pragma unroll
    for (int c = 0; c < R / 4; ++c) {
      const v4i *src = reinterpret_cast<const v4i *>(p + r0) + c;
      v4i x = nt ? __builtin_nontemporal_load(src) : *src;
      v[4 * c + 0] = x.x;
      v[4 * c + 1] = x.y;
      v[4 * c + 2] = x.z;
      v[4 * c + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = r0 + q < n ? p[r0 + q] : kNullI32;
  }
}

__device__ __forceinline__ bool in_range(const FragJoinDev &J, int32_t k) {
  return k != kNullI32 && k >= J.key_min && k <= J.key_max;
}

// NJ joins, NA aggregates, R rows per thread (4 or 8). F.late / F.nt are uniform flags.
template <int NJ, int NA, int R>
__global__ __launch_bounds__(kFragMaxThreads) void k_fragment(FragDev F, GroupTableDev T, int lcap_log2,
                                                           unsigned long long *joined) {
  extern __shared__ uint64_t lds64[];
  uint32_t *lbits = reinterpret_cast<uint32_t *>(lds64);
  uint64_t *ltab = lds64 + F.lds_bitmap_words / 2;
  const bool use_lds = lcap_log2 >= 0;
  const bool late = F.late != 0, nt = F.nt != 0;
  const int lstride = use_lds ? (1 << lcap_log2) + 1 : 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (F.joins[j].mode == kJoinLdsBits)
      for (uint32_t w = threadIdx.x; w < F.joins[j].nwords; w += blockDim.x)
        lbits[F.joins[j].lds_word_off + w] = F.joins[j].bitmap[w];
  if (use_lds) lds_table_init(ltab, T, lcap_log2);
  __syncthreads();

  constexpr uint32_t kAll = (1u << R) - 1u;
  const uint64_t tile_rows = (uint64_t)blockDim.x * R;
  const uint64_t n = F.n;
  const uint64_t ntiles = (n + tile_rows - 1) / tile_rows;
  uint32_t njoined = 0;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t r0 = tile * tile_rows + (uint64_t)threadIdx.x * R;
    uint32_t alive = r0 + R <= n ? kAll : 0u;
    if (!alive) {
#pragma unroll
      for (int q = 0; q < R; ++q) alive |= (r0 + q < n ? 1u : 0u) << q;
    }
    int32_t fk[NJ > 0 ? NJ : 1][R];
    int32_t av[NA][R], bv[NA][R];
    if (!late) {  // every column load of the tile issued up front: many bytes in flight
#pragma unroll
      for (int j = 0; j < NJ; ++j) loadR<R>(F.joins[j].fk, r0, n, fk[j], nt);
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        if (F.aggs[a].expr != GHIVE_EXPR_ONE) loadR<R>(F.aggs[a].a, r0, n, av[a], nt);
        if (F.aggs[a].expr == GHIVE_EXPR_MUL) loadR<R>(F.aggs[a].b, r0, n, bv[a], nt);
      }
    }
    if (F.has_filter && alive) {
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (((alive >> q) & 1u) && !eval_pred(F.filter, r0 + q)) alive &= ~(1u << q);
    }
    int32_t pay[NJ > 0 ? NJ : 1][R];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const FragJoinDev &J = F.joins[j];
      if (late && alive) loadR<R>(J.fk, r0, n, fk[j], nt);
      if (J.mode == kJoinLdsBits) {
        const uint32_t *bits = lbits + J.lds_word_off;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (!((alive >> q) & 1u)) continue;
          const int32_t k = fk[j][q];
          bool hit = false;
          if (in_range(J, k)) {
            const uint32_t u = (uint32_t)((int64_t)k - J.key_min);
            hit = (bits[u >> 5] >> (u & 31)) & 1u;
          }
          if (!hit) alive &= ~(1u << q);
        }
      } else if (J.mode == kJoinGlobalBits) {
        uint32_t w[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {  // independent loads, all in flight together
          w[q] = 0;
          const int32_t k = fk[j][q];
          if (((alive >> q) & 1u) && in_range(J, k)) {
            const uint32_t u = (uint32_t)((int64_t)k - J.key_min);
            w[q] = J.bitmap[u >> 5] >> (u & 31);
          }
        }
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (!(w[q] & 1u)) alive &= ~(1u << q);
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if (!((alive >> q) & 1u)) continue;
          const int32_t k = fk[j][q];
          int32_t p = 0;
          bool hit = in_range(J, k) && probe_unique8(J.slots, J.mask, J.log2, k, p);
          if (!hit) alive &= ~(1u << q);
          pay[j][q] = p;
        }
      }
      if (J.mode != kJoinHash && J.need_payload) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if ((alive >> q) & 1u) pay[j][q] = J.dense[(uint32_t)((int64_t)fk[j][q] - J.key_min)];
      }
    }
    if (!alive) continue;
    if (late) {
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        if (F.aggs[a].expr != GHIVE_EXPR_ONE) loadR<R>(F.aggs[a].a, r0, n, av[a], nt);
        if (F.aggs[a].expr == GHIVE_EXPR_MUL) loadR<R>(F.aggs[a].b, r0, n, bv[a], nt);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (!((alive >> q) & 1u)) continue;
      ++njoined;
      uint64_t key = 0;
      for (int k = 0; k < F.nkeys; ++k) {
        int32_t kv = 0;
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (F.key_src[k] == GHIVE_SRC_JOIN && F.key_idx[k] == j) kv = pay[j][q];
        if (F.key_src[k] != GHIVE_SRC_JOIN) kv = F.key_col[k][r0 + q];
        key |= (uint64_t)(uint32_t)kv << (32 * k);
      }
      const int ls = use_lds ? lds_find_insert(ltab, key, lcap_log2) : -1;
      int64_t g = -1;
      if (ls < 0) {
        g = gt_find_insert(T, key);
        if (g < 0) continue;
      }
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const int expr = F.aggs[a].expr;
        AggVal v;
        v.neg = false;
        v.b = 0;
        if (expr == GHIVE_EXPR_ONE) {
          v.valid = true;
          v.a = 1;
        } else {
          const int32_t x = av[a][q];
          v.valid = x != kNullI32;
          int64_t val = x;
          if (expr == GHIVE_EXPR_MUL) {
            const int32_t y = bv[a][q];
            v.valid = v.valid && y != kNullI32;
            val = (int64_t)((uint64_t)(int64_t)x * (uint64_t)(int64_t)y);
          }
          v.a = (uint64_t)val;
        }
        const AggSpec &A = T.aggs[a];
        if (ls >= 0)
          acc_apply(A, [&](int w) { return ltab + (size_t)(1 + A.word0 + w) * lstride + ls; }, v);
        else
          acc_apply(A, [&](int w) { return T.acc + (size_t)(A.word0 + w) * T.stride + g; }, v);
      }
    }
  }
  __syncthreads();
  if (use_lds) lds_table_flush(ltab, T, lcap_log2);
  unsigned long long s = njoined;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
  if (lane_id() == 0 && s) atomicAdd(joined, s);
}

using FragKernel = void (*)(FragDev, GroupTableDev, int, unsigned long long *);

FragKernel pick_kernel_r4(int nj, int na);
FragKernel pick_kernel_r8(int nj, int na);

}  // namespace ghive
