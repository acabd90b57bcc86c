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

// R consecutive int32 values at p: 16-byte loads with no per-lane guard, so the loads of all
// columns of a tile issue back to back.
template <int R>
__device__ __forceinline__ void load_rows(const int32_t *p, int32_t *v) {
#pragma unroll
  for (int c = 0; c < R / 4; ++c) {
    const v4i x = reinterpret_cast<const v4i *>(p)[c];
    v[4 * c + 0] = x.x;
    v[4 * c + 1] = x.y;
    v[4 * c + 2] = x.z;
    v[4 * c + 3] = x.w;
  }
}

// Shared state of one workgroup of the fused kernel.
struct FragBlock {
  uint32_t *lbits;  // LDS bitmaps
  uint64_t *ltab;   // LDS group table (lcap_log2 >= 0)
  int lcap_log2;
  int lstride;
  uint32_t njoined;
  uint64_t tile_rows;
};

// Copies the rows of the partial last tile (from row `base`) of every streamed column into the
// staging buffer, NULL past the end of the table, so that the tile is read like any other.
template <int NJ, int NA>
__device__ __forceinline__ void frag_stage_tail(const FragDev &F, const FragBlock &B, uint64_t base) {
  const uint64_t rem = F.n - base;
  auto stage = [&](const int32_t *col, int slot) {
    int32_t *dst = F.stage + (uint64_t)slot * B.tile_rows;
    for (uint32_t i = threadIdx.x; i < B.tile_rows; i += blockDim.x) dst[i] = i < rem ? col[base + i] : kNullI32;
  };
#pragma unroll
  for (int j = 0; j < NJ; ++j) stage(F.joins[j].fk, j);
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (F.aggs[a].a) stage(F.aggs[a].a, NJ + 2 * a);
    if (F.aggs[a].b) stage(F.aggs[a].b, NJ + 2 * a + 1);
  }
}

// Rows of a tile that survived every join (bit q of mask): dimension attributes of the dense
// joins, group key, aggregate update in the LDS table (else the global one).
template <int NJ, int NA, int R>
__device__ __forceinline__ void frag_emit(const FragDev &F, const GroupTableDev &T, FragBlock &B, uint64_t r0,
                                          uint32_t mask, const int32_t (&fk)[NJ > 0 ? NJ : 1][R],
                                          int32_t (&pay)[NJ > 0 ? NJ : 1][R], const int32_t (&av)[NA][R],
                                          const int32_t (&bv)[NA][R]) {
  const bool use_lds = B.lcap_log2 >= 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const FragJoinDev &J = F.joins[j];
    if (J.mode != kJoinHash && J.need_payload) {
#pragma unroll
      for (int q = 0; q < R; ++q)
        if ((mask >> q) & 1u) pay[j][q] = J.dense[(uint32_t)fk[j][q] - (uint32_t)J.key_min];
    }
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    if (!((mask >> q) & 1u)) continue;
    ++B.njoined;
    uint64_t key = 0;
    for (int k = 0; k < F.nkeys; ++k) {
      int32_t kv = 0;
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        if (F.key_src[k] == GHIVE_SRC_JOIN && F.key_idx[k] == j) kv = pay[j][q];
      if (F.key_src[k] != GHIVE_SRC_JOIN) kv = F.key_col[k][r0 + q];
      key |= (uint64_t)(uint32_t)kv << (32 * k);
    }
    const int ls = use_lds ? lds_find_insert(B.ltab, key, B.lcap_log2) : -1;
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
        acc_apply(A, [&](int w) { return B.ltab + (size_t)(1 + A.word0 + w) * B.lstride + ls; }, v);
      else
        acc_apply(A, [&](int w) { return acc_word(T, A.word0 + w, (uint64_t)g); }, v);
    }
  }
}

// One tile: R consecutive rows per lane starting at row r0 (lane_row within the tile). tail: the
// table's partial last tile, read from the staging copy.
template <int NJ, int NA, int R>
__device__ __forceinline__ void frag_tile(const FragDev &F, const GroupTableDev &T, FragBlock &B, uint64_t r0,
                                          uint64_t lane_row, bool tail) {
  const uint64_t n = F.n;
  const bool late = F.late != 0;
  auto at = [&](const int32_t *col, int slot) -> const int32_t * {
    return tail ? F.stage + (uint64_t)slot * B.tile_rows + lane_row : col + r0;
  };
  bool alive[R];
#pragma unroll
  for (int q = 0; q < R; ++q) alive[q] = r0 + q < n;
  int32_t fk[NJ > 0 ? NJ : 1][R];
  int32_t av[NA][R], bv[NA][R];
  if (!late) {  // every column of the tile requested before any is used
#pragma unroll
    for (int j = 0; j < NJ; ++j) load_rows<R>(at(F.joins[j].fk, j), fk[j]);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (F.aggs[a].expr != GHIVE_EXPR_ONE) load_rows<R>(at(F.aggs[a].a, NJ + 2 * a), av[a]);
      if (F.aggs[a].expr == GHIVE_EXPR_MUL) load_rows<R>(at(F.aggs[a].b, NJ + 2 * a + 1), bv[a]);
    }
  }
  if (F.has_filter) {
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (alive[q]) alive[q] = eval_pred(F.filter, r0 + q);
  }
  int32_t pay[NJ > 0 ? NJ : 1][R];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const FragJoinDev &J = F.joins[j];
    if (late) {
      bool any = false;
#pragma unroll
      for (int q = 0; q < R; ++q) any |= alive[q];
      if (any) {
        load_rows<R>(at(J.fk, j), fk[j]);
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) fk[j][q] = kNullI32;
      }
    }
    const uint32_t kmin = (uint32_t)J.key_min, span = J.span;
    if (J.mode == kJoinLdsBits) {  // branch-free: an out-of-range key reads a clamped word, masked
      const uint32_t *bits = B.lbits + J.lds_word_off;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const uint32_t u = (uint32_t)fk[j][q] - kmin;
        const uint32_t w = bits[(u < span ? u : span) >> 5];
        alive[q] = alive[q] & (u <= span) & (((w >> (u & 31u)) & 1u) != 0);
      }
    } else if (J.mode == kJoinGlobalBits) {
      uint32_t w[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {  // independent loads, all in flight together
        const uint32_t u = (uint32_t)fk[j][q] - kmin;
        w[q] = 0;
        if (alive[q] && u <= span) w[q] = J.bitmap[u >> 5] >> (u & 31u);
      }
#pragma unroll
      for (int q = 0; q < R; ++q) alive[q] = alive[q] && (w[q] & 1u);
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        int32_t p = 0;
        if (alive[q]) {
          const int32_t k = fk[j][q];
          alive[q] = (uint32_t)k - kmin <= span && probe_unique8(J.slots, J.mask, J.log2, k, p);
        }
        pay[j][q] = p;
      }
    }
  }
  uint32_t mask = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) mask |= (uint32_t)alive[q] << q;
  if (!mask) return;
  if (late) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (F.aggs[a].expr != GHIVE_EXPR_ONE) load_rows<R>(at(F.aggs[a].a, NJ + 2 * a), av[a]);
      if (F.aggs[a].expr == GHIVE_EXPR_MUL) load_rows<R>(at(F.aggs[a].b, NJ + 2 * a + 1), bv[a]);
    }
  }
  frag_emit<NJ, NA, R>(F, T, B, r0, mask, fk, pay, av, bv);
}

// NJ joins, NA aggregates, R rows per thread (4 or 8); F.late is a uniform flag. Every workgroup
// streams one contiguous range of tiles (consecutive DRAM pages per column); a tile is
// blockDim.x * R rows, R consecutive rows per lane. Only the table's last tile can be partial.
template <int NJ, int NA, int R>
__global__ __launch_bounds__(kFragMaxThreads) void k_fragment(FragDev F, GroupTableDev T, int lcap_log2,
                                                           unsigned long long *joined) {
  extern __shared__ uint64_t lds64[];
  FragBlock B;
  B.lbits = reinterpret_cast<uint32_t *>(lds64);
  B.ltab = lds64 + F.lds_bitmap_words / 2;
  B.lcap_log2 = lcap_log2;
  B.lstride = lcap_log2 >= 0 ? (1 << lcap_log2) + 1 : 0;
  B.njoined = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (F.joins[j].mode == kJoinLdsBits)
      for (uint32_t w = threadIdx.x; w < F.joins[j].nwords; w += blockDim.x)
        B.lbits[F.joins[j].lds_word_off + w] = F.joins[j].bitmap[w];
  if (lcap_log2 >= 0) lds_table_init(B.ltab, T, lcap_log2);
  __syncthreads();

  const uint64_t tile_rows = (uint64_t)blockDim.x * R;
  B.tile_rows = tile_rows;
  const uint64_t n = F.n;
  const uint64_t ntiles = (n + tile_rows - 1) / tile_rows;
  const uint64_t tile_begin = ntiles * blockIdx.x / gridDim.x;
  const uint64_t tile_end = ntiles * (blockIdx.x + 1) / gridDim.x;
  const uint64_t full_tiles = n / tile_rows;
  const uint64_t lane_row = (uint64_t)threadIdx.x * R;
  for (uint64_t tile = tile_begin; tile < tile_end; ++tile) {
    const bool tail = tile >= full_tiles;  // uniform: only the partial last tile of the table
    if (tail) {
      frag_stage_tail<NJ, NA>(F, B, full_tiles * tile_rows);
      __syncthreads();
    }
    frag_tile<NJ, NA, R>(F, T, B, tile * tile_rows + lane_row, lane_row, tail);
  }
  __syncthreads();
  if (lcap_log2 >= 0) lds_table_flush(B.ltab, T, lcap_log2);
  unsigned long long s = B.njoined;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
  if (lane_id() == 0 && s) atomicAdd(joined, s);
}

// The fused kernel for the common star shape: no fact filter, every column read up front, every
// join probed through a dense existence bitmap (LDS or HBM), group keys from join attributes.
// Same results as k_fragment; the leaner body keeps the per-join constants in registers, and the
// LDS probes of a tile are branch-free so that they issue together.
template <int NJ, int NA, int R>
__global__ __launch_bounds__(kFragMaxThreads) void k_fragment_dense(FragDev F, GroupTableDev T, int lcap_log2,
                                                                 unsigned long long *joined) {
  extern __shared__ uint64_t lds64[];
  FragBlock B;
  B.lbits = reinterpret_cast<uint32_t *>(lds64);
  B.ltab = lds64 + F.lds_bitmap_words / 2;
  B.lcap_log2 = lcap_log2;
  B.lstride = lcap_log2 >= 0 ? (1 << lcap_log2) + 1 : 0;
  B.njoined = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (F.joins[j].mode == kJoinLdsBits)
      for (uint32_t w = threadIdx.x; w < F.joins[j].nwords; w += blockDim.x)
        B.lbits[F.joins[j].lds_word_off + w] = F.joins[j].bitmap[w];
  if (lcap_log2 >= 0) lds_table_init(B.ltab, T, lcap_log2);
  __syncthreads();

  const uint64_t tile_rows = (uint64_t)blockDim.x * R;
  B.tile_rows = tile_rows;
  const uint64_t n = F.n;
  const uint64_t ntiles = (n + tile_rows - 1) / tile_rows;
  const uint64_t tile_begin = ntiles * blockIdx.x / gridDim.x;
  const uint64_t tile_end = ntiles * (blockIdx.x + 1) / gridDim.x;
  const uint64_t full_tiles = n / tile_rows;
  const uint64_t lane_row = (uint64_t)threadIdx.x * R;
  for (uint64_t tile = tile_begin; tile < tile_end; ++tile) {
    const bool tail = tile >= full_tiles;  // uniform: only the partial last tile of the table
    if (tail) {
      frag_stage_tail<NJ, NA>(F, B, full_tiles * tile_rows);
      __syncthreads();
    }
    const uint64_t r0 = tile * tile_rows + lane_row;
    auto at = [&](const int32_t *col, int slot) -> const int32_t * {
      return tail ? F.stage + (uint64_t)slot * tile_rows + lane_row : col + r0;
    };
    int32_t fk[NJ][R];
    int32_t av[NA][R], bv[NA][R];
#pragma unroll
    for (int j = 0; j < NJ; ++j) load_rows<R>(at(F.joins[j].fk, j), fk[j]);
#pragma unroll
    for (int a = 0; a < NA; ++a) {  // both inputs always valid here (the plan fills unused ones)
      load_rows<R>(at(F.aggs[a].a, NJ + 2 * a), av[a]);
      load_rows<R>(at(F.aggs[a].b, NJ + 2 * a + 1), bv[a]);
    }
    uint32_t alive = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) alive |= (uint32_t)(r0 + q < n) << q;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const FragJoinDev &J = F.joins[j];
      const uint32_t kmin = (uint32_t)J.key_min, span = J.span;
      uint32_t hit = 0;
      if (J.mode == kJoinLdsBits) {  // every row probes; an out-of-range key reads a clamped word
        const uint32_t *bits = B.lbits + J.lds_word_off;
        uint32_t u[R], w[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          u[q] = (uint32_t)fk[j][q] - kmin;
          w[q] = bits[(u[q] < span ? u[q] : span) >> 5];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) hit |= ((uint32_t)(u[q] <= span) & (w[q] >> (u[q] & 31u))) << q;
      } else {  // HBM bitmap (L2-resident): only rows still alive issue a load
        uint32_t u[R], w[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
          u[q] = (uint32_t)fk[j][q] - kmin;
          w[q] = 0;
          if (((alive >> q) & 1u) && u[q] <= span) w[q] = J.bitmap[u[q] >> 5];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) hit |= ((w[q] >> (u[q] & 31u)) & 1u) << q;
      }
      alive &= hit;
    }
    if (!alive) continue;
    int32_t pay[NJ][R];
    frag_emit<NJ, NA, R>(F, T, B, r0, alive, fk, pay, av, bv);
  }
  __syncthreads();
  if (lcap_log2 >= 0) lds_table_flush(B.ltab, T, lcap_log2);
  unsigned long long s = B.njoined;
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, kWave);
  if (lane_id() == 0 && s) atomicAdd(joined, s);
}

using FragKernel = void (*)(FragDev, GroupTableDev, int, unsigned long long *);

FragKernel pick_kernel_r4(int nj, int na);
FragKernel pick_kernel_r8(int nj, int na);
// k_fragment_dense instantiations (1 <= nj)
FragKernel pick_dense_r4(int nj, int na);
FragKernel pick_dense_r8(int nj, int na);

}  // namespace ghive
