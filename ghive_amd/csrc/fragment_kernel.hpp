// The fused fragment kernel templates (see fragment.hip for the design). Instantiated in
// fragment_r4.hip, fragment_r8.hip, fragment_dense_r4.hip, fragment_dense_r8.hip and
// fragment_pipe.hip so that the instantiations compile in parallel.
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
// columns of a tile issue back to back. Fact columns are read once: non-temporal loads, so that
// they do not displace the dimension bitmaps and group tables from the caches.
template <int R>
__device__ __forceinline__ void load_rows(const int32_t *p, int32_t *v) {
#pragma unroll
  for (int c = 0; c < R / 4; ++c) {
    const v4i x = __builtin_nontemporal_load(reinterpret_cast<const v4i *>(p) + c);
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
    if (F.aggs[a].dbl) continue;  // read per surviving row, from the column itself
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
      if (F.aggs[a].dbl) {
        // Hive's double arithmetic: the product rounded to a double, then the aggregate over it
        const double x = reinterpret_cast<const double *>(F.aggs[a].a)[r0 + q];
        const double y = expr == GHIVE_EXPR_MUL ? reinterpret_cast<const double *>(F.aggs[a].b)[r0 + q] : 1.0;
        const double p = expr == GHIVE_EXPR_MUL ? x * y : x;
        v.valid = !__builtin_isnan(x) && !__builtin_isnan(y);
        const AggSpec &A = T.aggs[a];
        if (A.kind == AK_SUMD) {
          uint64_t lo, hi;
          v.neg = to_fixed(p, A.scale_e, lo, hi);
          v.a = lo;
          v.b = hi;
        } else {
          v.a = dbl_order(p);
        }
      } else if (expr == GHIVE_EXPR_ONE) {
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
      if (F.aggs[a].dbl) continue;  // DOUBLE inputs: read for the surviving rows only
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
      for (int q = 0; q < R; ++q) alive[q] = alive[q] && (uint32_t)fk[j][q] - kmin <= span;
      probe_unique8_rows<R>(J.slots, J.mask, J.log2, fk[j], alive, pay[j]);
    }
  }
  uint32_t mask = 0;
#pragma unroll
  for (int q = 0; q < R; ++q) mask |= (uint32_t)alive[q] << q;
  if (!mask) return;
  if (late) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (F.aggs[a].dbl) continue;  // DOUBLE inputs: read for the surviving rows only
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

// ---- the pipelined star-shape kernel ---------------------------------------------------------
// Same shape as k_fragment_dense (every join through a bitmap, group keys from join attributes),
// for group domains too large for a workgroup's LDS table but small enough to be addressed
// directly (FragDev::direct: no hashing, no returning atomic). Vector memory operations complete
// in issue order, so a wave that waits for one load also waits for every older one. One tile step
// is ordered so that the next tile's column loads are in flight while the current tile waits for
// everything else it needs:
//   1. wait for the current tile's columns (and for any global updates of the previous step);
//      apply the wave's survivor queue if it is half full (see pipe_emit);
//   2. LDS bitmap probes;
//   3. HBM bitmap words, loaded only by lanes whose rows are still alive;
//   4. the next tile's columns, into the other register set;
//   5. HBM bitmap hits: waits for 3 only;
//   6. survivors into the wave's LDS queue.
// Global writes are issued only in steps 1 and 6, and step 1 waits for all memory operations:
// with writes pending as well as loads, loads can no longer be counted in order, so step 5 would
// wait for the next tile's columns too.
constexpr int kPipeRows = 4;
// s_waitcnt immediate of "vmcnt(0)" on gfx9-family chips (expcnt and lgkmcnt left at their maxima);
// OR-ing a count below 16 into the low bits gives "vmcnt(count)"
constexpr int kWaitVmcnt0 = 0x0f70;
static_assert(kMaxFragJoins + 2 * kMaxFragAggs < 16, "the prefetch's load count must fit vmcnt's low bits");

template <int NJ, int NA>
struct PipeTile {
  int32_t c[NJ + 2 * NA][kPipeRows];  // joins' fact keys, then each aggregate's two inputs
};

template <int NJ, int NA>
__device__ __forceinline__ void pipe_load(const int32_t *const (&col)[NJ + 2 * NA], uint64_t off, PipeTile<NJ, NA> &t) {
#pragma unroll
  for (int i = 0; i < NJ + 2 * NA; ++i) load_rows<kPipeRows>(col[i] + off, t.c[i]);
}

// Steps 2-3. Joins 0 .. NJ-NH-1 are probed in LDS, joins NJ-NH .. NJ-1 through HBM bitmaps (the
// plan orders them so). Returns the rows alive after the LDS probes whose keys are in range of
// every HBM bitmap; sh[q] packs the bit positions (5 bits per HBM join) that step 5 tests.
template <int NJ, int NH, int NA>
__device__ __forceinline__ uint32_t pipe_probe_issue(const FragDev &F, const uint32_t *lbits, const PipeTile<NJ, NA> &t,
                                                     uint32_t (&w)[NH > 0 ? NH : 1][kPipeRows],
                                                     uint32_t (&sh)[kPipeRows]) {
  constexpr int R = kPipeRows, NL = NJ - NH;
  uint32_t alive = (1u << R) - 1;
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const FragJoinDev &J = F.joins[j];
    const uint32_t kmin = (uint32_t)J.key_min, span = J.span;
    const uint32_t *bits = lbits + J.lds_word_off;
    uint32_t u[R], x[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      u[q] = (uint32_t)t.c[j][q] - kmin;
      x[q] = bits[(u[q] < span ? u[q] : span) >> 5];
    }
    uint32_t hit = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) hit |= ((uint32_t)(u[q] <= span) & (x[q] >> (u[q] & 31u))) << q;
    alive &= hit;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) sh[q] = 0;
#pragma unroll
  for (int h = 0; h < NH; ++h) {  // out-of-range keys of the HBM bitmaps leave first
    const FragJoinDev &J = F.joins[NL + h];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const uint32_t u = (uint32_t)t.c[NL + h][q] - (uint32_t)J.key_min;
      alive &= ~((uint32_t)(u > J.span) << q);
      sh[q] |= (u & 31u) << (5 * h);
    }
  }
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const FragJoinDev &J = F.joins[NL + h];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      w[h][q] = 0;
      if ((alive >> q) & 1u)  // exec-masked: a dead row issues no request
        w[h][q] = J.bitmap[((uint32_t)t.c[NL + h][q] - (uint32_t)J.key_min) >> 5];
    }
  }
  return alive;
}

// Step 5: the HBM bitmap tests.
template <int NH>
__device__ __forceinline__ uint32_t pipe_probe_finish(uint32_t alive, const uint32_t (&w)[NH > 0 ? NH : 1][kPipeRows],
                                                      const uint32_t (&sh)[kPipeRows]) {
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    uint32_t hit = 0;
#pragma unroll
    for (int q = 0; q < kPipeRows; ++q) hit |= ((w[h][q] >> ((sh[q] >> (5 * h)) & 31u)) & 1u) << q;
    alive &= hit;
  }
  return alive;
}

// Step 6. Each wave appends its survivors -- the offsets of their keys in the key joins' dense
// attribute arrays and their raw aggregate inputs -- to a small queue of its own in LDS. The queue
// is applied to the global direct table, one entry per lane, at the start of a tile step once it
// is half full (and once at the end): the attribute loads (one per survivor, not one per row
// alive after the LDS probes) and the fire-and-forget atomics then make a wave wait a few times
// in its life rather than on every tile, and run with every lane busy. A tile with more
// survivors than the queue has room for sends them straight to the global table.
struct PipeState {
  const FragDev *F;
  const GroupTableDev *T;
  uint32_t *qoff;       // this wave's queue: 2 rows of qcap key offsets, then
  int32_t *qin;         // 2 * NA rows of qcap aggregate inputs
  uint32_t qcap;
  uint32_t qn;          // entries queued (wave-uniform)
  uint32_t njoined;
  const int32_t *kdense[2];  // dense attribute array of each group key's join
  int32_t kmin[2];           // that join's smallest key
};

__device__ __forceinline__ void pipe_apply_global(int kind, uint64_t *w0, uint64_t *w1, int64_t val) {
  typedef __attribute__((address_space(1))) unsigned long long g_u64;
  typedef __attribute__((address_space(1))) long long g_i64;
  g_u64 *a0 = (g_u64 *)w0, *a1 = (g_u64 *)w1;
  if (kind == AK_CNT) {
    __hip_atomic_fetch_add(a0, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  if (kind == AK_SUMI) __hip_atomic_fetch_add(a0, (unsigned long long)val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else if (kind == AK_MINI) __hip_atomic_fetch_min((g_i64 *)w0, (long long)val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else __hip_atomic_fetch_max((g_i64 *)w0, (long long)val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(a1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// value and validity of an aggregate for the inputs (x, y)
__device__ __forceinline__ bool pipe_agg_value(int expr, int32_t x, int32_t y, int64_t &val) {
  if (expr == GHIVE_EXPR_ONE) {
    val = 1;
    return true;
  }
  val = x;
  if (expr != GHIVE_EXPR_MUL) return x != kNullI32;
  val = (int64_t)((uint64_t)(int64_t)x * (uint64_t)(int64_t)y);
  return x != kNullI32 && y != kNullI32;
}

// one surviving row into the global table: key attribute offsets off[k], aggregate inputs x, y
template <int NA>
__device__ __forceinline__ void pipe_global_row(const PipeState &S, const uint32_t (&off)[2], const int32_t (&x)[NA],
                                                const int32_t (&y)[NA]) {
  const FragDev &F = *S.F;
  const GroupTableDev &T = *S.T;
  int32_t v[2] = {0, 0};
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k < F.nkeys) v[k] = S.kdense[k][off[k]];
  uint32_t g = 0;
  uint64_t key = 0;
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k < F.nkeys) {
      g += (uint32_t)(v[k] - F.dmin[k]) * F.dmul[k];
      key |= (uint64_t)(uint32_t)v[k] << (32 * k);
    }
  *(__attribute__((address_space(1))) uint64_t *)(T.keys + g) = key;  // same key for every writer
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    int64_t val;
    if (pipe_agg_value(F.aggs[a].expr, x[a], y[a], val))
      pipe_apply_global(T.aggs[a].kind, acc_word(T, T.aggs[a].word0, (uint64_t)g),
                        acc_word(T, T.aggs[a].word0 + 1, (uint64_t)g), val);
  }
}

template <int NA>
__device__ __forceinline__ void pipe_queue_flush(PipeState &S) {
  for (uint32_t i = lane_id(); i < S.qn; i += kWave) {
    uint32_t off[2];
    int32_t x[NA], y[NA];
#pragma unroll
    for (int k = 0; k < 2; ++k) off[k] = S.qoff[(size_t)k * S.qcap + i];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      x[a] = S.qin[(size_t)(2 * a) * S.qcap + i];
      y[a] = S.qin[(size_t)(2 * a + 1) * S.qcap + i];
    }
    pipe_global_row<NA>(S, off, x, y);
  }
  S.qn = 0;
}

// offset of row q's key k in its join's dense attribute array
template <int NJ, int NA>
__device__ __forceinline__ uint32_t pipe_key_off(const PipeState &S, const PipeTile<NJ, NA> &t, int k, int q) {
  int32_t fk = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (S.F->key_idx[k] == j) fk = t.c[j][q];
  return (uint32_t)fk - (uint32_t)S.kmin[k];
}

template <int NJ, int NA>
__device__ __forceinline__ void pipe_emit(PipeState &S, uint32_t mask, const PipeTile<NJ, NA> &t) {
  const FragDev &F = *S.F;
  // this lane's survivors (0..4) and the number queued by the lanes below it: a prefix sum of the
  // 3-bit counts from one ballot per bit
  const uint32_t c = (uint32_t)__builtin_popcount(mask);
  uint32_t total = 0, below = 0;
#pragma unroll
  for (int bit = 0; bit < 3; ++bit) {
    const uint64_t b = __ballot((c >> bit) & 1u);
    total += (uint32_t)__popcll(b) << bit;
    below += __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u)) << bit;
  }
  S.njoined += c;
  if (S.qn + total > S.qcap) {  // more rows than the queue has room for: straight to the global table
#pragma unroll
    for (int q = 0; q < kPipeRows; ++q) {
      if (!((mask >> q) & 1u)) continue;
      uint32_t off[2] = {0, 0};
      int32_t x[NA], y[NA];
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (k < F.nkeys) off[k] = pipe_key_off<NJ, NA>(S, t, k, q);
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        x[a] = t.c[NJ + 2 * a][q];
        y[a] = t.c[NJ + 2 * a + 1][q];
      }
      pipe_global_row<NA>(S, off, x, y);
    }
    return;
  }
  uint32_t i = S.qn + below;
#pragma unroll
  for (int q = 0; q < kPipeRows; ++q) {
    if (!((mask >> q) & 1u)) continue;
#pragma unroll
    for (int k = 0; k < 2; ++k) S.qoff[(size_t)k * S.qcap + i] = k < F.nkeys ? pipe_key_off<NJ, NA>(S, t, k, q) : 0u;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      S.qin[(size_t)(2 * a) * S.qcap + i] = t.c[NJ + 2 * a][q];
      S.qin[(size_t)(2 * a + 1) * S.qcap + i] = t.c[NJ + 2 * a + 1][q];
    }
    ++i;
  }
  S.qn += total;
}

// One tile step: `cur` holds the tile, `nxt` receives the tile at element offset next_off (the
// tile itself when it is the workgroup's last); prefetch = false for the staged partial tile.
template <int NJ, int NH, int NA>
__device__ __forceinline__ void pipe_step(PipeState &S, const uint32_t *lbits, const int32_t *const (&col)[NJ + 2 * NA],
                                          uint64_t next_off, const PipeTile<NJ, NA> &cur, PipeTile<NJ, NA> &nxt,
                                          bool prefetch) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  // the queue is applied here, where the other register set is free, once it is half full
  if (S.qn > S.qcap / 2) pipe_queue_flush<NA>(S);
  uint32_t w[NH > 0 ? NH : 1][kPipeRows], sh[kPipeRows];
  uint32_t alive = pipe_probe_issue<NJ, NH, NA>(*S.F, lbits, cur, w, sh);
  if (prefetch) {
    pipe_load<NJ, NA>(col, next_off, nxt);
    // everything older than the next tile's loads: the bitmap words and payloads of this tile
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0 | (NJ + 2 * NA));
  } else {
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  }
  alive = pipe_probe_finish<NH>(alive, w, sh);
  if (__ballot(alive != 0)) pipe_emit<NJ, NA>(S, alive, cur);
}

// NJ joins of which the last NH probe HBM bitmaps, NA aggregates.
template <int NJ, int NH, int NA>
__global__ __launch_bounds__(kFragThreads) void k_fragment_pipe(FragDev F, GroupTableDev T, int lcap_log2,
                                                             unsigned long long *joined) {
  constexpr int R = kPipeRows;
  extern __shared__ uint64_t lds64[];
  uint32_t *lbits = reinterpret_cast<uint32_t *>(lds64);
  (void)lcap_log2;  // no LDS group table: the direct slots live in the global table
  PipeState S;
  S.F = &F;
  S.T = &T;
  // queues after the bitmaps, one per wave: 2 rows of qcap key offsets, 2 * NA rows of inputs
  S.qcap = F.queue_cap;
  S.qoff = reinterpret_cast<uint32_t *>(lds64 + F.lds_bitmap_words / 2) +
           (size_t)(threadIdx.x / kWave) * F.queue_cap * (2 + 2 * NA);
  S.qin = reinterpret_cast<int32_t *>(S.qoff + 2 * (size_t)F.queue_cap);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const FragJoinDev &J = F.joins[k < F.nkeys ? F.key_idx[k] : 0];
    S.kdense[k] = J.dense;
    S.kmin[k] = J.key_min;
  }
  S.qn = 0;
  S.njoined = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
    if (F.joins[j].mode == kJoinLdsBits)
      for (uint32_t i = threadIdx.x; i < F.joins[j].nwords; i += blockDim.x)
        lbits[F.joins[j].lds_word_off + i] = F.joins[j].bitmap[i];
  __syncthreads();

  const int32_t *col[NJ + 2 * NA];
#pragma unroll
  for (int j = 0; j < NJ; ++j) col[j] = F.joins[j].fk;
#pragma unroll
  for (int a = 0; a < NA; ++a) {  // both inputs always valid here (the plan fills unused ones)
    col[NJ + 2 * a] = F.aggs[a].a;
    col[NJ + 2 * a + 1] = F.aggs[a].b;
  }
  const uint64_t tile_rows = (uint64_t)blockDim.x * R;
  const uint64_t n = F.n;
  const uint64_t ntiles = (n + tile_rows - 1) / tile_rows;
  const uint64_t tile_begin = ntiles * blockIdx.x / gridDim.x;
  const uint64_t tile_end = ntiles * (blockIdx.x + 1) / gridDim.x;
  const uint64_t full_tiles = n / tile_rows;
  const uint64_t full_end = tile_end < full_tiles ? tile_end : full_tiles;
  const uint64_t lane_row = (uint64_t)threadIdx.x * R;
  PipeTile<NJ, NA> ta, tb;
  if (tile_begin < full_end) {
    // two steps per iteration with the register sets' roles swapped, so that no tile is copied
    pipe_load<NJ, NA>(col, tile_begin * tile_rows + lane_row, ta);
    for (uint64_t tile = tile_begin;; tile += 2) {
      const uint64_t n1 = tile + 1 < full_end ? tile + 1 : tile;
      pipe_step<NJ, NH, NA>(S, lbits, col, n1 * tile_rows + lane_row, ta, tb, true);
      if (tile + 1 >= full_end) break;
      const uint64_t n2 = tile + 2 < full_end ? tile + 2 : tile + 1;
      pipe_step<NJ, NH, NA>(S, lbits, col, n2 * tile_rows + lane_row, tb, ta, true);
      if (tile + 2 >= full_end) break;
    }
  }
  if (tile_end > full_end) {  // uniform: this workgroup owns the table's partial last tile
    FragBlock B;
    B.tile_rows = tile_rows;
    frag_stage_tail<NJ, NA>(F, B, full_tiles * tile_rows);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NJ + 2 * NA; ++i) load_rows<R>(F.stage + (uint64_t)i * tile_rows + lane_row, ta.c[i]);
    pipe_step<NJ, NH, NA>(S, lbits, col, 0, ta, tb, false);
  }
  pipe_queue_flush<NA>(S);
  unsigned long long s = S.njoined;
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
// k_fragment_pipe instantiations (1 <= nj, 0 <= nh <= nj); one file per join count
FragKernel pick_pipe_j1(int nh, int na);
FragKernel pick_pipe_j2(int nh, int na);
FragKernel pick_pipe_j3(int nh, int na);
FragKernel pick_pipe_j4(int nh, int na);

}  // namespace ghive
