// ORDER BY as a stable LSD radix sort (replaces SinkOperator's thrust::sort of row ids with a
// type-generic comparator plus a CPU permutation gather, gpu_impl/src/Operator/SinkOperator.cu:
// 135-288; the sort itself is thrust::sequence + thrust::sort(sort_comparator) at :218-222). Keys are turned into one order-preserving unsigned bit string per row (first key in
// the most significant bits; DESC inverts its field; NaN is canonical and sorts after +inf), then
// sorted 8 bits at a time together with the row id. The first pass builds those bit strings from
// the key columns as it scatters them (on the lowest key byte) and collects which bytes vary;
// later digits that are equal in every row are skipped. Each pass: per-tile LDS histogram -> scan
// of the digit x tile matrix -> stable scatter
// ranked with wave64 ballots and staged through LDS in digit order, so that each digit's run is
// written with consecutive addresses. The scatter also writes the next pass's digit as a one-byte
// plane, so every histogram after the first reads one byte per row instead of a key word. Equal keys keep ascending row ids, so the result is a total
// order.
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "groupby.hpp"
#include "scan.hpp"

namespace ghive {

// Tile shapes of the LSD passes: TH threads x PER elements per thread (option sort.shape). The
// default, 512 x 16, keeps the 8192-element tile of 256 x 32 (each digit's run on write-out is as
// long) with twice the waves per workgroup and half the registers per lane: 0.9e9 rows, 12 passes
// in 119 ms against 144 ms. 1024 x 8 measured the same; 512 x 32 (16384 elements) 140 ms.
constexpr int kSortThreads = 256;
constexpr int kSortPerThread = 32;

struct SortKeySpec {
  int32_t nkeys;
  int32_t type[8];
  int32_t desc[8];
  int32_t shift[8];  // bit offset of the field inside the composite (at most 256 bits)
  int32_t width[8];
  const void *data[8];
};

// The order-preserving bits of key k (its column's type) for R rows at once: the type is chosen
// once and the R loads issue together. INT / LONG: sign bit flipped; FLOAT / DOUBLE: sign-magnitude
// made monotone, -0 as +0, every NaN canonical (above +inf).
template <int R>
__device__ __forceinline__ void field_bits_rows(const SortKeySpec &S, int k, const uint64_t (&r)[R],
                                                uint64_t (&b)[R]) {
  const void *p = S.data[k];
  switch (S.type[k]) {
    case GHIVE_INT: {
      int32_t x[R];
#pragma unroll
      for (int j = 0; j < R; ++j) x[j] = static_cast<const int32_t *>(p)[r[j]];
#pragma unroll
      for (int j = 0; j < R; ++j) b[j] = (uint32_t)x[j] ^ 0x80000000u;
      break;
    }
    case GHIVE_LONG: {
      int64_t x[R];
#pragma unroll
      for (int j = 0; j < R; ++j) x[j] = static_cast<const int64_t *>(p)[r[j]];
#pragma unroll
      for (int j = 0; j < R; ++j) b[j] = (uint64_t)x[j] ^ 0x8000000000000000ull;
      break;
    }
    case GHIVE_FLOAT: {
      float x[R];
#pragma unroll
      for (int j = 0; j < R; ++j) x[j] = static_cast<const float *>(p)[r[j]];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint32_t u = __builtin_isnan(x[j]) ? 0x7fc00000u : (x[j] == 0.0f ? 0u : __float_as_uint(x[j]));
        b[j] = (u >> 31) ? (uint32_t)~u : (u | 0x80000000u);
      }
      break;
    }
    default: {
      double x[R];
#pragma unroll
      for (int j = 0; j < R; ++j) x[j] = static_cast<const double *>(p)[r[j]];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint64_t u = __builtin_isnan(x[j]) ? 0x7ff8000000000000ull
                                                 : (x[j] == 0.0 ? 0ull : (uint64_t)__double_as_longlong(x[j]));
        b[j] = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
      }
      break;
    }
  }
}

// The composite keys of R rows as KW words each (w[j][0] least significant), key by key with each
// key's R loads in flight together; embed_id: the row id fills the low 32 bits of word 0.
template <int KW, int R>
__device__ __forceinline__ void sort_words_rows(const SortKeySpec &S, const uint64_t (&r)[R], int embed_id,
                                                uint64_t (&w)[R][KW]) {
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int q = 0; q < KW; ++q) w[j][q] = 0;
  for (int k = 0; k < S.nkeys; ++k) {
    uint64_t b[R];
    field_bits_rows<R>(S, k, r, b);
    const uint64_t mask = S.width[k] < 64 ? (1ull << S.width[k]) - 1 : ~0ull, flip = S.desc[k] ? ~0ull : 0ull;
    const int sh = S.shift[k], word = sh / 64, off = sh % 64;
    const bool spill = off + S.width[k] > 64;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const uint64_t v = (b[j] ^ flip) & mask;
#pragma unroll
      for (int q = 0; q < KW; ++q) {
        if (q == word) w[j][q] |= v << off;
        if (q == word + 1 && spill) w[j][q] |= v >> (64 - off);
      }
    }
  }
  if (embed_id)
#pragma unroll
    for (int j = 0; j < R; ++j) w[j][0] |= (uint32_t)r[j];
}

// hist[digit * nblocks + block] (scanned afterwards into global offsets) and the same counts
// block-major, counts[block * 256 + digit] (a tile's local offsets in the scatter).
template <int TH, int PER>
__global__ __launch_bounds__(TH) void k_radix_hist(const uint64_t *keys, uint64_t n, int word, int shift,
                                                   uint32_t *hist, uint32_t *counts, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  for (uint32_t d = threadIdx.x; d < 256; d += TH) h[d] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * (uint64_t)(TH * PER);
  const uint64_t *kw = keys + (uint64_t)word * n;
  for (int j = 0; j < PER; ++j) {
    uint64_t e = t0 + (uint64_t)j * TH + threadIdx.x;
    if (e < n) atomicAdd(&h[(kw[e] >> shift) & 255u], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < 256; d += TH) {
    hist[(uint64_t)d * nblocks + blockIdx.x] = h[d];
    counts[(uint64_t)blockIdx.x * 256 + d] = h[d];
  }
}

// The same counts for the first pass of a sort, whose digit is the lowest byte of the last key
// field (column S.data[k]), read from that column: no composite keys exist yet. Every load of the
// tile is issued before its LDS counts.
template <int TH, int PER>
__global__ __launch_bounds__(TH) void k_radix_hist_key(SortKeySpec S, int k, uint64_t n, uint32_t *hist,
                                                       uint32_t *counts, uint32_t nblocks) {
  __shared__ uint32_t h[256];
  for (uint32_t d = threadIdx.x; d < 256; d += TH) h[d] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * (uint64_t)(TH * PER);
  uint64_t r[PER], b[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint64_t e = t0 + (uint64_t)j * TH + threadIdx.x;
    r[j] = e < n ? e : n - 1;
  }
  field_bits_rows<PER>(S, k, r, b);
  const uint32_t flip = S.desc[k] ? 255u : 0u;
#pragma unroll
  for (int j = 0; j < PER; ++j)
    if (t0 + (uint64_t)j * TH + threadIdx.x < n) atomicAdd(&h[((uint32_t)b[j] ^ flip) & 255u], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < 256; d += TH) {
    hist[(uint64_t)d * nblocks + blockIdx.x] = h[d];
    counts[(uint64_t)blockIdx.x * 256 + d] = h[d];
  }
}

// The same counts from a digit plane: one byte per row, the digit this pass sorts on, written by
// the previous pass's scatter (8x fewer bytes than the key word). The plane is padded to whole
// tiles; bytes past n are not counted.
template <int TH, int PER>
__global__ __launch_bounds__(TH) void k_radix_hist_plane(const uint8_t *plane, uint64_t n, uint32_t *hist,
                                                         uint32_t *counts, uint32_t nblocks) {
  constexpr uint32_t kTile = (uint32_t)TH * PER;
  __shared__ uint32_t h[256];
  for (uint32_t d = threadIdx.x; d < 256; d += TH) h[d] = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
  for (uint32_t c = threadIdx.x; c < kTile / 16; c += TH) {
    const uint4 v = reinterpret_cast<const uint4 *>(plane + t0)[c];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    const uint64_t e0 = t0 + 16ull * c;
    const uint32_t m = e0 >= n ? 0u : (n - e0 >= 16 ? 16u : (uint32_t)(n - e0));  // rows of this lane
    // equal neighbours are counted together (after the lower digits' passes equal keys are adjacent,
    // and one LDS atomic per run replaces one per row); a wave whose 1024 rows all carry one digit
    // adds them with a single atomic
    const uint32_t d0 = w[0] & 255u;
    const bool same = m == 16 && w[0] == d0 * 0x01010101u && w[1] == w[0] && w[2] == w[0] && w[3] == w[0];
    const uint64_t all = ballot(true);
    const uint32_t f = __builtin_amdgcn_readfirstlane(d0);
    if (ballot(same && d0 == f) == all) {  // uniform across the wave
      if (lane_id() == (uint32_t)__builtin_ctzll(all)) atomicAdd(&h[f], 16u * (uint32_t)__popcll(all));
      continue;
    }
    uint32_t cur = d0, run = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const uint32_t d = (w[b >> 2] >> (8 * (b & 3))) & 255u;
      if ((uint32_t)b < m) {
        if (d != cur) {
          atomicAdd(&h[cur], run);
          cur = d;
          run = 0;
        }
        ++run;
      }
    }
    if (run) atomicAdd(&h[cur], run);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < 256; d += TH) {
    hist[(uint64_t)d * nblocks + blockIdx.x] = h[d];
    counts[(uint64_t)blockIdx.x * 256 + d] = h[d];
  }
}

// One stable pass over a tile of TH x PER elements. (1) Every element's position within the tile
// in digit order: wave64 ballots give the rank among equal digits of a wave, LDS counters carry it
// across waves and iterations. (2) The tile is rewritten through an LDS buffer in digit order, so
// each digit's run leaves as consecutive addresses (coalesced stores) at its global offset.
// DROP0: the last pass over word 0, whose low 32 bits hold the row id: the later passes need only
// the id, so word 0 leaves as a uint32 id array (idx_out) and words 1.. as a (KW - 1)-word array
// (KW 1: the sort's last pass, which writes the permutation itself).
// plane_out (optional): the next pass's digit, byte (next_shift / 8) of input word next_word, is
// written beside every row as a one-byte plane for that pass's histogram.
// COLS: the sort's first pass. It builds each row's composite key from the key columns (S, embed)
// instead of reading key words, its row ids are the row numbers, and every workgroup leaves the OR
// and AND of each key word over its tile in orand_part[block * 2 KW ..] (which later digits vary).
// PACK: a row's two words are staged as one 16-byte LDS record (one staging round, not one per word)
// where that fits and was measured faster (option sort.pack): one key word and its id, and the pass
// that drops word 0 (two key words, the id inside word 0). Two key words that both leave as key
// words (10.3 ms against 9.8 ms per pass over 0.9e9 rows) and the other shapes stage one 8-byte
// word at a time.
// MERGE (one key word with the ids beside it, the pass on the last digit of the word's low half):
// the later digits all lie in the high half, so each row leaves as one word, its high half above its
// id, and the passes after it move 8 bytes a row instead of 12 (option sort.merge).
// OCC: workgroups per CU the kernel is compiled for (2: at most 128 VGPRs; the 8-byte staging of a
// 512 x 16 tile holds under 80 KiB of LDS, so two workgroups share a CU and one's loads and stores
// overlap the other's ranking and barriers; option sort.occupancy).
// (The launch bound's second figure is waves per SIMD: a workgroup of TH threads puts TH / 256 on each.)
template <int KW, bool IDS, bool DROP0, int TH, int PER, bool COLS = false, bool PACK = true, bool MERGE = false,
          int OCC = 1>
__global__ __launch_bounds__(TH, TH / 256 * OCC) void k_radix_scatter(const uint64_t *keys_in, const uint32_t *idx_in,
                                                      uint64_t *keys_out, uint32_t *idx_out, uint64_t n,
                                                      int word, int shift, const uint32_t *offs,
                                                      const uint32_t *counts, uint32_t nblocks,
                                                      uint8_t *plane_out, int next_word, int next_shift,
                                                      SortKeySpec S = SortKeySpec{}, int embed = 0,
                                                      unsigned long long *orand_part = nullptr) {
  constexpr uint32_t kTile = (uint32_t)TH * PER, kWaves = TH / kWave;
  // 16-byte records when the tile's records fit LDS that way and the workgroup holds a CU's LDS
  // anyway (512 threads or more: one workgroup per CU either way; at 256 threads the 8-byte staging
  // keeps two)
  static_assert(!MERGE || (KW == 1 && IDS && !DROP0 && !COLS), "MERGE: one key word with the ids beside it");
  constexpr bool kPacked = !MERGE && PACK && ((KW == 1 && IDS) || (KW == 2 && DROP0)) && kTile <= 8192 && TH >= 512;
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) uint64_t stage[kPacked ? 2 * kTile : kTile];
  __shared__ uint8_t sdigit[kTile];
  __shared__ uint32_t lbase[256], goff[256];
  __shared__ uint16_t wcnt[kWaves][256];  // per-wave digit counts, then their prefix over waves (< 2^16)
  __shared__ uint32_t scan_tmp[kWaves];
  constexpr uint32_t kSeg = PER * kWave;  // each wave owns one contiguous segment of the tile
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  uint64_t kv[PER][KW];
  // lpos: a row's rank among its wave's equal digits, with the digit in bits 16..23, then its
  // position in the tile (one register a row: registers decide whether two workgroups share a CU)
  uint32_t id[PER], lpos[PER];
  const uint32_t tile = blockIdx.x;
  const uint64_t t0 = (uint64_t)tile * kTile;
  const uint32_t tn = (uint32_t)(n - t0 < kTile ? n - t0 : kTile);
  uint32_t total;
  // threads past the 256 digits add zeros at the end of the scan
  const uint32_t lb = block_exclusive_scan_256(tid < 256 ? counts[(uint64_t)tile * 256 + tid] : 0u, scan_tmp, total);
  if (tid < 256) {
    lbase[tid] = lb;
    goff[tid] = offs[(uint64_t)tid * nblocks + tile];
  }
#pragma unroll
  for (int q = 0; q < 256 / kWave; ++q) wcnt[wave][lane + q * kWave] = 0;
  // all loads of the tile up front, element wave * kSeg + j * 64 + lane (coalesced per j); past the
  // end a lane re-reads the last element (ignored)
  if constexpr (COLS) {
    uint64_t rows[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t el = wave * kSeg + (uint32_t)j * kWave + lane;
      rows[j] = t0 + (el < tn ? el : tn - 1);
      id[j] = (uint32_t)rows[j];
    }
    sort_words_rows<KW, PER>(S, rows, embed, kv);
  } else {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t el = wave * kSeg + (uint32_t)j * kWave + lane;
      const uint64_t e = t0 + (el < tn ? el : tn - 1);
#pragma unroll
      for (int q = 0; q < KW; ++q) kv[j][q] = keys_in[(uint64_t)q * n + e];
      id[j] = IDS ? idx_in[e] : 0;
    }
  }
  if constexpr (COLS) {  // OR / AND of every key word over the tile's rows
    __shared__ unsigned long long ored[kWaves][2 * KW];
    unsigned long long vor[KW], vand[KW];
#pragma unroll
    for (int q = 0; q < KW; ++q) {
      vor[q] = 0ull;
      vand[q] = ~0ull;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (wave * kSeg + (uint32_t)j * kWave + lane < tn)
#pragma unroll
        for (int q = 0; q < KW; ++q) {
          vor[q] |= kv[j][q];
          vand[q] &= kv[j][q];
        }
#pragma unroll
    for (int q = 0; q < KW; ++q) {
#pragma unroll
      for (int d = kWave / 2; d > 0; d >>= 1) {
        vor[q] |= __shfl_xor(vor[q], d, kWave);
        vand[q] &= __shfl_xor(vand[q], d, kWave);
      }
      if (lane == 0) {
        ored[wave][2 * q] = vor[q];
        ored[wave][2 * q + 1] = vand[q];
      }
    }
    __syncthreads();
    if (tid < 2 * KW) {
      unsigned long long r = ored[0][tid];
      for (uint32_t w = 1; w < kWaves; ++w) r = (tid & 1) ? (r & ored[w][tid]) : (r | ored[w][tid]);
      orand_part[(uint64_t)tile * 2 * KW + tid] = r;
    }
  }
  __syncthreads();
  // Rank every element among the equal digits of its wave's segment, in element order: ballots give
  // the rank inside one load of 64, the wave's own LDS row carries the count across its loads (one
  // wave reads and then updates its row: no barrier between loads).
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const bool valid = wave * kSeg + (uint32_t)j * kWave + lane < tn;
    uint32_t digit = 0;
#pragma unroll
    for (int q = 0; q < KW; ++q)
      if (q == word) digit = (uint32_t)(kv[j][q] >> shift) & 255u;
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = ballot((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t before = wcnt[wave][digit];
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wcnt[wave][digit] = (uint16_t)(before + (uint32_t)__popcll(peers));
    __builtin_amdgcn_wave_barrier();
    lpos[j] = digit << 16 | (before + rank);  // before + rank < kSeg <= 2^16
  }
  __syncthreads();
  if (tid < 256) {  // exclusive prefix of each digit's counts over the waves (segments are in wave order)
    uint32_t acc = 0;
#pragma unroll
    for (int w = 0; w < (int)kWaves; ++w) {
      const uint32_t c = wcnt[w][tid];
      wcnt[w][tid] = (uint16_t)acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    if (wave * kSeg + (uint32_t)j * kWave + lane < tn) {
      const uint32_t digit = lpos[j] >> 16;
      const uint32_t pos = lbase[digit] + wcnt[wave][digit] + (lpos[j] & 0xFFFFu);
      lpos[j] = pos;
      sdigit[pos] = (uint8_t)digit;
    }
  }
  if constexpr (kPacked) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (wave * kSeg + (uint32_t)j * kWave + lane < tn) {
        u64x2 rec;
        if constexpr (IDS) rec = u64x2{kv[j][0], (unsigned long long)id[j]};
        else rec = u64x2{kv[j][0], kv[j][KW - 1]};
        *reinterpret_cast<u64x2 *>(stage + 2 * lpos[j]) = rec;
      }
    __syncthreads();
    const bool plane = plane_out && next_word >= 0 && next_word < KW;
    for (uint32_t p = tid; p < tn; p += TH) {
      const uint32_t d = sdigit[p];
      const uint64_t g = (uint64_t)goff[d] + (p - lbase[d]);
      const u64x2 v = *reinterpret_cast<const u64x2 *>(stage + 2 * p);
      if constexpr (DROP0) {  // word 0 (the id in its low half) -> ids, word 1 -> the one-word keys
        idx_out[g] = (uint32_t)v.x;
        keys_out[g] = v.y;
      } else if constexpr (IDS) {
        keys_out[g] = v.x;
        idx_out[g] = (uint32_t)v.y;
      } else {
        keys_out[g] = v.x;
        keys_out[n + g] = v.y;
      }
      if (plane) plane_out[g] = (uint8_t)((next_word == 0 ? v.x : v.y) >> next_shift);
    }
    return;
  }
  // write-out, one word at a time through the stage buffer (then the ids, when carried; MERGE: one
  // round of the merged words)
#pragma unroll
  for (int q = 0; q < (MERGE ? 1 : IDS ? KW + 1 : KW); ++q) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (wave * kSeg + (uint32_t)j * kWave + lane < tn)
        stage[lpos[j]] = MERGE ? (kv[j][0] & 0xFFFFFFFF00000000ull) | id[j] : q < KW ? kv[j][q] : (uint64_t)id[j];
    __syncthreads();
    for (uint32_t p = tid; p < tn; p += TH) {
      const uint32_t d = sdigit[p];
      const uint64_t g = (uint64_t)goff[d] + (p - lbase[d]);
      if (DROP0 && q == 0)
        idx_out[g] = (uint32_t)stage[p];
      else if (MERGE || q < KW)
        keys_out[(uint64_t)(DROP0 ? q - 1 : q) * n + g] = stage[p];
      else
        idx_out[g] = (uint32_t)stage[p];
      if (plane_out && q == next_word) plane_out[g] = (uint8_t)(stage[p] >> next_shift);
    }
  }
  return;
}

// out[2q] = OR, out[2q + 1] = AND of key word q over the first pass's per-tile partials (one
// workgroup; the loading lanes are a multiple of 2 KW, so each always meets the same word of the
// contiguous partials array; eight loads in flight).
constexpr int kOrandThreads = 1024;
__global__ __launch_bounds__(kOrandThreads) void k_orand_reduce(const unsigned long long *part, uint32_t nblocks,
                                                                int words2, unsigned long long *out) {
  __shared__ unsigned long long red[kOrandThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t nt = kOrandThreads - kOrandThreads % (uint32_t)words2;  // lanes that load: a multiple of 2 KW
  const bool is_and = (t % (uint32_t)words2) & 1u;
  unsigned long long acc = is_and ? ~0ull : 0ull;
  const uint64_t total = (uint64_t)nblocks * (uint32_t)words2;
  constexpr int U = 8;
  for (uint64_t i0 = 0; i0 < total; i0 += (uint64_t)nt * U) {
    unsigned long long x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = i0 + (uint64_t)u * nt + t;
      x[u] = t < nt && i < total ? part[i] : (is_and ? ~0ull : 0ull);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc = is_and ? (acc & x[u]) : (acc | x[u]);
  }
  red[t] = acc;
  __syncthreads();
  if ((int)t < words2) {
    unsigned long long r = red[t];
    for (uint32_t j = t + (uint32_t)words2; j < kOrandThreads; j += (uint32_t)words2) r = is_and ? (r & red[j]) : (r | red[j]);
    out[t] = r;
  }
}

// One scatter pass over key words (not the first pass): the shapes that stage 16-byte LDS records
// take them when sort.pack is on.
template <int KW, bool IDS, bool DROP0, int TH, int PER, bool MERGE = false>
static int scatter_pass(ghive_ctx *ctx, const uint64_t *kin, const uint32_t *iin, uint64_t *kout, uint32_t *iout,
                        uint64_t n, int word, int shift, const uint32_t *offs, const uint32_t *counts,
                        uint32_t nblocks, uint8_t *pout, int nw, int ns) {
  constexpr bool kPackable = !MERGE && ((KW == 1 && IDS) || (KW == 2 && DROP0)) && TH >= 512 && TH * PER <= 8192;
  if constexpr (kPackable) {
    if (ctx->tune.sort_pack) {
      k_radix_scatter<KW, IDS, DROP0, TH, PER, false, true><<<nblocks, TH, 0, ctx->stream>>>(
          kin, iin, kout, iout, n, word, shift, offs, counts, nblocks, pout, nw, ns);
      GH_LAUNCH(ctx);
      return GHIVE_OK;
    }
  }
  // 8-byte staging of a 512 x 16 tile: two workgroups per CU fit its LDS
  constexpr bool kTwo = TH == 512 && PER == 16 && KW <= 2;
  if constexpr (kTwo) {
    if (ctx->tune.sort_occupancy == 2) {
      k_radix_scatter<KW, IDS, DROP0, TH, PER, false, false, MERGE, 2><<<nblocks, TH, 0, ctx->stream>>>(
          kin, iin, kout, iout, n, word, shift, offs, counts, nblocks, pout, nw, ns);
      GH_LAUNCH(ctx);
      return GHIVE_OK;
    }
  }
  k_radix_scatter<KW, IDS, DROP0, TH, PER, false, false, MERGE><<<nblocks, TH, 0, ctx->stream>>>(
      kin, iin, kout, iout, n, word, shift, offs, counts, nblocks, pout, nw, ns);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

// The sort's first pass: a stable pass on `digit` (the lowest byte of the last key field), reading
// the key columns themselves (no key-building pass before it): keys_out receives the KW-word
// composite keys in that digit's order, ids_out the row ids when they are not embedded, plane_out
// (optional) the plane of digit `next` for the next pass, and oa[2 KW] the OR / AND of every word.
template <int KW, int TH, int PER>
static int sort_first_pass(ghive_ctx *ctx, const SortKeySpec &S, bool embed, uint64_t n, int digit, uint64_t *keys_out,
                           uint32_t *ids_out, uint8_t *plane_out, int next, unsigned long long *oa) {
  const uint32_t nblocks = ceil_div(n, (uint64_t)TH * PER);
  DevBuf hist, counts, part;
  GH_TRY(hist.alloc(ctx, sizeof(uint32_t) * 256ull * nblocks));
  GH_TRY(counts.alloc(ctx, sizeof(uint32_t) * 256ull * nblocks));
  GH_TRY(part.alloc(ctx, sizeof(unsigned long long) * 2ull * KW * nblocks));
  k_radix_hist_key<TH, PER><<<nblocks, TH, 0, ctx->stream>>>(S, S.nkeys - 1, n, hist.as<uint32_t>(),
                                                             counts.as<uint32_t>(), nblocks);
  GH_LAUNCH(ctx);
  GH_TRY(exclusive_scan_u32(ctx, hist.as<uint32_t>(), hist.as<uint32_t>(), 256ull * nblocks, nullptr));
  const int word = digit / 8, shift = 8 * (digit % 8);
  const int nw = next >= 0 ? next / 8 : -1, ns = next >= 0 ? 8 * (next % 8) : 0;
  if (embed)
    k_radix_scatter<KW, false, false, TH, PER, true, false><<<nblocks, TH, 0, ctx->stream>>>(
        nullptr, nullptr, keys_out, nullptr, n, word, shift, hist.as<uint32_t>(), counts.as<uint32_t>(), nblocks,
        plane_out, nw, ns, S, 1, part.as<unsigned long long>());
  else
    k_radix_scatter<KW, true, false, TH, PER, true, false><<<nblocks, TH, 0, ctx->stream>>>(
        nullptr, nullptr, keys_out, ids_out, n, word, shift, hist.as<uint32_t>(), counts.as<uint32_t>(), nblocks,
        plane_out, nw, ns, S, 0, part.as<unsigned long long>());
  GH_LAUNCH(ctx);
  k_orand_reduce<<<1, kOrandThreads, 0, ctx->stream>>>(part.as<unsigned long long>(), nblocks, 2 * KW, oa);
  GH_LAUNCH(ctx);
  int in_b = 0;
  for (int k = 0; k < S.nkeys; ++k) in_b += S.width[k] / 8;
  ctx->sort_info.passes += 1;
  ctx->sort_info.scatter_bytes += n * ((uint64_t)in_b + 8ull * KW + (embed ? 0 : 4) + (plane_out ? 1 : 0));
  return GHIVE_OK;
}

template <int KW>
static int sort_first(ghive_ctx *ctx, const SortKeySpec &S, bool embed, uint64_t n, int digit, uint64_t *keys_out,
                      uint32_t *ids_out, uint8_t *plane_out, int next, unsigned long long *oa) {
  switch (ctx->tune.sort_shape) {
    case 0: return sort_first_pass<KW, kSortThreads, kSortPerThread>(ctx, S, embed, n, digit, keys_out, ids_out, plane_out, next, oa);
    case 2: return sort_first_pass<KW, 512, 32>(ctx, S, embed, n, digit, keys_out, ids_out, plane_out, next, oa);
    case 3: return sort_first_pass<KW, 1024, 8>(ctx, S, embed, n, digit, keys_out, ids_out, plane_out, next, oa);
    default: return sort_first_pass<KW, 512, 16>(ctx, S, embed, n, digit, keys_out, ids_out, plane_out, next, oa);
  }
}

// ids: carry the uint32 id arrays i0/i1 (false: the ids live inside the keys; i0/i1 unused).
// drop0_last: the last pass writes word 0 as the uint32 id array (k_radix_scatter DROP0; needs
// !ids): its outputs are then *result (ids) and *keys_result (KW - 1 words).
// planes (optional): two byte buffers of at least n rounded up to 8192 x 2 rows. *plane_cur: the
// one that holds the first digit's plane (-1: none, the first histogram reads the keys); every
// pass writes the next digit's plane into the other, and on return *plane_cur names the plane of
// next_after (a digit of the same KW-word layout that a later call sorts on; -1: none).
// merge_last (KW 1 with ids, the digits all in the low half): the last pass writes each word's high
// half above its id (k_radix_scatter MERGE); the result is then *keys_result alone.
template <int KW, int TH, int PER>
static int radix_passes_shape(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                              const std::vector<int> &digits, uint32_t **result, uint64_t **keys_result, bool ids,
                              bool drop0_last, uint8_t *const *planes, int *plane_cur, int next_after,
                              bool merge_last) {
  const uint32_t nblocks = ceil_div(n, (uint64_t)TH * PER);
  DevBuf hist, counts;
  GH_TRY(hist.alloc(ctx, sizeof(uint32_t) * 256ull * nblocks));
  GH_TRY(counts.alloc(ctx, sizeof(uint32_t) * 256ull * nblocks));
  uint64_t *kin = k0, *kout = k1;
  uint32_t *iin = i0, *iout = i1;
  int cur = planes && plane_cur ? *plane_cur : -1;
  for (size_t di = 0; di < digits.size(); ++di) {
    const int d = digits[di];
    const int word = d / 8, shift = 8 * (d % 8);
    if (cur >= 0)
      k_radix_hist_plane<TH, PER><<<nblocks, TH, 0, ctx->stream>>>(planes[cur], n, hist.as<uint32_t>(),
                                                                   counts.as<uint32_t>(), nblocks);
    else
      k_radix_hist<TH, PER><<<nblocks, TH, 0, ctx->stream>>>(kin, n, word, shift, hist.as<uint32_t>(),
                                                             counts.as<uint32_t>(), nblocks);
    GH_LAUNCH(ctx);
    GH_TRY(exclusive_scan_u32(ctx, hist.as<uint32_t>(), hist.as<uint32_t>(), 256ull * nblocks, nullptr));
    const bool drop = drop0_last && !ids && di + 1 == digits.size();
    const int next = di + 1 < digits.size() ? digits[di + 1] : next_after;
    uint8_t *pout = planes && next >= 0 ? planes[cur == 0 ? 1 : 0] : nullptr;
    const int nw = next >= 0 ? next / 8 : -1, ns = next >= 0 ? 8 * (next % 8) : 0;
    const uint32_t *ho = hist.as<uint32_t>(), *co = counts.as<uint32_t>();
    const bool merge = KW == 1 && ids && merge_last && di + 1 == digits.size();
    if constexpr (KW == 1) {
      if (merge)
        GH_TRY((scatter_pass<1, true, false, TH, PER, true>(ctx, kin, iin, kout, iout, n, word, shift, ho, co, nblocks,
                                                            pout, nw, ns)));
    }
    if (drop)
      GH_TRY((scatter_pass<KW, false, true, TH, PER>(ctx, kin, iin, kout, iout, n, word, shift, ho, co, nblocks, pout,
                                                      nw, ns)));
    else if (ids && !merge)
      GH_TRY((scatter_pass<KW, true, false, TH, PER>(ctx, kin, iin, kout, iout, n, word, shift, ho, co, nblocks, pout,
                                                      nw, ns)));
    else if (!ids)
      GH_TRY((scatter_pass<KW, false, false, TH, PER>(ctx, kin, iin, kout, iout, n, word, shift, ho, co, nblocks,
                                                       pout, nw, ns)));
    cur = pout ? (cur == 0 ? 1 : 0) : -1;
    const uint64_t in_b = 8ull * KW + (ids ? 4 : 0);
    const uint64_t out_b = (drop ? 8ull * (KW - 1) + 4 : merge ? 8ull : in_b) + (pout ? 1 : 0);
    ctx->sort_info.passes += 1;
    ctx->sort_info.scatter_bytes += n * (in_b + out_b);
    std::swap(kin, kout);
    std::swap(iin, iout);
  }
  *result = iin;
  if (keys_result) *keys_result = kin;
  if (plane_cur) *plane_cur = cur;
  return GHIVE_OK;
}

template <int KW>
static int radix_passes(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                        const std::vector<int> &digits, uint32_t **result, uint64_t **keys_result = nullptr,
                        bool ids = true, bool drop0_last = false, uint8_t *const *planes = nullptr,
                        int *plane_cur = nullptr, int next_after = -1, bool merge_last = false) {
#define GH_PASSES(TH, PER) \
  radix_passes_shape<KW, TH, PER>(ctx, n, k0, k1, i0, i1, digits, result, keys_result, ids, drop0_last, planes, \
                                  plane_cur, next_after, merge_last)
  switch (ctx->tune.sort_shape) {
    case 0: return GH_PASSES(kSortThreads, kSortPerThread);
    case 2: return GH_PASSES(512, 32);
    case 3: return GH_PASSES(1024, 8);
    default: return GH_PASSES(512, 16);
  }
#undef GH_PASSES
}

// Stable LSD passes over the given 8-bit digits of one uint64 word per row, carrying a uint32 id
// (the group-by's hash partitioning). The ping-pong buffers hold n entries each; the sorted
// words and ids end in whichever buffer the last pass wrote.
int radix_sort_u64_ids(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                       const std::vector<int> &digits, uint64_t **keys_out, uint32_t **ids_out, bool ids) {
  return radix_passes<1>(ctx, n, k0, k1, i0, i1, digits, ids_out, keys_out, ids);
}

int radix_sort_u64x2_ids(ghive_ctx *ctx, uint64_t n, uint64_t *k0, uint64_t *k1, uint32_t *i0, uint32_t *i1,
                         const std::vector<int> &digits, uint64_t **keys_out, uint32_t **ids_out, bool ids) {
  return radix_passes<2>(ctx, n, k0, k1, i0, i1, digits, ids_out, keys_out, ids);
}

// The two id buffers in the order that makes the last of p passes (p >= 1) write its ids into
// `target`: a pass reads the first and writes the second, and they swap after every pass.
static void ids_last_into(uint32_t *target, uint32_t *spare, size_t p, uint32_t **i0, uint32_t **i1) {
  const bool swapped = (p - 1) % 2 == 1;
  *i0 = swapped ? target : spare;
  *i1 = swapped ? spare : target;
}

// The ORDER BY's passes over one key word with the row ids beside it (keys K0 / ids I0, ping-pong
// partners K1 / I1; out is I0 or I1). When digits remain in both halves of the word (and option
// sort.merge is on), the pass on the low half's last digit writes each row as one word, its high
// half above its id, and the passes over the high half then move 8 bytes a row instead of 12, the
// last of them writing the ids alone into out. *ids: where the permutation ends (out after a merge).
static int passes_one_word_ids(ghive_ctx *ctx, uint64_t n, uint64_t *K0, uint64_t *K1, uint32_t *I0, uint32_t *I1,
                               uint32_t *out, const std::vector<int> &digits, uint8_t *const *planes, int *pcur,
                               uint32_t **ids) {
  std::vector<int> lo, hi;
  for (int d : digits) (d < 4 ? lo : hi).push_back(d);
  uint64_t *kres = nullptr;
  if (!ctx->tune.sort_merge || lo.empty() || hi.empty())
    return radix_passes<1>(ctx, n, K0, K1, I0, I1, digits, ids, &kres, true, false, planes, pcur);
  uint64_t *words = nullptr;
  uint32_t *unused = nullptr;
  GH_TRY(radix_passes<1>(ctx, n, K0, K1, I0, I1, lo, &unused, &words, true, false, planes, pcur, hi.front(),
                         /*merge_last=*/true));
  uint32_t *j0 = nullptr, *j1 = nullptr;  // the ids now ride in the words: both buffers are free
  ids_last_into(out, out == I0 ? I1 : I0, hi.size(), &j0, &j1);
  return radix_passes<1>(ctx, n, words, words == K0 ? K1 : K0, j0, j1, hi, ids, &kres, false, /*drop0_last=*/true,
                         planes, pcur);
}

// ---- STRING keys: dense ranks in Hive's string order -------------------------------------------
// A STRING key sorts by the rank of its value among the column's distinct values. The key
// dictionary of the group-by names each distinct value once (a representative row); the distinct
// values are sorted by stable LSD passes over one 64-bit word each: first the length, then the
// 8-byte big-endian chunks from the last to the first (zero padded), which is the order of
// unsigned bytes, then length (StringExpr.compare). NULL takes all-ones words: it ranks last.
constexpr uint64_t kDictEmpty = ~0ull;

// The used dictionary slots, compacted: dslot[i] = slot, drep[i] = representative row.
__global__ void k_dict_list(const uint64_t *dict, uint64_t slots, unsigned long long *cnt, uint32_t *dslot,
                            uint32_t *drep) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t iters = (slots + stride - 1) / stride;  // the same trip count in every lane
  for (uint64_t it = 0; it < iters; ++it) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + it * stride;
    const uint64_t w = s < slots ? dict[s] : kDictEmpty;
    const bool used = w != kDictEmpty;
    const uint64_t idx = wave_append(cnt, used);
    if (used) {
      dslot[idx] = (uint32_t)s;
      drep[idx] = (uint32_t)w;
    }
  }
}

__global__ void k_str_maxlen(const int32_t *aux, const uint32_t *drep, uint64_t d, unsigned int *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned int m = 0;
  for (; i < d; i += stride) {
    const uint64_t r = drep[i];
    const int32_t s = aux[2 * r], e = aux[2 * r + 1];
    if (s >= 0 && e > s && (unsigned)(e - s) > m) m = (unsigned)(e - s);
  }
  if (m) atomicMax(out, m);
}

// key[i] = word `word` of the i-th value in the current order (word -1: the length).
__global__ void k_str_word(const uint8_t *blob, const int32_t *aux, const uint32_t *drep, const uint32_t *ids,
                           uint64_t d, int word, uint64_t *key) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < d; i += stride) {
    const uint64_t r = drep[ids[i]];
    const int32_t s = aux[2 * r], e = aux[2 * r + 1];
    uint64_t w;
    if (s < 0) {
      w = ~0ull;
    } else if (word < 0) {
      w = (uint64_t)(uint32_t)(e - s);
    } else {
      w = 0;
      const int64_t p0 = (int64_t)s + 8 * (int64_t)word;
      for (int b = 0; b < 8; ++b) w = (w << 8) | (p0 + b < e ? blob[p0 + b] : 0u);
    }
    key[i] = w;
  }
}

__global__ void k_rank_of_slot(const uint32_t *dslot, const uint32_t *ids, uint64_t d, uint32_t *rank_of_slot) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < d; i += stride) rank_of_slot[dslot[ids[i]]] = (uint32_t)i;
}

__global__ void k_row_rank(const uint32_t *codes, uint64_t n, const uint32_t *rank_of_slot, int32_t *rank) {
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; r < n; r += stride) rank[r] = (int32_t)rank_of_slot[codes[r]];
}

__global__ void k_low32(const uint64_t *w, uint64_t n, uint32_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (uint32_t)w[i];
}

__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (uint32_t)i;
}

// rank[r] (n rows) = rank of row r's value of STRING column c among its distinct values.
int string_ranks(ghive_ctx *ctx, const ghive_column &c, uint64_t n, int32_t *rank) {
  WideKeySpec K;
  memset(&K, 0, sizeof K);
  K.nkeys = 1;
  K.type[0] = GHIVE_STRING;
  K.data[0] = c.data;
  K.aux[0] = static_cast<const int32_t *>(c.aux);
  DevBuf dict, codes;
  uint64_t slots = 0, d = 0;
  GH_TRY(key_dictionary(ctx, K, n, nullptr, n, n, dict, codes, &slots, &d));
  DevBuf dslot, drep, k0, k1, i0, i1, ros;
  GH_TRY(dslot.alloc(ctx, 4 * d));
  GH_TRY(drep.alloc(ctx, 4 * d));
  GH_TRY(k0.alloc(ctx, 8 * d));
  GH_TRY(k1.alloc(ctx, 8 * d));
  GH_TRY(i0.alloc(ctx, 4 * d));
  GH_TRY(i1.alloc(ctx, 4 * d));
  GH_TRY(ros.alloc(ctx, 4 * slots));
  auto grid_for = [&](uint64_t m) {
    const uint64_t g = (m + 255) / 256, cap = (uint64_t)ctx->cus * 16;
    return (uint32_t)(g > cap ? cap : (g ? g : 1));
  };
  GH_TRY(fill_u64(ctx, ctx->d_counters, 0, 2));
  k_dict_list<<<grid_for(slots), 256, 0, ctx->stream>>>(dict.as<uint64_t>(), slots,
                                                       reinterpret_cast<unsigned long long *>(ctx->d_counters),
                                                       dslot.as<uint32_t>(), drep.as<uint32_t>());
  GH_LAUNCH(ctx);
  unsigned int *d_max = reinterpret_cast<unsigned int *>(ctx->d_counters + 1);
  k_str_maxlen<<<grid_for(d), 256, 0, ctx->stream>>>(K.aux[0], drep.as<uint32_t>(), d, d_max);
  GH_LAUNCH(ctx);
  unsigned int maxlen = 0;
  GH_TRY(d2h(ctx, &maxlen, d_max, sizeof maxlen));
  k_iota_u32<<<grid_for(d), 256, 0, ctx->stream>>>(i0.as<uint32_t>(), d);
  GH_LAUNCH(ctx);
  std::vector<int> digits;
  for (int b = 0; b < 8; ++b) digits.push_back(b);
  uint32_t *ids = i0.as<uint32_t>(), *other = i1.as<uint32_t>();
  const int nwords = (int)((maxlen + 7) / 8);
  for (int word = -1; word < nwords; ++word) {  // least significant first: the length, then chunks last to first
    const int w = word < 0 ? -1 : nwords - 1 - word;
    k_str_word<<<grid_for(d), 256, 0, ctx->stream>>>(static_cast<const uint8_t *>(c.data), K.aux[0],
                                                     drep.as<uint32_t>(), ids, d, w, k0.as<uint64_t>());
    GH_LAUNCH(ctx);
    uint64_t *kout = nullptr;
    uint32_t *iout = nullptr;
    GH_TRY(radix_sort_u64_ids(ctx, d, k0.as<uint64_t>(), k1.as<uint64_t>(), ids, other, digits, &kout, &iout));
    other = iout == ids ? other : ids;
    ids = iout;
  }
  k_rank_of_slot<<<grid_for(d), 256, 0, ctx->stream>>>(dslot.as<uint32_t>(), ids, d, ros.as<uint32_t>());
  GH_LAUNCH(ctx);
  k_row_rank<<<grid_for(n), 256, 0, ctx->stream>>>(codes.as<uint32_t>(), n, ros.as<uint32_t>(), rank);
  GH_LAUNCH(ctx);
  return GHIVE_OK;
}

}  // namespace ghive

using namespace ghive;

extern "C" {

int ghive_sort(ghive_ctx *ctx, const ghive_table *table, const uint32_t *key_cols, const char *orders,
               uint32_t nkeys, uint32_t *out_perm) {
  GH_CHECK_CTX(ctx);
  GH_TRY(check_table(ctx, table, "sort"));
  if (nkeys == 0 || nkeys > 8 || !key_cols)
    return set_error(ctx, GHIVE_ERR_INVALID, "sort: %u key columns (1..8)", nkeys);
  const uint64_t n = table->row_num;
  if (n && !out_perm) return set_error(ctx, GHIVE_ERR_INVALID, "sort: out_perm is NULL");
  SortKeySpec S;
  memset(&S, 0, sizeof S);
  S.nkeys = (int)nkeys;
  int bits = 0;
  DevBuf ranks[8];  // STRING keys sort by the rank of their value
  for (uint32_t k = 0; k < nkeys; ++k) {
    if (key_cols[k] >= table->ncols) return set_error(ctx, GHIVE_ERR_INVALID, "sort: key column %u", key_cols[k]);
    const ghive_column &c = table->columns[key_cols[k]];
    char o = orders ? orders[k] : '+';
    if (o != '+' && o != '-') return set_error(ctx, GHIVE_ERR_INVALID, "sort: order '%c' (use '+' or '-')", o);
    S.desc[k] = o == '-';
    if (c.type == GHIVE_STRING) {
      if (n > (1ull << 31)) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "sort: STRING keys over 2^31 rows");
      GH_TRY(ranks[k].alloc(ctx, 4 * (n ? n : 1)));
      if (n) GH_TRY(string_ranks(ctx, c, n, ranks[k].as<int32_t>()));
      S.type[k] = GHIVE_INT;
      S.width[k] = 32;
      S.data[k] = ranks[k].get();
    } else {
      S.type[k] = c.type;
      S.width[k] = 8 * type_width(c.type);
      S.data[k] = c.data;
    }
    bits += S.width[k];
  }
  if (bits > 256) return set_error(ctx, GHIVE_ERR_UNSUPPORTED, "sort: keys are %d bits wide (max 256)", bits);
  memset(&ctx->sort_info, 0, sizeof ctx->sort_info);
  ctx->sort_info.rows = n;
  if (n == 0) return GHIVE_OK;
  // when the last key word has 32 spare bits the row id rides in them (below every field): the
  // passes then move no separate id array, and equal keys stay in row-id order as before
  const int KW = (bits + 63) / 64;
  const bool embed = KW * 64 - bits >= 32;
  ctx->sort_info.key_words = (uint32_t)KW;
  ctx->sort_info.embedded_ids = embed ? 1u : 0u;
  int top = bits + (embed ? 32 : 0);
  for (uint32_t k = 0; k < nkeys; ++k) {
    top -= S.width[k];
    S.shift[k] = top;
  }
  DevBuf k0, k1, i1, orand;
  GH_TRY(k0.alloc(ctx, 8ull * KW * n));
  GH_TRY(k1.alloc(ctx, 8ull * KW * n));
  GH_TRY(i1.alloc(ctx, 4ull * n));
  GH_TRY(orand.alloc(ctx, 8 * sizeof(uint64_t)));
  uint64_t g = (n + 255) / 256, cap = (uint64_t)ctx->cus * 8;
  uint32_t grid = (uint32_t)(g > cap ? cap : g);
  unsigned long long *oa_d = orand.as<unsigned long long>();
  // every pass after the first reads its histogram from a one-byte digit plane that the pass
  // before it wrote (two planes, alternating)
  DevBuf plane_buf;
  uint8_t *planes[2] = {nullptr, nullptr};
  const uint64_t pad = (n + 16383) / 16384 * 16384;
  GH_TRY(plane_buf.alloc(ctx, 2 * pad));
  planes[0] = plane_buf.as<uint8_t>();
  planes[1] = planes[0] + pad;
  // The first pass sorts on the lowest byte of the last key field, reading the key columns: it
  // builds the composite keys as it scatters them, so no pass writes them first. Whether the later
  // digits vary is known only after it (the OR / AND of every word): a first digit equal in every
  // row costs one pass that keeps the order, never a wrong result. Its plane is the next byte's.
  const int first = S.shift[nkeys - 1] / 8;
  const int guess = first + 1 < 8 * KW ? first + 1 : -1;
  switch (KW) {
    case 1: GH_TRY(sort_first<1>(ctx, S, embed, n, first, k0.as<uint64_t>(), out_perm, planes[0], guess, oa_d)); break;
    case 2: GH_TRY(sort_first<2>(ctx, S, embed, n, first, k0.as<uint64_t>(), out_perm, planes[0], guess, oa_d)); break;
    case 3: GH_TRY(sort_first<3>(ctx, S, embed, n, first, k0.as<uint64_t>(), out_perm, planes[0], guess, oa_d)); break;
    default: GH_TRY(sort_first<4>(ctx, S, embed, n, first, k0.as<uint64_t>(), out_perm, planes[0], guess, oa_d)); break;
  }
  uint64_t oa[8];
  GH_TRY(d2h(ctx, oa, orand.get(), sizeof oa));
  std::vector<int> digits;  // the digits still to sort on, lowest first
  for (int q = 0; q < KW; ++q) {
    uint64_t varying = oa[2 * q] ^ oa[2 * q + 1];
    for (int b = 0; b < 8; ++b)
      if (((varying >> (8 * b)) & 255u) && !(embed && q == 0 && b < 4) && 8 * q + b != first)
        digits.push_back(8 * q + b);
  }
  int pcur = !digits.empty() && digits.front() == guess ? 0 : -1;
  uint32_t *res = out_perm;
  uint64_t *K0 = k0.as<uint64_t>(), *K1 = k1.as<uint64_t>(), *kres = K0;
  uint32_t *I1 = i1.as<uint32_t>();
  // With the row id inside word 0 and more words above it, word 0 is needed only until its last
  // digit: that pass writes the ids out as their own array, and the passes over the higher words
  // move (KW - 1) key words plus a 4-byte id instead of KW words.
  std::vector<int> d0, dhi;
  for (int d : digits) (d < 8 ? d0 : dhi).push_back(d);
  if (embed && KW >= 2 && d0.empty() && !dhi.empty()) {
    // the first pass was word 0's only varying digit: one pass on the next (constant) byte of word
    // 0 keeps the order and leaves the ids as their own array
    d0.push_back(first + 1);
    pcur = -1;
  }
  if (embed && KW >= 2 && !d0.empty() && !dhi.empty()) {
    uint32_t *ids = nullptr;
    uint64_t *keys = nullptr;
    const int first_hi = dhi.front();  // in the KW-word layout of phase A
    switch (KW) {
      case 2: GH_TRY(radix_passes<2>(ctx, n, K0, K1, out_perm, I1, d0, &ids, &keys, false, true, planes, &pcur, first_hi)); break;
      case 3: GH_TRY(radix_passes<3>(ctx, n, K0, K1, out_perm, I1, d0, &ids, &keys, false, true, planes, &pcur, first_hi)); break;
      default: GH_TRY(radix_passes<4>(ctx, n, K0, K1, out_perm, I1, d0, &ids, &keys, false, true, planes, &pcur, first_hi)); break;
    }
    // the higher words now live at `keys` (KW - 1 words) with the ids beside them; the free key
    // buffer and the other id buffer take the next ping-pong
    uint64_t *other = keys == K0 ? K1 : K0;
    uint32_t *iother = ids == out_perm ? I1 : out_perm;
    for (int &d : dhi) d -= 8;
    switch (KW) {
      case 2: GH_TRY(passes_one_word_ids(ctx, n, keys, other, ids, iother, out_perm, dhi, planes, &pcur, &res)); break;
      case 3: GH_TRY(radix_passes<2>(ctx, n, keys, other, ids, iother, dhi, &res, &kres, true, false, planes, &pcur)); break;
      default: GH_TRY(radix_passes<3>(ctx, n, keys, other, ids, iother, dhi, &res, &kres, true, false, planes, &pcur)); break;
    }
    if (res != out_perm) GH_HIP(ctx, hipMemcpyAsync(out_perm, res, 4ull * n, hipMemcpyDeviceToDevice, ctx->stream));
    GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GHIVE_OK;
  }
  // ids inside the keys: the last pass writes them alone, as the permutation, into out_perm
  const bool drop_last = embed && !digits.empty();
  uint32_t *j0 = out_perm, *j1 = I1;
  if (drop_last) ids_last_into(out_perm, I1, digits.size(), &j0, &j1);
  switch (KW) {
    case 1:
      if (embed) GH_TRY(radix_passes<1>(ctx, n, K0, K1, j0, j1, digits, &res, &kres, false, drop_last, planes, &pcur));
      else GH_TRY(passes_one_word_ids(ctx, n, K0, K1, out_perm, I1, out_perm, digits, planes, &pcur, &res));
      break;
    case 2: GH_TRY(radix_passes<2>(ctx, n, K0, K1, j0, j1, digits, &res, &kres, !embed, drop_last, planes, &pcur)); break;
    case 3: GH_TRY(radix_passes<3>(ctx, n, K0, K1, j0, j1, digits, &res, &kres, !embed, drop_last, planes, &pcur)); break;
    default: GH_TRY(radix_passes<4>(ctx, n, K0, K1, j0, j1, digits, &res, &kres, !embed, drop_last, planes, &pcur)); break;
  }
  if (embed && !drop_last) {  // no pass after the first: the permutation is the low 32 bits of word 0
    k_low32<<<grid, 256, 0, ctx->stream>>>(kres, n, out_perm);
    GH_LAUNCH(ctx);
  } else if (res != out_perm) {
    GH_HIP(ctx, hipMemcpyAsync(out_perm, res, 4ull * n, hipMemcpyDeviceToDevice, ctx->stream));
  }
  GH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GHIVE_OK;
}

int ghive_last_sort_info(ghive_ctx *ctx, ghive_sort_info *out) {
  if (!ctx || !out) return GHIVE_ERR_INVALID;
  *out = ctx->sort_info;
  return GHIVE_OK;
}

}  // extern "C"
