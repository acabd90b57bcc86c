/*
 * ghive_oracle.c -- CPU ORACLE (test infrastructure, see ghive_oracle.h). Plain C99 + pthreads.
 *
 * Not part of the product: libghive_mi355x.so neither links nor loads this file. It is used by
 * tests/ as the parity checker and by bench.py as the timed CPU baseline ("C restatement of
 * Hive 3.1 vectorized operators", never labelled "Hive").
 */
#define _GNU_SOURCE
#include "ghive_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------------------ */
/* Philox4x32-10, Salmon et al. SC'11 (Random123 reference constants).                     */
/* ------------------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_row(uint64_t seed, uint32_t col, uint64_t row, uint32_t out[4]) {
  uint32_t ctr[4] = {(uint32_t)row, (uint32_t)(row >> 32), col, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, out);
}

void orc_gen_uniform_i32(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         int32_t lo, int32_t hi, int32_t *out) {
  uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1u;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r[4];
    philox_row(seed, column_id, row_begin + i, r);
    out[i] = (int32_t)((int64_t)lo + (int64_t)(((uint64_t)r[0] * range) >> 32));
  }
}

void orc_gen_uniform_i64(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         int64_t lo, int64_t hi, int64_t *out) {
  uint64_t range = (uint64_t)hi - (uint64_t)lo + 1u; /* 0 means the full 2^64 range */
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r[4];
    philox_row(seed, column_id, row_begin + i, r);
    uint64_t x = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    uint64_t off = range ? (uint64_t)(((unsigned __int128)x * range) >> 64) : x;
    out[i] = (int64_t)((uint64_t)lo + off);
  }
}

void orc_gen_uniform_f32(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         float *out) {
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r[4];
    philox_row(seed, column_id, row_begin + i, r);
    out[i] = (float)(r[0] >> 8) * (1.0f / 16777216.0f);
  }
}

void orc_gen_normal_f64(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                        double *out) {
  const double two_pi = 6.283185307179586476925286766559;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t r[4];
    philox_row(seed, column_id, row_begin + i, r);
    uint64_t a = ((uint64_t)r[0] << 32) | r[1];
    uint64_t b = ((uint64_t)r[2] << 32) | r[3];
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0); /* (0,1] */
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);         /* [0,1) */
    out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
}

/* Bijection of [0, 2^62): offset by a seed constant, then xorshift / odd-multiply rounds. */
uint64_t orc_key62(uint64_t seed, uint64_t rank) {
  const uint64_t M = (1ull << 62) - 1;
  uint64_t x = (rank + seed * 0x9E3779B97F4A7C15ull) & M;
  x ^= x >> 31;
  x = (x * 0x7fb5d329728ea185ull) & M;
  x ^= x >> 27;
  x = (x * 0x81dadef4bc2dd44dull) & M;
  x ^= x >> 33;
  return x;
}

/* Zipf(s) over [1, n] by rejection-inversion (Hoermann & Derflinger, ACM TOMACS 6(3), 1996). */
static double z_h1(double x) { return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x)); }
static double z_h2(double x) { return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x)); }
static double z_H(double s, double x) { double lx = log(x); return z_h2((1.0 - s) * lx) * lx; }
static double z_h(double s, double x) { return exp(-s * log(x)); }
static double z_Hinv(double s, double y) {
  double t = y * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(z_h1(t) * y);
}

void orc_gen_zipf_keys(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n, uint64_t nranks,
                       double theta, double miss_frac, int64_t *out_keys, uint64_t *out_ranks) {
  const double s = theta, N = (double)nranks;
  const double h_x1 = z_H(s, 1.5) - 1.0, h_n = z_H(s, N + 0.5);
  const double sx = 2.0 - z_Hinv(s, z_H(s, 2.5) - z_h(s, 2.0));
  const uint32_t thr = miss_frac >= 1.0 ? 0xffffffffu : (uint32_t)(miss_frac * 4294967296.0);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t r = row_begin + i;
    uint32_t w[4];
    uint32_t ctr[4] = {(uint32_t)r, (uint32_t)(r >> 32), column_id, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, w);
    uint64_t rank = 0;
    if (w[0] < thr) {
      rank = nranks + 1 + (((uint64_t)w[1] << 8) | (w[2] & 0xffu));
    } else {
      for (uint32_t a = 0; a < 64 && rank == 0; ++a) {
        if (a) {
          ctr[3] = a;
          orc_philox4x32_10(ctr, key, w);
        }
        uint64_t bits = ((uint64_t)w[2] << 32) | w[3];
        double u = (double)(bits >> 11) * (1.0 / 9007199254740992.0);
        double uu = h_n + u * (h_x1 - h_n);
        double x = z_Hinv(s, uu);
        double kf = floor(x + 0.5);
        if (kf < 1.0) kf = 1.0;
        if (kf > N) kf = N;
        if (kf - x <= sx || uu >= z_H(s, kf + 0.5) - z_h(s, kf)) rank = (uint64_t)kf;
      }
      if (rank == 0) rank = 1;
    }
    if (out_ranks) out_ranks[i] = rank;
    out_keys[i] = (int64_t)orc_key62(seed, rank);
  }
}

/* ------------------------------------------------------------------------------------ */
/* Column access and NULL sentinels (gpu_impl/inc/DataFlow/Table.hpp:29-52).               */
/* ------------------------------------------------------------------------------------ */
static inline int col_is_int(const ghive_column *c) {
  return c->type == GHIVE_INT || c->type == GHIVE_LONG;
}
static inline int col_is_fp(const ghive_column *c) {
  return c->type == GHIVE_DOUBLE || c->type == GHIVE_FLOAT;
}
static inline int64_t col_i64(const ghive_column *c, uint64_t r) {
  return c->type == GHIVE_INT ? (int64_t)((const int32_t *)c->data)[r]
                              : ((const int64_t *)c->data)[r];
}
static inline double col_f64(const ghive_column *c, uint64_t r) {
  switch (c->type) {
    case GHIVE_DOUBLE: return ((const double *)c->data)[r];
    case GHIVE_FLOAT: return (double)((const float *)c->data)[r];
    case GHIVE_INT: return (double)((const int32_t *)c->data)[r];
    default: return (double)((const int64_t *)c->data)[r];
  }
}
static inline int col_null(const ghive_column *c, uint64_t r) {
  switch (c->type) {
    case GHIVE_INT: return ((const int32_t *)c->data)[r] == INT32_MAX;
    case GHIVE_LONG: return ((const int64_t *)c->data)[r] == INT64_MAX;
    case GHIVE_DOUBLE: return isnan(((const double *)c->data)[r]);
    case GHIVE_FLOAT: return isnan(((const float *)c->data)[r]);
    case GHIVE_STRING: return ((const int32_t *)c->aux)[2 * r] < 0;
    default: return 0;
  }
}

/* ------------------------------------------------------------------------------------ */
/* Filter: Hive's in-place selected[] refinement per 1024-row VectorizedRowBatch          */
/* (VectorizedRowBatch.java:36-60; FilterColumnCompareScalar.txt:73-101; AND = children   */
/* applied in sequence, FilterExprAndExpr; OR = union of children, FilterExprOrExpr).      */
/* ------------------------------------------------------------------------------------ */
#define ORC_BATCH 1024

static int cmp_mode_i(int mode, int64_t x, int64_t a, int64_t b) {
  switch (mode) {
    case GHIVE_FILTER_EQ: return x == a;
    case GHIVE_FILTER_LE: return x <= a;
    case GHIVE_FILTER_GE: return x >= a;
    case GHIVE_FILTER_LT: return x < a;
    case GHIVE_FILTER_GT: return x > a;
    case GHIVE_FILTER_RANGE: return x > a && x < b;
    case GHIVE_FILTER_EQ_RANGE: return x >= a && x <= b;
    case GHIVE_FILTER_NOT_EQ: return x != a;
    case GHIVE_FILTER_NOT_RANGE: return x < a || x > b;
    case GHIVE_FILTER_NOT_EQ_RANGE: return x <= a || x >= b;
    case GHIVE_FILTER_NOT_NULL: return 1;
    default: return 0;
  }
}
static int cmp_mode_d(int mode, double x, double a, double b) {
  switch (mode) {
    case GHIVE_FILTER_EQ: return x == a;
    case GHIVE_FILTER_LE: return x <= a;
    case GHIVE_FILTER_GE: return x >= a;
    case GHIVE_FILTER_LT: return x < a;
    case GHIVE_FILTER_GT: return x > a;
    case GHIVE_FILTER_RANGE: return x > a && x < b;
    case GHIVE_FILTER_EQ_RANGE: return x >= a && x <= b;
    case GHIVE_FILTER_NOT_EQ: return x != a;
    case GHIVE_FILTER_NOT_RANGE: return x < a || x > b;
    case GHIVE_FILTER_NOT_EQ_RANGE: return x <= a || x >= b;
    case GHIVE_FILTER_NOT_NULL: return 1;
    default: return 0;
  }
}

/* StringExpr.compare (hive storage-api .../vector/expressions/StringExpr.java:37-47): unsigned
 * bytes, then length. */
static int str_cmp(const uint8_t *a, int la, const uint8_t *b, int lb) {
  for (int i = 0; i < la && i < lb; ++i)
    if (a[i] != b[i]) return (int)a[i] - (int)b[i];
  return la - lb;
}

/* A STRING row against the node: string literals (FilterStringGroupColumnCompareStringGroupScalarBase.txt,
 * FilterStringColumnBetween.txt:79-99 -- BETWEEN inclusive, NOT BETWEEN its negation) or
 * another STRING column. */
static int str_leaf_pass(const ghive_table *t, const ghive_pred *p, uint64_t r) {
  const ghive_column *c = &t->columns[p->filter_col[0]];
  const int32_t *ix = (const int32_t *)c->aux;
  const uint8_t *x = (const uint8_t *)c->data + ix[2 * r];
  const int lx = ix[2 * r + 1] - ix[2 * r];
  if (p->pred_col_num == 2) {
    const ghive_column *d = &t->columns[p->filter_col[1]];
    const int32_t *iy = (const int32_t *)d->aux;
    const int cy = str_cmp(x, lx, (const uint8_t *)d->data + iy[2 * r], iy[2 * r + 1] - iy[2 * r]);
    return cmp_mode_i(p->mode, cy, 0, 0);
  }
  const int c0 = str_cmp(x, lx, (const uint8_t *)p->string_params[0], p->string_sizes[0]);
  const int c1 = p->param_num == 2 ? str_cmp(x, lx, (const uint8_t *)p->string_params[1], p->string_sizes[1]) : 0;
  switch (p->mode) {
    case GHIVE_FILTER_EQ: return c0 == 0;
    case GHIVE_FILTER_LE: return c0 <= 0;
    case GHIVE_FILTER_GE: return c0 >= 0;
    case GHIVE_FILTER_LT: return c0 < 0;
    case GHIVE_FILTER_GT: return c0 > 0;
    case GHIVE_FILTER_RANGE: return c0 > 0 && c1 < 0;
    case GHIVE_FILTER_EQ_RANGE: return c0 >= 0 && c1 <= 0;
    case GHIVE_FILTER_NOT_EQ: return c0 != 0;
    case GHIVE_FILTER_NOT_RANGE: return !(c0 >= 0 && c1 <= 0);
    case GHIVE_FILTER_NOT_EQ_RANGE: return !(c0 > 0 && c1 < 0);
    default: return 0;
  }
}

/* One row against a leaf node; value semantics (an int literal on a LONG column is widened,
 * FilterPredicate.cu:484-487; a double literal compares the converted value, as Hive's
 * FilterLongColCompareDoubleScalar does). */
static int leaf_pass(const ghive_table *t, const ghive_pred *p, uint64_t r) {
  const ghive_column *c = &t->columns[p->filter_col[0]];
  if (col_null(c, r)) return 0;
  if (p->pred_col_num == 2 && col_null(&t->columns[p->filter_col[1]], r)) return 0;
  if (c->type == GHIVE_STRING) return p->mode == GHIVE_FILTER_NOT_NULL ? 1 : str_leaf_pass(t, p, r);
  if (p->pred_col_num == 2) {
    const ghive_column *d = &t->columns[p->filter_col[1]];
    if (col_is_int(c) && col_is_int(d))
      return cmp_mode_i(p->mode, col_i64(c, r), col_i64(d, r), 0);
    return cmp_mode_d(p->mode, col_f64(c, r), col_f64(d, r), 0.0);
  }
  if (p->mode == GHIVE_FILTER_NOT_NULL) return 1;
  int dt = p->data_type;
  if (col_is_int(c) && (dt == 0 || dt == 2)) {
    int64_t a = dt == 0 ? p->long_params[0] : p->int_params[0];
    int64_t b = dt == 0 ? p->long_params[1] : p->int_params[1];
    return cmp_mode_i(p->mode, col_i64(c, r), a, b);
  }
  double a, b;
  if (dt == 1) { a = p->double_params[0]; b = p->double_params[1]; }
  else if (dt == 0) { a = (double)p->long_params[0]; b = (double)p->long_params[1]; }
  else { a = p->int_params[0]; b = p->int_params[1]; }
  return cmp_mode_d(p->mode, col_f64(c, r), a, b);
}

static int pred_supported(const ghive_table *t, const ghive_pred *nodes, uint32_t nnodes,
                          uint32_t idx, int depth) {
  if (idx >= nnodes || depth > 64) return 0;
  const ghive_pred *p = &nodes[idx];
  if (p->mode == GHIVE_FILTER_AND || p->mode == GHIVE_FILTER_OR)
    return p->children[0] >= 0 && p->children[1] >= 0 &&
           pred_supported(t, nodes, nnodes, (uint32_t)p->children[0], depth + 1) &&
           pred_supported(t, nodes, nnodes, (uint32_t)p->children[1], depth + 1);
  if (p->mode < 0 || p->mode > GHIVE_FILTER_NOT_NULL) return 0;
  if (p->filter_col[0] >= t->ncols) return 0;
  const int str0 = t->columns[p->filter_col[0]].type == GHIVE_STRING;
  if (p->pred_col_num == 2) {
    if (p->filter_col[1] >= t->ncols) return 0;
    if ((t->columns[p->filter_col[1]].type == GHIVE_STRING) != str0) return 0;
  } else if (p->mode != GHIVE_FILTER_NOT_NULL) {
    if (str0 != (p->data_type == 3)) return 0;
    if (p->data_type < 0 || p->data_type > 3) return 0;
    if (str0 && (p->param_num < 1 || p->param_num > 2 || p->string_sizes[0] < 0 ||
                 p->string_sizes[0] > GHIVE_PRED_STR_MAX || p->string_sizes[1] < 0 ||
                 p->string_sizes[1] > GHIVE_PRED_STR_MAX))
      return 0;
  }
  return 1;
}

/* Refines sel[0..size) in place (ascending batch-relative indexes); returns the new size. */
static int eval_node(const ghive_table *t, const ghive_pred *nodes, uint32_t idx,
                     uint64_t base, int *sel, int size) {
  const ghive_pred *p = &nodes[idx];
  if (p->mode == GHIVE_FILTER_AND) {
    size = eval_node(t, nodes, (uint32_t)p->children[0], base, sel, size);
    return eval_node(t, nodes, (uint32_t)p->children[1], base, sel, size);
  }
  if (p->mode == GHIVE_FILTER_OR) {
    int a[ORC_BATCH], b[ORC_BATCH];
    memcpy(a, sel, sizeof(int) * (size_t)size);
    int na = eval_node(t, nodes, (uint32_t)p->children[0], base, a, size);
    /* rows not taken by the first child go to the second */
    int nb = 0, j = 0;
    for (int i = 0; i < size; ++i) {
      if (j < na && a[j] == sel[i]) { ++j; continue; }
      b[nb++] = sel[i];
    }
    nb = eval_node(t, nodes, (uint32_t)p->children[1], base, b, nb);
    int i = 0, k = 0, n = 0;
    while (i < na || k < nb) {
      if (k >= nb || (i < na && a[i] < b[k])) sel[n++] = a[i++];
      else sel[n++] = b[k++];
    }
    return n;
  }
  int n = 0;
  for (int i = 0; i < size; ++i)
    if (leaf_pass(t, p, base + (uint64_t)sel[i])) sel[n++] = sel[i];
  return n;
}

int64_t orc_filter(const ghive_table *table, const ghive_pred *nodes, uint32_t nnodes,
                   uint32_t root, uint32_t *out_sel) {
  if (!pred_supported(table, nodes, nnodes, root, 0)) return -2;
  int64_t count = 0;
  int sel[ORC_BATCH];
  for (uint64_t base = 0; base < table->row_num; base += ORC_BATCH) {
    int size = (int)((table->row_num - base) < ORC_BATCH ? (table->row_num - base) : ORC_BATCH);
    for (int i = 0; i < size; ++i) sel[i] = i;
    size = eval_node(table, nodes, root, base, sel, size);
    for (int i = 0; i < size; ++i) out_sel[count++] = (uint32_t)(base + (uint64_t)sel[i]);
  }
  return count;
}

/* ------------------------------------------------------------------------------------ */
/* Hive VectorMapJoinFastLongHashMap (VectorMapJoinFastLongHashTable.java:99-268).          */
/* ------------------------------------------------------------------------------------ */
int32_t orc_hive_long_hash(int64_t key) {
  uint64_t k = (uint64_t)key; /* Java >>> is a logical shift */
  k ^= (k >> 7) ^ (k >> 4);
  k ^= (k >> 20) ^ (k >> 12);
  return (int32_t)(uint32_t)k;
}

struct orc_mj_table {
  int64_t *slot_pairs; /* [2*bucket]: valueRef (entry+1, 0 = empty), key */
  int64_t buckets, mask, keys_assigned, resize_threshold;
  float load_factor;
  int32_t largest_steps, expands;
  int64_t min, max;
  int64_t *head, *tail; /* per key entry: first / last value index */
  int64_t nentries, cap_entries;
  int64_t *val, *next;
  int64_t nvals, cap_vals;
};

#define ORC_ONE_QUARTER_LIMIT (1073741824 / 4)

static void mj_expand(orc_mj_table *t) {
  int64_t nb = t->buckets * 2, nm = nb - 1;
  int64_t *np = (int64_t *)calloc((size_t)(2 * nb), sizeof(int64_t));
  int32_t steps = 0;
  for (int64_t s = 0; s < t->buckets; ++s) {
    int64_t ref = t->slot_pairs[2 * s];
    if (!ref) continue;
    int64_t key = t->slot_pairs[2 * s + 1];
    int64_t slot = (int64_t)orc_hive_long_hash(key) & nm;
    int64_t probe = slot;
    int32_t i = 0;
    while (np[2 * slot]) {
      probe += ++i;
      slot = probe & nm;
    }
    if (steps < i) steps = i;
    np[2 * slot] = ref;
    np[2 * slot + 1] = key;
  }
  free(t->slot_pairs);
  t->slot_pairs = np;
  t->buckets = nb;
  t->mask = nm;
  t->largest_steps = steps;
  t->resize_threshold = (int64_t)((float)nb * t->load_factor);
  t->expands++;
}

static void mj_add(orc_mj_table *t, int64_t key, int64_t value) {
  if (t->resize_threshold <= t->keys_assigned) mj_expand(t);
  int64_t slot = (int64_t)orc_hive_long_hash(key) & t->mask;
  int64_t probe = slot;
  int32_t i = 0;
  int is_new;
  for (;;) {
    int64_t ref = t->slot_pairs[2 * slot];
    if (ref == 0) { is_new = 1; break; }
    if (t->slot_pairs[2 * slot + 1] == key) { is_new = 0; break; }
    probe += ++i;
    slot = probe & t->mask;
  }
  if (t->largest_steps < i) t->largest_steps = i;
  if (t->nvals == t->cap_vals) {
    t->cap_vals = t->cap_vals ? 2 * t->cap_vals : 1024;
    t->val = (int64_t *)realloc(t->val, sizeof(int64_t) * (size_t)t->cap_vals);
    t->next = (int64_t *)realloc(t->next, sizeof(int64_t) * (size_t)t->cap_vals);
  }
  int64_t v = t->nvals++;
  t->val[v] = value;
  t->next[v] = -1;
  if (is_new) {
    if (t->nentries == t->cap_entries) {
      t->cap_entries = t->cap_entries ? 2 * t->cap_entries : 1024;
      t->head = (int64_t *)realloc(t->head, sizeof(int64_t) * (size_t)t->cap_entries);
      t->tail = (int64_t *)realloc(t->tail, sizeof(int64_t) * (size_t)t->cap_entries);
    }
    int64_t e = t->nentries++;
    t->head[e] = t->tail[e] = v;
    t->slot_pairs[2 * slot] = e + 1;
    t->slot_pairs[2 * slot + 1] = key;
    t->keys_assigned++;
    if (key < t->min) t->min = key;
    if (key > t->max) t->max = key;
  } else {
    int64_t e = t->slot_pairs[2 * slot] - 1;
    t->next[t->tail[e]] = v;
    t->tail[e] = v;
  }
}

/* findReadSlot (VectorMapJoinFastLongHashTable.java:221-251): entry index or -1. */
static int64_t mj_find(const orc_mj_table *t, int64_t key) {
  int64_t slot = (int64_t)orc_hive_long_hash(key) & t->mask;
  int64_t probe = slot;
  int32_t i = 0;
  for (;;) {
    int64_t ref = t->slot_pairs[2 * slot];
    if (ref == 0) return -1;
    if (t->slot_pairs[2 * slot + 1] == key) return ref - 1;
    probe += ++i;
    if (i > t->largest_steps) return -1;
    slot = probe & t->mask;
  }
}

orc_mj_table *orc_mj_build(const ghive_column *key, const uint32_t *sel, uint64_t n,
                           const ghive_column *payload, float load_factor) {
  if (!col_is_int(key)) return NULL;
  if (payload && !col_is_int(payload)) return NULL;
  orc_mj_table *t = (orc_mj_table *)calloc(1, sizeof(orc_mj_table));
  t->load_factor = load_factor > 0.f ? load_factor : 0.75f; /* hive.hashtable.loadfactor */
  /* pre-sized from the key-count estimate, as VectorMapJoinFastTableContainer does */
  int64_t want = (int64_t)((double)n / t->load_factor) + 1, b = 16;
  while (b < want && b < ORC_ONE_QUARTER_LIMIT) b <<= 1;
  t->buckets = b;
  t->mask = b - 1;
  t->resize_threshold = (int64_t)((float)b * t->load_factor);
  t->slot_pairs = (int64_t *)calloc((size_t)(2 * b), sizeof(int64_t));
  t->min = INT64_MAX;
  t->max = INT64_MIN;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t r = sel ? sel[i] : i;
    if (col_null(key, r)) continue; /* NULL keys never join */
    int64_t v = payload ? col_i64(payload, r) : (int64_t)r;
    mj_add(t, col_i64(key, r), v);
  }
  return t;
}

void orc_mj_free(orc_mj_table *t) {
  if (!t) return;
  free(t->slot_pairs); free(t->head); free(t->tail); free(t->val); free(t->next);
  free(t);
}

void orc_mj_stats(const orc_mj_table *t, int64_t *keys, int64_t *buckets, int64_t *values,
                  int32_t *largest_steps, int32_t *expands) {
  *keys = t->keys_assigned; *buckets = t->buckets; *values = t->nvals;
  *largest_steps = t->largest_steps; *expands = t->expands;
}

/* VectorMapJoinInnerLongOperator.process: NULL keys dropped, min/max pre-check, then
 * findReadSlot; every value of a matched key is emitted. */
int64_t orc_mj_probe(const orc_mj_table *t, const ghive_column *key, const uint32_t *sel,
                     uint64_t n, uint32_t *out_probe_rows, int64_t *out_payload,
                     uint64_t capacity, uint64_t *needed) {
  uint64_t count = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t r = sel ? sel[i] : i;
    if (col_null(key, r)) continue;
    int64_t k = col_i64(key, r);
    if (k < t->min || k > t->max) continue;
    int64_t e = mj_find(t, k);
    if (e < 0) continue;
    for (int64_t v = t->head[e]; v >= 0; v = t->next[v]) {
      if (count < capacity) {
        out_probe_rows[count] = (uint32_t)r;
        out_payload[count] = t->val[v];
      }
      ++count;
    }
  }
  if (needed) *needed = count;
  return count > capacity ? -1 : (int64_t)count;
}

/* ------------------------------------------------------------------------------------ */
/* N-way join with result_offsets (SortMergeJoinPredicate.cu:8-461 contract).              */
/* ------------------------------------------------------------------------------------ */
int64_t orc_join(const ghive_table *tables, uint32_t ntables, const uint32_t *join_cols,
                 const ghive_join_cond *conds, uint32_t nconds, int32_t *out_offsets,
                 uint64_t capacity, uint64_t *needed) {
  if (ntables < 2 || nconds != ntables - 1) return -2;
  for (uint32_t t = 0; t < ntables; ++t)
    if (join_cols[t] >= tables[t].ncols || !col_is_int(&tables[t].columns[join_cols[t]]))
      return -2;
  int *semi = (int *)calloc(ntables, sizeof(int));
  for (uint32_t c = 0; c < nconds; ++c) {
    int ty = conds[c].type;
    if (ty != GHIVE_INNER_JOIN && ty != GHIVE_LEFT_OUTER_JOIN && ty != GHIVE_LEFT_SEMI_JOIN) {
      free(semi);
      return -2;
    }
    if (conds[c].left >= ntables || conds[c].right >= ntables) { free(semi); return -2; }
    if (ty == GHIVE_LEFT_SEMI_JOIN) semi[conds[c].right] = 1;
  }
  /* tuples: ntables ids per row; -1 = NULL side */
  uint64_t nt = tables[0].row_num, cap = nt ? nt : 1;
  int32_t *tup = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap * ntables));
  for (uint64_t r = 0; r < nt; ++r) {
    for (uint32_t t = 0; t < ntables; ++t) tup[r * ntables + t] = -1;
    tup[r * ntables] = (int32_t)r;
  }
  for (uint32_t c = 0; c < nconds; ++c) {
    uint32_t L = conds[c].left, R = conds[c].right;
    const ghive_column *rk = &tables[R].columns[join_cols[R]];
    const ghive_column *lk = &tables[L].columns[join_cols[L]];
    orc_mj_table *ht = orc_mj_build(rk, NULL, tables[R].row_num, NULL, 0.75f);
    uint64_t ncap = nt > 0 ? nt : 1, nn = 0;
    int32_t *nxt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ncap * ntables));
    for (uint64_t i = 0; i < nt; ++i) {
      int32_t *row = &tup[i * ntables];
      int64_t e = -1;
      if (row[L] >= 0 && !col_null(lk, (uint64_t)row[L])) {
        int64_t k = col_i64(lk, (uint64_t)row[L]);
        if (k >= ht->min && k <= ht->max) e = mj_find(ht, k);
      }
      int emitted = 0;
      if (e >= 0) {
        for (int64_t v = ht->head[e]; v >= 0; v = ht->next[v]) {
          if (nn == ncap) {
            ncap *= 2;
            nxt = (int32_t *)realloc(nxt, sizeof(int32_t) * (size_t)(ncap * ntables));
          }
          memcpy(&nxt[nn * ntables], row, sizeof(int32_t) * ntables);
          if (conds[c].type != GHIVE_LEFT_SEMI_JOIN) nxt[nn * ntables + R] = (int32_t)ht->val[v];
          ++nn;
          emitted = 1;
          if (conds[c].type == GHIVE_LEFT_SEMI_JOIN) break;
        }
      }
      if (!emitted && conds[c].type == GHIVE_LEFT_OUTER_JOIN) {
        if (nn == ncap) {
          ncap *= 2;
          nxt = (int32_t *)realloc(nxt, sizeof(int32_t) * (size_t)(ncap * ntables));
        }
        memcpy(&nxt[nn * ntables], row, sizeof(int32_t) * ntables);
        nxt[nn * ntables + R] = -1;
        ++nn;
      }
    }
    orc_mj_free(ht);
    free(tup);
    tup = nxt;
    nt = nn;
  }
  uint32_t nout = 0;
  for (uint32_t t = 0; t < ntables; ++t) nout += !semi[t];
  if (needed) *needed = nt;
  int64_t ret = (int64_t)nt;
  if (nt > capacity) {
    ret = -1;
  } else {
    for (uint64_t i = 0; i < nt; ++i) {
      uint32_t o = 0;
      for (uint32_t t = 0; t < ntables; ++t)
        if (!semi[t]) out_offsets[i * nout + o++] = tup[i * ntables + t];
    }
  }
  free(tup);
  free(semi);
  return ret;
}

/* ------------------------------------------------------------------------------------ */
/* Exact (correctly rounded) double sum: Shewchuk's non-overlapping partials with the      */
/* final half-way correction (J. R. Shewchuk, "Adaptive Precision Floating-Point           */
/* Arithmetic and Fast Robust Geometric Predicates", DCG 18(3), 1997).                      */
/* ------------------------------------------------------------------------------------ */
typedef struct { double *p; int n, cap; } orc_partials;

static void partials_add(orc_partials *s, double x) {
  int i = 0;
  for (int j = 0; j < s->n; ++j) {
    double y = s->p[j];
    if (fabs(x) < fabs(y)) { double tmp = x; x = y; y = tmp; }
    double hi = x + y;
    double lo = y - (hi - x);
    if (lo != 0.0) s->p[i++] = lo;
    x = hi;
  }
  if (i + 1 > s->cap) {
    s->cap = s->cap ? 2 * s->cap : 8;
    s->p = (double *)realloc(s->p, sizeof(double) * (size_t)s->cap);
  }
  s->p[i] = x;
  s->n = i + 1;
}

static double partials_sum(const orc_partials *s) {
  int n = s->n;
  double hi = 0.0;
  if (n > 0) {
    double lo = 0.0;
    hi = s->p[--n];
    while (n > 0) {
      double x = hi, y = s->p[--n];
      hi = x + y;
      double yr = hi - x;
      lo = y - yr;
      if (lo != 0.0) break;
    }
    if (n > 0 && ((lo < 0 && s->p[n - 1] < 0) || (lo > 0 && s->p[n - 1] > 0))) {
      double y = lo * 2, x = hi + y, yr = x - hi;
      if (y == yr) hi = x;
    }
  }
  return hi;
}

double orc_exact_sum(const double *x, uint64_t n) {
  orc_partials s = {0, 0, 0};
  for (uint64_t i = 0; i < n; ++i) partials_add(&s, x[i]);
  double r = partials_sum(&s);
  free(s.p);
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* Hash group-by (VectorGroupByOperator.ProcessingModeHashAggregate, :282-650).             */
/* ------------------------------------------------------------------------------------ */
static uint64_t fnv1a(const uint8_t *b, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
  return h ^ (h >> 29);
}

/* Serialises the key of row r; returns its length. NaN is canonical (one NULL group). */
static size_t key_bytes(const ghive_table *t, const uint32_t *keys, uint32_t nkeys, uint64_t r,
                        uint8_t *buf, size_t cap) {
  size_t o = 0;
  for (uint32_t k = 0; k < nkeys; ++k) {
    const ghive_column *c = &t->columns[keys[k]];
    switch (c->type) {
      case GHIVE_INT: memcpy(buf + o, (const int32_t *)c->data + r, 4); o += 4; break;
      case GHIVE_LONG: memcpy(buf + o, (const int64_t *)c->data + r, 8); o += 8; break;
      case GHIVE_DOUBLE: {
        double d = ((const double *)c->data)[r];
        if (isnan(d)) d = NAN;
        memcpy(buf + o, &d, 8); o += 8; break;
      }
      case GHIVE_FLOAT: {
        float f = ((const float *)c->data)[r];
        if (isnan(f)) f = NAN;
        memcpy(buf + o, &f, 4); o += 4; break;
      }
      case GHIVE_STRING: {
        const int32_t *ix = (const int32_t *)c->aux;
        int32_t s = ix[2 * r], e = ix[2 * r + 1];
        int32_t len = s < 0 ? -1 : e - s;
        memcpy(buf + o, &len, 4); o += 4;
        if (len > 0) {
          size_t l = (size_t)len;
          if (o + l > cap) l = cap - o;
          memcpy(buf + o, (const char *)c->data + s, l); o += l;
        }
        break;
      }
      default: break;
    }
  }
  return o;
}

typedef struct {
  int64_t i;   /* integer SUM / MIN / MAX; integer part of AVG */
  double d;    /* double SUM (arrival order) / MIN / MAX       */
  int64_t cnt; /* non-NULL inputs (COUNT(*): rows)              */
} orc_acc;

int64_t orc_groupby(const ghive_table *table, const uint32_t *sel, uint64_t n,
                    const uint32_t *keys, uint32_t nkeys, const ghive_agg *aggs, uint32_t naggs,
                    int exact_double_sum, uint32_t *rep_row, int64_t *out_i64, double *out_f64,
                    uint8_t *out_null, uint64_t capacity) {
  for (uint32_t k = 0; k < nkeys; ++k)
    if (keys[k] >= table->ncols) return -2;
  for (uint32_t a = 0; a < naggs; ++a) {
    if (aggs[a].type < 0 || aggs[a].type > GHIVE_AGG_CNT) return -2;
    if (!(aggs[a].type == GHIVE_AGG_CNT && aggs[a].index == GHIVE_COUNT_STAR)) {
      if (aggs[a].index >= table->ncols) return -2;
      if (table->columns[aggs[a].index].type == GHIVE_STRING && aggs[a].type != GHIVE_AGG_CNT)
        return -2;
    }
  }
  uint64_t hcap = 16;
  while (hcap < 2 * (n + 1)) hcap <<= 1;
  int64_t *slots = (int64_t *)malloc(sizeof(int64_t) * (size_t)hcap);
  for (uint64_t i = 0; i < hcap; ++i) slots[i] = -1;
  uint64_t *hashes = NULL;
  size_t *koff = NULL, *klen = NULL;
  uint8_t *kblob = NULL;
  size_t kblob_n = 0, kblob_cap = 0;
  uint32_t *reps = NULL;
  orc_acc *acc = NULL;
  int64_t ng = 0, gcap = 0;
  uint32_t *grp_of = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  uint8_t kb[4096];

  uint64_t loop_n = (nkeys == 0) ? n : n;
  if (nkeys == 0) { /* global aggregate: exactly one group even for empty input */
    gcap = 1;
    reps = (uint32_t *)malloc(sizeof(uint32_t));
    acc = (orc_acc *)calloc(naggs ? naggs : 1, sizeof(orc_acc));
    reps[0] = n ? (sel ? sel[0] : 0) : 0;
    ng = 1;
  }
  for (uint64_t i = 0; i < loop_n; ++i) {
    uint64_t r = sel ? sel[i] : i;
    int64_t g = 0;
    if (nkeys) {
      size_t len = key_bytes(table, keys, nkeys, r, kb, sizeof kb);
      uint64_t h = fnv1a(kb, len), s = h & (hcap - 1);
      for (;;) {
        int64_t e = slots[s];
        if (e < 0) {
          if (ng == gcap) {
            gcap = gcap ? 2 * gcap : 64;
            hashes = (uint64_t *)realloc(hashes, sizeof(uint64_t) * (size_t)gcap);
            koff = (size_t *)realloc(koff, sizeof(size_t) * (size_t)gcap);
            klen = (size_t *)realloc(klen, sizeof(size_t) * (size_t)gcap);
            reps = (uint32_t *)realloc(reps, sizeof(uint32_t) * (size_t)gcap);
            acc = (orc_acc *)realloc(acc, sizeof(orc_acc) * (size_t)gcap * (naggs ? naggs : 1));
          }
          if (kblob_n + len > kblob_cap) {
            kblob_cap = (kblob_cap + len) * 2;
            kblob = (uint8_t *)realloc(kblob, kblob_cap);
          }
          memcpy(kblob + kblob_n, kb, len);
          koff[ng] = kblob_n; klen[ng] = len; kblob_n += len;
          hashes[ng] = h;
          reps[ng] = (uint32_t)r;
          memset(&acc[(size_t)ng * naggs], 0, sizeof(orc_acc) * naggs);
          slots[s] = ng;
          g = ng++;
          break;
        }
        if (hashes[e] == h && klen[e] == len && memcmp(kblob + koff[e], kb, len) == 0) {
          g = e;
          break;
        }
        s = (s + 1) & (hcap - 1);
      }
    }
    grp_of[i] = (uint32_t)g;
    for (uint32_t a = 0; a < naggs; ++a) {
      orc_acc *A = &acc[(size_t)g * naggs + a];
      if (aggs[a].type == GHIVE_AGG_CNT && aggs[a].index == GHIVE_COUNT_STAR) {
        A->cnt++;
        continue;
      }
      const ghive_column *c = &table->columns[aggs[a].index];
      if (col_null(c, r)) continue;
      int first = A->cnt == 0;
      A->cnt++;
      if (aggs[a].type == GHIVE_AGG_CNT) continue;
      if (col_is_int(c)) {
        int64_t v = col_i64(c, r);
        switch (aggs[a].type) {
          case GHIVE_AGG_SUM:
          case GHIVE_AGG_AVG: A->i = (int64_t)((uint64_t)A->i + (uint64_t)v); break; /* Java long */
          case GHIVE_AGG_MAX: if (first || v > A->i) A->i = v; break;
          case GHIVE_AGG_MIN: if (first || v < A->i) A->i = v; break;
        }
      } else {
        double v = col_f64(c, r);
        switch (aggs[a].type) {
          case GHIVE_AGG_SUM:
          case GHIVE_AGG_AVG: A->d += v; break; /* arrival order, VectorUDAFSum.txt */
          case GHIVE_AGG_MAX: if (first || v > A->d) A->d = v; break;
          case GHIVE_AGG_MIN: if (first || v < A->d) A->d = v; break;
        }
      }
    }
  }
  int64_t ret = ng;
  if ((uint64_t)ng > capacity) {
    ret = -1;
  } else {
    for (int64_t g = 0; g < ng; ++g) rep_row[g] = reps[g];
    for (uint32_t a = 0; a < naggs; ++a) {
      int itype = (aggs[a].type == GHIVE_AGG_CNT && aggs[a].index == GHIVE_COUNT_STAR)
                      ? GHIVE_LONG : table->columns[aggs[a].index].type;
      int is_int = itype == GHIVE_INT || itype == GHIVE_LONG || itype == GHIVE_STRING;
      double *exact = NULL;
      if (exact_double_sum && !is_int && (aggs[a].type == GHIVE_AGG_SUM || aggs[a].type == GHIVE_AGG_AVG)) {
        /* bucket the values by group, then a correctly rounded sum per group */
        uint64_t *start = (uint64_t *)calloc((size_t)ng + 1, sizeof(uint64_t));
        const ghive_column *c = &table->columns[aggs[a].index];
        for (uint64_t i = 0; i < n; ++i) {
          uint64_t r = sel ? sel[i] : i;
          if (!col_null(c, r)) start[grp_of[i] + 1]++;
        }
        for (int64_t g = 0; g < ng; ++g) start[g + 1] += start[g];
        double *vals = (double *)malloc(sizeof(double) * (size_t)(start[ng] ? start[ng] : 1));
        uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(ng ? ng : 1));
        for (int64_t g = 0; g < ng; ++g) fill[g] = start[g];
        for (uint64_t i = 0; i < n; ++i) {
          uint64_t r = sel ? sel[i] : i;
          if (!col_null(c, r)) vals[fill[grp_of[i]]++] = col_f64(c, r);
        }
        exact = (double *)malloc(sizeof(double) * (size_t)(ng ? ng : 1));
        for (int64_t g = 0; g < ng; ++g) exact[g] = orc_exact_sum(vals + start[g], start[g + 1] - start[g]);
        free(start); free(vals); free(fill);
      }
      for (int64_t g = 0; g < ng; ++g) {
        orc_acc *A = &acc[(size_t)g * naggs + a];
        size_t o = (size_t)g * naggs + a;
        out_i64[o] = 0;
        out_f64[o] = 0.0;
        out_null[o] = 0;
        switch (aggs[a].type) {
          case GHIVE_AGG_CNT: out_i64[o] = A->cnt; break;
          case GHIVE_AGG_SUM:
            if (!A->cnt) { out_null[o] = 1; break; }
            if (is_int) out_i64[o] = A->i;
            else out_f64[o] = exact ? exact[g] : A->d;
            break;
          case GHIVE_AGG_AVG:
            if (!A->cnt) { out_null[o] = 1; break; }
            out_f64[o] = (is_int ? (double)A->i : (exact ? exact[g] : A->d)) / (double)A->cnt;
            break;
          default: /* MIN / MAX */
            if (!A->cnt) { out_null[o] = 1; break; }
            if (is_int) out_i64[o] = A->i;
            else out_f64[o] = A->d;
            break;
        }
      }
      free(exact);
    }
  }
  free(slots); free(hashes); free(koff); free(klen); free(kblob); free(reps); free(acc);
  free(grp_of);
  return ret;
}

/* ------------------------------------------------------------------------------------ */
/* ORDER BY as a total order: keys, then row id (sort_comparator, GroupByPredicate.hpp).   */
/* ------------------------------------------------------------------------------------ */
typedef struct {
  const ghive_table *t;
  const uint32_t *keys;
  const char *orders;
  uint32_t nkeys;
} orc_sort_ctx;

static int sort_cmp(const void *pa, const void *pb, void *pc) {
  const orc_sort_ctx *c = (const orc_sort_ctx *)pc;
  uint32_t i = *(const uint32_t *)pa, j = *(const uint32_t *)pb;
  for (uint32_t k = 0; k < c->nkeys; ++k) {
    const ghive_column *col = &c->t->columns[c->keys[k]];
    int asc = c->orders[k] != '-';
    int r = 0;
    if (col->type == GHIVE_STRING) { /* Hive's string order; NULL sorts last ascending */
      const int32_t *ix = (const int32_t *)col->aux;
      int nx = ix[2 * i] < 0, ny = ix[2 * j] < 0;
      if (nx || ny) r = nx == ny ? 0 : (nx ? 1 : -1);
      else r = str_cmp((const uint8_t *)col->data + ix[2 * i], ix[2 * i + 1] - ix[2 * i],
                       (const uint8_t *)col->data + ix[2 * j], ix[2 * j + 1] - ix[2 * j]);
      r = r < 0 ? -1 : (r > 0 ? 1 : 0);
    } else if (col_is_int(col)) {
      int64_t x = col_i64(col, i), y = col_i64(col, j);
      r = x < y ? -1 : (x > y ? 1 : 0);
    } else {
      double x = col_f64(col, i), y = col_f64(col, j);
      int nx = isnan(x), ny = isnan(y);
      if (nx || ny) r = nx == ny ? 0 : (nx ? 1 : -1); /* NULL (NaN) sorts last ascending */
      else r = x < y ? -1 : (x > y ? 1 : 0);
    }
    if (r) return asc ? r : -r;
  }
  return i < j ? -1 : (i > j ? 1 : 0);
}

int orc_sort(const ghive_table *table, const uint32_t *key_cols, const char *orders,
             uint32_t nkeys, uint32_t *out_perm) {
  for (uint32_t k = 0; k < nkeys; ++k) {
    if (key_cols[k] >= table->ncols) return -2;
  }
  for (uint64_t i = 0; i < table->row_num; ++i) out_perm[i] = (uint32_t)i;
  orc_sort_ctx c = {table, key_cols, orders, nkeys};
  qsort_r(out_perm, (size_t)table->row_num, sizeof(uint32_t), sort_cmp, &c);
  return 0;
}

/* ORDER BY as Tez runs it: each of nruns map tasks sorts its contiguous slice of the rows (the
 * map-side sort of the shuffle), then the single ORDER BY reducer merges the sorted runs through a
 * binary heap of run heads. The order is the total order of orc_sort, so the result is the same. */
typedef struct {
  const orc_sort_ctx *c;
  uint32_t *perm;
  uint64_t n;
} orc_run_task;

static void *run_sort_task(void *p) {
  orc_run_task *t = (orc_run_task *)p;
  qsort_r(t->perm, (size_t)t->n, sizeof(uint32_t), sort_cmp, (void *)t->c);
  return NULL;
}

static double secs_between(struct timespec a, struct timespec b) {
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

int orc_sort_runs(const ghive_table *table, const uint32_t *key_cols, const char *orders,
                  uint32_t nkeys, int nruns, uint32_t *out_perm, double *run_secs,
                  double *merge_secs) {
  for (uint32_t k = 0; k < nkeys; ++k) {
    if (key_cols[k] >= table->ncols) return -2;
  }
  uint64_t n = table->row_num;
  if (nruns < 1) nruns = 1;
  if ((uint64_t)nruns > n) nruns = n ? (int)n : 1;
  orc_sort_ctx c = {table, key_cols, orders, nkeys};
  uint32_t *runs = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  uint64_t *beg = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nruns + 1));
  orc_run_task *tasks = (orc_run_task *)malloc(sizeof(orc_run_task) * (size_t)nruns);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nruns);
  for (uint64_t i = 0; i < n; ++i) runs[i] = (uint32_t)i;
  for (int r = 0; r <= nruns; ++r) beg[r] = n * (uint64_t)r / (uint64_t)nruns;
  struct timespec t0, t1, t2;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < nruns; ++r) {
    tasks[r].c = &c;
    tasks[r].perm = runs + beg[r];
    tasks[r].n = beg[r + 1] - beg[r];
    if (nruns == 1) run_sort_task(&tasks[r]);
    else pthread_create(&th[r], NULL, run_sort_task, &tasks[r]);
  }
  if (nruns > 1)
    for (int r = 0; r < nruns; ++r) pthread_join(th[r], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  /* heap[0..hn) holds run indices ordered by their current head row */
  int *heap = (int *)malloc(sizeof(int) * (size_t)nruns);
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)nruns);
  int hn = 0;
  for (int r = 0; r < nruns; ++r) {
    pos[r] = beg[r];
    if (pos[r] == beg[r + 1]) continue;
    int i = hn++;
    heap[i] = r;
    while (i > 0) { /* sift up */
      int p = (i - 1) / 2;
      if (sort_cmp(&runs[pos[heap[i]]], &runs[pos[heap[p]]], &c) >= 0) break;
      int t = heap[i]; heap[i] = heap[p]; heap[p] = t;
      i = p;
    }
  }
  for (uint64_t o = 0; o < n; ++o) {
    int r = heap[0];
    out_perm[o] = runs[pos[r]++];
    if (pos[r] == beg[r + 1]) heap[0] = heap[--hn];
    int i = 0;
    for (;;) { /* sift down */
      int l = 2 * i + 1, m = i;
      if (l < hn && sort_cmp(&runs[pos[heap[l]]], &runs[pos[heap[m]]], &c) < 0) m = l;
      if (l + 1 < hn && sort_cmp(&runs[pos[heap[l + 1]]], &runs[pos[heap[m]]], &c) < 0) m = l + 1;
      if (m == i) break;
      int t = heap[i]; heap[i] = heap[m]; heap[m] = t;
      i = m;
    }
  }
  clock_gettime(CLOCK_MONOTONIC, &t2);
  if (run_secs) *run_secs = secs_between(t0, t1);
  if (merge_secs) *merge_secs = secs_between(t1, t2);
  free(heap); free(pos); free(th); free(tasks); free(beg); free(runs);
  return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Fragment = CS1 map task, run as T Tez-style tasks plus one mergepartial step.           */
/* ------------------------------------------------------------------------------------ */
typedef struct {
  int64_t *keys; /* packed int32 group keys */
  int64_t *acc;  /* naggs accumulators per group */
  int64_t *cnt;  /* naggs non-NULL counts per group */
  int64_t *slot; /* open addressing: group index or -1 */
  int64_t ng, gcap, hcap;
} orc_gtab;

static void gtab_init(orc_gtab *g, uint32_t naggs) {
  g->ng = 0; g->gcap = 256; g->hcap = 1024;
  g->keys = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->gcap);
  g->acc = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->gcap * (naggs ? naggs : 1));
  g->cnt = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->gcap * (naggs ? naggs : 1));
  g->slot = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->hcap);
  for (int64_t i = 0; i < g->hcap; ++i) g->slot[i] = -1;
}
static void gtab_free(orc_gtab *g) { free(g->keys); free(g->acc); free(g->cnt); free(g->slot); }

static int64_t gtab_find(orc_gtab *g, int64_t key, uint32_t naggs, const ghive_fragment_agg *aggs) {
  if (2 * (g->ng + 1) > g->hcap) {
    g->hcap *= 2;
    g->slot = (int64_t *)realloc(g->slot, sizeof(int64_t) * (size_t)g->hcap);
    for (int64_t i = 0; i < g->hcap; ++i) g->slot[i] = -1;
    for (int64_t e = 0; e < g->ng; ++e) {
      int64_t s = (int64_t)orc_hive_long_hash(g->keys[e]) & (g->hcap - 1);
      while (g->slot[s] >= 0) s = (s + 1) & (g->hcap - 1);
      g->slot[s] = e;
    }
  }
  int64_t s = (int64_t)orc_hive_long_hash(key) & (g->hcap - 1);
  for (;;) {
    int64_t e = g->slot[s];
    if (e < 0) break;
    if (g->keys[e] == key) return e;
    s = (s + 1) & (g->hcap - 1);
  }
  if (g->ng == g->gcap) {
    g->gcap *= 2;
    g->keys = (int64_t *)realloc(g->keys, sizeof(int64_t) * (size_t)g->gcap);
    g->acc = (int64_t *)realloc(g->acc, sizeof(int64_t) * (size_t)g->gcap * (naggs ? naggs : 1));
    g->cnt = (int64_t *)realloc(g->cnt, sizeof(int64_t) * (size_t)g->gcap * (naggs ? naggs : 1));
  }
  int64_t e = g->ng++;
  g->keys[e] = key;
  for (uint32_t a = 0; a < naggs; ++a) {
    g->acc[e * naggs + a] = 0;
    g->cnt[e * naggs + a] = 0;
    (void)aggs;
  }
  g->slot[s] = e;
  return e;
}

static inline void agg_update(int type, int64_t *acc, int64_t *cnt, int64_t v) {
  switch (type) {
    case GHIVE_AGG_SUM: *acc = (int64_t)((uint64_t)*acc + (uint64_t)v); break;
    case GHIVE_AGG_CNT: *acc += 1; break;
    case GHIVE_AGG_MAX: if (*cnt == 0 || v > *acc) *acc = v; break;
    case GHIVE_AGG_MIN: if (*cnt == 0 || v < *acc) *acc = v; break;
  }
  *cnt += 1;
}

typedef struct {
  const orc_fragment *f;
  const uint32_t *const *dim_sel; /* filtered dimension rows */
  const uint64_t *dim_nsel;
  uint64_t begin, end;
  orc_gtab g;
  uint64_t joined;
  int failed;
} orc_task;

static void *fragment_task(void *arg) {
  orc_task *T = (orc_task *)arg;
  const orc_fragment *f = T->f;
  uint32_t nd = f->ndims;
  /* each map task loads its own copy of every broadcast table (HashTableLoader) */
  orc_mj_table *ht[16];
  for (uint32_t d = 0; d < nd; ++d) {
    const orc_dim *D = &f->dims[d];
    const ghive_column *pc = D->payload_col >= 0 ? &D->table.columns[D->payload_col] : NULL;
    ht[d] = orc_mj_build(&D->table.columns[D->key_col], T->dim_sel[d], T->dim_nsel[d], pc, 0.75f);
  }
  gtab_init(&T->g, f->naggs);
  int sel[ORC_BATCH];
  int64_t pay[16][ORC_BATCH];
  for (uint64_t base = T->begin; base < T->end; base += ORC_BATCH) {
    int size = (int)((T->end - base) < ORC_BATCH ? (T->end - base) : ORC_BATCH);
    for (int i = 0; i < size; ++i) sel[i] = i;
    if (f->filter_nnodes) size = eval_node(&f->fact, f->filter, f->filter_root, base, sel, size);
    for (uint32_t d = 0; d < nd && size; ++d) {
      const ghive_column *fk = &f->fact.columns[f->dims[d].fact_key_col];
      int n = 0;
      for (int i = 0; i < size; ++i) {
        uint64_t r = base + (uint64_t)sel[i];
        if (col_null(fk, r)) continue;
        int64_t k = col_i64(fk, r);
        if (k < ht[d]->min || k > ht[d]->max) continue;
        int64_t e = mj_find(ht[d], k);
        if (e < 0) continue;
        /* payloads of earlier joins travel with the surviving rows */
        for (uint32_t q = 0; q < d; ++q) pay[q][n] = pay[q][i];
        pay[d][n] = ht[d]->val[ht[d]->head[e]];
        sel[n++] = sel[i];
      }
      size = n;
    }
    T->joined += (uint64_t)size;
    for (int i = 0; i < size; ++i) {
      uint64_t r = base + (uint64_t)sel[i];
      int64_t key = 0;
      for (uint32_t k = 0; k < f->nkeys; ++k) {
        int64_t v = f->keys[k].source == GHIVE_SRC_JOIN ? pay[f->keys[k].index][i]
                                                        : col_i64(&f->fact.columns[f->keys[k].index], r);
        key |= (int64_t)((uint64_t)(uint32_t)(int32_t)v << (32 * k));
      }
      int64_t g = gtab_find(&T->g, key, f->naggs, f->aggs);
      for (uint32_t a = 0; a < f->naggs; ++a) {
        const ghive_fragment_agg *A = &f->aggs[a];
        int64_t v;
        if (A->expr == GHIVE_EXPR_ONE) {
          v = 1;
        } else {
          const ghive_column *ca = &f->fact.columns[A->a];
          if (col_null(ca, r)) continue;
          v = col_i64(ca, r);
          if (A->expr == GHIVE_EXPR_MUL) {
            const ghive_column *cb = &f->fact.columns[A->b];
            if (col_null(cb, r)) continue;
            v = (int64_t)((uint64_t)v * (uint64_t)col_i64(cb, r));
          }
        }
        agg_update(A->type, &T->g.acc[g * f->naggs + a], &T->g.cnt[g * f->naggs + a], v);
      }
    }
  }
  for (uint32_t d = 0; d < nd; ++d) orc_mj_free(ht[d]);
  return NULL;
}

int64_t orc_fragment_run(const orc_fragment *f, int nthreads, int32_t *out_keys,
                         int64_t *out_aggs, uint8_t *out_null, uint64_t capacity,
                         uint64_t *joined, double *seconds) {
  if (f->ndims > 16 || f->nkeys > 2) return -2;
  for (uint32_t d = 0; d < f->ndims; ++d) {
    const orc_dim *D = &f->dims[d];
    if (D->key_col >= D->table.ncols || !col_is_int(&D->table.columns[D->key_col])) return -2;
    if (D->fact_key_col >= f->fact.ncols) return -2;
    if (D->filter_nnodes && !pred_supported(&D->table, D->filter, D->filter_nnodes, D->filter_root, 0))
      return -2;
  }
  if (f->filter_nnodes && !pred_supported(&f->fact, f->filter, f->filter_nnodes, f->filter_root, 0))
    return -2;
  if (nthreads < 1) nthreads = 1;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  /* the dimension vertices: filter each dimension once */
  uint32_t *dsel[16];
  uint64_t dn[16];
  for (uint32_t d = 0; d < f->ndims; ++d) {
    const orc_dim *D = &f->dims[d];
    dsel[d] = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(D->table.row_num ? D->table.row_num : 1));
    if (D->filter_nnodes) {
      dn[d] = (uint64_t)orc_filter(&D->table, D->filter, D->filter_nnodes, D->filter_root, dsel[d]);
    } else {
      for (uint64_t i = 0; i < D->table.row_num; ++i) dsel[d][i] = (uint32_t)i;
      dn[d] = D->table.row_num;
    }
  }
  orc_task *tasks = (orc_task *)calloc((size_t)nthreads, sizeof(orc_task));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  uint64_t rows = f->fact.row_num;
  for (int t = 0; t < nthreads; ++t) {
    tasks[t].f = f;
    tasks[t].dim_sel = (const uint32_t *const *)dsel;
    tasks[t].dim_nsel = dn;
    tasks[t].begin = rows * (uint64_t)t / (uint64_t)nthreads;
    tasks[t].end = rows * (uint64_t)(t + 1) / (uint64_t)nthreads;
    if (nthreads == 1) fragment_task(&tasks[t]);
    else pthread_create(&th[t], NULL, fragment_task, &tasks[t]);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  /* mergepartial */
  orc_gtab m;
  gtab_init(&m, f->naggs);
  uint64_t jn = 0;
  for (int t = 0; t < nthreads; ++t) {
    orc_gtab *g = &tasks[t].g;
    jn += tasks[t].joined;
    for (int64_t e = 0; e < g->ng; ++e) {
      int64_t me = gtab_find(&m, g->keys[e], f->naggs, f->aggs);
      for (uint32_t a = 0; a < f->naggs; ++a) {
        int64_t c = g->cnt[e * f->naggs + a];
        if (!c) continue;
        int64_t *acc = &m.acc[me * f->naggs + a], *cnt = &m.cnt[me * f->naggs + a];
        int64_t v = g->acc[e * f->naggs + a];
        switch (f->aggs[a].type) {
          case GHIVE_AGG_SUM:
          case GHIVE_AGG_CNT: *acc = (int64_t)((uint64_t)*acc + (uint64_t)v); break;
          case GHIVE_AGG_MAX: if (*cnt == 0 || v > *acc) *acc = v; break;
          case GHIVE_AGG_MIN: if (*cnt == 0 || v < *acc) *acc = v; break;
        }
        *cnt += c;
      }
    }
    gtab_free(g);
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  if (joined) *joined = jn;
  int64_t ret = m.ng;
  if ((uint64_t)m.ng > capacity) {
    ret = -1;
  } else {
    for (int64_t e = 0; e < m.ng; ++e) {
      for (uint32_t k = 0; k < f->nkeys; ++k)
        out_keys[e * f->nkeys + k] = (int32_t)(uint32_t)((uint64_t)m.keys[e] >> (32 * k));
      for (uint32_t a = 0; a < f->naggs; ++a) {
        int64_t c = m.cnt[e * f->naggs + a];
        out_aggs[e * f->naggs + a] = f->aggs[a].type == GHIVE_AGG_CNT ? c : m.acc[e * f->naggs + a];
        out_null[e * f->naggs + a] = (f->aggs[a].type != GHIVE_AGG_CNT && c == 0);
      }
    }
  }
  gtab_free(&m);
  for (uint32_t d = 0; d < f->ndims; ++d) free(dsel[d]);
  free(tasks);
  free(th);
  return ret;
}
