/*
 * Test infrastructure: drives every function of the CPU oracle (ghive_oracle.c) over small seeded
 * inputs, edge cases included (empty inputs, NULL sentinels, duplicate keys, STRING columns,
 * capacity overflow), and checks each result against brute force or an invariant. Built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_oracle_asan.py (oracle/Makefile,
 * target oracle_asan); a sanitizer report or a failed check exits non-zero.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ghive_oracle.h"

static int failures = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

static const int64_t LNULL = INT64_MAX;
static const int32_t INULL = INT32_MAX;

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static uint64_t next_u64(void) {  /* xorshift64* */
  rng ^= rng >> 12;
  rng ^= rng << 25;
  rng ^= rng >> 27;
  return rng * 2685821657736338717ull;
}
static int64_t rnd(int64_t lo, int64_t hi) { return lo + (int64_t)(next_u64() % (uint64_t)(hi - lo + 1)); }

static ghive_column col_of(int type, void *data, uint64_t rows) {
  ghive_column c;
  memset(&c, 0, sizeof c);
  c.type = type;
  c.data = data;
  c.row_num = rows;
  return c;
}

static void test_filter(uint64_t n) {
  int64_t *a = malloc(8 * (n ? n : 1));
  double *b = malloc(8 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    a[i] = i % 11 == 0 ? LNULL : rnd(-50, 50);
    b[i] = i % 13 == 0 ? NAN : (double)rnd(-1000, 1000) / 10.0;
  }
  ghive_column cols[2] = {col_of(GHIVE_LONG, a, n), col_of(GHIVE_DOUBLE, b, n)};
  ghive_table t = {cols, 2, 0, n};
  ghive_pred p[3];
  memset(p, 0, sizeof p);
  p[0].mode = GHIVE_FILTER_EQ_RANGE;  /* -10 <= a <= 20 */
  p[0].pred_col_num = 1;
  p[0].filter_col[0] = 0;
  p[0].param_num = 2;
  p[0].data_type = 0;
  p[0].long_params[0] = -10;
  p[0].long_params[1] = 20;
  p[0].children[0] = p[0].children[1] = -1;
  p[1].mode = GHIVE_FILTER_LT;  /* b < 12.5 */
  p[1].pred_col_num = 1;
  p[1].filter_col[0] = 1;
  p[1].param_num = 1;
  p[1].data_type = 1;
  p[1].double_params[0] = 12.5;
  p[1].children[0] = p[1].children[1] = -1;
  p[2].mode = GHIVE_FILTER_AND;
  p[2].pred_col_num = 0;
  p[2].children[0] = 0;
  p[2].children[1] = 1;
  uint32_t *sel = malloc(4 * (n ? n : 1));
  const int64_t k = orc_filter(&t, p, 3, 2, sel);
  int64_t want = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const int pass = a[i] != LNULL && a[i] >= -10 && a[i] <= 20 && !isnan(b[i]) && b[i] < 12.5;
    if (pass) CHECK(want < k && sel[want] == i);
    want += pass;
  }
  CHECK(k == want);
  free(a);
  free(b);
  free(sel);
}

static void test_mapjoin(uint64_t nb, uint64_t np) {
  int64_t *bk = malloc(8 * (nb ? nb : 1)), *bp = malloc(8 * (nb ? nb : 1));
  int64_t *pk = malloc(8 * (np ? np : 1));
  for (uint64_t i = 0; i < nb; ++i) {
    bk[i] = i % 17 == 0 ? LNULL : rnd(0, (int64_t)nb / 2 + 1);  /* duplicates and NULLs */
    bp[i] = (int64_t)i * 3;
  }
  for (uint64_t i = 0; i < np; ++i) pk[i] = i % 19 == 0 ? LNULL : rnd(-5, (int64_t)nb / 2 + 5);
  ghive_column kc = col_of(GHIVE_LONG, bk, nb), pc = col_of(GHIVE_LONG, bp, nb), qc = col_of(GHIVE_LONG, pk, np);
  orc_mj_table *ht = orc_mj_build(&kc, NULL, nb, &pc, 0.75f);
  CHECK(ht != NULL);
  uint64_t want = 0;
  for (uint64_t i = 0; i < np; ++i)
    for (uint64_t j = 0; j < nb; ++j) want += pk[i] != LNULL && bk[j] == pk[i];
  uint64_t needed = 0;
  if (want > 0) {  /* too small a capacity: -1 and the count needed */
    const int64_t r = orc_mj_probe(ht, &qc, NULL, np, NULL, NULL, 0, &needed);
    CHECK(r == -1 && needed == want);
  }
  uint32_t *rows = malloc(4 * (want ? want : 1));
  int64_t *pay = malloc(8 * (want ? want : 1));
  const int64_t got = orc_mj_probe(ht, &qc, NULL, np, rows, pay, want, &needed);
  CHECK(got == (int64_t)want);
  for (int64_t m = 0; m < got; ++m) CHECK(bk[pay[m] / 3] == pk[rows[m]]);
  int64_t keys, buckets, values;
  int32_t steps, expands;
  orc_mj_stats(ht, &keys, &buckets, &values, &steps, &expands);
  CHECK(buckets > 0 && (buckets & (buckets - 1)) == 0 && keys <= buckets);
  orc_mj_free(ht);
  free(bk);
  free(bp);
  free(pk);
  free(rows);
  free(pay);
}

static void test_join(uint64_t n) {
  int32_t *a = malloc(4 * (n ? n : 1)), *b = malloc(4 * (n ? n : 1)), *c = malloc(4 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    a[i] = i % 9 == 0 ? INULL : (int32_t)rnd(0, 30);
    b[i] = (int32_t)rnd(0, 30);
    c[i] = (int32_t)rnd(0, 30);
  }
  ghive_column ca = col_of(GHIVE_INT, a, n), cb = col_of(GHIVE_INT, b, n), cc = col_of(GHIVE_INT, c, n);
  ghive_table t[3] = {{&ca, 1, 0, n}, {&cb, 1, 0, n}, {&cc, 1, 0, n}};
  const uint32_t jc[3] = {0, 0, 0};
  const ghive_join_cond conds[2] = {{GHIVE_INNER_JOIN, 0, 1}, {GHIVE_INNER_JOIN, 1, 2}};
  uint64_t want = 0;
  for (uint64_t i = 0; i < n; ++i)
    for (uint64_t j = 0; j < n; ++j)
      if (a[i] != INULL && a[i] == b[j])
        for (uint64_t k = 0; k < n; ++k) want += b[j] == c[k];
  uint64_t needed = 0;
  int32_t *off = malloc(12 * (want ? want : 1));
  const int64_t got = orc_join(t, 3, jc, conds, 2, off, want, &needed);
  CHECK(got == (int64_t)want);
  for (int64_t r = 0; r < got; ++r) CHECK(a[off[3 * r]] == b[off[3 * r + 1]] && b[off[3 * r + 1]] == c[off[3 * r + 2]]);
  free(a);
  free(b);
  free(c);
  free(off);
}

static void test_groupby(uint64_t n) {
  int32_t *k = malloc(4 * (n ? n : 1));
  int64_t *v = malloc(8 * (n ? n : 1));
  double *d = malloc(8 * (n ? n : 1));
  char *blob = malloc(4 * (n ? n : 1) + 1);
  int32_t *pairs = malloc(8 * (n ? n : 1));
  int64_t total = 0, nonnull = 0;
  for (uint64_t i = 0; i < n; ++i) {
    k[i] = i % 7 == 0 ? INULL : (int32_t)rnd(0, 20);
    v[i] = i % 5 == 0 ? LNULL : rnd(-100, 100);
    d[i] = i % 6 == 0 ? NAN : (double)rnd(-100, 100) * 0.25;
    if (v[i] != LNULL) {
      total += v[i];
      ++nonnull;
    }
    const int len = (int)(i % 4);
    for (int c = 0; c < len; ++c) blob[4 * i + c] = (char)('a' + rnd(0, 2));
    pairs[2 * i] = i % 10 == 0 ? -1 : (int32_t)(4 * i);
    pairs[2 * i + 1] = i % 10 == 0 ? -1 : (int32_t)(4 * i + len);
  }
  ghive_column cols[4] = {col_of(GHIVE_INT, k, n), col_of(GHIVE_LONG, v, n), col_of(GHIVE_DOUBLE, d, n),
                          col_of(GHIVE_STRING, blob, n)};
  cols[3].aux = pairs;
  cols[3].char_size = 4 * n;
  ghive_table t = {cols, 4, 0, n};
  const ghive_agg aggs[6] = {{GHIVE_AGG_SUM, 1}, {GHIVE_AGG_CNT, GHIVE_COUNT_STAR}, {GHIVE_AGG_MIN, 1},
                             {GHIVE_AGG_MAX, 2}, {GHIVE_AGG_AVG, 2}, {GHIVE_AGG_CNT, 1}};
  const uint64_t cap = n + 1;
  uint32_t *rep = malloc(4 * cap);
  int64_t *oi = malloc(8 * 6 * cap);
  double *of = malloc(8 * 6 * cap);
  uint8_t *on = malloc(6 * cap);
  for (int pass = 0; pass < 2; ++pass) {  /* an INT key, then an INT + STRING composite key */
    const uint32_t keys[2] = {0, 3};
    const int64_t g = orc_groupby(&t, NULL, n, keys, pass ? 2 : 1, aggs, 6, 1, rep, oi, of, on, cap);
    CHECK(g >= 0 && (uint64_t)g <= n + (n == 0));
    int64_t cnt = 0, sum = 0, cnt_nonnull = 0;
    for (int64_t x = 0; x < g; ++x) {
      cnt += oi[6 * x + 1];
      cnt_nonnull += oi[6 * x + 5];
      if (!on[6 * x]) sum += oi[6 * x];
    }
    CHECK(cnt == (int64_t)n && sum == total && cnt_nonnull == nonnull);
    if (g > 1) CHECK(orc_groupby(&t, NULL, n, keys, pass ? 2 : 1, aggs, 6, 1, rep, oi, of, on, 1) == -1);
  }
  free(k);
  free(v);
  free(d);
  free(blob);
  free(pairs);
  free(rep);
  free(oi);
  free(of);
  free(on);
}

static void test_exact_sum(void) {
  const double x[5] = {1e100, 1.0, -1e100, 0.5, 1e-300};
  CHECK(orc_exact_sum(x, 5) == 1.5);
  CHECK(orc_exact_sum(x, 0) == 0.0);
}

static void test_sort(uint64_t n) {
  int64_t *a = malloc(8 * (n ? n : 1));
  double *b = malloc(8 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {
    a[i] = i % 9 == 0 ? LNULL : rnd(-20, 20);
    b[i] = i % 8 == 0 ? NAN : (double)rnd(-5, 5);
  }
  ghive_column cols[2] = {col_of(GHIVE_LONG, a, n), col_of(GHIVE_DOUBLE, b, n)};
  ghive_table t = {cols, 2, 0, n};
  const uint32_t keys[2] = {0, 1};
  uint32_t *perm = malloc(4 * (n ? n : 1));
  CHECK(orc_sort(&t, keys, "+-", 2, perm) == 0);
  uint8_t *seen = calloc(n ? n : 1, 1);
  for (uint64_t i = 0; i < n; ++i) {
    CHECK(perm[i] < n && !seen[perm[i]]);
    if (perm[i] < n) seen[perm[i]] = 1;
    if (i) CHECK(a[perm[i - 1]] <= a[perm[i]]);  /* LONG NULL sorts as its sentinel: last */
  }
  uint32_t *perm2 = malloc(4 * (n ? n : 1));
  double rs = -1, ms = -1;
  for (int runs = 1; runs <= 5; runs += 2) { /* run sorts + merge: the same permutation */
    CHECK(orc_sort_runs(&t, keys, "+-", 2, runs, perm2, &rs, &ms) == 0);
    CHECK(rs >= 0 && ms >= 0);
    for (uint64_t i = 0; i < n; ++i) CHECK(perm2[i] == perm[i]);
  }
  free(perm2);
  free(a);
  free(b);
  free(perm);
  free(seen);
}

static void test_fragment(uint64_t n) {
  const uint64_t nd = 200;
  int32_t *fk = malloc(4 * (n ? n : 1)), *m = malloc(4 * (n ? n : 1)), *p = malloc(4 * (n ? n : 1));
  int32_t *dk = malloc(4 * nd), *da = malloc(4 * nd), *dc = malloc(4 * nd);
  for (uint64_t i = 0; i < n; ++i) {
    fk[i] = i % 23 == 0 ? INULL : (int32_t)rnd(1, (int64_t)nd + 3);
    m[i] = (int32_t)rnd(1, 100);
    p[i] = (int32_t)rnd(1000, 100000);
  }
  for (uint64_t i = 0; i < nd; ++i) {
    dk[i] = (int32_t)(i + 1);
    da[i] = (int32_t)rnd(0, 6);
    dc[i] = (int32_t)rnd(0, 24);
  }
  ghive_column fc[3] = {col_of(GHIVE_INT, fk, n), col_of(GHIVE_INT, m, n), col_of(GHIVE_INT, p, n)};
  ghive_column dcols[3] = {col_of(GHIVE_INT, dk, nd), col_of(GHIVE_INT, da, nd), col_of(GHIVE_INT, dc, nd)};
  ghive_pred dp;
  memset(&dp, 0, sizeof dp);
  dp.mode = GHIVE_FILTER_EQ;
  dp.pred_col_num = 1;
  dp.filter_col[0] = 1;
  dp.param_num = 1;
  dp.data_type = 2;
  dp.int_params[0] = 3;
  dp.children[0] = dp.children[1] = -1;
  orc_dim dim = {{dcols, 3, 0, nd}, &dp, 1, 0, 0, 2, 0};
  const ghive_fragment_key key = {GHIVE_SRC_JOIN, 0};
  const ghive_fragment_agg agg = {GHIVE_AGG_SUM, GHIVE_EXPR_MUL, 1, 2};
  orc_fragment f;
  memset(&f, 0, sizeof f);
  f.fact = (ghive_table){fc, 3, 0, n};
  f.dims = &dim;
  f.ndims = 1;
  f.keys = &key;
  f.nkeys = 1;
  f.aggs = &agg;
  f.naggs = 1;
  int32_t k1[64], k3[64];
  int64_t a1[64], a3[64];
  uint8_t n1[64], n3[64];
  uint64_t j1 = 0, j3 = 0;
  double s;
  const int64_t g1 = orc_fragment_run(&f, 1, k1, a1, n1, 64, &j1, &s);
  const int64_t g3 = orc_fragment_run(&f, 3, k3, a3, n3, 64, &j3, &s);
  CHECK(g1 >= 0 && g1 == g3 && j1 == j3);
  int64_t t1 = 0, t3 = 0, want = 0;
  for (int64_t x = 0; x < g1; ++x) t1 += a1[x];
  for (int64_t x = 0; x < g3; ++x) t3 += a3[x];
  for (uint64_t i = 0; i < n; ++i)
    if (fk[i] != INULL && fk[i] >= 1 && (uint64_t)fk[i] <= nd && da[fk[i] - 1] == 3) want += (int64_t)m[i] * p[i];
  CHECK(t1 == want && t3 == want);
  free(fk);
  free(m);
  free(p);
  free(dk);
  free(da);
  free(dc);
}

int main(void) {
  const uint64_t sizes[4] = {0, 1, 37, 2000};
  for (int s = 0; s < 4; ++s) {
    test_filter(sizes[s]);
    test_mapjoin(sizes[s], sizes[s] + 5);
    test_join(sizes[s] < 200 ? sizes[s] : 200);
    test_groupby(sizes[s]);
    test_sort(sizes[s]);
    test_fragment(sizes[s]);
  }
  test_exact_sum();
  uint32_t out[4];
  const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
  orc_philox4x32_10(ctr, key, out);
  CHECK(out[0] == 0x6627e8d5u);  /* Random123 known answer (Salmon et al., SC'11) */
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle under the sanitizers: all checks passed\n");
  return 0;
}
