/*
 * ghive_oracle.h -- CPU ORACLE for the GHive offloaded-operator path. TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or the timed CPU baseline. The product (libghive_mi355x.so) never calls it.
 *
 * It restates, in plain C, Apache Hive 3.1's vectorized CPU operators -- the semantics GHive's
 * GPU path must reproduce -- plus the reference GPU library's operator contracts:
 *   filter    hive ql/.../exec/vector/VectorFilterOperator.java:110-130,
 *             ql/src/gen/vectorization/ExpressionTemplates/FilterColumnCompareScalar.txt:73-101,
 *             FilterColumnBetween.txt:39-48,89-98, PredicateMode gpu_impl/inc/Operator/FilterPredicate.hpp:17-58
 *   map join  ql/.../vector/mapjoin/VectorMapJoinInnerLongOperator.java:104-340,
 *             ql/.../vector/mapjoin/fast/VectorMapJoinFastLongHashTable.java:99-268,
 *             common/src/java/org/apache/hive/common/util/HashCodeUtil.java:36-45
 *   group-by  ql/.../exec/vector/VectorGroupByOperator.java:282-650,
 *             ql/src/gen/vectorization/UDAFTemplates/VectorUDAFSum.txt:48-81
 *   sort      gpu_impl/src/Operator/SinkOperator.cu:135-288 with sort_comparator
 *             gpu_impl/inc/Operator/GroupByPredicate.hpp:22-101 (restated as a total order)
 *   join N-way gpu_impl/src/Operator/SortMergeJoinPredicate.cu:8-461 (result_offsets contract)
 *
 * Parity pins for each function are listed in tests/ (golden KATs, brute force, closed forms).
 */
#ifndef GHIVE_ORACLE_H
#define GHIVE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/ghive/ghive.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter (row_lo, row_hi, column, 0),
 * key (seed_lo, seed_hi). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_gen_uniform_i32(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         int32_t lo, int32_t hi, int32_t *out);
void orc_gen_uniform_i64(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         int64_t lo, int64_t hi, int64_t *out);
void orc_gen_uniform_f32(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                         float *out);
void orc_gen_normal_f64(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                        double *out);

/* Skewed-join workload generators (same definitions as the device generators). */
uint64_t orc_key62(uint64_t seed, uint64_t rank);
void orc_gen_zipf_keys(uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n, uint64_t nranks,
                       double theta, double miss_frac, int64_t *out_keys, uint64_t *out_ranks);

/* HashCodeUtil.calculateLongHashCode (HashCodeUtil.java:36-45). */
int32_t orc_hive_long_hash(int64_t key);

/* Filter: ids of passing rows in ascending order; returns the count. Columns are host. */
int64_t orc_filter(const ghive_table *table, const ghive_pred *nodes, uint32_t nnodes,
                   uint32_t root, uint32_t *out_sel);

/* Hive VectorMapJoinFastLongHashMap restated (multi-valued long-key map). */
typedef struct orc_mj_table orc_mj_table;
orc_mj_table *orc_mj_build(const ghive_column *key, const uint32_t *sel, uint64_t n,
                           const ghive_column *payload, float load_factor);
void orc_mj_free(orc_mj_table *t);
void orc_mj_stats(const orc_mj_table *t, int64_t *keys, int64_t *buckets, int64_t *values,
                  int32_t *largest_steps, int32_t *expands);
/* Inner probe in Hive batch order; returns the number of (probe row, payload) pairs or -1
 * when more than `capacity` (the count is then written to *needed). */
int64_t orc_mj_probe(const orc_mj_table *t, const ghive_column *key, const uint32_t *sel,
                     uint64_t n, uint32_t *out_probe_rows, int64_t *out_payload,
                     uint64_t capacity, uint64_t *needed);

/* N-way join with the reference's result_offsets contract. Returns cardinality or -1 on
 * capacity overflow (*needed set) / -2 on unsupported input. */
int64_t orc_join(const ghive_table *tables, uint32_t ntables, const uint32_t *join_cols,
                 const ghive_join_cond *conds, uint32_t nconds, int32_t *out_offsets,
                 uint64_t capacity, uint64_t *needed);

/* Hash group-by (Hive semantics). Keys may be INT/LONG/DOUBLE/FLOAT/STRING. Output per group
 * g (first-appearance order): rep_row[g] = first input row of the group (to gather keys);
 * for aggregate a: out_i64[g*naggs+a] (LONG outputs) or out_f64[g*naggs+a] (DOUBLE outputs)
 * and out_null[g*naggs+a] = 1 when the aggregate is NULL. exact_double_sum: 1 = correctly
 * rounded SUM/AVG of doubles (Shewchuk / fsum), 0 = Hive's arrival-order double sum.
 * Returns groups, or -1 if more than capacity. */
int64_t orc_groupby(const ghive_table *table, const uint32_t *sel, uint64_t n,
                    const uint32_t *keys, uint32_t nkeys, const ghive_agg *aggs, uint32_t naggs,
                    int exact_double_sum, uint32_t *rep_row, int64_t *out_i64, double *out_f64,
                    uint8_t *out_null, uint64_t capacity);

/* Correctly rounded sum of doubles (Shewchuk 1997 partials + round-half-even correction). */
double orc_exact_sum(const double *x, uint64_t n);

/* ORDER BY: permutation in key order with row-id tie-break. orders[k] is '+' or '-'. */
int orc_sort(const ghive_table *table, const uint32_t *key_cols, const char *orders,
             uint32_t nkeys, uint32_t *out_perm);

/* The same ORDER BY as nruns map-side run sorts (one thread each) and one k-way merge (the single
 * ORDER BY reducer); the wall time of each phase is returned. Used as the multi-core CPU baseline. */
int orc_sort_runs(const ghive_table *table, const uint32_t *key_cols, const char *orders,
                  uint32_t nkeys, int nruns, uint32_t *out_perm, double *run_secs,
                  double *merge_secs);

/* Fragment (CS1 map task): fact filter -> chain of broadcast map joins -> GROUP BY, run as
 * T "map tasks" over contiguous fact slices, each with its own copy of every dimension hash
 * table and its own partial aggregates, then one merge (the "mergepartial" reducer).
 * This is the CPU baseline: a C restatement of Hive 3.1 vectorized operators. */
typedef struct orc_dim {
  ghive_table table;        /* host columns of the dimension                          */
  const ghive_pred *filter; /* dimension filter (nnodes 0 = none)                      */
  uint32_t filter_nnodes;
  uint32_t filter_root;
  uint32_t key_col;         /* dimension key column (INT/LONG, unique)                  */
  int32_t payload_col;      /* dimension attribute carried into the join (-1 = none)    */
  uint32_t fact_key_col;    /* fact column probed                                       */
} orc_dim;

typedef struct orc_fragment {
  ghive_table fact;
  const ghive_pred *filter;
  uint32_t filter_nnodes;
  uint32_t filter_root;
  const orc_dim *dims;
  uint32_t ndims;
  const ghive_fragment_key *keys; /* at most 2, INT */
  uint32_t nkeys;
  const ghive_fragment_agg *aggs;
  uint32_t naggs;
} orc_fragment;

/* Returns groups (or -1 on capacity). out_keys[g*nkeys+k] (int32), out_aggs[g*naggs+a]
 * (int64), out_null[g*naggs+a]. *seconds = wall time of the operator pipeline (hash-table
 * loads included, input generation excluded). *joined = rows that survive every join. */
int64_t orc_fragment_run(const orc_fragment *f, int nthreads, int32_t *out_keys,
                         int64_t *out_aggs, uint8_t *out_null, uint64_t capacity,
                         uint64_t *joined, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
