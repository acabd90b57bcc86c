/*
 * ghive.h -- C ABI of libghive_mi355x.so, the MI355X (gfx950) implementation of the
 * operators GHive offloads from Hive's ql operator tree: filter with stream compaction,
 * hash-join build and probe, hash group-by aggregation and key sort, plus the fused
 * filter -> join -> group-by fragment those operators form inside one Tez map task.
 *
 * Reference interfaces replaced (paths relative to the GHive tree):
 *   gpu_impl/inc/ghive_JNIInterface.h:17-18   Java_..._JNIInterface_GPUProcess (JNI bridge; the
 *                                             JNI shim binds to the entry points below)
 *   gpu_impl/inc/DataFlow/Column.hpp:5-35     ColumnType / Column       -> ghive_column
 *   gpu_impl/inc/DataFlow/Table.hpp:7-53      Table, null_value/is_null -> ghive_table, sentinels
 *   gpu_impl/inc/Operator/FilterPredicate.hpp:17-90   PredicateMode / FilterPredicate -> ghive_pred
 *   gpu_impl/src/Operator/FilterPredicate.cu:434      FilterPredicate::process        -> ghive_filter
 *   gpu_impl/src/Operator/FilterOperator.cu:36-93     FilterOperator::clean_data      -> ghive_filter_compact
 *   gpu_impl/src/Operator/FilterOperator.cu:227-487   FilterOperator::generatePredicate -> ghive_parse_predicate
 *   gpu_impl/inc/Operator/BaseJoinPredicate.hpp:11-88 BaseJoinType / BaseJoinPredicate::process -> ghive_join
 *   gpu_impl/inc/Operator/AggregationDesc.hpp:6-36    AggregationType / AggregationDesc -> ghive_agg
 *   gpu_impl/inc/Operator/GroupByPredicate.hpp:172    GroupByPredicate::gpu_execute   -> ghive_groupby
 *   gpu_impl/src/Operator/SinkOperator.cu:135-288     SinkOperator ORDER BY sort      -> ghive_sort
 *
 * Conventions
 *   - Every entry point returns an int status (GHIVE_OK == 0). On failure the message is
 *     available from ghive_last_error(ctx). Nothing is printed.
 *   - Column data pointers are DEVICE pointers (hipMalloc / ghive_malloc / any HIP allocation)
 *     unless a parameter says otherwise. Counts returned through uint64_t* are host pointers.
 *   - All work of a context is ordered on the context's stream; entry points that return a
 *     count or a group number synchronise that stream before returning.
 *   - A context is a handle: calls on different contexts may run concurrently from different
 *     host threads; calls on one context must not overlap. There are no global statics.
 *   - NULL is an in-band sentinel, as in the reference (Table.hpp:29-52): LONG_MAX for LONG,
 *     INT_MAX for INT, NaN for DOUBLE/FLOAT, (start,end) = (-1,-1) for STRING.
 *   - Row ids are uint32 (the reference's row_num is uint32); tables must have < 2^32-1 rows.
 */
#ifndef GHIVE_GHIVE_H
#define GHIVE_GHIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GHIVE_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------------------- */
enum {
  GHIVE_OK = 0,
  GHIVE_ERR_INVALID = 1,     /* malformed descriptor or argument                           */
  GHIVE_ERR_UNSUPPORTED = 2, /* legal in the reference, not implemented by this library    */
  GHIVE_ERR_OOM = 3,         /* device memory manager could not satisfy an allocation       */
  GHIVE_ERR_CAPACITY = 4,    /* caller's output buffer too small: *out_count holds the need */
  GHIVE_ERR_HIP = 5,         /* a HIP runtime call failed                                  */
  GHIVE_ERR_ADMISSION = 6,   /* serving: request larger than the HBM admission budget      */
  GHIVE_ERR_NO_DEVICE = 7    /* no usable gfx950 device                                    */
};

/* ---- column types: ColumnType codes (gpu_impl/inc/DataFlow/Column.hpp:5-11) ------------ */
enum {
  GHIVE_LONG = 0,   /* int64                                                       */
  GHIVE_DOUBLE = 1, /* IEEE binary64                                               */
  GHIVE_INT = 2,    /* int32                                                       */
  GHIVE_STRING = 3, /* UTF-8 blob + int32 (start,end) per row                      */
  GHIVE_FLOAT = 4   /* IEEE binary32 -- extension: the reference has no FLOAT type  */
};

/* Column (gpu_impl/inc/DataFlow/Column.hpp:13-35). data/aux are device pointers. */
typedef struct ghive_column {
  int32_t type;       /* GHIVE_LONG .. GHIVE_FLOAT                      */
  uint32_t reserved;  /* must be 0                                      */
  void *data;         /* row_num values (STRING: char blob)             */
  void *aux;          /* STRING only: 2*row_num int32 (start,end) pairs */
  uint64_t row_num;
  uint64_t char_size; /* STRING only                                    */
} ghive_column;

/* Table (gpu_impl/inc/DataFlow/Table.hpp:7-53). */
typedef struct ghive_table {
  const ghive_column *columns;
  uint32_t ncols;
  uint32_t key_num; /* keys-first column count, as Table::key_num */
  uint64_t row_num;
} ghive_table;

/* ---- context, errors, streams ------------------------------------------------------- */
typedef struct ghive_ctx ghive_ctx;

/* Opens device `device`. mem_limit_bytes caps the device memory manager (0 = 90% of the
 * HBM free at init). Replaces the implicit CUDA context of the reference library. */
int ghive_init(int device, uint64_t mem_limit_bytes, ghive_ctx **out);
void ghive_shutdown(ghive_ctx *ctx);
const char *ghive_last_error(const ghive_ctx *ctx);
int ghive_abi_version(void);
/* Order all work of the context on a caller-owned hipStream_t (NULL = the HIP null stream).
 * A new context uses its own non-blocking stream; ghive_use_own_stream goes back to it. */
int ghive_set_stream(ghive_ctx *ctx, void *hip_stream);
int ghive_use_own_stream(ghive_ctx *ctx);
void *ghive_get_stream(ghive_ctx *ctx);
int ghive_sync(ghive_ctx *ctx);

typedef struct ghive_device_info {
  char name[64];
  char arch[32];          /* e.g. "gfx950"                 */
  int32_t compute_units;
  int32_t wavefront_size;
  uint64_t lds_per_block; /* bytes                          */
  uint64_t hbm_bytes;
  uint64_t l2_bytes;
  int32_t clock_khz;
  int32_t mem_clock_khz;
  int32_t mem_bus_width;
} ghive_device_info;
int ghive_device_query(ghive_ctx *ctx, ghive_device_info *out);

/* ---- device memory manager (the reference has none: gpu_impl leaks every cudaMalloc) -- */
int ghive_malloc(ghive_ctx *ctx, uint64_t bytes, void **out);
int ghive_free(ghive_ctx *ctx, void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. Synchronous w.r.t. the stream. */
int ghive_memcpy(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind);
int ghive_mem_stats(ghive_ctx *ctx, uint64_t *in_use, uint64_t *peak, uint64_t *limit,
                    uint64_t *reserved);
/* STREAM-style probe of achievable HBM bandwidth (no counterpart in the reference): mode 0
 * reads `bytes` from src, mode 1 copies src -> dst, mode 2 reads src as five equal columns with
 * every workgroup streaming the same row range of each (the fused star-join kernel's pattern;
 * bytes a multiple of 80); 16-byte aligned, iters launches, mean time per launch from HIP events.
 * Calibrates roofline fractions and PMC byte counters. */
int ghive_bench_stream(ghive_ctx *ctx, void *dst, const void *src, uint64_t bytes, int mode,
                       int iters, float *ms_per_iter);

/* ---- synthetic data: counter-based Philox4x32-10 (Salmon et al., SC'11) -------------- *
 * Value of row r in column c is a pure function of (seed, c, r), so any slice can be     *
 * regenerated anywhere. Uniform integers: lo + floor(u32 * (hi-lo+1) / 2^32).            */
int ghive_gen_uniform_i32(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, int32_t lo, int32_t hi, int32_t *out);
int ghive_gen_uniform_i64(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, int64_t lo, int64_t hi, int64_t *out);
/* out[i] = start + row_begin + i (dense unique dimension keys). */
int ghive_gen_sequence_i32(ghive_ctx *ctx, uint64_t row_begin, uint64_t n, int32_t start,
                           int32_t *out);
int ghive_gen_sequence_i64(ghive_ctx *ctx, uint64_t row_begin, uint64_t n, int64_t start,
                           int64_t *out);
/* Uniform [0,1) float32 with 24 random bits; standard normal float64 (Box-Muller). */
int ghive_gen_uniform_f32(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                          uint64_t n, float *out);
int ghive_gen_normal_f64(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin,
                         uint64_t n, double *out);
/* Unique 62-bit keys: out[i] = K(seed, rank_begin + i), K a bijection of [0, 2^62) (so keys
 * never equal LONG_MAX, the NULL sentinel). */
int ghive_gen_rank_keys(ghive_ctx *ctx, uint64_t seed, uint64_t rank_begin, uint64_t n, int64_t *out);
/* Skewed probe keys: with probability miss_frac the key of a rank above nranks (absent from a
 * build side holding ranks 1..nranks), else the key of a Zipf(theta) rank over [1, nranks]
 * drawn by rejection-inversion (Hoermann & Derflinger 1996). */
int ghive_gen_zipf_keys(ghive_ctx *ctx, uint64_t seed, uint32_t column_id, uint64_t row_begin, uint64_t n,
                        uint64_t nranks, double theta, double miss_frac, int64_t *out);

/* ---- filter ------------------------------------------------------------------------- */
/* PredicateMode (gpu_impl/inc/Operator/FilterPredicate.hpp:17-58), same numbering. */
enum {
  GHIVE_FILTER_EQ = 0,
  GHIVE_FILTER_LE = 1,
  GHIVE_FILTER_GE = 2,
  GHIVE_FILTER_LT = 3,
  GHIVE_FILTER_GT = 4,
  GHIVE_FILTER_RANGE = 5,        /* lo <  x <  hi                 */
  GHIVE_FILTER_EQ_RANGE = 6,     /* lo <= x <= hi   (BETWEEN)     */
  GHIVE_FILTER_NOT_EQ = 7,
  GHIVE_FILTER_NOT_RANGE = 8,    /* x <  lo || x >  hi (NOT BETWEEN) */
  GHIVE_FILTER_NOT_EQ_RANGE = 9, /* x <= lo || x >= hi             */
  GHIVE_FILTER_NOT_NULL = 10,
  GHIVE_FILTER_AND = 11,
  GHIVE_FILTER_OR = 12
};

/* One node of a predicate tree: the fields of FilterPredicate (FilterPredicate.hpp:61-90).
 * pred_col_num 0 = AND/OR over children[0..1]; 1 = column vs literal(s); 2 = column vs
 * column (filter_col[0] OP filter_col[1]). data_type: 0 long, 1 double, 2 int, 3 string,
 * 4 none. The comparison is done on values (an INT column against a long or double
 * literal compares the converted values); a NULL operand makes the node false. A STRING
 * column compares with string literals (data_type 3) or another STRING column in Hive's order:
 * unsigned bytes, then length (StringExpr.compare); all modes but RANGE / NOT_EQ_RANGE. */
#define GHIVE_PRED_STR_MAX 64
typedef struct ghive_pred {
  int32_t mode;
  int32_t pred_col_num;
  uint32_t filter_col[2];
  int32_t param_num;
  int32_t data_type;
  int64_t long_params[2];
  int32_t int_params[2];
  double double_params[2];
  int32_t children[2];                          /* node indexes (AND/OR), else -1 */
  int32_t string_sizes[2];                      /* data_type 3: literal lengths in bytes */
  char string_params[2][GHIVE_PRED_STR_MAX];    /* data_type 3: literal bytes (stringFilterParams) */
} ghive_pred;

/* FilterOperator::generatePredicate (FilterOperator.cu:227-487): parses Hive's EXPLAIN
 * predicate text against the operator's output column names. Host-only, no device. */
int ghive_parse_predicate(const char *text, const char *const *col_names, uint32_t ncols,
                          ghive_pred *nodes, uint32_t capacity, uint32_t *nnodes,
                          uint32_t *root, char *err, size_t err_cap);
/* FilterPredicate::toString(0) (FilterPredicate.cu:670-710). Returns the length needed. */
size_t ghive_predicate_to_string(const ghive_pred *nodes, uint32_t nnodes, uint32_t root,
                                 char *buf, size_t cap);

/* FilterPredicate::process + FilterOperator::clean_data, selection form: writes the ids of
 * the rows that pass, in ascending order, to out_sel (device, capacity row_num). */
int ghive_filter(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes,
                 uint32_t nnodes, uint32_t root, uint32_t *out_sel, uint64_t *out_count);
/* ghive_filter with no wait on the host (no counterpart in the reference, whose operators each
 * return to the host): the passing rows' ids go to out_sel as above and their count to
 * *count_dev (device memory, 8 bytes), ready for a consumer on the same stream such as
 * ghive_hash_build_async. */
int ghive_filter_async(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes,
                       uint32_t nnodes, uint32_t root, uint32_t *out_sel, uint64_t *count_dev);
/* clean_data compaction form: every column of `table` gathered by the passing rows into
 * out_cols[i].data (device buffers of capacity table->row_num, caller-owned). Unlike the
 * reference the input is not modified in place. A STRING column gathers its (start,end) pairs
 * into out_cols[i].aux (2 x row_num int32) and shares the input blob (out data = input data). */
int ghive_filter_compact(ghive_ctx *ctx, const ghive_table *table, const ghive_pred *nodes,
                         uint32_t nnodes, uint32_t root, ghive_column *out_cols,
                         uint64_t *out_count);
/* Gather rows `sel[0..n)` of every column into out_cols (device); STRING as ghive_filter_compact. */
int ghive_gather(ghive_ctx *ctx, const ghive_table *table, const uint32_t *sel, uint64_t n,
                 ghive_column *out_cols);

/* ---- hash join ---------------------------------------------------------------------- */
/* BaseJoinType (gpu_impl/inc/Operator/BaseJoinPredicate.hpp:11-25), same numbering. */
enum {
  GHIVE_INNER_JOIN = 0,
  GHIVE_LEFT_OUTER_JOIN = 1,
  GHIVE_RIGHT_OUTER_JOIN = 2, /* UNSUPPORTED (empty stub in the reference) */
  GHIVE_FULL_OUTER_JOIN = 3,  /* UNSUPPORTED (empty stub in the reference) */
  GHIVE_LEFT_SEMI_JOIN = 4,
  GHIVE_CROSS_JOIN = 5        /* UNSUPPORTED (empty stub in the reference) */
};

typedef struct ghive_hash_table ghive_hash_table;

/* Open-addressing hash table on device over an INT or LONG key column. Rows `sel[0..n)`
 * (sel == NULL: all rows) are inserted; NULL keys are skipped (they never join). Each slot
 * carries a payload: the value of `payload` (INT/LONG column, may be NULL) at that row, or
 * the build row id when payload == NULL. load_factor in (0, 0.95]; 0 = 0.5. */
int ghive_hash_build(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel, uint64_t n,
                     const ghive_column *payload, float load_factor, ghive_hash_table **out);
/* ghive_hash_build over rows sel[0 .. *n_dev) where the count is on the device (at most n_max,
 * e.g. from ghive_filter_async), with no wait on the host: the table is sized for n_max and its
 * statistics (entry count, key range, dense companion) are read when it is first used (probe,
 * info, fragment, batch), so that a chain filter -> build -> fused fragment waits once. key,
 * payload, sel and n_dev must stay valid until that first use. No counterpart in the reference
 * (HashTableLoader builds from rows already on the host). */
int ghive_hash_build_async(ghive_ctx *ctx, const ghive_column *key, const uint32_t *sel,
                           const uint64_t *n_dev, uint64_t n_max, const ghive_column *payload,
                           float load_factor, ghive_hash_table **out);
/* One build side of ghive_hash_build_filtered: a table's filter (as ghive_filter takes it; nnodes
 * 0: every row) and the hash build over the rows that pass. */
typedef struct ghive_filtered_build {
  const ghive_table *table;  /* device columns */
  const ghive_pred *nodes;
  uint32_t nnodes;
  uint32_t root;
  uint32_t key_col;     /* INT or LONG column of table */
  int32_t payload_col;  /* INT or LONG column of table, or -1: the build row id */
  float load_factor;    /* <= 0: 0.5 */
  uint32_t reserved;
} ghive_filtered_build;
#define GHIVE_MAX_FILTERED_BUILDS 8
/* Filters and builds n (1..GHIVE_MAX_FILTERED_BUILDS) tables at once: one pass over each table
 * evaluates the filter and inserts the passing rows with non-NULL keys, and every table's fill
 * and statistics run as one launch for all of them, with no wait on the host. Each out[i] equals
 * ghive_hash_build_async over ghive_filter_async's selection of jobs[i] (same entries, payloads,
 * statistics and dense companion; sized for the table's rows) and is finished the same way, on
 * its first use. The tables, their predicates' columns and the key and payload columns must stay
 * valid until that first use. No counterpart in the reference, which filters on the device and
 * builds its map-join tables on the host (FilterOperator.cu:36-93, exec/tez/HashTableLoader.java:277,440-468). */
int ghive_hash_build_filtered(ghive_ctx *ctx, const ghive_filtered_build *jobs, uint32_t n,
                              ghive_hash_table **out);
int ghive_hash_table_info(const ghive_hash_table *ht, uint64_t *entries, uint64_t *slots,
                          uint64_t *bytes, int32_t *unique_keys);
void ghive_hash_table_free(ghive_ctx *ctx, ghive_hash_table *ht);
/* Inner-join probe: for every probe row (sel or all) and every matching build entry emits
 * (probe row id, payload). Order is unspecified (compare as a multiset). out_* are device
 * arrays of `capacity`; on GHIVE_ERR_CAPACITY *out_count is the number needed. */
int ghive_hash_probe(ghive_ctx *ctx, const ghive_hash_table *ht, const ghive_column *key,
                     const uint32_t *sel, uint64_t n, uint32_t *out_probe_rows,
                     int64_t *out_payload, uint64_t capacity, uint64_t *out_count);

/* One entry of join_types (BaseJoinPredicate.hpp:66-72): join `left` to `right`. */
typedef struct ghive_join_cond {
  int32_t type;
  uint32_t left;
  uint32_t right;
} ghive_join_cond;

/* BaseJoinPredicate::process (SortMergeJoinPredicate.cu:8-461), as hash joins: tables[t]
 * joins on its column join_cols[t] (INT or LONG). Output is result_offsets: per result
 * row, one int32 row id per non-semi table in table order (-1 = NULL side of a LEFT
 * OUTER join), row-major, `capacity` rows. Rows come in unspecified order. */
int ghive_join(ghive_ctx *ctx, const ghive_table *tables, uint32_t ntables,
               const uint32_t *join_cols, const ghive_join_cond *conds, uint32_t nconds,
               int32_t *out_offsets, uint64_t capacity, uint64_t *out_cardinality);

/* BaseJoinPredicate::generate_result (BaseJoinPredicate.cu:4-97): the join's output table from
 * ghive_join's result_offsets (`cardinality` rows, same tables and conds). For each result table
 * (the non-semi inputs in table order) its columns maintain_cols[t][0..maintain_counts[t]) are
 * gathered, in that order; when the first result table maintains nothing its column 0 is kept,
 * as in the reference. An offset of -1 gives the type's NULL sentinel; a STRING column gathers its
 * (start,end) pairs into out aux ((-1,-1) for NULL) and shares the input blob (out data = input
 * data). out_cols: caller device buffers of `cardinality` rows; *out_ncols = the column count
 * (GHIVE_ERR_CAPACITY when it exceeds out_capacity_cols). */
int ghive_join_result(ghive_ctx *ctx, const ghive_table *tables, uint32_t ntables,
                      const ghive_join_cond *conds, uint32_t nconds,
                      const uint32_t *const *maintain_cols, const uint32_t *maintain_counts,
                      const int32_t *offsets, uint64_t cardinality, ghive_column *out_cols,
                      uint32_t out_capacity_cols, uint32_t *out_ncols);

/* ---- group-by ----------------------------------------------------------------------- */
/* AggregationType (gpu_impl/inc/Operator/AggregationDesc.hpp:6-27), same numbering. */
enum {
  GHIVE_AGG_SUM = 0,
  GHIVE_AGG_MAX = 1,
  GHIVE_AGG_MIN = 2,
  GHIVE_AGG_AVG = 3,
  GHIVE_AGG_CNT = 4,
  GHIVE_AGG_RANK = 5,   /* UNSUPPORTED */
  GHIVE_AGG_UNKNOWN = 6 /* UNSUPPORTED */
};
#define GHIVE_COUNT_STAR 0xFFFFFFFFu /* ghive_agg.index for COUNT(*) */

typedef struct ghive_agg {
  int32_t type;
  uint32_t index; /* input column */
} ghive_agg;

/* Output type of an aggregate over an input column type, following Hive (not the
 * reference's int32 SUM/COUNT): SUM(INT|LONG) -> LONG, SUM(DOUBLE|FLOAT) -> DOUBLE,
 * COUNT -> LONG, AVG -> DOUBLE, MIN/MAX -> input type. */
int ghive_agg_output_type(int32_t agg_type, int32_t input_type);

/* GroupByPredicate::gpu_execute (GroupByPredicate.cu:19-359) as a hash aggregation. Rows
 * sel[0..n) (sel == NULL: all rows). Keys: up to 8 columns of any type, STRING included
 * (string keys as GroupByPredicate.hpp:69-96); keys that pack into 64 bits aggregate on the
 * packed key, wider ones through a key dictionary (group bound <= 2^31). NULL keys form one
 * group. A STRING key's output gathers (start,end) pairs into the caller's aux buffer (2 x
 * capacity int32) and shares the input blob (out data = input data). Output: keys then aggregates into out_cols
 * (types as ghive_agg_output_type; device buffers of `capacity` rows), groups in
 * unspecified order. An aggregate with no non-NULL input is NULL (Hive semantics).
 * SUM(LONG) wraps modulo 2^64 like a Java long. SUM/AVG(DOUBLE): every value is rounded once to a
 * multiple of 2^E, with 2^E about 2^-124 x max|x| x the row count, and these multiples are added
 * exactly as 128-bit integers per group, so the result does not depend on row order and is
 * reproducible bit for bit; its error is at most rows x 2^(E-1) before the final rounding to the
 * nearest double (for 1e9 rows at most max|x| x 2^-64). A DOUBLE column with an infinite value
 * is rejected (GHIVE_ERR_UNSUPPORTED); NaN is NULL. */
int ghive_groupby(ghive_ctx *ctx, const ghive_table *table, const uint32_t *sel, uint64_t n,
                  const uint32_t *keys, uint32_t nkeys, const ghive_agg *aggs, uint32_t naggs,
                  ghive_column *out_cols, uint64_t capacity, uint64_t *out_groups);

/* ---- projection arithmetic ----------------------------------------------------------- */
/* CalculationSelectPredicate::process (SelectPredicate.cu:84-409; functors
 * SelectPredicate.hpp:193-234): out[i] = a[i] op b[i], op one of '+', '-', '*', '/'. A column of
 * row_num 1 broadcasts (a literal). '/' gives DOUBLE; otherwise INT op INT gives INT (wraps like a
 * Java int), other integer pairs LONG (wraps), any FLOAT/DOUBLE operand DOUBLE -- the reference
 * accepts only equal operand types for + - * and these are Hive's widenings. A NULL operand gives
 * NULL; x / 0 gives NULL (Hive; the reference divides). out->data: device buffer of n values of
 * the result type (written to out->type). */
int ghive_select_arith(ghive_ctx *ctx, const ghive_column *a, const ghive_column *b, char op,
                       uint64_t n, ghive_column *out);

/* CAST (FunctorSelectPredicate with string_convert_int_functor, SelectPredicate.hpp:22-66, as
 * Hive's UDFToInteger / UDFToLong of Text: ql/.../udf/UDFToInteger.java:169-186 with
 * serde2/lazy/LazyInteger.java parseInt(trim)): a STRING to INT or LONG -- ASCII whitespace
 * trimmed, an optional sign, decimal digits, a '.' fraction truncated; malformed, empty or
 * overflowing text is NULL (the reference traps on a non-digit). A numeric column to any numeric
 * type as Java converts (narrowing integers keep the low bits, floating to integer truncates and
 * saturates). NULL stays NULL. out->data: device buffer of n values of to_type. */
int ghive_select_cast(ghive_ctx *ctx, const ghive_column *in, int32_t to_type, uint64_t n,
                      ghive_column *out);
/* SUBSTR(str, pos, len) (substr_functor, SelectPredicate.hpp:68-113, as Hive's UDFSubstr,
 * ql/.../udf/UDFSubstr.java:61-105): positions in UTF-16 units of the UTF-8 text, pos 1-based
 * (negative: from the end, 0: from the start), len <= 0 or |pos| > length gives ''. The result
 * shares the input blob: out->aux receives n (start,end) pairs (caller buffer), out->data is set to
 * the input's. NULL stays NULL. */
int ghive_select_substr(ghive_ctx *ctx, const ghive_column *str, int32_t pos, int32_t len, uint64_t n,
                        ghive_column *out);

/* ---- sort --------------------------------------------------------------------------- */
/* SinkOperator ORDER BY (SinkOperator.cu:135-288; orders as SortOrder '+'/'-',
 * GroupByPredicate.hpp:17-20): LSD radix sort; writes the permutation of row ids so that
 * rows are in key order, ties in ascending row id (a total order). Keys: INT, LONG,
 * DOUBLE, FLOAT columns, at most 8 keys and 256 bits in total; NULL sorts as its sentinel
 * value (NaN last ascending). */
int ghive_sort(ghive_ctx *ctx, const ghive_table *table, const uint32_t *key_cols,
               const char *orders, uint32_t nkeys, uint32_t *out_perm);

/* What the last ghive_sort on this context did (no counterpart in the reference; for measurement):
 * the 64-bit key words built per row, the LSD passes run (an 8-bit digit whose byte is the same
 * in every row is skipped), whether the row ids rode in spare key bits, and the bytes the scatter
 * passes read and wrote in total (rows x bytes moved per row, summed over the passes). */
typedef struct ghive_sort_info {
  uint64_t rows;
  uint32_t key_words;
  uint32_t passes;
  uint32_t embedded_ids;
  uint32_t reserved;
  uint64_t scatter_bytes;
} ghive_sort_info;
int ghive_last_sort_info(ghive_ctx *ctx, ghive_sort_info *out);

/* ---- fused fragment: filter -> star joins -> expression -> group-by ------------------- *
 * One Tez map-task fragment of GHive's CS1 call stack (TableScan -> Filter -> MapJoin* ->
 * Select -> GroupBy -> ReduceSink), executed as one streaming pass over the fact columns. */
enum { GHIVE_SRC_FACT = 0, GHIVE_SRC_JOIN = 1 };
enum { GHIVE_EXPR_COL = 0, GHIVE_EXPR_MUL = 1, GHIVE_EXPR_ONE = 2 };

typedef struct ghive_fragment_join {
  uint32_t fact_key_col;        /* fact column probed (INT/LONG)                       */
  const ghive_hash_table *table; /* built with unique keys; payload = dimension attribute */
} ghive_fragment_join;

typedef struct ghive_fragment_key {
  int32_t source; /* GHIVE_SRC_FACT: fact column `index`; GHIVE_SRC_JOIN: payload of join */
  uint32_t index;
} ghive_fragment_key;

/* Aggregate over a fact expression: EXPR_COL a; EXPR_MUL a*b; EXPR_ONE counts rows. Inputs are
 * INT columns (16-byte aligned; the product is the int64 product, as SelectPredicate's LONG
 * multiply) or DOUBLE columns (both, for a product: the product rounded to a double, as Hive's
 * double multiply; read only for the rows that survive the joins). SUM over DOUBLE inputs is the
 * 128-bit fixed-point sum of ghive_groupby (each value rounded once to the LSB chosen from the
 * columns' largest magnitudes and the row count, the rounded values added exactly), rounded once
 * to DOUBLE; MIN / MAX over DOUBLE give DOUBLE; every other aggregate gives LONG. */
typedef struct ghive_fragment_agg {
  int32_t type; /* GHIVE_AGG_SUM / CNT / MIN / MAX */
  int32_t expr;
  uint32_t a;
  uint32_t b;
} ghive_fragment_agg;

typedef struct ghive_fragment {
  ghive_table fact;
  const ghive_pred *filter; /* optional fact filter; nnodes 0 = none */
  uint32_t filter_nnodes;
  uint32_t filter_root;
  const ghive_fragment_join *joins;
  uint32_t njoins;
  const ghive_fragment_key *keys; /* INT payload/columns; at most 2 keys */
  uint32_t nkeys;
  const ghive_fragment_agg *aggs;
  uint32_t naggs;
} ghive_fragment;

/* Output: keys (INT each) then aggregates (LONG, or DOUBLE as above) into out_cols of `capacity`
 * rows (8-byte buffers for the aggregates); out_cols[i].type is set. */
int ghive_fragment_execute(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                           uint64_t capacity, uint64_t *out_groups);

/* ghive_fragment_execute that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The group table, its counters and the partial-tile staging
 * come from a static array of GHIVE_NOALLOC_SCRATCH_BYTES on each device (held until the process
 * ends); a fragment whose group table does not fit is rejected with GHIVE_ERR_UNSUPPORTED. The join
 * tables must be finished (ghive_hash_build, or an asynchronous build already used once). Results
 * equal ghive_fragment_execute's. Every call on a device shares that device's array and
 * re-initialises it: calls that overlap on one device race. */
#define GHIVE_NOALLOC_SCRATCH_BYTES (8u << 20)
int ghive_fragment_execute_noalloc(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                                   uint64_t capacity, uint64_t *out_groups);

/* ---- batched form: many fragments of one query shape in one persistent kernel ----------- *
 * No counterpart in the reference (it runs one Tez task per JNI call). Query per item:
 *   SELECT g, SUM(a*b) FROM fact JOIN dim ON fact.key = dim.key
 *   WHERE filter_lo <= f <= filter_hi GROUP BY g
 * where f and g are bit fields of the dimension payload (f = (payload >> filter_shift) &
 * (2^filter_bits - 1), likewise g with at most 12 bits). The dimension table is shared and
 * stays resident; every item brings its own fact columns (16-byte aligned INT), rows and filter
 * constants. Results are dense: sums[item * 2^group_bits + g] and counts[...] (count 0 = no
 * group). */
typedef struct ghive_batch ghive_batch;
typedef struct ghive_batch_item {
  const int32_t *key;
  const int32_t *a;
  const int32_t *b; /* NULL: SUM(a) */
  uint64_t rows;
  int32_t filter_lo;
  int32_t filter_hi;
} ghive_batch_item;
typedef struct ghive_batch_spec {
  const ghive_hash_table *dim;
  uint32_t filter_shift, filter_bits;
  uint32_t group_shift, group_bits;
} ghive_batch_spec;
/* Autotuned knobs: rows per tile (multiple of 1024; 0 = 8192), workgroups per CU (0 = 4),
 * probe mode (0 auto, 1 LDS bitmap, 2 HBM bitmap, 3 hash table, 4 LDS bitmap with the group
 * field of every passing key ranked beside it, for group_bits <= 8: no per-row read of the
 * dimension payload). */
typedef struct ghive_batch_tuning {
  uint32_t tile_rows;
  uint32_t blocks_per_cu;
  int32_t probe_mode;
  int32_t reserved;
} ghive_batch_tuning;
/* Uploads the item descriptors; with use_graph the zeroing and the persistent kernel are
 * captured once in a HIP graph that every run replays. */
int ghive_batch_create(ghive_ctx *ctx, const ghive_batch_spec *spec, const ghive_batch_item *items,
                       uint32_t nitems, const ghive_batch_tuning *tuning, int use_graph,
                       ghive_batch **out);
/* Runs the batch (spec needed only without a graph) and waits for it. */
int ghive_batch_run(ghive_ctx *ctx, ghive_batch *bt, const ghive_batch_spec *spec);
/* Host copies of the dense results (nitems * 2^group_bits each). */
int ghive_batch_results(ghive_ctx *ctx, ghive_batch *bt, int64_t *sums, int64_t *counts);
int ghive_batch_info(const ghive_batch *bt, uint32_t *groups_per_item, uint64_t *total_rows,
                     uint64_t *total_tiles, int32_t *probe_mode, uint32_t *grid, float *last_ms);
void ghive_batch_free(ghive_ctx *ctx, ghive_batch *bt);

/* Same operator with late materialisation: a fact column is read only for the lanes that
 * still hold a row after the previous join (measured beside the default form). */
int ghive_fragment_execute_late(ghive_ctx *ctx, const ghive_fragment *frag, ghive_column *out_cols,
                                uint64_t capacity, uint64_t *out_groups);

/* Tunables of a context: "fragment.lds_budget" (bytes, <= 160 KiB), "fragment.late" (0/1),
 * "fragment.grid" (workgroups, 0 = one resident wave), "fragment.lds_bits" (0/1: LDS copies of
 * small dimension bitmaps), "fragment.dense" (0/1: probe compact-key dimensions through their
 * bitmap + payload array instead of the hash table), "fragment.rows_per_thread" (4 or 8),
 * "fragment.block" (workgroup size 256, 512 or 1024), "fragment.generic" (0/1: always the general
 * fused kernel, not its dense-probe form), "fragment.pipe" (0/1: the pipelined dense-probe kernel
 * with direct group slots where the group keys' ranges allow it), "groupby.lds_budget" (bytes),
 * "groupby.partition" (0 never, 1 by size, 2 always: the hash-partitioned high-cardinality form),
 * "groupby.part_threads" (256, 512 or 1024: workgroup size of the partition passes; default 512),
 * "groupby.part_occupancy" (1/2: workgroups per CU its 512-thread 8-bit passes are compiled for; default 2),
 * "groupby.partition_fallbacks" (read only, ghive_get_option: how many partitioned group-bys of this
 * context had a partition overflow its LDS table and ran the direct form instead),
 * "hash.regional_rows" (synchronous hash builds of at least this many rows sort the rows by table
 * region and build each region of 8192 slots in one workgroup's LDS, with no fill pass and no
 * device-scope atomics but for keys that run past their region; 0 = never; default 4194304),
 * "plan.fuse" (0/1: run map-join vertices of a plan as one fused fragment), "sort.shape" (LSD tile
 * of ghive_sort: 0 = 256 threads x 32 rows, 1 = 512 x 16 (default), 2 = 512 x 32, 3 = 1024 x 8),
 * "sort.pack" (0/1: the passes that carry one key word and a row id, and the pass that turns two
 * key words with the row id inside into one key word and an id, stage a row as one 16-byte LDS
 * record, on tiles of 8192 rows and 512 threads or more, one workgroup per CU; default 0: the
 * 8-byte staging lets two workgroups share a CU), "sort.occupancy" (1/2: workgroups per CU the
 * 8-byte-staging passes of the 512 x 16 tile are compiled for; 2 holds them to 128 VGPRs; default
 * 2), "sort.merge" (0/1: a sort on
 * one key word with the row ids beside it, whose digits lie in both halves of the word, writes each
 * row as one word, the high half above the id, after the low half's last digit, so the passes over
 * the high half move 8 bytes a row instead of 12; default 1). */
int ghive_set_option(ghive_ctx *ctx, const char *name, int64_t value);
int ghive_get_option(ghive_ctx *ctx, const char *name, int64_t *value);

/* ---- serving form: request loop with HBM-footprint admission -------------------------- *
 * No counterpart in the reference (one Tez task per JNI call, GPUProcess.java:20-143). A
 * server owns `slots` contexts (each with its own stream and memory manager) on one device and
 * a dispatcher thread. Requests are admitted in arrival order as soon as a slot is free and
 * their HBM footprint fits hbm_budget_bytes beside the requests in flight; admitted requests
 * run concurrently and overlap on the device. Dimension hash tables referenced by fragments
 * stay resident (built once by the caller, shared read-only by all slots). */
enum { GHIVE_REQ_FRAGMENT = 1, GHIVE_REQ_GROUPBY = 2, GHIVE_REQ_SORT = 3 };
typedef struct ghive_server ghive_server;
typedef struct ghive_server_config {
  uint32_t slots;                /* concurrent requests on the device, 1..16            */
  uint32_t reserved;
  uint64_t hbm_budget_bytes;     /* admission budget over all requests in flight        */
  uint64_t slot_mem_limit_bytes; /* memory-manager limit of each slot (0 = 90% of free) */
} ghive_server_config;
/* One request. FRAGMENT: fragment + out_cols/capacity (as ghive_fragment_execute); GROUPBY:
 * table + keys/aggs + out_cols/capacity (as ghive_groupby over all rows); SORT: table +
 * keys/orders + out_perm (as ghive_sort). host_io = 1: the fact table / table columns and the
 * outputs are host memory; the slot uploads the inputs after admission and downloads the
 * outputs (pinned host memory lets transfers of different slots overlap). host_io = 0: all
 * device memory. footprint_bytes = 0 lets the server estimate inputs + intermediates. All
 * pointed-to descriptors must stay valid until ghive_server_wait returns for the request. */
typedef struct ghive_request {
  int32_t kind;
  int32_t host_io;
  const ghive_fragment *fragment;
  const ghive_table *table;
  const uint32_t *keys;
  const char *orders;
  uint32_t nkeys;
  uint32_t naggs;
  const ghive_agg *aggs;
  ghive_column *out_cols;
  uint64_t capacity;
  uint32_t *out_perm;
  uint64_t footprint_bytes;
} ghive_request;
typedef struct ghive_response {
  int32_t status; /* the operator's status code */
  uint32_t slot;
  uint64_t groups; /* groups written (FRAGMENT/GROUPBY) or rows sorted */
  double submit_us, start_us, end_us; /* host steady clock; latency = end - submit */
  char error[256];
} ghive_response;
typedef struct ghive_server_stats {
  uint64_t completed, rejected, queued;
  uint32_t inflight, peak_inflight;
  uint64_t inflight_bytes, peak_inflight_bytes;
} ghive_server_stats;
int ghive_server_create(int device, const ghive_server_config *cfg, ghive_server **out);
/* Queues a request (thread-safe). GHIVE_ERR_ADMISSION when its footprint exceeds the whole
 * budget, so it could never be admitted. */
int ghive_server_submit(ghive_server *srv, const ghive_request *req, uint64_t *id);
/* Blocks until request `id` has finished; each id can be waited for once. */
int ghive_server_wait(ghive_server *srv, uint64_t id, ghive_response *out);
int ghive_server_get_stats(ghive_server *srv, ghive_server_stats *out);
/* Context of a slot (to set options on it before submitting work). */
ghive_ctx *ghive_server_slot_context(ghive_server *srv, uint32_t slot);
/* Finishes the requests already admitted, drops queued ones, releases the slots. */
void ghive_server_destroy(ghive_server *srv);

/* ---- query plans: the operator tree of every vertex ------------------------------------- *
 * Parser::getOperatorTree + linkOperatorTreeAndExtendedPlan (gpu_impl/src/Parser/Parser.cu:
 * 58-677) and the operators' parseExtended (JoinOperator.cu:21-167, GroupByOperator.cu:16-99,
 * SinkOperator.cu:17-96): Hive's EXPLAIN text (`plan`, formatted tree) and EXPLAIN EXTENDED text
 * (`extended`, may be NULL) of one query, plus the Java side's maintain-values JSON (Gson of
 * Map<String, Map<Byte, List<String>>>: join operator id -> input tag -> column names; may be
 * NULL), become the operator tree with each operator's resolved descriptor: filter predicate
 * text, select expressions, join input order / key columns / join types (BaseJoinType numbering)
 * / maintained columns, group-by key and aggregate columns (AggregationType numbering), sink key
 * columns and sort orders. With `plan` NULL or empty, the EXPLAIN EXTENDED text alone gives the
 * tree (its "Edges:" and each vertex's operator tree; operator ids are then numbered in order of
 * appearance) -- for plans whose formatted text comes from another optimizer run. Host only;
 * needs no context. On failure `err` receives the message. */
typedef struct ghive_plan ghive_plan;
int ghive_plan_parse(const char *plan, const char *extended, const char *maintain_json, char *err,
                     uint64_t errcap, ghive_plan **out);
/* The parsed plan as JSON: {"root": id, "vertices": {vertex: sink id}, "operators": {id: {...}}}.
 * *len = bytes needed (without the terminator); GHIVE_ERR_CAPACITY when cap is too small. */
int ghive_plan_describe(const ghive_plan *plan, char *buf, uint64_t cap, uint64_t *len);
void ghive_plan_free(ghive_plan *plan);

/* One input of a vertex: a base table named as the TableScan names it ("db@table") or by the
 * vertex being run (GTables names the scanned table so, ghive_JNIInterface.cu:200-204), or the
 * result of another vertex named by that vertex ("Map 3"). column_names NULL: the columns are
 * positional, in the order GTables delivers them -- the TableScan's output columns, or the
 * producing sink's outputs with its key columns first. */
typedef struct ghive_plan_input {
  const char *name;
  const ghive_table *table; /* device columns */
  const char *const *column_names;
} ghive_plan_input;
typedef struct ghive_result ghive_result;
/* Operator::execute over one vertex's operator chain (the execute() of each gpu_impl/src/Operator/ source; the
 * reference's GPUProcess runs one vertex per call): TableScan binds its input by name and column
 * names; Filter parses its predicate (ghive_parse_predicate) and compacts (ghive_filter_compact);
 * Select aliases columns, reads KEY.reducesinkkeyN / VALUE._colN of a reduce input, or computes
 * (x op y) with ghive_select_arith; MapJoin / MergeJoin run ghive_join + ghive_join_result
 * (single-column keys; the outputs are the maintained columns, else the join outputs found by
 * name); GroupBy runs ghive_groupby; the vertex's sink orders its columns keys first; at the
 * plan's root the final ORDER BY (ghive_sort + ghive_gather) and LIMIT apply. The result holds
 * device columns owned by it (release with ghive_result_free on the same context); rows of a
 * group-by or join come in unspecified order unless the root sorts them. */
int ghive_plan_execute(ghive_ctx *ctx, const ghive_plan *plan, const char *vertex,
                       const ghive_plan_input *inputs, uint32_t ninputs, char *err, uint64_t errcap,
                       ghive_result **out);
/* What the last ghive_plan_execute on this context did with its vertex's group-by. A map vertex
 * of the CS1 shape -- TableScan -> [Filter] -> [Select] -> MapJoin+ (inner, one broadcast
 * dimension each, single-column keys) -> [Select] -> GroupBy (hash mode; keys: dimension attributes
 * or fact columns, at most 2; SUM / COUNT / MIN / MAX of a fact column or of a product of two) ->
 * ReduceSink -- runs as one ghive_fragment_execute over the fact columns instead of the operator
 * chain (option "plan.fuse", default 1). LONG columns the fused kernel streams are given to it as
 * 32-bit copies when every value fits; IS NOT NULL conjuncts on probed keys are dropped (an inner
 * join drops NULL keys anyway). The result is the same as the operator chain's. */
typedef struct ghive_plan_exec_info {
  int32_t fused;             /* 1: the vertex's group-by ran as one fused fragment              */
  uint32_t joins;            /* map joins folded into it                                       */
  uint32_t narrowed_columns; /* LONG columns it read as 32-bit copies                          */
  int32_t has_filter;        /* 1: a fact predicate was evaluated inside the fused kernel      */
  float fused_kernel_ms;     /* the fused kernel (HIP events)                                  */
  uint32_t reserved;
  char not_fused[160];       /* why a vertex with a group-by ran as the operator chain         */
} ghive_plan_exec_info;
int ghive_last_plan_info(ghive_ctx *ctx, ghive_plan_exec_info *out);
/* ---- the JNI boundary's marshalling, without a JVM ---------------------------------------- *
 * GTables (ghive/GTables.java:13-31) as the addresses of its direct ByteBuffers, in GTables
 * order: per table its column type codes (0 LONG, 1 DOUBLE, 2 INT, 3 STRING; other codes are
 * skipped), row and key counts and vertex name; the column buffers of each type concatenated
 * across tables and consumed in table and column order (ghive_JNIInterface.cu:103-200). A STRING
 * column is a UTF-8 buffer plus an int32 (start,end) buffer; the blob's size is the largest value
 * of its pairs, as the reference computes it. Buffers are host memory (little-endian). */
typedef struct ghive_gtables {
  uint32_t ntables;
  const int32_t *const *types;      /* types[t][0..ncols[t]) */
  const uint32_t *ncols;
  const int32_t *row_cnt;           /* rowCnt */
  const int32_t *key_cnt;           /* keyCnt (may be NULL) */
  const char *const *vertex_name;   /* vertexName; a NULL entry means a missing input */
  const void *const *long_cols;
  uint32_t nlong;
  const void *const *double_cols;
  uint32_t ndouble;
  const void *const *int_cols;
  uint32_t nint;
  const void *const *string_cols;
  const void *const *string_idx_cols;
  uint32_t nstring;
} ghive_gtables;
typedef struct ghive_staged ghive_staged;
/* ghive_JNIInterface.cu:88-204: copies every table to device memory (two pinned buffers, so the
 * host copy of one chunk overlaps the transfer of the previous one) and describes each as a plan
 * input named by its vertex name, columns positional. A NULL vertex name (the reference's missing
 * input) or too few buffers of a type: GHIVE_ERR_INVALID. */
int ghive_gtables_stage(ghive_ctx *ctx, const ghive_gtables *g, ghive_staged **out);
/* The staged tables as ghive_plan_execute inputs (valid until ghive_staged_free). */
int ghive_staged_inputs(const ghive_staged *s, const ghive_plan_input **inputs, uint32_t *n);
void ghive_staged_free(ghive_ctx *ctx, ghive_staged *s);
/* generate_tbl_result numbering (ghive_JNIInterface.cu:223-342): GPUResult holds the LONG columns,
 * then DOUBLE, INT and STRING, each group in column order; sequence[c] = the position of column c
 * in that concatenation, counts[4] = the columns of each group. FLOAT is numbered as DOUBLE. Host
 * only. */
int ghive_gpuresult_sequence(const int32_t *types, uint32_t ncols, int32_t *sequence, uint32_t counts[4]);

/* GPUProcess in one call (Java_..._JNIInterface_GPUProcess, gpu_impl/src/ghive_JNIInterface.cu:13-221,
 * with generate_tbl_result :223-342): GTables in, the vertex `this_vertex` of the plan run, a GPUResult
 * out -- so that a JNI shim is a pure translation of Java arrays. The table whose vertex name is
 * this_vertex is the vertex's own scan input. Staging sends the LONG columns a fused form of the
 * vertex streams as 32-bit integers when every value fits (half the bytes over PCIe and through HBM;
 * they are LONG again to every other operator). A table with a NULL vertex name is a missing input:
 * GHIVE_OK with *out = NULL and nothing run, as the reference returns nullptr. The result is host
 * memory grouped as GPUResult's constructor takes it: counts[4] columns of LONG, DOUBLE (FLOAT
 * widened), INT and STRING, each group in column order, and sequence[c] = the position of output
 * column c in that concatenation; a STRING column is a UTF-8 blob plus int32 (start,end) per row.
 * Everything staged on the device is released before the call returns. */
typedef struct ghive_gpu_result ghive_gpu_result;
int ghive_gpu_process(ghive_ctx *ctx, const ghive_plan *plan, const ghive_gtables *g, const char *this_vertex,
                      char *err, uint64_t errcap, ghive_gpu_result **out);
int ghive_gpu_result_info(const ghive_gpu_result *r, uint64_t *rows, uint32_t *ncols, uint32_t counts[4],
                          const int32_t **sequence);
/* Column i of a group (i < counts[group]); NULL when out of range. */
const int64_t *ghive_gpu_result_long(const ghive_gpu_result *r, uint32_t i);
const double *ghive_gpu_result_double(const ghive_gpu_result *r, uint32_t i);
const int32_t *ghive_gpu_result_int(const ghive_gpu_result *r, uint32_t i);
int ghive_gpu_result_string(const ghive_gpu_result *r, uint32_t i, const char **blob, uint64_t *blob_bytes,
                            const int32_t **pairs);
void ghive_gpu_result_free(ghive_gpu_result *r);
/* Wall-clock split of the last ghive_gpu_process on this context (host steady clock, ms). */
typedef struct ghive_process_times {
  double stage_ms;           /* GTables buffers -> device columns (host copies + PCIe)  */
  double execute_ms;         /* the vertex's operators on the device                    */
  double result_ms;          /* device result -> host GPUResult arrays                  */
  double total_ms;
  uint64_t h2d_bytes;        /* bytes sent to the device                                */
  uint32_t narrowed_columns; /* LONG columns sent as 32-bit integers                    */
  int32_t fused;             /* the vertex ran as one fused fragment                    */
} ghive_process_times;
int ghive_last_process_times(ghive_ctx *ctx, ghive_process_times *out);

int ghive_result_info(const ghive_result *r, uint64_t *rows, uint32_t *ncols);
/* Column i (device data, owned by the result) and its name. */
int ghive_result_column(const ghive_result *r, uint32_t i, ghive_column *col, const char **name);
void ghive_result_free(ghive_ctx *ctx, ghive_result *r);

/* Per-stage timing of the last fragment on this context (HIP events, milliseconds). */
typedef struct ghive_stage_times {
  float total_ms;
  float main_kernel_ms; /* the fused scan/probe/aggregate kernel */
  float finalize_ms;
  uint64_t rows_in;
  uint64_t rows_joined;
} ghive_stage_times;
int ghive_last_stage_times(ghive_ctx *ctx, ghive_stage_times *out);

#ifdef __cplusplus
}
#endif
#endif /* GHIVE_GHIVE_H */
