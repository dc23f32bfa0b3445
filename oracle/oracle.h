/*
 * oracle.h — CPU restatement of vexq's hot-path operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against and the timed CPU baseline in bench.py; the product path
 * never links, loads or calls it.
 *
 * It reuses the plain-C query description types of include/vexq_gpu.h so a
 * test can hand the same predicate / aggregate description to both sides.
 */
#ifndef VEXQ_ORACLE_H
#define VEXQ_ORACLE_H

#include <stdint.h>

#include "../include/vexq_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* exec/parallel_join.go:91-99 (murmur3 fmix64), :73-82, exec/join_table.go:461-476 */
uint64_t or_radix_hash(int64_t key);
int32_t or_radix_bits_for(int64_t est_rows);
int64_t or_join_table_capacity(int64_t keys);

/* bench/simd_filter/cpp/main.cpp:28-45 data generator. */
void or_xorshift_i64(int64_t n, uint64_t seed, int64_t* out);

/* Correctly rounded sum of n doubles (exact superaccumulator). */
double or_exact_sum(const double* x, int64_t n);

/* The strings of the dictionary columns, for MIN/MAX over strings (accum
 * VQ_STR): strs[col][code], lens[col][code] bytes; strs[col] NULL for a
 * column without one.  Set before a call that folds such an aggregate and
 * cleared with ncols 0 after it (one caller at a time). */
void or_set_dicts(int32_t ncols, const char* const* const* strs, const int64_t* const* lens);

/* storage/writer.go:205-348 zone statistics, nrg*ncols row-group-major. */
void or_zone_maps(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                  const int64_t* rg_rows, vq_zone* out);

/* planner/physical.go:631-750 over a conjunction: keep[rg] = 1 unless pruned. */
void or_zone_prune(int32_t ncols, const vq_col_desc* cols, int32_t nrg, const vq_zone* zones,
                   const vq_pred* pred, uint8_t* keep);

/* TableScan(zone-pruned) + Filter + BoolToSelVec: ascending passing rows. */
int64_t or_filter(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                  const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                  uint32_t* out_rows);
/* The same selection by nthreads workers over row-group morsels (the CPU
 * baseline); the result is the serial one. */
int64_t or_filter_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                           const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                           int32_t nthreads, uint32_t* out_rows);

/* Serial Scan->Filter->Project->HashAggregate over dictionary keys.
 * exact = 0: float sums are the reference's row-order left fold (bit pattern
 * of the serial plan); exact = 1: float sums are correctly rounded. */
int or_agg_dict(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                int32_t nkeys, const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                int32_t exact, vq_agg_result* out);

/* Serial GROUP BY one integer column (HashAggregate's general key path):
 * groups in first-seen order, key values in out->keys. */
int or_agg_hash(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                int32_t key_col, int32_t naggs, const vq_agg* aggs, int32_t exact,
                vq_agg_result* out);

/* Serial HashJoin -> HashAggregate(GROUP BY join key) [-> Sort DESC -> Limit]. */
int or_join_agg(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols, int32_t b_key,
                const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t naggs,
                const vq_agg* aggs, int32_t topk, int32_t exact, vq_agg_result* out);

/* Serial HashJoin -> HashAggregate over the joined rows (the Q12 shape of
 * ParallelHashJoinAggregate): GROUP BY nkeys probe columns (codes packed as
 * the int-key path packs them when nkeys > 1), aggregates over columns of
 * either side (col | VQ_BUILD_COL names a build column); groups in the order
 * of their first joined probe row. */
int or_join_group(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols, int32_t b_key,
                  const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                  const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t nkeys,
                  const int32_t* key_cols, int32_t naggs, const vq_agg* aggs, int32_t exact,
                  vq_agg_result* out);

/* A join whose build side is a join (the nested build, Q3's chain): inner
 * rows passing i_pred hashed on i_key; build rows passing b_pred, each once
 * per inner row its b_fk meets, in build-row order, are the outer build side
 * keyed by b_key; probe rows join it as or_join_agg does, GROUP BY the join
 * key [-> top-k]; npay build columns ride with each group into
 * pay_out[g * npay + c] (int64: a date sign-extended, a float64's bits);
 * VQ_E_UNSUPPORTED when rows of one key disagree on them. */
int or_join_agg_nested(int64_t i_nrows, int32_t i_ncols, const vq_col_desc* i_cols, int32_t i_key,
                       const vq_pred* i_pred, int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols,
                       int32_t b_fk, int32_t b_key, const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                       const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t npay,
                       const int32_t* pay_cols, int32_t naggs, const vq_agg* aggs, int32_t topk, int32_t exact,
                       vq_agg_result* out, int64_t* pay_out);

/* Multithreaded morsel restatement (CPU baseline): ParallelHashAggregate and
 * ParallelHashJoinAggregate with 64K-row morsels on an atomic cursor,
 * thread-local partial aggregates merged serially, and the two-pass radix
 * build.  Float sums follow the reference's parallel order (not
 * deterministic across thread counts). */
int or_agg_dict_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                         const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                         int32_t nkeys, const int32_t* key_cols, int32_t naggs,
                         const vq_agg* aggs, int32_t nthreads, vq_agg_result* out);

/* The same partial-aggregate plan over one integer key column (the general
 * key path, exec/aggregate.go:361-559, partials merged as parallel.go:186-209). */
int or_agg_hash_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                         const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                         int32_t key_col, int32_t naggs, const vq_agg* aggs, int32_t nthreads,
                         vq_agg_result* out);

int or_join_agg_parallel(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols,
                         int32_t b_key, const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                         const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred,
                         int32_t naggs, const vq_agg* aggs, int32_t topk, int32_t nthreads,
                         vq_agg_result* out);

#ifdef __cplusplus
}
#endif
#endif
