/*
 * vexq_gpu.h — C ABI of the MI355X (gfx950) execution library for vexq's
 * data-parallel hot path: selection-vector Filter, zone-map-pruned TableScan,
 * dictionary-code GROUP BY (HashAggregate / ParallelHashAggregate) and the
 * radix-partitioned hash join feeding an aggregate (ParallelHashJoinAggregate).
 *
 * Plain C: no C++ types, no HIP types, no torch types cross this boundary.
 * Every entry point returns an int status (VQ_OK == 0); on failure the
 * calling thread's vq_last_error() names the cause.  Host pointers passed in
 * are never retained after the call returns (the cgo pointer rule).
 *
 * Ownership: the library owns every device allocation and every handle it
 * returns; the caller owns host output buffers and frees handles with the
 * matching *_destroy call.  A vq_table is immutable after upload and may be
 * shared by concurrent calls.  A vq_ctx may be used from several host threads;
 * each call enqueues on the context's stream and returns when its result is on
 * the host.
 *
 * Reference interfaces each group replaces are cited file:line in the vexq
 * checkout (Go).
 */
#ifndef VEXQ_GPU_H
#define VEXQ_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (wrapped Go-side as fmt.Errorf("exec: gpu ...: %w")) ---- */
#define VQ_OK 0
#define VQ_E_ARG 1         /* invalid argument (mirrors the exec: ... errors)   */
#define VQ_E_DEVICE 2      /* HIP runtime / device failure                      */
#define VQ_E_BUDGET 3      /* query exceeds the context's device-memory budget  */
#define VQ_E_UNSUPPORTED 4 /* shape the GPU operators do not cover: fall back   */
#define VQ_E_NOMEM 5       /* host allocation failure                           */
#define VQ_E_CORRUPT 6     /* a .vxq checksum or structure does not verify
                              (storage.ErrChecksum / ErrCorruptFooter /
                              ErrBadMagic, storage/errors.go)                   */
#define VQ_E_CANCELED 7    /* the query was canceled (vq_ticket_cancel,
                              vq_ctx_cancel): the reference's ctx.Err() check
                              per batch (exec/scan.go:199-203)                   */

/* Column physical types as they sit in HBM.  Mirrors storage type codes
 * (storage/format.go:20-26) after the decode in exec/scan.go:294-409:
 * dictionary STRING columns become global codes (u8/u16), DATE is int32 days
 * since 1970-01-01, INT64 and FLOAT64 keep their width. */
#define VQ_U8 1
#define VQ_U16 2
#define VQ_DATE 3
#define VQ_I64 4
#define VQ_F64 5

/* Rows per row group and per batch (storage/format.go:8-10). */
#define VQ_ROW_GROUP_ROWS 65536
#define VQ_BLOCK_ROWS 1024

/* Per-row-group, per-column zone map, raw bits exactly as storage.ZoneMap
 * (storage/schema.go:26-35) and storage/writer.go:205-348 compute them. */
typedef struct vq_zone {
  uint64_t min_bits;
  uint64_t max_bits;
  int64_t null_count;
  int32_t has_minmax;
  int32_t pad_;
} vq_zone;

/* One column of a bulk upload.  data points at nrows host values of the
 * type's width; dict_size is the number of global dictionary codes for
 * VQ_U8/VQ_U16 dictionary columns (0 for a plain integer column). */
typedef struct vq_col_desc {
  int32_t type;
  int32_t dict_size;
  const void* data;
  /* Validity (exec.Vector's NullBitmap, exec/batch.go:113-121): bit r, LSB
   * first in byte r / 8, set when row r holds a value and clear for NULL.
   * NULL: every row holds a value.  The dictionary GROUP BY, its scan and
   * the filter take columns with NULLs (a predicate term on a NULL is not
   * TRUE, so the row is excluded; aggregates skip NULL inputs; a NULL group
   * key is a group of its own); every other operator returns
   * VQ_E_UNSUPPORTED for a query that reads one (fall back). */
  const uint8_t* valid;
} vq_col_desc;

/* ---- predicates: a conjunction of `column OP literal` terms -------------
 * Replaces Filter(BinOp compare, AndExpr, BetweenExpr) — exec/filter.go:49-78,
 * exec/expr.go:320-896,1030-1067,1319-1336.  BETWEEN lo AND hi is passed as
 * the two terms GE lo, LE hi (the reference's own rewrite, expr.go:1319-1336).
 * ival is the literal for VQ_U8/VQ_U16/VQ_DATE/VQ_I64 columns, fval for
 * VQ_F64 columns (an integer literal next to a FLOAT64 column is coerced to
 * float by the caller, planner/physical.go:773-818). */
#define VQ_EQ 0
#define VQ_NE 1
#define VQ_LT 2
#define VQ_LE 3
#define VQ_GT 4
#define VQ_GE 5
/* col IN (set) over dictionary codes or small integers (InExpr,
 * exec/expr.go:1188-1238, with the planner's string literals already turned
 * into global codes): ival is a 64-bit mask, a row passes when its value v
 * is in [0, 64) and bit v is set.  Integer-typed columns only.  Never prunes
 * a row group. */
#define VQ_IN 6
/* op | VQ_COLCMP compares two columns of the row, col OP column ival (a
 * BinOp compare of two column references, exec/expr.go:320-896, e.g.
 * l_commitdate < l_receiptdate): both float64, or both integer-typed (dates
 * and int64 compared as int64).  A dictionary-coded (string) column on either
 * side returns VQ_E_UNSUPPORTED: strings compare by value, which codes do not
 * order, and two dictionaries need not share codes.  Never prunes a row
 * group. */
#define VQ_COLCMP 16

#define VQ_MAX_TERMS 8
typedef struct vq_term {
  int32_t col;
  int32_t op;
  int64_t ival;
  double fval;
} vq_term;

typedef struct vq_pred {
  int32_t nterms; /* 0 = every row passes */
  int32_t pad_;
  vq_term terms[VQ_MAX_TERMS];
} vq_pred;

/* ---- row-local arithmetic: a left fold of up to four factors -------------
 *   value = ((f0 c1 f1) c2 f2) c3 f3
 * where each factor is a column, a literal, (literal op column) or
 * (column op literal), and c_k / op are + - * /.  Every operation is rounded
 * to float64 on its own (no fused multiply-add), as Project materialises each
 * BinOp result before the next (exec/project.go:64-110, expr.go:956-1023).
 * is_float = 1 evaluates in float64 (integer columns cast, CastIntToFloatExpr
 * expr.go:194-222; x/0 = NaN); is_float = 0 evaluates in int64 (x/0 = 0). */
#define VQ_F_COL 0
#define VQ_F_LIT 1
#define VQ_F_LIT_OP_COL 2
#define VQ_F_COL_OP_LIT 3

#define VQ_ADD 0
#define VQ_SUB 1
#define VQ_MUL 2
#define VQ_DIV 3

#define VQ_MAX_FACTORS 4
/* CASE WHEN <when> THEN <the fold> ELSE <literal> END: with nwhen > 0 the
 * fold is guarded by a conjunction of nwhen predicate terms (vq_term: a
 * compare with a literal, VQ_IN over codes, op | VQ_COLCMP) on any columns,
 * and a row failing it takes the ELSE literal (else_f when is_float, else_i
 * otherwise).  CaseExpr with one WHEN and a literal ELSE (exec/expr.go:
 * 1386-1435), e.g. TPC-H Q12's SUM(CASE WHEN o_orderpriority IN ('1-URGENT',
 * '2-HIGH') THEN 1 ELSE 0 END) with the string literals as global codes
 * (StringEqExpr as a code compare, exec/expr.go:1342-1384).  A CASE with
 * several WHENs, or with no ELSE (NULL), is not representable: the caller
 * keeps the Go operator. */
#define VQ_MAX_WHEN 2
typedef struct vq_factor {
  int32_t kind;
  int32_t col;
  int32_t op;      /* op inside the factor (LIT_OP_COL / COL_OP_LIT)      */
  int32_t combine; /* how this factor folds into the running value (k>0) */
  double fval;     /* literal when is_float                              */
  int64_t ival;    /* literal when !is_float                             */
} vq_factor;

typedef struct vq_expr {
  int32_t nfactors; /* 0 = no argument (COUNT(*)) */
  int32_t is_float;
  vq_factor f[VQ_MAX_FACTORS];
  int32_t nwhen;    /* 0 = no CASE */
  int32_t pad_;
  vq_term when[VQ_MAX_WHEN];
  double else_f;
  int64_t else_i;
} vq_expr;

/* Aggregates: exec.AggExpr (exec/aggregate.go:25-37).  accum is the
 * AccumType rule of AccumTypeFor (aggregate.go:51-72): VQ_F64 for AVG and for
 * SUM/MIN/MAX over a float argument, VQ_STR for MIN/MAX over a dictionary
 * (STRING) column, VQ_I64 otherwise. */
#define VQ_AGG_COUNT 0
#define VQ_AGG_SUM 1
#define VQ_AGG_AVG 2
#define VQ_AGG_MIN 3
#define VQ_AGG_MAX 4

/* accum of MIN/MAX over a STRING column (AccumType TypeString, the strAccs
 * of exec/aggregate.go:98-110, folded by minMaxString, aggregate.go:670-730):
 * the argument is that one dictionary column; the result, in ivals, is the
 * global code of the group's smallest / largest string as Go compares
 * strings (bytewise), valid 0 for a group with no value.  The dictionary
 * GROUP BY and its global aggregates take it once the column's byte order
 * is known (vq_table_dict_order; a table from vq_table_upload_vxq knows its
 * dictionaries); every other operator returns VQ_E_UNSUPPORTED. */
#define VQ_STR 6

#define VQ_MAX_AGGS 8
typedef struct vq_agg {
  int32_t kind;
  int32_t accum;
  vq_expr arg;
} vq_agg;

/* Aggregate output, one row per group.  The caller sizes the arrays for
 * `capacity` groups.  When the result needs more rows the call returns
 * VQ_E_ARG with ngroups set to the rows needed and nothing written to the
 * arrays (the two-call size query: retry with that capacity); on any other
 * failure ngroups is 0, so a reused result buffer never shows a failed
 * call's rows.  Groups come out in first-occurrence order (the order the
 * serial HashAggregate inserts keys, exec/aggregate.go:83-129), or for a
 * top-k join in rank order.  Per aggregate j of group g:
 *   ivals[g*naggs+j] for VQ_I64 accumulators, fvals[g*naggs+j] for VQ_F64,
 *   valid[g*naggs+j] = 0 where the SQL result is NULL (buildOutputBatch,
 *   aggregate.go:780-876).  keys: nkeys global codes per group (dictionary
 *   GROUP BY) or the int64 join key (join aggregate). */
typedef struct vq_agg_result {
  int64_t capacity;
  int64_t ngroups;
  int64_t* keys;      /* [capacity * nkeys] (codes widened to int64) */
  int64_t* counts;    /* [capacity] input rows folded into the group  */
  int64_t* first_row; /* [capacity] lowest contributing row index     */
  int64_t* ivals;     /* [capacity * naggs]                           */
  double* fvals;      /* [capacity * naggs]                           */
  uint8_t* valid;     /* [capacity * naggs]                           */
} vq_agg_result;

/* Per-call measurements, filled when a non-NULL pointer is passed. */
typedef struct vq_stats {
  double device_ms;      /* HIP-event time of the call's kernel sequence  */
  double main_kernel_ms; /* HIP-event time of its dominant kernel (the
                            dictionary GROUP BY with kernel repeats 1: of
                            its two kernels together)                   */
  int64_t rows_scanned;  /* rows in row groups that survived zone pruning */
  int64_t rg_total;
  int64_t rg_scanned;
  int64_t bytes_alg;     /* algorithmic bytes: referenced columns once     */
  int64_t build_rows;    /* join: build rows inserted (JoinBuildStats.Rows) */
  int64_t build_keys;    /* join: distinct build keys   (JoinBuildStats.Keys) */
  int32_t partitions;    /* join: radix partitions (JoinBuildStats.Partitions) */
  int32_t radix_bits;
  int32_t grid;          /* workgroups of the dominant kernel */
  int32_t tile_rows;     /* rows per tile of the scan */
} vq_stats;

typedef struct vq_ctx vq_ctx;
typedef struct vq_table vq_table;
typedef struct vq_selvec vq_selvec;

/* ---- context --------------------------------------------------------------
 * budget_bytes bounds the device memory this context may hold for tables and
 * query arenas (0 = 90% of the device's free memory).  No reference
 * counterpart: vexq's operators allocate from the Go heap. */
int vq_ctx_create(int device, uint64_t budget_bytes, vq_ctx** out);
int vq_ctx_destroy(vq_ctx* ctx);
/* Thread-local text of the last error on this thread ("" if none). */
const char* vq_last_error(void);
/* Timing aid (no counterpart in the reference): each later vq_agg_dict,
 * vq_filter and vq_table_upload_vxq call launches its main kernel `repeats`
 * times back to back between the HIP events that time it, and reports the
 * mean duration per launch in vq_stats.main_kernel_ms, so the events' own
 * overhead is amortised.  Results are unchanged (the kernel's output is the
 * same each time).  1 (the default) to 1000. */
int vq_ctx_set_kernel_repeats(vq_ctx* ctx, int32_t repeats);
/* Device bytes currently held by the context (tables + arenas). */
uint64_t vq_ctx_bytes_in_use(const vq_ctx* ctx);

/* ---- tables: the one bulk upload from storage.Reader ----------------------
 * Replaces the per-batch decode of TableScan (exec/scan.go:254-409) with one
 * upload; rg_rows gives the row count of each row group (storage.RowGroupMeta
 * .NumRows, storage/schema.go:47-52).  zones may be NULL, in which case the
 * library computes them the way storage/writer.go:205-348 does. */
int vq_table_upload(vq_ctx* ctx, int64_t nrows, int32_t ncols, const vq_col_desc* cols,
                    int32_t nrg, const int64_t* rg_rows, const vq_zone* zones,
                    vq_table** out);
int vq_table_destroy(vq_table* t);
int64_t vq_table_rows(const vq_table* t);
/* Copy the zone maps (nrg * ncols entries, row-group-major) into out. */
int vq_table_zones(const vq_table* t, vq_zone* out, int64_t cap);

/* The bulk upload straight from a .vxq file image (the bytes of the file as
 * storage.Writer wrote it, storage/writer.go): the footer is parsed and its
 * checksum verified on the host (storage/reader.go:295-408), each STRING
 * column's per-row-group dictionaries are read and merged into one global
 * dictionary in first-seen order (getOrBuildRemap, exec/agg_intkey.go:
 * 118-134); then the row-group sections are copied to the device in one
 * transfer, and the device checks every block's CRC-32 (ColumnReader.NextBlock,
 * storage/rowgroup.go:101-151) and decodes it into the resident columns
 * (TableScan's decode, exec/scan.go:254-409), STRING codes remapped to the
 * global codes and narrowed to VQ_U8 (up to 256 strings) or VQ_U16 (up to
 * 65,536).  cols: the file columns to keep, in table-column order.  Returns
 * VQ_E_CORRUPT naming the row group, column and offset of a block whose
 * checksum fails; VQ_E_UNSUPPORTED for a BOOL column, a column with NULLs or
 * more than 65,536 strings.  stats (may be NULL): device_ms and
 * main_kernel_ms of the one check-and-decode kernel (per launch, see
 * vq_ctx_set_kernel_repeats), bytes_alg the block bytes read plus the column
 * bytes written. */
int vq_table_upload_vxq(vq_ctx* ctx, const void* file, uint64_t file_len, int32_t ncols,
                        const int32_t* cols, vq_table** out, vq_stats* stats);
/* A column's type and dictionary size (0: not a dictionary column). */
int vq_table_column(const vq_table* t, int32_t col, int32_t* type, int32_t* dict_size);
/* The string of a global dictionary code of a table ingested from a .vxq
 * image (Dictionary.Get, storage/codec.go:160-194): *str points at len bytes
 * owned by the table, valid until it is destroyed. */
int vq_table_dict_entry(const vq_table* t, int32_t col, int32_t code, const char** str,
                        int64_t* len);
/* The byte order of a dictionary column's strings, for MIN/MAX with accum
 * VQ_STR: rank[code] is the position of code's string among the column's
 * dict_size strings sorted as Go compares strings (bytewise), a permutation
 * of 0 .. dict_size - 1.  The library keeps a column of ranks on the device
 * (2 bytes a row, under the context's budget) until the table is destroyed;
 * calling again replaces the order (not while a query on the table runs:
 * the rank column is rewritten in place).  Not needed for a table from
 * vq_table_upload_vxq (its first VQ_STR query orders the dictionary it
 * carries).  No counterpart in the reference, whose strAccs compare the
 * strings themselves (exec/aggregate.go:670-730). */
int vq_table_dict_order(vq_table* t, int32_t col, const uint32_t* rank);

/* Zone-map pruning alone: writes 1/0 per row group into keep (nrg entries).
 * Restates TableScan's zonePred check (exec/scan.go:211-214) with the
 * predicate evaluation of planner/physical.go:624-750. */
int vq_zone_prune(const vq_table* t, const vq_pred* pred, uint8_t* keep);

/* ---- Filter -> selection vector -------------------------------------------
 * Filter.Next + BoolToSelVecInto (exec/filter.go:49-78, expr.go:1545-1562)
 * over the zone-pruned TableScan: every row that passes, as ascending global
 * row indices (u32), left on the device in *out.  The same semantics as the
 * AVX2 kernel filter_gt_i64_avx2 (bench/simd_filter/cpp/filter_kernel.h). */
int vq_filter(vq_ctx* ctx, const vq_table* t, const vq_pred* pred, vq_selvec** out,
              vq_stats* stats);
int64_t vq_selvec_count(const vq_selvec* sv);
/* Copies min(count, cap) row indices to host memory. */
int vq_selvec_copy(const vq_selvec* sv, uint32_t* host_rows, int64_t cap);
int vq_selvec_destroy(vq_selvec* sv);

/* ---- dictionary-code GROUP BY ---------------------------------------------
 * ParallelHashAggregate over Scan -> Filter -> Project -> HashAggregate with
 * the int-key fast path (exec/parallel.go:78-86,124-314, exec/agg_intkey.go:
 * 206-361, exec/aggregate.go:361-876).  key_cols name 0..8 dictionary columns
 * (VQ_U8/VQ_U16 with dict_size > 0); nkeys = 0 is the global aggregate
 * (accumulateDirect, aggregate.go:567-745; one row COUNT 0 / NULLs on empty
 * input, aggregate.go:961-982).  Up to 4 columns whose code space (the
 * product of their dict_size) is at most 128 slots, with room for the
 * accumulators in the kernel's per-lane LDS slots, run the dense kernel;
 * any other key set has its codes packed into one 64-bit key as the int-key
 * path packs them (32 bits per column for up to 2 columns, 16 for up to 4, 8
 * for up to 8; exec/agg_intkey.go:67-100) and runs through vq_agg_hash's
 * kernels, float sums then exact; codes wider than their packed width return
 * VQ_E_UNSUPPORTED (the reference's string-key path).  A key column may also
 * be an integer column (VQ_U8/VQ_U16 without a dictionary, VQ_DATE,
 * VQ_I64): a composite key (the string-key path, aggregate.go:409-557) packs
 * its values as they are when its zone maps put them in [0, 2^width), its
 * result keys being those values; one integer key alone has any values (the
 * hash GROUP BY).  Composite keys with NULLs, or outside the width: VQ_E_UNSUPPORTED. */
int vq_agg_dict(vq_ctx* ctx, const vq_table* t, const vq_pred* pred, int32_t nkeys,
                const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                vq_agg_result* out, vq_stats* stats);

/* Several global aggregates (no GROUP BY) over one table in one pass: n
 * queries (at most 16) with the same aggregates and predicates of one shape
 * (the same columns and comparisons; the literals may differ), preds[i] the
 * i-th query's predicate (nterms 0: every row), outs[i] / stats[i] (stats may
 * be NULL) its result, as vq_agg_dict with no group-by columns gives it.  Every
 * row's columns are read once for all of them.  No counterpart in the
 * reference (its operators run one query per call): this is the serving
 * path's continuous batching of queued scans.  Queries not of one shape run
 * one after another. */
int vq_agg_dict_multi(vq_ctx* ctx, const vq_table* t, int32_t n, const vq_pred* preds,
                      int32_t naggs, const vq_agg* aggs, vq_agg_result* outs, vq_stats* stats);

/* The same dictionary-code GROUP BY with no device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, semantics and result as vq_agg_dict;
 * it runs the precompiled kernels only (no run-time compilation) and keeps
 * its partials and final words in a fixed 2 MiB device array that the library
 * holds on each device for the life of the process (no arena, no pinned
 * buffer; fewer waves when their partials would not fit).  Only the dense
 * kernel's shapes: a key set beyond its slots (which vq_agg_dict packs into
 * the hash kernels) returns VQ_E_UNSUPPORTED.  Every call
 * on one device shares that device's array: calls from this library take
 * turns on it (a host lock held until the result is on the host), and work
 * that overlaps it on the same device in any other way races. */
int vq_agg_dict_noalloc(vq_ctx* ctx, const vq_table* t, const vq_pred* pred, int32_t nkeys,
                        const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                        vq_agg_result* out, vq_stats* stats);

/* ---- radix-partitioned hash join feeding an aggregate ---------------------
 * ParallelHashJoinAggregate (exec/parallel_join.go:665-758,780-988) over
 * BuildSharedHashTableParallel (parallel_join.go:389-552) and the flat table
 * of exec/join_table.go:400-569: inner equi-join of build.build_key with
 * probe.probe_key (int64 columns), each side filtered by its own predicate,
 * GROUP BY the join key, aggregates over probe-side expressions, each join
 * output row (a probe row times every build row carrying its key) folded
 * once.  radix_bits < 0 lets the library choose.  topk > 0 keeps the topk
 * groups by aggregate 0 descending (ExternalSort + Limit peeled above the
 * merge, planner/parallel_join.go:479-506; ties in first-occurrence order,
 * the stable-sort order of the serial plan). */
int vq_join_agg(vq_ctx* ctx, const vq_table* build, int32_t build_key, const vq_pred* build_pred,
                const vq_table* probe, int32_t probe_key, const vq_pred* probe_pred,
                int32_t naggs, const vq_agg* aggs, int32_t radix_bits, int32_t topk,
                vq_agg_result* out, vq_stats* stats);

/* ---- join + GROUP BY columns of the joined rows (the Q12 shape) -----------
 * ParallelHashJoinAggregate with its group keys and aggregates over the
 * joined schema (exec/parallel_join.go:665-758,780-988; the Q12 benchmark,
 * planner/parallel_join_bench_test.go:19-30): inner equi-join of
 * build.build_key with probe.probe_key (integer columns, compared as int64),
 * each side filtered by its own predicate (on its own columns), GROUP BY
 * nkeys columns of the probe side (dictionary codes packed as the int-key
 * path packs them, exec/agg_intkey.go:67-100, or one integer column),
 * aggregates over expressions of either side: a build column is written
 * col | VQ_BUILD_COL in a factor, a CASE WHEN term, or the other column of a
 * VQ_COLCMP term.  Each join output row (a probe row times every build row
 * with its key) is folded once.  Groups come out in the order of their first
 * joined probe row (first_row); keys as vq_agg_dict's (nkeys codes) or
 * vq_agg_hash's (one key).  Float sums are exact (fixed point, one rounding),
 * so the result does not depend on the order matches are found in.  The
 * build side goes into a device hash table (key -> its rows); the probe side
 * runs through the few-key GROUP BY's per-workgroup partials, so the group
 * keys are limited as there (8 passes of 256-1024 keys): more returns
 * VQ_E_UNSUPPORTED.  A query touches at most 8 columns of the two tables. */
#define VQ_BUILD_COL 0x10000
int vq_join_group(vq_ctx* ctx, const vq_table* build, int32_t build_key, const vq_pred* build_pred,
                  const vq_table* probe, int32_t probe_key, const vq_pred* probe_pred,
                  int32_t nkeys, const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                  vq_agg_result* out, vq_stats* stats);

/* ---- a join whose build side is itself a join (Q3's chain) ----------------
 * The nested build of ParallelHashJoinAggregate (planner/parallel_join.go:
 * 341-400; exec/parallel_join.go:389-552): the inner join
 * inner.inner_key = build.build_fk (customer JOIN orders, each filtered by its
 * own predicate: a build row meeting m inner rows is m rows of its output, in
 * build-row order) is the build side of the outer join
 * build.build_key = probe.probe_key (JOIN lineitem), GROUP BY the join key,
 * aggregates over probe expressions, topk as vq_join_agg (TPC-H Q3's ORDER BY
 * revenue DESC LIMIT 10).  npay (0..7) build columns that are functions of
 * the key (Q3's o_orderdate, o_shippriority in its GROUP BY) come with each
 * output group: pay_out[g * npay + c], int64 (a float64's raw bits); rows of
 * one key that disagree on them return VQ_E_UNSUPPORTED (they would be
 * separate groups).  Keys compare as int64; a float key is unsupported. */
int vq_join_agg_nested(vq_ctx* ctx, const vq_table* inner, int32_t inner_key, const vq_pred* inner_pred,
                       const vq_table* build, int32_t build_fk, int32_t build_key, const vq_pred* build_pred,
                       const vq_table* probe, int32_t probe_key, const vq_pred* probe_pred, int32_t npay,
                       const int32_t* pay_cols, int32_t naggs, const vq_agg* aggs, int32_t radix_bits,
                       int32_t topk, vq_agg_result* out, int64_t* pay_out, vq_stats* stats);

/* ---- concurrent and asynchronous queries (serving) ------------------------
 * No counterpart in the reference's exec package: its operators run one
 * query per call on the caller's goroutine.  A served query stream
 * (planner.Parallel called from many request goroutines) needs queries that
 * overlap on the device and an admission bound on their device memory.
 *
 * vq_ctx_set_slots: nslots (1..16) queries of this context may run at once,
 * each on its own stream with its own scratch arena (calls from several host
 * threads, or vq_submit, use them); queries started through vq_submit are
 * admitted in arrival order while the device memory they are estimated to
 * need, summed over the running ones, stays within serve_budget bytes
 * (0: half the context budget); a query larger than the bound runs alone.
 * Call with no query in flight. */
int vq_ctx_set_slots(vq_ctx* ctx, int32_t nslots, uint64_t serve_budget);

#define VQ_Q_AGG_DICT 1
#define VQ_Q_JOIN_AGG 2
#define VQ_Q_AGG_HASH 3 /* vq_agg_hash: table, pred, key_cols[0] (nkeys 1), naggs, aggs */

/* One query for vq_submit: the arguments of vq_agg_dict (kind
 * VQ_Q_AGG_DICT: table, pred, nkeys, key_cols, naggs, aggs) or of
 * vq_join_agg (kind VQ_Q_JOIN_AGG: build, build_key, build_pred, table as the
 * probe side, probe_key, pred as the probe predicate, naggs, aggs,
 * radix_bits, topk).  The arrays are copied by vq_submit; the tables, out and
 * stats must stay valid until vq_wait returns. */
typedef struct vq_query {
  int32_t kind;
  int32_t nkeys;
  const vq_table* table;
  const vq_pred* pred;
  const int32_t* key_cols;
  const vq_table* build;
  const vq_pred* build_pred;
  int32_t build_key;
  int32_t probe_key;
  int32_t naggs;
  int32_t radix_bits;
  const vq_agg* aggs;
  int32_t topk;
  int32_t pad_;
} vq_query;

typedef struct vq_ticket vq_ticket;

/* Queue a query; returns at once with a ticket.  The result is written to
 * out/stats by the time vq_wait returns. */
int vq_submit(vq_ctx* ctx, const vq_query* q, vq_agg_result* out, vq_stats* stats,
              vq_ticket** ticket);
/* 1 once the query has finished (successfully or not), else 0. */
int vq_ticket_done(const vq_ticket* ticket);
/* Block until the query has finished; returns its status (its message is then
 * this thread's vq_last_error) and frees the ticket. */
int vq_wait(vq_ticket* ticket);
/* Cancel a query (the Go operator's ctx.Done(), as TableScan checks ctx.Err()
 * per batch, exec/scan.go:199-203, and a failing worker cancels its siblings,
 * exec/parallel_join.go:436-476): still queued, it is dropped and finishes
 * with VQ_E_CANCELED at once; running, it stops at its next launch boundary
 * (a query's kernels run for well under a millisecond each) and finishes
 * with VQ_E_CANCELED, unless it finished first (then its result stands).
 * The ticket stays valid until vq_wait. */
int vq_ticket_cancel(vq_ticket* ticket);
/* Cancel every query queued or running on the context, as vq_ticket_cancel. */
int vq_ctx_cancel(vq_ctx* ctx);

/* GROUP BY one integer column of any cardinality: HashAggregate's general
 * key path (exec/aggregate.go:361-559; the int-key path's output order,
 * exec/agg_intkey.go:206-361) for keys that are not small dictionary codes —
 * u8/u16/date/int64 columns, up to millions of groups.  Up to 8 x 256-1024
 * distinct keys (fewer for more accumulator words): every workgroup
 * aggregates an even share of the rows into its own partials, in passes of
 * up to 256-1024 keys, merged per group (ParallelHashAggregate's partials and
 * mergePartialAgg, exec/parallel.go:18-314); more: rows passing pred are
 * radix-partitioned on radixHash(key) and each partition aggregated in an
 * LDS hash table.  out->keys holds the key values, groups in first-occurrence
 * order; float sums are exact (fixed point) and rounded once, so both ways
 * give the same bits. */
int vq_agg_hash(vq_ctx* ctx, const vq_table* t, const vq_pred* pred, int32_t key_col,
                int32_t naggs, const vq_agg* aggs, vq_agg_result* out, vq_stats* stats);

/* ---- batches of independent queries ----------------------------------------
 * No counterpart in the reference (its operators run one query per call).
 * Runs n independent queries as one unit: every join and hash GROUP BY is
 * planned on the host, its scratch carved from one arena, and all their
 * launches captured into one hipGraph of eight concurrent branches (instance
 * i on branch i mod 8), which is instantiated once and replayed while the
 * batch's launch arguments stay the same.  Dictionary GROUP BYs
 * (VQ_Q_AGG_DICT) run one after another after the graph.  An instance whose
 * partition tables overflow is run again on its own.  outs[i] / stats[i]
 * (stats may be NULL) receive instance i's result; stats[i].device_ms is the
 * whole graph's time.  Returns the first failing instance's status.
 *
 * Flags: VQ_BATCH_TUNE — before planning, autotune each instance class not
 * yet tuned in this context (radix bits around the planner's choice, the
 * scatter's LDS buffer size, the tile size; the fastest device time wins,
 * results are identical under every choice) and keep the choice (also
 * appended to the file named by the VQ_TUNE_FILE environment variable,
 * which a new context reads back); VQ_BATCH_NO_GRAPH — enqueue the same
 * launches on the eight branch streams directly instead of through a graph. */
#define VQ_BATCH_TUNE 1
#define VQ_BATCH_NO_GRAPH 2
int vq_batch_run(vq_ctx* ctx, int32_t n, const vq_query* qs, vq_agg_result* outs,
                 vq_stats* stats, int32_t flags);

/* RadixBitsFor (exec/parallel_join.go:73-82), radixHash (91-99) and
 * joinTableCapacity (exec/join_table.go:461-476), exported so host planners
 * and tests share the library's definitions. */
int32_t vq_radix_bits_for(int64_t est_rows);
uint64_t vq_radix_hash(int64_t key);
int64_t vq_join_table_capacity(int64_t keys);

#ifdef __cplusplus
}
#endif
#endif /* VEXQ_GPU_H */
