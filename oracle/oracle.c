/*
 * oracle.c — CPU restatement of vexq's hot-path operators.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * timed by bench.py.  Nothing in the product path links or calls it.
 *
 * Built with -ffp-contract=off and without -ffast-math, so every float64
 * operation below is one IEEE round-to-nearest operation, as in Go.
 *
 * Parity status of each part (what pins it) is listed in tests/test_oracle.py;
 * every function cites the reference code it restates.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* radix partitioning and table sizing                                       */
/* ------------------------------------------------------------------------ */

/* exec/parallel_join.go:91-99 — murmur3 64-bit finalizer. */
uint64_t or_radix_hash(int64_t key) {
  uint64_t x = (uint64_t)key;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

/* exec/parallel_join.go:26-82 — 0 below 65,536 rows, else 4..6 bits aiming at
 * 8,192 rows per partition. */
int32_t or_radix_bits_for(int64_t est_rows) {
  if (est_rows < 65536) return 0;
  int32_t bits = 4;
  while (bits < 6 && ((int64_t)1 << bits) * 8192 < est_rows) bits++;
  return bits;
}

/* exec/join_table.go:461-476 — smallest power of two >= keys/0.7, min 8. */
int64_t or_join_table_capacity(int64_t keys) {
  int64_t slots = 8;
  if (keys <= 0) return slots;
  int64_t need = (keys * 10 + 7 - 1) / 7;
  while (slots < need) slots *= 2;
  return slots;
}

/* ------------------------------------------------------------------------ */
/* exact summation: a superaccumulator of 32-bit digits                      */
/* ------------------------------------------------------------------------ */

/* Value = sum_i d[i] * 2^(32 i - 1074).  A double m*2^(p-1074) (m < 2^53,
 * p in [0,2045]) lands in three digits; int64 digits holding 32-bit chunks
 * leave 31 bits of headroom, so digits are renormalised every 2^30 adds. */
#define SA_DIGITS 72
typedef struct sacc {
  int64_t d[SA_DIGITS];
  int64_t nadds;
  int32_t nan, pinf, ninf;
  uint64_t nan_bits; /* the first NaN added (an IEEE sum carries its payload) */
} sacc;

static void sa_norm(sacc* a) {
  for (int i = 0; i < SA_DIGITS - 1; i++) {
    int64_t carry = a->d[i] >> 32; /* floor division by 2^32 */
    a->d[i] -= carry * 4294967296LL;
    a->d[i + 1] += carry;
  }
  a->nadds = 0;
}

static void sa_add(sacc* a, double x) {
  uint64_t bits;
  memcpy(&bits, &x, 8);
  int sign = (int)(bits >> 63);
  int ex = (int)((bits >> 52) & 0x7FF);
  uint64_t frac = bits & ((1ULL << 52) - 1);
  if (ex == 0x7FF) {
    if (frac) {
      if (!a->nan) a->nan_bits = bits | (1ULL << 51); /* quieted */
      a->nan = 1;
    } else if (sign) a->ninf = 1;
    else a->pinf = 1;
    return;
  }
  uint64_t m;
  int p;
  if (ex == 0) {
    m = frac;
    p = 0;
  } else {
    m = frac | (1ULL << 52);
    p = ex - 1;
  }
  if (m == 0) return;
  int i = p >> 5, s = p & 31;
  unsigned __int128 w = (unsigned __int128)m << s;
  int64_t c0 = (int64_t)(uint32_t)w;
  int64_t c1 = (int64_t)(uint32_t)(w >> 32);
  int64_t c2 = (int64_t)(uint32_t)(w >> 64);
  if (sign) {
    a->d[i] -= c0;
    a->d[i + 1] -= c1;
    a->d[i + 2] -= c2;
  } else {
    a->d[i] += c0;
    a->d[i + 1] += c1;
    a->d[i + 2] += c2;
  }
  if (++a->nadds >= (1LL << 30)) sa_norm(a);
}

static int sa_bit(const int64_t* d, int64_t pos) {
  if (pos < 0) return 0;
  return (int)((d[pos >> 5] >> (pos & 31)) & 1);
}

/* Correctly rounded (nearest, ties to even) value of the accumulator. */
static double sa_round(const sacc* src) {
  if (src->nan) {
    double d;
    memcpy(&d, &src->nan_bits, 8);
    return d;
  }
  if (src->pinf && src->ninf) return NAN;
  if (src->pinf) return INFINITY;
  if (src->ninf) return -INFINITY;
  sacc a = *src;
  sa_norm(&a);
  int neg = a.d[SA_DIGITS - 1] < 0;
  if (neg) {
    for (int i = 0; i < SA_DIGITS; i++) a.d[i] = -a.d[i];
    sa_norm(&a);
  }
  int h = SA_DIGITS - 1;
  while (h >= 0 && a.d[h] == 0) h--;
  if (h < 0) return 0.0;
  int top = 63 - __builtin_clzll((unsigned long long)a.d[h]);
  int64_t H = (int64_t)h * 32 + top;
  double r;
  if (H < 53) {
    uint64_t n = 0;
    for (int64_t b = H; b >= 0; b--) n = (n << 1) | (uint64_t)sa_bit(a.d, b);
    r = ldexp((double)n, -1074);
  } else {
    uint64_t mant = 0;
    for (int64_t b = H; b >= H - 52; b--) mant = (mant << 1) | (uint64_t)sa_bit(a.d, b);
    int rb = sa_bit(a.d, H - 53);
    int sticky = 0;
    int64_t lim = H - 54;
    for (int64_t dgt = 0; dgt <= (lim >> 5) && !sticky; dgt++) {
      int64_t lo = dgt * 32, hi = lo + 31;
      if (hi <= lim) {
        if (a.d[dgt]) sticky = 1;
      } else {
        for (int64_t b = lo; b <= lim; b++)
          if (sa_bit(a.d, b)) sticky = 1;
      }
    }
    if (rb && (sticky || (mant & 1))) mant++;
    r = ldexp((double)mant, (int)(H - 52 - 1074));
  }
  return neg ? -r : r;
}

double or_exact_sum(const double* x, int64_t n) {
  sacc* a = (sacc*)calloc(1, sizeof(sacc));
  for (int64_t i = 0; i < n; i++) sa_add(a, x[i]);
  double r = sa_round(a);
  free(a);
  return r;
}

/* ------------------------------------------------------------------------ */
/* column access, predicates, arithmetic                                     */
/* ------------------------------------------------------------------------ */

/* extractInt64 (exec/aggregate.go:1083-1099): integers widen, dates sign-
 * extend, codes zero-extend, floats yield their raw bits. */
static inline int64_t col_i(const vq_col_desc* c, int64_t r) {
  switch (c->type) {
    case VQ_U8: return ((const uint8_t*)c->data)[r];
    case VQ_U16: return ((const uint16_t*)c->data)[r];
    case VQ_DATE: return ((const int32_t*)c->data)[r];
    case VQ_I64: return ((const int64_t*)c->data)[r];
    default: {
      int64_t v;
      memcpy(&v, &((const double*)c->data)[r], 8);
      return v;
    }
  }
}

/* Validity (exec/batch.go:113-121, storage.IsNullBit): bit r of the LSB-first
 * bitmap set when row r holds a value; no bitmap, every row does. */
static inline int col_valid(const vq_col_desc* c, int64_t r) {
  return !c->valid || ((c->valid[r >> 3] >> (r & 7)) & 1);
}

/* CastIntToFloatExpr (exec/expr.go:194-222) for integer columns. */
static inline double col_f(const vq_col_desc* c, int64_t r) {
  if (c->type == VQ_F64) return ((const double*)c->data)[r];
  return (double)col_i(c, r);
}

static inline int cmp_i(int op, int64_t a, int64_t b) {
  switch (op) {
    case VQ_EQ: return a == b;
    case VQ_NE: return a != b;
    case VQ_LT: return a < b;
    case VQ_LE: return a <= b;
    case VQ_GT: return a > b;
    default: return a >= b;
  }
}

/* evalCmpFloat64 (exec/expr.go:513-): IEEE compares, NaN compares false
 * except for NE. */
static inline int cmp_f(int op, double a, double b) {
  switch (op) {
    case VQ_EQ: return a == b;
    case VQ_NE: return a != b;
    case VQ_LT: return a < b;
    case VQ_LE: return a <= b;
    case VQ_GT: return a > b;
    default: return a >= b;
  }
}

/* Filter over a conjunction (AndExpr, exec/expr.go:1030-1067). */
/* One predicate term on row r.  A compare that reads a NULL is NULL
 * (mergeNullBitmapsInto, exec/expr.go:290-299, 964-980), and the filter keeps
 * only rows whose result is TRUE and valid, so the term fails. */
static inline int term_pass(const vq_col_desc* cols, const vq_term* tm, int64_t r) {
  const vq_col_desc* c = &cols[tm->col];
  if (!col_valid(c, r)) return 0;
  if ((tm->op & VQ_COLCMP) && tm->op != VQ_IN && !col_valid(&cols[tm->ival], r)) return 0;
  if (tm->op == VQ_IN) {
    /* InExpr (exec/expr.go:1188-1238) over codes: the value is one of the
     * set's members, given as a mask over 0..63 */
    const int64_t v = col_i(c, r);
    return v >= 0 && v < 64 && (((uint64_t)tm->ival >> v) & 1u);
  }
  if (tm->op & VQ_COLCMP) {
    /* a BinOp compare of two columns (exec/expr.go:320-896) */
    const vq_col_desc* c2 = &cols[tm->ival];
    const int op = tm->op & ~VQ_COLCMP;
    return c->type == VQ_F64 ? cmp_f(op, col_f(c, r), col_f(c2, r)) : cmp_i(op, col_i(c, r), col_i(c2, r));
  }
  return c->type == VQ_F64 ? cmp_f(tm->op, col_f(c, r), tm->fval) : cmp_i(tm->op, col_i(c, r), tm->ival);
}

static inline int pred_pass(const vq_col_desc* cols, const vq_pred* p, int64_t r) {
  if (!p) return 1;
  for (int t = 0; t < p->nterms; t++)
    if (!term_pass(cols, &p->terms[t], r)) return 0;
  return 1;
}

/* CaseExpr with one WHEN and a literal ELSE (exec/expr.go:1386-1435): the
 * WHEN's conjunction on row r (no CASE: always).  mergeVectors takes the THEN
 * value only where the condition is TRUE and valid, so a WHEN reading a NULL
 * selects the ELSE literal. */
static inline int when_pass(const vq_col_desc* cols, const vq_expr* e, int64_t r) {
  for (int t = 0; t < e->nwhen; t++)
    if (!term_pass(cols, &e->when[t], r)) return 0;
  return 1;
}

/* Is the expression's value on row r non-NULL?  Arithmetic over a NULL is
 * NULL (the operands' validity ANDed, exec/expr.go:290-299); literals always
 * hold a value, and so does the ELSE literal of a CASE whose WHEN fails. */
static inline int expr_valid(const vq_col_desc* cols, const vq_expr* e, int64_t r) {
  if (!when_pass(cols, e, r)) return 1;
  for (int k = 0; k < e->nfactors; k++)
    if (e->f[k].kind != VQ_F_LIT && !col_valid(&cols[e->f[k].col], r)) return 0;
  return 1;
}

/* Go's math.NaN(): the bits of uvnan (0x7FF8000000000001). */
static inline double go_nan(void) {
  const uint64_t b = 0x7FF8000000000001ULL;
  double d;
  memcpy(&d, &b, 8);
  return d;
}

/* applyArithFloat64 (exec/expr.go:1005-1020): x/0 is math.NaN(). */
static inline double arith_f(int op, double a, double b) {
  switch (op) {
    case VQ_ADD: return a + b;
    case VQ_SUB: return a - b;
    case VQ_MUL: return a * b;
    default: return b == 0 ? go_nan() : a / b;
  }
}

/* applyArithInt64 (exec/expr.go:990-1003): wrapping, x/0 is 0. */
static inline int64_t arith_i(int op, int64_t a, int64_t b) {
  switch (op) {
    case VQ_ADD: return (int64_t)((uint64_t)a + (uint64_t)b);
    case VQ_SUB: return (int64_t)((uint64_t)a - (uint64_t)b);
    case VQ_MUL: return (int64_t)((uint64_t)a * (uint64_t)b);
    default:
      if (b == 0) return 0;
      if (a == INT64_MIN && b == -1) return INT64_MIN;
      return a / b;
  }
}

static inline double expr_f(const vq_col_desc* cols, const vq_expr* e, int64_t r) {
  if (!when_pass(cols, e, r)) return e->else_f;
  double v = 0;
  for (int k = 0; k < e->nfactors; k++) {
    const vq_factor* f = &e->f[k];
    double x;
    switch (f->kind) {
      case VQ_F_COL: x = col_f(&cols[f->col], r); break;
      case VQ_F_LIT: x = f->fval; break;
      case VQ_F_LIT_OP_COL: x = arith_f(f->op, f->fval, col_f(&cols[f->col], r)); break;
      default: x = arith_f(f->op, col_f(&cols[f->col], r), f->fval); break;
    }
    v = k == 0 ? x : arith_f(f->combine, v, x);
  }
  return v;
}

static inline int64_t expr_i(const vq_col_desc* cols, const vq_expr* e, int64_t r) {
  if (!when_pass(cols, e, r)) return e->else_i;
  int64_t v = 0;
  for (int k = 0; k < e->nfactors; k++) {
    const vq_factor* f = &e->f[k];
    int64_t x;
    switch (f->kind) {
      case VQ_F_COL: x = col_i(&cols[f->col], r); break;
      case VQ_F_LIT: x = f->ival; break;
      case VQ_F_LIT_OP_COL: x = arith_i(f->op, f->ival, col_i(&cols[f->col], r)); break;
      default: x = arith_i(f->op, col_i(&cols[f->col], r), f->ival); break;
    }
    v = k == 0 ? x : arith_i(f->combine, v, x);
  }
  return v;
}

/* ------------------------------------------------------------------------ */
/* zone maps and pruning                                                     */
/* ------------------------------------------------------------------------ */

/* storage/writer.go:205-348 (and 390-395 for dictionary columns): min and
 * max over the rows that hold a value, the NULL count, and no min/max for a
 * row group with no value at all. */
void or_zone_maps(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                  const int64_t* rg_rows, vq_zone* out) {
  (void)nrows;
  int64_t base = 0;
  for (int32_t g = 0; g < nrg; g++) {
    int64_t n = rg_rows[g];
    for (int32_t c = 0; c < ncols; c++) {
      vq_zone z;
      memset(&z, 0, sizeof z);
      const vq_col_desc* cd = &cols[c];
      int64_t seen = 0, imn = 0, imx = 0;
      double fmn = 0, fmx = 0;
      for (int64_t r = base; r < base + n; r++) {
        if (!col_valid(cd, r)) {
          z.null_count++;
          continue;
        }
        if (cd->type == VQ_F64) {
          const double v = ((const double*)cd->data)[r];
          if (!seen || v < fmn) fmn = v;
          if (!seen || v > fmx) fmx = v;
        } else {
          const int64_t v = col_i(cd, r);
          if (!seen || v < imn) imn = v;
          if (!seen || v > imx) imx = v;
        }
        seen++;
      }
      if (seen > 0) {
        z.has_minmax = 1;
        if ((cd->type == VQ_U8 || cd->type == VQ_U16) && cd->dict_size > 0) {
          z.min_bits = 0;
          z.max_bits = (uint64_t)(cd->dict_size - 1);
        } else if (cd->type == VQ_F64) {
          memcpy(&z.min_bits, &fmn, 8);
          memcpy(&z.max_bits, &fmx, 8);
        } else if (cd->type == VQ_DATE) {
          z.min_bits = (uint64_t)(uint32_t)(int32_t)imn;
          z.max_bits = (uint64_t)(uint32_t)(int32_t)imx;
        } else {
          z.min_bits = (uint64_t)imn;
          z.max_bits = (uint64_t)imx;
        }
      }
      out[(int64_t)g * ncols + c] = z;
    }
    base += n;
  }
}

/* planner/physical.go:631-750: zoneRangePred compares the literal with the
 * row group's min/max as raw signed int64 bits; a float literal contributes
 * its IEEE bits.  Dictionary-code columns never prune (a string literal is
 * not an int64 literal there).  AND keeps a row group only if every term can
 * match; NE never prunes. */
static int zone_term_keep(const vq_col_desc* c, const vq_zone* z, const vq_term* t) {
  if (!z->has_minmax) return 1;
  if (t->op == VQ_IN || (t->op & VQ_COLCMP)) return 1; /* (never prunes) */
  if ((c->type == VQ_U8 || c->type == VQ_U16) && c->dict_size > 0) return 1;
  int64_t v;
  if (c->type == VQ_F64) memcpy(&v, &t->fval, 8);
  else v = t->ival;
  int64_t mn = (int64_t)z->min_bits, mx = (int64_t)z->max_bits;
  switch (t->op) {
    case VQ_EQ: return mn <= v && v <= mx;
    case VQ_NE: return 1;
    case VQ_LT: return mn < v;
    case VQ_LE: return mn <= v;
    case VQ_GT: return mx > v;
    default: return mx >= v;
  }
}

void or_zone_prune(int32_t ncols, const vq_col_desc* cols, int32_t nrg, const vq_zone* zones,
                   const vq_pred* pred, uint8_t* keep) {
  for (int32_t g = 0; g < nrg; g++) {
    int k = 1;
    if (pred)
      for (int t = 0; t < pred->nterms && k; t++)
        k = zone_term_keep(&cols[pred->terms[t].col],
                           &zones[(int64_t)g * ncols + pred->terms[t].col], &pred->terms[t]);
    keep[g] = (uint8_t)k;
  }
}

int64_t or_filter(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                  const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                  uint32_t* out_rows) {
  (void)nrows;
  uint8_t* keep = (uint8_t*)malloc((size_t)nrg + 1);
  or_zone_prune(ncols, cols, nrg, zones, pred, keep);
  int64_t n = 0, base = 0;
  for (int32_t g = 0; g < nrg; g++) {
    if (keep[g])
      for (int64_t r = base; r < base + rg_rows[g]; r++)
        if (pred_pass(cols, pred, r)) out_rows[n++] = (uint32_t)r;
    base += rg_rows[g];
  }
  free(keep);
  return n;
}

/* The same selection over row-group morsels claimed from an atomic cursor
 * by nthreads workers (the morsel plan of exec/parallel.go:18-36,143-183):
 * each morsel's passing rows go to that row group's own stretch of out_rows
 * (its rows are ascending and every row group starts after the previous
 * one's), then the stretches are closed up in row-group order, so the
 * result is the serial one, ascending. */
typedef struct filter_job {
  const vq_col_desc* cols;
  const vq_pred* pred;
  const int64_t* rg_base;
  const int64_t* rg_rows;
  const uint8_t* keep;
  int64_t* rg_count;
  uint32_t* out_rows;
  int64_t next, end; /* the shared cursor (first job's) */
  struct filter_job* shared;
} filter_job;

static void* filter_worker(void* arg) {
  filter_job* j = (filter_job*)arg;
  for (;;) {
    const int64_t g = __atomic_fetch_add(&j->shared->next, 1, __ATOMIC_RELAXED);
    if (g >= j->shared->end) break;
    int64_t n = 0;
    if (j->keep[g])
      for (int64_t r = j->rg_base[g]; r < j->rg_base[g] + j->rg_rows[g]; r++)
        if (pred_pass(j->cols, j->pred, r)) j->out_rows[j->rg_base[g] + n++] = (uint32_t)r;
    j->rg_count[g] = n;
  }
  return NULL;
}

int64_t or_filter_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                           const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                           int32_t nthreads, uint32_t* out_rows) {
  (void)nrows;
  if (nthreads < 1) nthreads = 1;
  uint8_t* keep = (uint8_t*)malloc((size_t)nrg + 1);
  or_zone_prune(ncols, cols, nrg, zones, pred, keep);
  int64_t* rg_base = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nrg + 1));
  int64_t* rg_count = (int64_t*)calloc((size_t)nrg + 1, sizeof(int64_t));
  rg_base[0] = 0;
  for (int32_t g = 0; g < nrg; g++) rg_base[g + 1] = rg_base[g] + rg_rows[g];
  filter_job* jobs = (filter_job*)calloc((size_t)nthreads, sizeof(filter_job));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  jobs[0].end = nrg;
  for (int i = 0; i < nthreads; i++) {
    jobs[i] = (filter_job){cols, pred, rg_base, rg_rows, keep, rg_count, out_rows, 0, nrg, &jobs[0]};
    pthread_create(&th[i], NULL, filter_worker, &jobs[i]);
  }
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  int64_t n = 0;
  for (int32_t g = 0; g < nrg; g++) {
    if (n != rg_base[g]) memmove(&out_rows[n], &out_rows[rg_base[g]], (size_t)rg_count[g] * 4);
    n += rg_count[g];
  }
  free(th);
  free(jobs);
  free(rg_count);
  free(rg_base);
  free(keep);
  return n;
}

/* ------------------------------------------------------------------------ */
/* aggregate state                                                           */
/* ------------------------------------------------------------------------ */

/* One group's accumulators: HashAggregate keeps []int64 per group with float64
 * stored as IEEE bits (exec/aggregate.go:83-129); newAccums seeds MIN/MAX with
 * the extreme sentinel (aggregate.go:255-279). */
typedef struct group {
  int64_t key[8];
  int64_t count;
  int64_t first_row;
  int64_t acc[VQ_MAX_AGGS];
  int64_t nonnull[VQ_MAX_AGGS];
  sacc* ex[VQ_MAX_AGGS]; /* exact float sums (exact mode only) */
} group;

static void group_init(group* g, int32_t naggs, const vq_agg* aggs, int exact) {
  memset(g, 0, sizeof *g);
  g->first_row = -1;
  for (int j = 0; j < naggs; j++) {
    const vq_agg* a = &aggs[j];
    if (a->kind == VQ_AGG_MIN) {
      if (a->accum == VQ_F64) {
        double v = 1.7976931348623157e308;
        memcpy(&g->acc[j], &v, 8);
      } else g->acc[j] = INT64_MAX;
    } else if (a->kind == VQ_AGG_MAX) {
      if (a->accum == VQ_F64) {
        double v = -1.7976931348623157e308;
        memcpy(&g->acc[j], &v, 8);
      } else g->acc[j] = INT64_MIN;
    }
    if (exact && (a->kind == VQ_AGG_AVG || (a->kind == VQ_AGG_SUM && a->accum == VQ_F64)))
      g->ex[j] = (sacc*)calloc(1, sizeof(sacc));
  }
}

static void group_free(group* g, int32_t naggs) {
  for (int j = 0; j < naggs; j++) free(g->ex[j]);
}

static inline double bits_f(int64_t b) {
  double d;
  memcpy(&d, &b, 8);
  return d;
}
static inline int64_t f_bits(double d) {
  int64_t b;
  memcpy(&b, &d, 8);
  return b;
}

/* ---- strings of dictionary columns (MIN/MAX over strings) --------------- */
static int32_t g_ndicts;
static const char* const* const* g_strs;
static const int64_t* const* g_lens;

void or_set_dicts(int32_t ncols, const char* const* const* strs, const int64_t* const* lens) {
  g_ndicts = ncols;
  g_strs = strs;
  g_lens = lens;
}

/* Go's string compare (bytewise, a prefix first) of two codes' strings. */
static int str_cmp(int32_t col, int64_t a, int64_t b) {
  if (col < 0 || col >= g_ndicts || !g_strs[col]) return (a > b) - (a < b); /* (no strings: codes) */
  const int64_t la = g_lens[col][a], lb = g_lens[col][b];
  const int c = memcmp(g_strs[col][a], g_strs[col][b], (size_t)(la < lb ? la : lb));
  if (c) return c;
  return (la > lb) - (la < lb);
}

/* minMaxString (exec/aggregate.go:670-730): the group's first value, then
 * the smaller (MIN) or larger (MAX) string; acc holds the winner's code. */
static void str_minmax(group* g, int j, const vq_agg* a, int64_t code) {
  const int c = g->nonnull[j] ? str_cmp(a->arg.f[0].col, code, g->acc[j]) : 0;
  if (!g->nonnull[j] || (a->kind == VQ_AGG_MIN ? c < 0 : c > 0)) g->acc[j] = code;
}

/* foldRow (exec/agg_intkey.go:303-361): COUNT(*) counts every row, COUNT(x)
 * the rows where x holds a value; SUM/AVG/MIN/MAX skip a NULL argument and
 * count the values they fold in nonNull (aggregate.go:458-546). */
static void fold_row(group* g, const vq_col_desc* cols, int32_t naggs, const vq_agg* aggs,
                     int64_t r) {
  g->count++;
  for (int j = 0; j < naggs; j++) {
    const vq_agg* a = &aggs[j];
    if (a->arg.nfactors > 0 && !expr_valid(cols, &a->arg, r)) continue;
    switch (a->kind) {
      case VQ_AGG_COUNT:
        g->acc[j]++;
        g->nonnull[j]++;
        break;
      case VQ_AGG_SUM:
        g->nonnull[j]++;
        if (a->accum == VQ_F64) {
          double v = expr_f(cols, &a->arg, r);
          if (g->ex[j]) sa_add(g->ex[j], v);
          g->acc[j] = f_bits(bits_f(g->acc[j]) + v);
        } else {
          g->acc[j] = (int64_t)((uint64_t)g->acc[j] + (uint64_t)expr_i(cols, &a->arg, r));
        }
        break;
      case VQ_AGG_AVG: {
        g->nonnull[j]++;
        double v = a->arg.is_float ? expr_f(cols, &a->arg, r) : (double)expr_i(cols, &a->arg, r);
        if (g->ex[j]) sa_add(g->ex[j], v);
        g->acc[j] = f_bits(bits_f(g->acc[j]) + v);
        break;
      }
      case VQ_AGG_MIN:
      case VQ_AGG_MAX: {
        if (a->accum == VQ_STR) {
          str_minmax(g, j, a, col_i(&cols[a->arg.f[0].col], r));
          g->nonnull[j]++;
          break;
        }
        g->nonnull[j]++;
        int less = a->kind == VQ_AGG_MIN;
        if (a->accum == VQ_F64) {
          double v = expr_f(cols, &a->arg, r), cur = bits_f(g->acc[j]);
          if ((less && v < cur) || (!less && v > cur)) g->acc[j] = f_bits(v);
        } else {
          int64_t v = expr_i(cols, &a->arg, r);
          if ((less && v < g->acc[j]) || (!less && v > g->acc[j])) g->acc[j] = v;
        }
        break;
      }
    }
  }
}

/* mergePartialAgg (exec/parallel.go:237-314). */
static void merge_group(group* dst, const group* src, int32_t naggs, const vq_agg* aggs) {
  if (dst->first_row < 0 || (src->first_row >= 0 && src->first_row < dst->first_row))
    dst->first_row = src->first_row;
  dst->count += src->count;
  for (int j = 0; j < naggs; j++) {
    const vq_agg* a = &aggs[j];
    switch (a->kind) {
      case VQ_AGG_COUNT: dst->acc[j] += src->acc[j]; break;
      case VQ_AGG_SUM:
      case VQ_AGG_AVG:
        if (a->accum == VQ_F64 || a->kind == VQ_AGG_AVG)
          dst->acc[j] = f_bits(bits_f(dst->acc[j]) + bits_f(src->acc[j]));
        else dst->acc[j] += src->acc[j];
        break;
      default: {
        int less = a->kind == VQ_AGG_MIN;
        if (a->accum == VQ_STR) {
          if (src->nonnull[j]) str_minmax(dst, j, a, src->acc[j]);
          break;
        }
        if (a->accum == VQ_F64) {
          double d = bits_f(dst->acc[j]), s = bits_f(src->acc[j]);
          if ((less && s < d) || (!less && s > d)) dst->acc[j] = src->acc[j];
        } else if ((less && src->acc[j] < dst->acc[j]) || (!less && src->acc[j] > dst->acc[j]))
          dst->acc[j] = src->acc[j];
      }
    }
    dst->nonnull[j] += src->nonnull[j];
  }
}

/* buildOutputBatch (exec/aggregate.go:780-876) for one group into row i. */
static void emit_group(const group* g, int32_t nkeys, int32_t naggs, const vq_agg* aggs,
                       vq_agg_result* out, int64_t i) {
  for (int k = 0; k < nkeys; k++) out->keys[i * nkeys + k] = g->key[k];
  out->counts[i] = g->count;
  out->first_row[i] = g->first_row;
  for (int j = 0; j < naggs; j++) {
    const vq_agg* a = &aggs[j];
    int64_t o = i * naggs + j;
    out->ivals[o] = 0;
    out->fvals[o] = 0;
    out->valid[o] = 1;
    if (a->kind == VQ_AGG_COUNT) {
      out->ivals[o] = g->acc[j];
      continue;
    }
    if (g->nonnull[j] == 0) {
      out->valid[o] = 0;
      continue;
    }
    if (a->kind == VQ_AGG_AVG) {
      double s = g->ex[j] ? sa_round(g->ex[j]) : bits_f(g->acc[j]);
      out->fvals[o] = s / (double)g->nonnull[j];
    } else if (a->accum == VQ_F64) {
      out->fvals[o] = (a->kind == VQ_AGG_SUM && g->ex[j]) ? sa_round(g->ex[j]) : bits_f(g->acc[j]);
    } else {
      out->ivals[o] = g->acc[j];
    }
  }
}

/* buildEmptyGlobalResult (exec/aggregate.go:961-982). */
static void emit_empty_global(int32_t naggs, const vq_agg* aggs, vq_agg_result* out) {
  out->counts[0] = 0;
  out->first_row[0] = -1;
  for (int j = 0; j < naggs; j++) {
    out->ivals[j] = 0;
    out->fvals[j] = 0;
    out->valid[j] = aggs[j].kind == VQ_AGG_COUNT;
  }
}

/* ------------------------------------------------------------------------ */
/* int64-keyed open-addressing map (insertion-ordered groups)                 */
/* ------------------------------------------------------------------------ */

typedef struct kmap {
  int64_t* keys;
  int32_t* idx; /* -1 = empty */
  int64_t cap;
  int64_t n;
} kmap;

static void kmap_init(kmap* m, int64_t cap) {
  m->cap = 16;
  while (m->cap < cap * 2) m->cap *= 2;
  m->keys = (int64_t*)malloc(sizeof(int64_t) * m->cap);
  m->idx = (int32_t*)malloc(sizeof(int32_t) * m->cap);
  memset(m->idx, 0xff, sizeof(int32_t) * m->cap);
  m->n = 0;
}

static void kmap_free(kmap* m) {
  free(m->keys);
  free(m->idx);
}

/* Returns the slot position; *found tells whether the key was present. */
static int64_t kmap_find(const kmap* m, int64_t key, int* found) {
  uint64_t i = or_radix_hash(key) & (uint64_t)(m->cap - 1);
  for (;;) {
    if (m->idx[i] < 0) {
      *found = 0;
      return (int64_t)i;
    }
    if (m->keys[i] == key) {
      *found = 1;
      return (int64_t)i;
    }
    i = (i + 1) & (uint64_t)(m->cap - 1);
  }
}

static void kmap_grow(kmap* m) {
  kmap nm;
  nm.cap = m->cap * 2;
  nm.keys = (int64_t*)malloc(sizeof(int64_t) * nm.cap);
  nm.idx = (int32_t*)malloc(sizeof(int32_t) * nm.cap);
  memset(nm.idx, 0xff, sizeof(int32_t) * nm.cap);
  nm.n = m->n;
  for (int64_t i = 0; i < m->cap; i++)
    if (m->idx[i] >= 0) {
      int f;
      int64_t s = kmap_find(&nm, m->keys[i], &f);
      nm.keys[s] = m->keys[i];
      nm.idx[s] = m->idx[i];
    }
  kmap_free(m);
  *m = nm;
}

/* Group list that grows; groups are kept in insertion order. */
typedef struct glist {
  group* g;
  int64_t n, cap;
} glist;

static group* glist_push(glist* l) {
  if (l->n == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 64;
    l->g = (group*)realloc(l->g, sizeof(group) * l->cap);
  }
  return &l->g[l->n++];
}

/* Look up key; create the group (insertion order) if absent. */
static group* kmap_group(kmap* m, glist* l, int64_t key, int32_t naggs, const vq_agg* aggs,
                         int exact) {
  int f;
  int64_t s = kmap_find(m, key, &f);
  if (f) return &l->g[m->idx[s]];
  if ((m->n + 1) * 2 > m->cap) {
    kmap_grow(m);
    s = kmap_find(m, key, &f);
  }
  m->keys[s] = key;
  m->idx[s] = (int32_t)l->n;
  m->n++;
  group* g = glist_push(l);
  group_init(g, naggs, aggs, exact);
  return g;
}

static int check_cap(vq_agg_result* out, int64_t n) {
  out->ngroups = n;
  return n > out->capacity ? VQ_E_ARG : VQ_OK;
}

/* Group-key packing of the int-key path (exec/agg_intkey.go:67-100): 32 bits
 * per column for <= 2 columns, 16 for <= 4, 8 for <= 8. */
static inline int key_width(int32_t nkeys) { return nkeys <= 2 ? 32 : nkeys <= 4 ? 16 : 8; }

/* A key column's code on row r; a NULL takes the code past the dictionary, a
 * group of its own (the isNull sample of exec/aggregate.go:424-427). */
static inline int64_t key_code(const vq_col_desc* c, int64_t r) {
  return col_valid(c, r) ? col_i(c, r) : c->dict_size;
}

static inline int64_t pack_key(const vq_col_desc* cols, int32_t nkeys, const int32_t* key_cols,
                               int64_t r) {
  const int shift = key_width(nkeys);
  uint64_t k = 0;
  for (int i = 0; i < nkeys; i++) k |= (uint64_t)key_code(&cols[key_cols[i]], r) << (shift * i);
  return (int64_t)k;
}

/* Result keys: a NULL group's code (past the dictionary) is reported as -1. */
static void null_keys(const vq_col_desc* cols, int32_t nkeys, const int32_t* key_cols,
                      vq_agg_result* out) {
  for (int k = 0; k < nkeys; k++) {
    const vq_col_desc* c = &cols[key_cols[k]];
    if (!c->valid) continue;
    for (int64_t i = 0; i < out->ngroups; i++)
      if (out->keys[i * nkeys + k] == c->dict_size) out->keys[i * nkeys + k] = -1;
  }
}

static void unpack_key(int64_t packed, int32_t nkeys, int64_t* key) {
  const int shift = key_width(nkeys);
  const uint64_t mask = shift == 32 ? 0xFFFFFFFFULL : ((1ULL << shift) - 1);
  for (int i = 0; i < nkeys; i++) key[i] = (int64_t)(((uint64_t)packed >> (shift * i)) & mask);
}

/* Every code of a key column fits its packed width (the reference leaves the
 * int-key path for the string-key one otherwise, agg_intkey.go:80-100; this
 * restatement then refuses the query).  An integer key column (no dictionary:
 * a composite key of the string-key path, aggregate.go:409-557) packs its
 * values as they are: every one in [0, 2^width), and no NULL (whose code
 * would collide with a value). */
static int keys_fit(int64_t nrows, const vq_col_desc* cols, int32_t nkeys, const int32_t* key_cols) {
  const int w = key_width(nkeys);
  for (int i = 0; i < nkeys; i++) {
    const vq_col_desc* c = &cols[key_cols[i]];
    if (c->dict_size > 0) {
      const int64_t codes = (int64_t)c->dict_size + (c->valid ? 1 : 0);
      if (w < 32 && codes > (1 << w)) return 0;
      continue;
    }
    if (c->type == VQ_F64 || c->valid) return 0;
    for (int64_t r = 0; r < nrows; r++) {
      const int64_t v = col_i(c, r);
      if (v < 0 || (uint64_t)v >= (1ULL << w)) return 0;
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* serial dictionary-code GROUP BY                                           */
/* ------------------------------------------------------------------------ */

/* Serial plan: TableScan(zone pruned, exec/scan.go:199-236) -> Filter ->
 * HashAggregate.accumulate over the int-key path (agg_intkey.go:206-288) or
 * accumulateDirect (aggregate.go:567-745).  Rows arrive in row order, so the
 * float sums are the serial plan's left fold bit for bit. */
int or_agg_dict(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                int32_t nkeys, const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                int32_t exact, vq_agg_result* out) {
  (void)nrows;
  if (nkeys < 0 || nkeys > 8 || naggs < 0 || naggs > VQ_MAX_AGGS) return VQ_E_ARG;
  if (!keys_fit(nrows, cols, nkeys, key_cols)) return VQ_E_UNSUPPORTED;
  uint8_t* keep = (uint8_t*)malloc((size_t)nrg + 1);
  or_zone_prune(ncols, cols, nrg, zones, pred, keep);
  kmap m;
  kmap_init(&m, 64);
  glist l = {0, 0, 0};
  int64_t base = 0;
  for (int32_t rg = 0; rg < nrg; rg++) {
    if (keep[rg])
      for (int64_t r = base; r < base + rg_rows[rg]; r++) {
        if (!pred_pass(cols, pred, r)) continue;
        int64_t pk = nkeys ? pack_key(cols, nkeys, key_cols, r) : 0;
        group* g = kmap_group(&m, &l, pk, naggs, aggs, exact);
        if (g->first_row < 0) {
          g->first_row = r;
          unpack_key(pk, nkeys, g->key);
        }
        fold_row(g, cols, naggs, aggs, r);
      }
    base += rg_rows[rg];
  }
  int rc;
  if (l.n == 0 && nkeys == 0) {
    rc = check_cap(out, 1);
    if (rc == VQ_OK) emit_empty_global(naggs, aggs, out);
  } else {
    rc = check_cap(out, l.n);
    if (rc == VQ_OK) {
      for (int64_t i = 0; i < l.n; i++) emit_group(&l.g[i], nkeys, naggs, aggs, out, i);
      null_keys(cols, nkeys, key_cols, out);
    }
  }
  for (int64_t i = 0; i < l.n; i++) group_free(&l.g[i], naggs);
  free(l.g);
  kmap_free(&m);
  free(keep);
  return rc;
}

/* Serial GROUP BY one integer column: TableScan(zone pruned) -> Filter ->
 * HashAggregate over the general key path (exec/aggregate.go:361-559), groups
 * in first-seen order (the order its output walks, aggregate.go:780-876). */
int or_agg_hash(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                int32_t key_col, int32_t naggs, const vq_agg* aggs, int32_t exact,
                vq_agg_result* out) {
  (void)nrows;
  if (key_col < 0 || key_col >= ncols || naggs < 0 || naggs > VQ_MAX_AGGS) return VQ_E_ARG;
  uint8_t* keep = (uint8_t*)malloc((size_t)nrg + 1);
  or_zone_prune(ncols, cols, nrg, zones, pred, keep);
  kmap m;
  kmap_init(&m, 1024);
  glist l = {0, 0, 0};
  int64_t base = 0;
  for (int32_t rg = 0; rg < nrg; rg++) {
    if (keep[rg])
      for (int64_t r = base; r < base + rg_rows[rg]; r++) {
        if (!pred_pass(cols, pred, r)) continue;
        const int64_t key = col_i(&cols[key_col], r);
        group* g = kmap_group(&m, &l, key, naggs, aggs, exact);
        if (g->first_row < 0) {
          g->first_row = r;
          g->key[0] = key;
        }
        fold_row(g, cols, naggs, aggs, r);
      }
    base += rg_rows[rg];
  }
  int rc = check_cap(out, l.n);
  if (rc == VQ_OK)
    for (int64_t i = 0; i < l.n; i++) emit_group(&l.g[i], 1, naggs, aggs, out, i);
  for (int64_t i = 0; i < l.n; i++) group_free(&l.g[i], naggs);
  free(l.g);
  kmap_free(&m);
  free(keep);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* flat join table (exec/join_table.go)                                      */
/* ------------------------------------------------------------------------ */

/* joinSlot {key, head, tail}; head == -1 marks an empty slot, so key 0 is an
 * ordinary key (join_table.go:39-53,411-415). */
typedef struct jslot {
  int64_t key;
  int32_t head, tail;
} jslot;

typedef struct jtable {
  jslot* slots;
  int64_t cap;
  int shift;
  int64_t keys, rows, grow_at;
  int32_t* next; /* shared chain links, indexed by build row */
} jtable;

static void jt_set_capacity(jtable* t, int64_t cap) {
  t->slots = (jslot*)malloc(sizeof(jslot) * cap);
  for (int64_t i = 0; i < cap; i++) t->slots[i].head = -1;
  t->cap = cap;
  t->shift = 64 - __builtin_ctzll((unsigned long long)cap);
  t->grow_at = cap * 7 / 10;
}

static void jt_init(jtable* t, int64_t expected_keys, int32_t* next) {
  memset(t, 0, sizeof *t);
  t->next = next;
  jt_set_capacity(t, or_join_table_capacity(expected_keys));
}

/* slotFor: the high log2(cap) bits of radixHash (join_table.go:488-492). */
static inline uint64_t jt_slot(const jtable* t, uint64_t h) {
  return t->shift >= 64 ? 0 : h >> t->shift;
}

static void jt_grow(jtable* t) {
  jslot* old = t->slots;
  int64_t oc = t->cap;
  jt_set_capacity(t, oc * 2);
  for (int64_t i = 0; i < oc; i++) {
    if (old[i].head == -1) continue;
    uint64_t s = jt_slot(t, or_radix_hash(old[i].key));
    while (t->slots[s].head != -1) s = (s + 1) & (uint64_t)(t->cap - 1);
    t->slots[s] = old[i];
  }
  free(old);
}

/* insert (join_table.go:496-517): append row to key's chain. */
static void jt_insert(jtable* t, int64_t key, int32_t row) {
  if (t->keys >= t->grow_at) jt_grow(t);
  t->next[row] = -1;
  t->rows++;
  uint64_t i = jt_slot(t, or_radix_hash(key));
  for (;;) {
    jslot* s = &t->slots[i];
    if (s->head == -1) {
      s->key = key;
      s->head = s->tail = row;
      t->keys++;
      return;
    }
    if (s->key == key) {
      t->next[s->tail] = row;
      s->tail = row;
      return;
    }
    i = (i + 1) & (uint64_t)(t->cap - 1);
  }
}

/* lookupHashed (join_table.go:529-541). */
static int32_t jt_lookup(const jtable* t, int64_t key, uint64_t h) {
  uint64_t i = jt_slot(t, h);
  for (;;) {
    const jslot* s = &t->slots[i];
    if (s->head == -1) return -1;
    if (s->key == key) return s->head;
    i = (i + 1) & (uint64_t)(t->cap - 1);
  }
}

/* ExternalSort.less (exec/sort.go:130-163) on aggregate 0, descending: the
 * sort compares the int64 a row carries — a float's raw IEEE bits — so ties
 * and ordering follow signed comparison of those bits.  The stable sort keeps
 * first-occurrence order among ties. */
typedef struct rank_item {
  int64_t v;
  int64_t pos;
} rank_item;

static int rank_cmp(const void* a, const void* b) {
  const rank_item* x = (const rank_item*)a;
  const rank_item* y = (const rank_item*)b;
  if (x->v != y->v) return x->v > y->v ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

static int64_t sort_value(const vq_agg_result* tmp, int32_t naggs, const vq_agg* aggs, int64_t i) {
  int64_t o = i * naggs;
  if (aggs[0].kind == VQ_AGG_COUNT) return tmp->ivals[o];
  if (aggs[0].kind == VQ_AGG_AVG || aggs[0].accum == VQ_F64) return f_bits(tmp->fvals[o]);
  return tmp->ivals[o];
}

/* Emit groups (insertion order) into out, optionally keeping the topk by
 * aggregate 0 descending. */
static int emit_join_groups(glist* l, int32_t naggs, const vq_agg* aggs, int32_t topk,
                            vq_agg_result* out) {
  int64_t n = l->n;
  vq_agg_result tmp;
  tmp.capacity = n ? n : 1;
  tmp.keys = (int64_t*)malloc(sizeof(int64_t) * tmp.capacity);
  tmp.counts = (int64_t*)malloc(sizeof(int64_t) * tmp.capacity);
  tmp.first_row = (int64_t*)malloc(sizeof(int64_t) * tmp.capacity);
  tmp.ivals = (int64_t*)malloc(sizeof(int64_t) * tmp.capacity * (naggs ? naggs : 1));
  tmp.fvals = (double*)malloc(sizeof(double) * tmp.capacity * (naggs ? naggs : 1));
  tmp.valid = (uint8_t*)malloc(tmp.capacity * (naggs ? naggs : 1));
  for (int64_t i = 0; i < n; i++) emit_group(&l->g[i], 1, naggs, aggs, &tmp, i);
  rank_item* ord = (rank_item*)malloc(sizeof(rank_item) * (n ? n : 1));
  for (int64_t i = 0; i < n; i++) {
    ord[i].pos = i;
    ord[i].v = (topk > 0 && naggs > 0) ? sort_value(&tmp, naggs, aggs, i) : 0;
  }
  if (topk > 0 && naggs > 0) qsort(ord, (size_t)n, sizeof(rank_item), rank_cmp);
  int64_t m = (topk > 0 && topk < n) ? topk : n;
  int rc = check_cap(out, m);
  if (rc == VQ_OK)
    for (int64_t i = 0; i < m; i++) {
      int64_t s = ord[i].pos;
      out->keys[i] = tmp.keys[s];
      out->counts[i] = tmp.counts[s];
      out->first_row[i] = tmp.first_row[s];
      for (int j = 0; j < naggs; j++) {
        out->ivals[i * naggs + j] = tmp.ivals[s * naggs + j];
        out->fvals[i * naggs + j] = tmp.fvals[s * naggs + j];
        out->valid[i * naggs + j] = tmp.valid[s * naggs + j];
      }
    }
  free(ord);
  free(tmp.keys);
  free(tmp.counts);
  free(tmp.first_row);
  free(tmp.ivals);
  free(tmp.fvals);
  free(tmp.valid);
  return rc;
}

/* Serial HashJoin (exec/join.go:84-186) -> HashAggregate(GROUP BY join key):
 * build rows inserted in row order (NULL keys would be dropped; columns here
 * carry none), probe rows visited in row order, each matching build row in
 * chain (insertion) order yielding one join row folded into its group. */
int or_join_agg(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols, int32_t b_key,
                const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t naggs,
                const vq_agg* aggs, int32_t topk, int32_t exact, vq_agg_result* out) {
  (void)b_ncols;
  (void)p_ncols;
  if (naggs < 0 || naggs > VQ_MAX_AGGS) return VQ_E_ARG;
  int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (b_nrows ? b_nrows : 1));
  jtable t;
  jt_init(&t, 8, next);
  for (int64_t r = 0; r < b_nrows; r++)
    if (pred_pass(b_cols, b_pred, r)) jt_insert(&t, col_i(&b_cols[b_key], r), (int32_t)r);
  kmap m;
  kmap_init(&m, 1024);
  glist l = {0, 0, 0};
  if (t.keys > 0)
    for (int64_t r = 0; r < p_nrows; r++) {
      if (!pred_pass(p_cols, p_pred, r)) continue;
      int64_t key = col_i(&p_cols[p_key], r);
      for (int32_t b = jt_lookup(&t, key, or_radix_hash(key)); b != -1; b = next[b]) {
        group* g = kmap_group(&m, &l, key, naggs, aggs, exact);
        if (g->first_row < 0) {
          g->first_row = r;
          g->key[0] = key;
        }
        fold_row(g, p_cols, naggs, aggs, r);
      }
    }
  int rc = emit_join_groups(&l, naggs, aggs, topk, out);
  for (int64_t i = 0; i < l.n; i++) group_free(&l.g[i], naggs);
  free(l.g);
  kmap_free(&m);
  free(t.slots);
  free(next);
  return rc;
}

static inline int64_t type_width(int32_t type) {
  return type == VQ_U8 ? 1 : type == VQ_U16 ? 2 : type == VQ_DATE ? 4 : 8;
}

/* A column reference of a joined-row expression in the joined column space
 * (probe columns, then build columns: col | VQ_BUILD_COL is column
 * p_ncols + col); -1 when out of range. */
static int32_t joined_col(int64_t col, int32_t p_ncols, int32_t b_ncols) {
  if (col & VQ_BUILD_COL) {
    const int64_t b = col & ~(int64_t)VQ_BUILD_COL;
    return b >= 0 && b < b_ncols ? (int32_t)(p_ncols + b) : -1;
  }
  return col >= 0 && col < p_ncols ? (int32_t)col : -1;
}

/* HashJoin (exec/join.go:84-186) feeding HashAggregate with its group keys
 * and aggregates over the joined schema, as ParallelHashJoinAggregate runs
 * them per probe morsel (exec/parallel_join.go:780-988): build rows passing
 * their predicate inserted in row order, probe rows visited in row order,
 * each matching build row yielding one joined row; a joined row's columns
 * are the probe row's and the build row's (the expression evaluated over a
 * one-row view of the two). */
int or_join_group(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols, int32_t b_key,
                  const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                  const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t nkeys,
                  const int32_t* key_cols, int32_t naggs, const vq_agg* aggs, int32_t exact,
                  vq_agg_result* out) {
  if (naggs < 0 || naggs > VQ_MAX_AGGS || nkeys < 1 || nkeys > 8) return VQ_E_ARG;
  if (nkeys > 1 && !keys_fit(p_nrows, p_cols, nkeys, key_cols)) return VQ_E_UNSUPPORTED;
  /* the aggregates over the joined column space */
  vq_agg ja[VQ_MAX_AGGS];
  for (int j = 0; j < naggs; j++) {
    ja[j] = aggs[j];
    vq_expr* e = &ja[j].arg;
    for (int k = 0; k < e->nfactors; k++) {
      if (e->f[k].kind == VQ_F_LIT) continue;
      if ((e->f[k].col = joined_col(e->f[k].col, p_ncols, b_ncols)) < 0) return VQ_E_ARG;
    }
    for (int k = 0; k < e->nwhen; k++) {
      if ((e->when[k].col = joined_col(e->when[k].col, p_ncols, b_ncols)) < 0) return VQ_E_ARG;
      if (e->when[k].op & VQ_COLCMP) {
        const int32_t c2 = joined_col(e->when[k].ival, p_ncols, b_ncols);
        if (c2 < 0) return VQ_E_ARG;
        e->when[k].ival = c2;
      }
    }
  }
  int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (b_nrows ? b_nrows : 1));
  jtable t;
  jt_init(&t, 8, next);
  for (int64_t r = 0; r < b_nrows; r++)
    if (pred_pass(b_cols, b_pred, r)) jt_insert(&t, col_i(&b_cols[b_key], r), (int32_t)r);
  const int32_t nj = p_ncols + b_ncols;
  vq_col_desc* view = (vq_col_desc*)malloc(sizeof(vq_col_desc) * (size_t)nj);
  for (int32_t c = 0; c < nj; c++) view[c] = c < p_ncols ? p_cols[c] : b_cols[c - p_ncols];
  kmap m;
  kmap_init(&m, 1024);
  glist l = {0, 0, 0};
  if (t.keys > 0)
    for (int64_t r = 0; r < p_nrows; r++) {
      if (!pred_pass(p_cols, p_pred, r)) continue;
      const int64_t key = col_i(&p_cols[p_key], r);
      const int64_t gk = nkeys > 1 ? pack_key(p_cols, nkeys, key_cols, r) : col_i(&p_cols[key_cols[0]], r);
      for (int32_t b = jt_lookup(&t, key, or_radix_hash(key)); b != -1; b = next[b]) {
        for (int32_t c = 0; c < p_ncols; c++)
          view[c].data = (const char*)p_cols[c].data + r * type_width(p_cols[c].type);
        for (int32_t c = 0; c < b_ncols; c++)
          view[p_ncols + c].data = (const char*)b_cols[c].data + (int64_t)b * type_width(b_cols[c].type);
        group* g = kmap_group(&m, &l, gk, naggs, ja, exact);
        if (g->first_row < 0) {
          g->first_row = r;
          if (nkeys > 1) unpack_key(gk, nkeys, g->key);
          else g->key[0] = gk;
        }
        fold_row(g, view, naggs, ja, 0);
      }
    }
  int rc = check_cap(out, l.n);
  if (rc == VQ_OK)
    for (int64_t i = 0; i < l.n; i++) emit_group(&l.g[i], nkeys, naggs, ja, out, i);
  for (int64_t i = 0; i < l.n; i++) group_free(&l.g[i], naggs);
  free(l.g);
  kmap_free(&m);
  free(view);
  free(t.slots);
  free(next);
  return rc;
}

/* The nested build of ParallelHashJoinAggregate (planner/parallel_join.go:
 * 341-400): the inner join's own build side hashed once (BuildSharedHashTable
 * over the inner rows), its probe side (the build table's rows, in order)
 * joined against it (HashJoinShared, exec/parallel_join.go:389-552), each
 * output row once per inner match; those rows are the outer join's build
 * side (HashJoin, exec/join.go:84-186, inserted in order), GROUP BY the outer
 * join key with the build columns that ride with it. */
int or_join_agg_nested(int64_t i_nrows, int32_t i_ncols, const vq_col_desc* i_cols, int32_t i_key,
                       const vq_pred* i_pred, int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols,
                       int32_t b_fk, int32_t b_key, const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                       const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred, int32_t npay,
                       const int32_t* pay_cols, int32_t naggs, const vq_agg* aggs, int32_t topk, int32_t exact,
                       vq_agg_result* out, int64_t* pay_out) {
  (void)i_ncols;
  (void)b_ncols;
  (void)p_ncols;
  if (naggs < 0 || naggs > VQ_MAX_AGGS || npay < 0 || npay > 7) return VQ_E_ARG;
  int32_t* inext = (int32_t*)malloc(sizeof(int32_t) * (i_nrows ? i_nrows : 1));
  jtable it;
  jt_init(&it, 8, inext);
  for (int64_t r = 0; r < i_nrows; r++)
    if (pred_pass(i_cols, i_pred, r)) jt_insert(&it, col_i(&i_cols[i_key], r), (int32_t)r);
  int64_t nn = 0, ncap = 1024;
  int32_t* nested = (int32_t*)malloc(sizeof(int32_t) * ncap);
  if (it.keys > 0)
    for (int64_t b = 0; b < b_nrows; b++) {
      if (!pred_pass(b_cols, b_pred, b)) continue;
      const int64_t fk = col_i(&b_cols[b_fk], b);
      for (int32_t x = jt_lookup(&it, fk, or_radix_hash(fk)); x != -1; x = inext[x]) {
        if (nn == ncap) nested = (int32_t*)realloc(nested, sizeof(int32_t) * (ncap *= 2));
        nested[nn++] = (int32_t)b;
      }
    }
  int32_t* onext = (int32_t*)malloc(sizeof(int32_t) * (nn ? nn : 1));
  jtable ot;
  jt_init(&ot, 8, onext);
  for (int64_t i = 0; i < nn; i++) jt_insert(&ot, col_i(&b_cols[b_key], nested[i]), (int32_t)i);
  kmap m;
  kmap_init(&m, 1024);
  glist l = {0, 0, 0};
  int rc = VQ_OK;
  if (ot.keys > 0)
    for (int64_t r = 0; r < p_nrows && rc == VQ_OK; r++) {
      if (!pred_pass(p_cols, p_pred, r)) continue;
      const int64_t key = col_i(&p_cols[p_key], r);
      for (int32_t x = jt_lookup(&ot, key, or_radix_hash(key)); x != -1; x = onext[x]) {
        group* g = kmap_group(&m, &l, key, naggs, aggs, exact);
        const int64_t b = nested[x];
        if (g->first_row < 0) {
          g->first_row = r;
          g->key[0] = key;
          for (int c = 0; c < npay; c++) g->key[1 + c] = col_i(&b_cols[pay_cols[c]], b);
        } else {
          for (int c = 0; c < npay; c++)
            if (g->key[1 + c] != col_i(&b_cols[pay_cols[c]], b)) rc = VQ_E_UNSUPPORTED;
        }
        fold_row(g, p_cols, naggs, aggs, r);
      }
    }
  if (rc == VQ_OK) rc = emit_join_groups(&l, naggs, aggs, topk, out);
  if (rc == VQ_OK)
    for (int64_t i = 0; i < out->ngroups; i++) {
      int found = 0;
      const int64_t s = kmap_find(&m, out->keys[i], &found);
      const group* g = &l.g[m.idx[s]];
      for (int c = 0; c < npay; c++) pay_out[i * npay + c] = g->key[1 + c];
    }
  for (int64_t i = 0; i < l.n; i++) group_free(&l.g[i], naggs);
  free(l.g);
  kmap_free(&m);
  free(ot.slots);
  free(it.slots);
  free(onext);
  free(inext);
  free(nested);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* multithreaded CPU baseline                                                */
/* ------------------------------------------------------------------------ */

/* morselQueue (exec/parallel.go:18-36): one atomic cursor. */
typedef struct morsel_queue {
  int64_t next;
  char pad[56];
  int64_t end;
} morsel_queue;

static int morsel_claim(morsel_queue* q, int64_t* start) {
  int64_t s = __atomic_fetch_add(&q->next, 1, __ATOMIC_RELAXED);
  if (s >= q->end) return 0;
  *start = s;
  return 1;
}

typedef struct agg_job {
  const vq_col_desc* cols;
  const int64_t* rg_base;
  const int64_t* rg_rows;
  const uint8_t* keep;
  const vq_pred* pred;
  int32_t nkeys;
  int32_t hash_col; /* >= 0: GROUP BY this one integer column (general key path) */
  const int32_t* key_cols;
  int32_t naggs;
  const vq_agg* aggs;
  morsel_queue* q;
  kmap m;
  glist l;
} agg_job;

/* One worker: claim row-group morsels, run Scan -> Filter -> partial
 * HashAggregate (exec/parallel.go:143-183). */
static void* agg_worker(void* arg) {
  agg_job* j = (agg_job*)arg;
  kmap_init(&j->m, 64);
  int64_t rg;
  while (morsel_claim(j->q, &rg)) {
    if (!j->keep[rg]) continue;
    for (int64_t r = j->rg_base[rg]; r < j->rg_base[rg] + j->rg_rows[rg]; r++) {
      if (!pred_pass(j->cols, j->pred, r)) continue;
      int64_t pk = j->hash_col >= 0 ? col_i(&j->cols[j->hash_col], r)
                   : j->nkeys    ? pack_key(j->cols, j->nkeys, j->key_cols, r)
                                 : 0;
      group* g = kmap_group(&j->m, &j->l, pk, j->naggs, j->aggs, 0);
      if (g->first_row < 0) {
        g->first_row = r;
        if (j->hash_col >= 0) g->key[0] = pk;
        else unpack_key(pk, j->nkeys, g->key);
      }
      fold_row(g, j->cols, j->naggs, j->aggs, r);
    }
  }
  return NULL;
}

/* ParallelHashAggregate (exec/parallel.go:124-314) over the dictionary-code
 * keys (hash_col < 0) or one integer key column (hash_col >= 0). */
static int agg_parallel(int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                        const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                        int32_t nkeys, const int32_t* key_cols, int32_t hash_col, int32_t naggs,
                        const vq_agg* aggs, int32_t nthreads, vq_agg_result* out) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nrg && nrg > 0) nthreads = nrg;
  uint8_t* keep = (uint8_t*)malloc((size_t)nrg + 1);
  or_zone_prune(ncols, cols, nrg, zones, pred, keep);
  int64_t* rg_base = (int64_t*)malloc(sizeof(int64_t) * (nrg + 1));
  rg_base[0] = 0;
  for (int32_t g = 0; g < nrg; g++) rg_base[g + 1] = rg_base[g] + rg_rows[g];
  morsel_queue q;
  memset(&q, 0, sizeof q);
  q.end = nrg;
  agg_job* jobs = (agg_job*)calloc((size_t)nthreads, sizeof(agg_job));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  for (int i = 0; i < nthreads; i++) {
    agg_job* j = &jobs[i];
    j->cols = cols;
    j->rg_base = rg_base;
    j->rg_rows = rg_rows;
    j->keep = keep;
    j->pred = pred;
    j->nkeys = nkeys;
    j->hash_col = hash_col;
    j->key_cols = key_cols;
    j->naggs = naggs;
    j->aggs = aggs;
    j->q = &q;
    pthread_create(&th[i], NULL, agg_worker, j);
  }
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  /* merge (exec/parallel.go:186-209) */
  kmap m;
  kmap_init(&m, 64);
  glist l = {0, 0, 0};
  for (int i = 0; i < nthreads; i++) {
    agg_job* j = &jobs[i];
    for (int64_t k = 0; k < j->l.n; k++) {
      group* src = &j->l.g[k];
      int64_t pk = 0;
      const int shift = key_width(nkeys);
      if (hash_col >= 0) pk = src->key[0];
      else
        for (int c = 0; c < nkeys; c++) pk |= (int64_t)((uint64_t)src->key[c] << (shift * c));
      int f;
      int64_t s = kmap_find(&m, pk, &f);
      if (f) merge_group(&l.g[m.idx[s]], src, naggs, aggs);
      else {
        group* g = kmap_group(&m, &l, pk, naggs, aggs, 0);
        *g = *src;
      }
    }
    kmap_free(&j->m);
    free(j->l.g);
  }
  /* first-occurrence order, as the serial plan emits */
  rank_item* ord = (rank_item*)malloc(sizeof(rank_item) * (l.n ? l.n : 1));
  for (int64_t i = 0; i < l.n; i++) {
    ord[i].v = -l.g[i].first_row;
    ord[i].pos = i;
  }
  qsort(ord, (size_t)l.n, sizeof(rank_item), rank_cmp);
  int rc;
  if (l.n == 0 && nkeys == 0) {
    rc = check_cap(out, 1);
    if (rc == VQ_OK) emit_empty_global(naggs, aggs, out);
  } else {
    rc = check_cap(out, l.n);
    if (rc == VQ_OK) {
      for (int64_t i = 0; i < l.n; i++) emit_group(&l.g[ord[i].pos], nkeys, naggs, aggs, out, i);
      if (hash_col < 0) null_keys(cols, nkeys, key_cols, out);
    }
  }
  free(ord);
  free(l.g);
  kmap_free(&m);
  free(jobs);
  free(th);
  free(rg_base);
  free(keep);
  return rc;
}

int or_agg_dict_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                         const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                         int32_t nkeys, const int32_t* key_cols, int32_t naggs,
                         const vq_agg* aggs, int32_t nthreads, vq_agg_result* out) {
  (void)nrows;
  if (nkeys < 0 || nkeys > 8 || naggs < 0 || naggs > VQ_MAX_AGGS) return VQ_E_ARG;
  if (!keys_fit(nrows, cols, nkeys, key_cols)) return VQ_E_UNSUPPORTED;
  return agg_parallel(ncols, cols, nrg, rg_rows, zones, pred, nkeys, key_cols, -1, naggs, aggs,
                      nthreads, out);
}

int or_agg_hash_parallel(int64_t nrows, int32_t ncols, const vq_col_desc* cols, int32_t nrg,
                         const int64_t* rg_rows, const vq_zone* zones, const vq_pred* pred,
                         int32_t key_col, int32_t naggs, const vq_agg* aggs, int32_t nthreads,
                         vq_agg_result* out) {
  (void)nrows;
  if (key_col < 0 || key_col >= ncols || naggs < 0 || naggs > VQ_MAX_AGGS) return VQ_E_ARG;
  return agg_parallel(ncols, cols, nrg, rg_rows, zones, pred, 1, NULL, key_col, naggs, aggs,
                      nthreads, out);
}

/* ---- parallel join: two-pass radix build, morsel-parallel probe --------- */

typedef struct keyed_row {
  int64_t key;
  int32_t row;
} keyed_row;

typedef struct join_shared {
  /* inputs */
  int64_t b_nrows, p_nrows;
  const vq_col_desc* b_cols;
  const vq_col_desc* p_cols;
  int32_t b_key, p_key;
  const vq_pred* b_pred;
  const vq_pred* p_pred;
  int32_t naggs;
  const vq_agg* aggs;
  /* build state */
  int32_t bits;
  int64_t nparts;
  int64_t nmorsels;
  keyed_row** bucket;   /* [morsel * nparts + p] */
  int64_t* bucket_n;    /* [morsel * nparts + p] */
  jtable* parts;
  int32_t* next;
  morsel_queue q1, q2, q3;
} join_shared;

typedef struct join_worker {
  join_shared* s;
  kmap m;
  glist l;
} join_worker;

/* Pass 1 (partitionMorsel, exec/parallel_join.go:322-347,422-476): each
 * claimed build morsel drains its filter and buckets (key,row) per
 * partition. */
static void* build_pass1(void* arg) {
  join_shared* s = ((join_worker*)arg)->s;
  int64_t mo;
  while (morsel_claim(&s->q1, &mo)) {
    int64_t lo = mo * 65536, hi = lo + 65536;
    if (hi > s->b_nrows) hi = s->b_nrows;
    int64_t* cnt = &s->bucket_n[mo * s->nparts];
    keyed_row** bk = &s->bucket[mo * s->nparts];
    int64_t* capv = (int64_t*)calloc((size_t)s->nparts, sizeof(int64_t));
    for (int64_t r = lo; r < hi; r++) {
      if (!pred_pass(s->b_cols, s->b_pred, r)) continue;
      int64_t key = col_i(&s->b_cols[s->b_key], r);
      int64_t p = s->nparts > 1 ? (int64_t)(or_radix_hash(key) & (uint64_t)(s->nparts - 1)) : 0;
      if (cnt[p] == capv[p]) {
        capv[p] = capv[p] ? capv[p] * 2 : 64;
        bk[p] = (keyed_row*)realloc(bk[p], sizeof(keyed_row) * capv[p]);
      }
      bk[p][cnt[p]].key = key;
      bk[p][cnt[p]].row = (int32_t)r;
      cnt[p]++;
    }
    free(capv);
  }
  return NULL;
}

/* Pass 2 (parallel_join.go:478-550): one owner per partition walks morsels in
 * index order, presizing from exact counts, so per-key order equals a serial
 * drain. */
static void* build_pass2(void* arg) {
  join_shared* s = ((join_worker*)arg)->s;
  int64_t p;
  while (morsel_claim(&s->q2, &p)) {
    int64_t total = 0;
    for (int64_t mo = 0; mo < s->nmorsels; mo++) total += s->bucket_n[mo * s->nparts + p];
    jt_init(&s->parts[p], total, s->next);
    for (int64_t mo = 0; mo < s->nmorsels; mo++) {
      keyed_row* b = s->bucket[mo * s->nparts + p];
      for (int64_t i = 0; i < s->bucket_n[mo * s->nparts + p]; i++)
        jt_insert(&s->parts[p], b[i].key, b[i].row);
    }
  }
  return NULL;
}

/* Probe workers (runMorselAggWorkers, parallel_join.go:911-988). */
static void* probe_worker(void* arg) {
  join_worker* w = (join_worker*)arg;
  join_shared* s = w->s;
  kmap_init(&w->m, 1024);
  int64_t mo;
  while (morsel_claim(&s->q3, &mo)) {
    int64_t lo = mo * 65536, hi = lo + 65536;
    if (hi > s->p_nrows) hi = s->p_nrows;
    for (int64_t r = lo; r < hi; r++) {
      if (!pred_pass(s->p_cols, s->p_pred, r)) continue;
      int64_t key = col_i(&s->p_cols[s->p_key], r);
      uint64_t h = or_radix_hash(key);
      const jtable* t = &s->parts[s->nparts > 1 ? (int64_t)(h & (uint64_t)(s->nparts - 1)) : 0];
      for (int32_t b = jt_lookup(t, key, h); b != -1; b = s->next[b]) {
        group* g = kmap_group(&w->m, &w->l, key, s->naggs, s->aggs, 0);
        if (g->first_row < 0) {
          g->first_row = r;
          g->key[0] = key;
        }
        fold_row(g, s->p_cols, s->naggs, s->aggs, r);
      }
    }
  }
  return NULL;
}

static void run_workers(void* (*fn)(void*), join_worker* w, int n) {
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * n);
  for (int i = 0; i < n; i++) pthread_create(&th[i], NULL, fn, &w[i]);
  for (int i = 0; i < n; i++) pthread_join(th[i], NULL);
  free(th);
}

int or_join_agg_parallel(int64_t b_nrows, int32_t b_ncols, const vq_col_desc* b_cols,
                         int32_t b_key, const vq_pred* b_pred, int64_t p_nrows, int32_t p_ncols,
                         const vq_col_desc* p_cols, int32_t p_key, const vq_pred* p_pred,
                         int32_t naggs, const vq_agg* aggs, int32_t topk, int32_t nthreads,
                         vq_agg_result* out) {
  (void)b_ncols;
  (void)p_ncols;
  if (nthreads < 1) nthreads = 1;
  join_shared s;
  memset(&s, 0, sizeof s);
  s.b_nrows = b_nrows;
  s.p_nrows = p_nrows;
  s.b_cols = b_cols;
  s.p_cols = p_cols;
  s.b_key = b_key;
  s.p_key = p_key;
  s.b_pred = b_pred;
  s.p_pred = p_pred;
  s.naggs = naggs;
  s.aggs = aggs;
  /* buildSideRadixOptions: the estimate is the build scan's row count
   * (planner/parallel_join.go:327-339). */
  s.bits = or_radix_bits_for(b_nrows);
  s.nparts = (int64_t)1 << s.bits;
  s.nmorsels = (b_nrows + 65535) / 65536;
  s.bucket = (keyed_row**)calloc((size_t)(s.nmorsels * s.nparts + 1), sizeof(keyed_row*));
  s.bucket_n = (int64_t*)calloc((size_t)(s.nmorsels * s.nparts + 1), sizeof(int64_t));
  s.parts = (jtable*)calloc((size_t)s.nparts, sizeof(jtable));
  s.next = (int32_t*)malloc(sizeof(int32_t) * (b_nrows ? b_nrows : 1));
  s.q1.end = s.nmorsels;
  s.q2.end = s.nparts;
  s.q3.end = (p_nrows + 65535) / 65536;
  join_worker* w = (join_worker*)calloc((size_t)nthreads, sizeof(join_worker));
  for (int i = 0; i < nthreads; i++) w[i].s = &s;
  run_workers(build_pass1, w, nthreads);
  run_workers(build_pass2, w, nthreads);
  int64_t nkeys = 0;
  for (int64_t p = 0; p < s.nparts; p++) nkeys += s.parts[p].keys;
  if (nkeys > 0) run_workers(probe_worker, w, nthreads);
  /* serial merge in worker order (parallel_join.go:821-825) */
  kmap m;
  kmap_init(&m, 1024);
  glist l = {0, 0, 0};
  for (int i = 0; i < nthreads; i++) {
    for (int64_t k = 0; k < w[i].l.n; k++) {
      group* src = &w[i].l.g[k];
      int f;
      int64_t sl = kmap_find(&m, src->key[0], &f);
      if (f) merge_group(&l.g[m.idx[sl]], src, naggs, aggs);
      else {
        group* g = kmap_group(&m, &l, src->key[0], naggs, aggs, 0);
        *g = *src;
      }
    }
    if (w[i].m.keys) kmap_free(&w[i].m);
    free(w[i].l.g);
  }
  /* first-occurrence order before the stable top-k */
  rank_item* ord = (rank_item*)malloc(sizeof(rank_item) * (l.n ? l.n : 1));
  for (int64_t i = 0; i < l.n; i++) {
    ord[i].v = -l.g[i].first_row;
    ord[i].pos = i;
  }
  qsort(ord, (size_t)l.n, sizeof(rank_item), rank_cmp);
  glist sorted = {0, 0, 0};
  sorted.g = (group*)malloc(sizeof(group) * (l.n ? l.n : 1));
  sorted.n = sorted.cap = l.n;
  for (int64_t i = 0; i < l.n; i++) sorted.g[i] = l.g[ord[i].pos];
  int rc = emit_join_groups(&sorted, naggs, aggs, topk, out);
  free(sorted.g);
  free(ord);
  free(l.g);
  kmap_free(&m);
  for (int64_t i = 0; i < s.nmorsels * s.nparts; i++) free(s.bucket[i]);
  free(s.bucket);
  free(s.bucket_n);
  for (int64_t p = 0; p < s.nparts; p++) free(s.parts[p].slots);
  free(s.parts);
  free(s.next);
  free(w);
  return rc;
}

/* The reference filter benchmark's data (bench/simd_filter/cpp/main.cpp:28-45):
 * xorshift64* from state `seed`, value (state * 0x2545F4914F6CDD1D) % 50 + 1. */
void or_xorshift_i64(int64_t n, uint64_t seed, int64_t* out) {
  uint64_t s = seed;
  for (int64_t i = 0; i < n; i++) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    out[i] = (int64_t)((s * 0x2545F4914F6CDD1DULL) % 50 + 1);
  }
}
