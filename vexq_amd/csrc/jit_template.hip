// jit_template.hip — the row program of one query shape, compiled at run
// time with hipRTC for gfx950 and cached by shape (vq_jit.cpp).
//
// A generated prologue fixes the shape as compile-time constants: the type of
// every referenced column, each predicate term's column and comparison, the
// group-key strides, and every aggregate's left fold of factors.  Literal
// values stay kernel arguments, so queries that differ only in constants share
// one compiled kernel.  With the shape known, every loop below unrolls and
// every switch folds: a row costs its loads, its compares and its arithmetic,
// with no descriptor interpretation on the scalar unit.
//
// The kernels follow the same contracts as the static ones in k_scan.hip:
//   vq_dense       per-wave partial aggregates (persistent grid, prefetch)
//   vq_filter_bits pass bits + per-tile counts for the selection vector
//   vq_part_count / vq_part_scatter   radix partitioning of a join side
//   vq_hash_local  GROUP BY of few keys into per-workgroup partials
//   vq_dj_probe / vq_dj_match  the direct join's probe and its matches
//   vq_dense_multi several one-group aggregates of one shape, one pass
#ifndef VQ_JIT_PROLOGUE
// Stand-alone build check (hipcc -c): the 8-aggregate GROUP BY shape.
#define VQ_JIT_PROLOGUE 1
#define VQ_NS 7
#define VQ_U 4
#define VQ_G 6
#define VQ_NA 5
#define VQ_NT 1
#define VQ_KEYSLOT 2
#define VQ_NV 5
#define VQ_PROBE 1
#define VQ_KERNELS 63
#define VQ_PB 8
#define VQ_BLOOM 1
#define VQ_LIVEBITS 1
#define VQ_DJ_SLOTS 6
#define VQ_NQ 4
#define VQ_NULLS 1
#define VQ_NKEY 1
#define VQ_LOOKUP 0
#define VQ_JKEY 0
static constexpr int kSB[8] = {0, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kSN[8] = {1, 0, 0, 1, 0, 0, 1, 0};
static constexpr int kNC[8] = {3, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kACN[8] = {0, 0, 0, 0, 1, 0, 0, 0};
static constexpr int kKSL[8] = {2, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kKSH[8] = {0, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kST[8] = {1, 1, 4, 5, 5, 5, 3, 0};
#define VQ_TERMS 24
static constexpr int kTS[24] = {6, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kTO[24] = {3, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kTS2[24] = {0, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kAW0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kAWN[8] = {0, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kKS[8] = {2, 1, 0, 0, 0, 0, 0, 0};
static constexpr int kAOP[8] = {1, 0, 0, 0, 0, 0, 0, 0};
static constexpr int kAISF[8] = {0, 1, 1, 1, 1, 0, 0, 0};
static constexpr int kANF[8] = {1, 1, 2, 3, 1, 0, 0, 0};
static constexpr int kFK[8][4] = {{0}, {0}, {0, 2}, {0, 2, 2}, {0}};
static constexpr int kFS[8][4] = {{2}, {3}, {3, 4}, {3, 4, 5}, {4}};
static constexpr int kFO[8][4] = {{0}, {0}, {0, 1}, {0, 1, 0}, {0}};
static constexpr int kFC[8][4] = {{0}, {0}, {0, 2}, {0, 2, 2}, {0}};
#define VQ_NAG 5
static constexpr int kLOP[8] = {0, 1, 0, 2, 5, 0, 0, 0};
static constexpr int kLSRC[8] = {0, 1, 2, 2, 1, 0, 0, 0};
#include <hip/hip_runtime.h>
#endif
// exact fixed-point chunks of a float64 (the host prepends this header's
// text to the run-time source; the stand-alone check includes it)
#ifndef VQ_FIXCHUNKS_H
#include "fixchunks.h"
#endif

// No headers: hipRTC compiles this with its built-in runtime declarations only.
typedef long long i64;
typedef unsigned long long u64;
typedef int i32;
typedef unsigned int u32;
typedef unsigned short u16;
typedef unsigned char u8;
constexpr i64 kI64Min = -0x7fffffffffffffffLL - 1;
constexpr i64 kI64Max = 0x7fffffffffffffffLL;
// The bits of Go's math.NaN() (uvnan), which applyArithFloat64 returns for
// x/0 (exec/expr.go:1016-1018).
constexpr i64 kGoNaNBits = 0x7FF8000000000001LL;

#define VQ_U8 1
#define VQ_U16 2
#define VQ_DATE 3
#define VQ_I64 4
#define VQ_F64 5

namespace vqj {

constexpr int kWave = 64;
constexpr int kMaxSpans = 24;

// compile-time index sequences (no <utility> under hipRTC)
template <int... I>
struct Seq {};
template <int N, int... I>
struct MakeSeq : MakeSeq<N - 1, N - 1, I...> {};
template <int... I>
struct MakeSeq<0, I...> {
  using type = Seq<I...>;
};
template <int N>
using seq = typename MakeSeq<N>::type;

struct Args {
  const void* col[8];
  i64 ti[VQ_TERMS];   // the predicate's terms [0, VQ_NT), then CASE WHEN terms
  double tf[VQ_TERMS];
  i64 fi[8][4];
  double ff[8][4];
  i64 ei[8];          // ELSE literals of aggregates with a CASE WHEN
  double ef[8];
  i64 sp_start[kMaxSpans];
  i64 sp_rows[kMaxSpans];
  i64 sp_tile0[kMaxSpans];
  i32 nspans;
  i32 tile_rows;
  i64 ntiles;
  u64* part;          // vq_dense: partials [G][NA+2][grid] (item-major)
  u64* bits;          // vq_filter_bits: pass words
  u32* tile_counts;   // vq_filter_bits: per-tile counts
  // join partitioning
  u32* counts;        // [G][P] rows per workgroup and partition, then positions
  const u32* pstart;  // [P + 1] first entry of each partition
  u32* perr;          // set to 1 if a write would pass its partition's end
  i64* keys;
  u32* rows;
  u64* vals;          // [NV][vstride]
  i64 vstride;
  i32 P;
  u32 lds_words;      // dynamic LDS the launch provides, in 4-byte words
  // GROUP BY through per-workgroup partials (vq_hash_local)
  const i64* lkeys;   // [lcap] the distinct keys (kI64Min = empty slot)
  u32* lflags;        // the pipeline's flags: [0] overflow bits, [2] keys,
                      // [10] key INT64_MIN present
  i64* lgkey;         // [groups] the key of each group index
  i32 lcap;           // table slots (a power of two)
  i32 lnw;            // accumulator words per group
  i32 vop[8];         // per aggregate with an argument (kLOP, kLSRC in the
                      // prologue: its fold and its argument's value index):
  i32 vword[8];       // its first accumulator word
  i32 vlsb[8];        // float sums: the fixed-point scale
  i32 vw[8];
  i32 vsrc[8];
  i32 nvop;
  // semi-join filter of the build side's keys (VQ_BLOOM 1: the probe side
  // drops rows whose two bits are not both set): a blocked Bloom filter,
  // 2^bloom_lpp 64-byte lines per partition
  const u32* bloom;
  u32 bloom_lpp;      // log2 of the lines per partition
  u32 bloom_pmask;    // the join's partitions - 1 (the filter's blocks; the
                      // kernels here may write fewer, coarser partitions)
  // VQ_LIVEBITS: the count pass writes one bit per row (predicate and filter
  // passed), word row / 64; the scatter reads them instead of testing again
  u64* livebits;
  int lk0;
  int lpad_;
  // VQ_LOOKUP (vq_join_group): the build side's lookup table -- open
  // addressing on radix_hash(key) >> 32, key INT64_MIN marks an empty slot;
  // head[slot] starts the slot's chain of build rows, next[row] continues it
  // (0xFFFFFFFF ends it); lk_min[0] starts the chain of key INT64_MIN
  const i64* lk_keys;
  const u32* lk_head;
  const u32* lk_next;
  const u32* lk_min;
  u32 lk_mask;
  u32 lk_packed;      // slots are words key << 32 | head (all ones: empty)
  i64 lk_base;        // lk_range > 0: a direct table, head[key - lk_base]
  u32 lk_range;
  u32 lk_pad;
  // the direct join's probe (vq_dj_probe, vq_dj_match): build keys in
  // [dj_base, dj_base + dj_range); bit o of dj_pres: a passing build row has
  // key dj_base + o; dj_rec[o][8]: its matches so far, 1 + its latest match
  // past kDjSlots, the entries of its first kDjSlots matches
  const u32* dj_pres;
  u32* dj_rec;
  u32* dj_mrow;       // the matched probe rows of each workgroup, at its region
  u32* dj_mcount;     // [grid] their count
  u64* dj_ent;        // [e][dj_ew]: (1 + previous further match) << 32 | row, values
  u32* dj_tlist;      // the keys each workgroup met first, at its region
  u32* dj_tcount;     // [grid] their count
  i64 dj_base;
  u64 dj_range;
  i32 dj_ew;
  i32 dj_pad;
  // vq_dense_multi: the predicate literals of each of the VQ_NQ queries,
  // [q][VQ_NT] (floats as their bits)
  const i64* mlit;
  // validity words of each slot's column (kSN: the slot can be NULL): bit r
  // of word r / 64 set when row r holds a value
  const u64* nul[8];
};

// ---- column reads: the element type is a compile-time constant -------------
template <int T>
__device__ __forceinline__ i64 ld(const void* p, i64 r) {
  if constexpr (T == VQ_U8) return ((const u8*)p)[r];
  else if constexpr (T == VQ_U16) return ((const u16*)p)[r];
  else if constexpr (T == VQ_DATE) return ((const i32*)p)[r];
  else return ((const i64*)p)[r];
}

__device__ __forceinline__ double as_f64(i64 b) { return __longlong_as_double(b); }
__device__ __forceinline__ i64 f64_bits(double d) { return __double_as_longlong(d); }

template <int OP, class V>
__device__ __forceinline__ bool cmp(V a, V b) {
  if constexpr (OP == 0) return a == b;
  else if constexpr (OP == 1) return a != b;
  else if constexpr (OP == 2) return a < b;
  else if constexpr (OP == 3) return a <= b;
  else if constexpr (OP == 4) return a > b;
  else return a >= b;
}

// applyArithFloat64 / applyArithInt64 (exec/expr.go:990-1020).
template <int OP>
__device__ __forceinline__ double arf(double a, double b) {
  if constexpr (OP == 0) return a + b;
  else if constexpr (OP == 1) return a - b;
  else if constexpr (OP == 2) return a * b;
  else return b == 0.0 ? __longlong_as_double(kGoNaNBits) : a / b;
}

template <int OP>
__device__ __forceinline__ i64 ari(i64 a, i64 b) {
  if constexpr (OP == 0) return (i64)((u64)a + (u64)b);
  else if constexpr (OP == 1) return (i64)((u64)a - (u64)b);
  else if constexpr (OP == 2) return (i64)((u64)a * (u64)b);
  else return b == 0 ? 0 : ((a == kI64Min && b == -1) ? kI64Min : a / b);
}

template <int S>
__device__ __forceinline__ double colf(i64 raw) {
  if constexpr (kST[S] == VQ_F64) return as_f64(raw);
  else return (double)raw;
}

// ---- the row program -----------------------------------------------------------
template <int U, int... C>
__device__ __forceinline__ void load_one(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u,
                                         Seq<C...>) {
  ((raw[C][u] = ld<kST[C]>(a.col[C], r)), ...);
}

template <int U>
__device__ __forceinline__ void load_rows(const Args& a, i64 base, i64 rend, int lane,
                                          i64 (&raw)[VQ_NS][U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    i64 r = base + u * kWave + lane;
    r = r < rend ? r : rend - 1;
    load_one(a, r, raw, u, seq<VQ_NS>{});
  }
}

// Row-blocked loads for the dense GROUP BY: lane l takes the U consecutive
// rows base + l*U .. base + l*U + U-1, read with 4- to 16-byte vector loads
// (a wave still reads one contiguous span of each column).  base is a
// multiple of U (tile_begin<U>), so every lane's r0 is too and the loads are
// aligned.  r0 is clamped to the last U-aligned start before rend: a lane
// whose rows all lie at or past rend re-reads rows inside the column (plus at
// most U-1 rows of the upload's tail padding) instead; the caller masks rows
// outside [lo, rend), so a clamped lane contributes nothing.
template <int T, int U>
__device__ __forceinline__ void ld_block(const void* p, i64 r0, i64* out) {
  if constexpr (T == VQ_U8) {
    static_assert(U == 4 || U == 8, "U");
    if constexpr (U == 4) {
      const u32 w = *(const u32*)((const u8*)p + r0);
#pragma unroll
      for (int i = 0; i < 4; i++) out[i] = (w >> (8 * i)) & 0xff;
    } else {
      const u64 w = *(const u64*)((const u8*)p + r0);
#pragma unroll
      for (int i = 0; i < 8; i++) out[i] = (i64)((w >> (8 * i)) & 0xff);
    }
  } else if constexpr (T == VQ_U16) {
#pragma unroll
    for (int h = 0; h < U / 4; h++) {
      const u64 w = *(const u64*)((const u16*)p + r0 + 4 * h);
#pragma unroll
      for (int i = 0; i < 4; i++) out[4 * h + i] = (i64)((w >> (16 * i)) & 0xffff);
    }
  } else if constexpr (T == VQ_DATE) {
#pragma unroll
    for (int h = 0; h < U / 4; h++) {
      const int4 w = *(const int4*)((const i32*)p + r0 + 4 * h);
      out[4 * h + 0] = w.x;
      out[4 * h + 1] = w.y;
      out[4 * h + 2] = w.z;
      out[4 * h + 3] = w.w;
    }
  } else {
#pragma unroll
    for (int h = 0; h < U / 2; h++) {
      const longlong2 w = *(const longlong2*)((const i64*)p + r0 + 2 * h);
      out[2 * h + 0] = w.x;
      out[2 * h + 1] = w.y;
    }
  }
}

template <int U, int... C>
__device__ __forceinline__ void load_block_one(const Args& a, i64 r0, i64 (&raw)[VQ_NS][U],
                                               Seq<C...>) {
  (ld_block<kST[C], U>(a.col[C], r0, raw[C]), ...);
}

template <int U>
__device__ __forceinline__ void load_rows_blocked(const Args& a, i64 base, i64 rend, int lane,
                                                  i64 (&raw)[VQ_NS][U]) {
  i64 r0 = base + (i64)lane * U;
  const i64 last = ((rend - 1) / U) * U;
  r0 = r0 < last ? r0 : last;
  load_block_one(a, r0, raw, seq<VQ_NS>{});
}

// One predicate term: column OP literal; column IN set (a mask over
// [0, 64)); or column OP column (op | 16).
template <int T>
__device__ __forceinline__ bool term(const Args& a, const i64* rawrow) {
  constexpr int S = kTS[T], OP = kTO[T];
  if constexpr (OP == 6) {
    const i64 v = rawrow[S];
    return (u64)v < 64 && (((u64)a.ti[T] >> v) & 1ull);
  } else if constexpr (OP >= 16) {
    constexpr int S2 = kTS2[T];
    if constexpr (kST[S] == VQ_F64) return cmp<OP - 16>(as_f64(rawrow[S]), as_f64(rawrow[S2]));
    else return cmp<OP - 16>(rawrow[S], rawrow[S2]);
  } else if constexpr (kST[S] == VQ_F64) {
    return cmp<OP>(as_f64(rawrow[S]), a.tf[T]);
  } else {
    return cmp<OP>(rawrow[S], a.ti[T]);
  }
}

template <int... T>
__device__ __forceinline__ bool pass_all(const Args& a, const i64* rawrow,
                                         Seq<T...>) {
  return (true && ... && term<T>(a, rawrow));
}

template <int... C>
__device__ __forceinline__ int group_of(const i64* rawrow, Seq<C...>) {
  return (0 + ... + (kKS[C] * (int)rawrow[C]));
}

// ---- NULLs (exec/batch.go:113-121: LSB-first validity, 1 = a value) ---------
// vm: bit C set when slot C holds a value in the row (always, for a slot
// whose column has no NULLs).  A predicate term on a NULL is not TRUE, so
// the row fails the conjunction; an aggregate argument reading a NULL is NULL
// and skipped (its count-valid accumulator, kACN, counts the others); a NULL
// group key takes the code kNC past the dictionary.
__host__ __device__ constexpr u32 null_slots() {
  u32 m = 0;
  for (int c = 0; c < VQ_NS; c++)
    if (kSN[c]) m |= 1u << c;
  return m;
}
__host__ __device__ constexpr u32 pred_mask() {
  u32 m = 0;
  for (int t = 0; t < VQ_NT; t++) {
    if (kSN[kTS[t]]) m |= 1u << kTS[t];
    if (kTO[t] >= 16 && kSN[kTS2[t]]) m |= 1u << kTS2[t];
  }
  return m;
}
template <int J>
__host__ __device__ constexpr u32 acc_mask() {
  u32 m = 0;
  for (int k = 0; k < kANF[J]; k++)
    if (kFK[J][k] != 1 && kSN[kFS[J][k]]) m |= 1u << kFS[J][k];
  return m;
}

// Slot validity of row r (one word load per nullable slot).
template <int... C>
__device__ __forceinline__ u32 row_valid(const Args& a, i64 r, Seq<C...>) {
  return (~null_slots() | ... | (kSN[C] ? (u32)((a.nul[C][r >> 6] >> (r & 63)) & 1ull) << C : 0u));
}

template <int... C>
__device__ __forceinline__ int group_of_v(const i64* rawrow, u32 vm, Seq<C...>) {
  return (0 + ... + (kKS[C] * (kSN[C] && !((vm >> C) & 1u) ? kNC[C] : (int)rawrow[C])));
}

// One factor of aggregate J (float fold).
template <int J, int K>
__device__ __forceinline__ double factor_f(const Args& a, const i64* rawrow) {
  constexpr int kind = kFK[J][K];
  if constexpr (kind == 1) {
    return a.ff[J][K];
  } else {
    const double c = colf<kFS[J][K]>(rawrow[kFS[J][K]]);
    if constexpr (kind == 0) return c;
    else if constexpr (kind == 2) return arf<kFO[J][K]>(a.ff[J][K], c);
    else return arf<kFO[J][K]>(c, a.ff[J][K]);
  }
}

template <int J, int K>
__device__ __forceinline__ i64 factor_i(const Args& a, const i64* rawrow) {
  constexpr int kind = kFK[J][K];
  if constexpr (kind == 1) {
    return a.fi[J][K];
  } else {
    const i64 c = rawrow[kFS[J][K]];
    if constexpr (kind == 0) return c;
    else if constexpr (kind == 2) return ari<kFO[J][K]>(a.fi[J][K], c);
    else return ari<kFO[J][K]>(c, a.fi[J][K]);
  }
}

// The left fold ((f0 c1 f1) c2 f2) c3 f3, one rounding per operation.
template <int J, int K>
__device__ __forceinline__ double fold_f(const Args& a, const i64* rawrow, double v) {
  if constexpr (K >= kANF[J]) {
    return v;
  } else {
    const double x = factor_f<J, K>(a, rawrow);
    return fold_f<J, K + 1>(a, rawrow, K == 0 ? x : arf<kFC[J][K]>(v, x));
  }
}

template <int J, int K>
__device__ __forceinline__ i64 fold_i(const Args& a, const i64* rawrow, i64 v) {
  if constexpr (K >= kANF[J]) {
    return v;
  } else {
    const i64 x = factor_i<J, K>(a, rawrow);
    return fold_i<J, K + 1>(a, rawrow, K == 0 ? x : ari<kFC[J][K]>(v, x));
  }
}

// CASE WHEN of aggregate J (CaseExpr with one WHEN and a literal ELSE,
// exec/expr.go:1386-1435): its terms are kAW0[J] .. kAW0[J] + kAWN[J] - 1 of
// the term arrays; a row failing them takes the ELSE literal.
template <int J, int... T>
__device__ __forceinline__ bool when_all(const Args& a, const i64* rawrow, Seq<T...>) {
  return (true && ... && term<kAW0[J] + T>(a, rawrow));
}

template <int J>
__device__ __forceinline__ double val_f(const Args& a, const i64* rawrow) {
  const double v = fold_f<J, 0>(a, rawrow, 0.0);
  if constexpr (kAWN[J] == 0) return v;
  else return when_all<J>(a, rawrow, seq<kAWN[J]>{}) ? v : a.ef[J];
}

template <int J>
__device__ __forceinline__ i64 val_i(const Args& a, const i64* rawrow) {
  const i64 v = fold_i<J, 0>(a, rawrow, 0);
  if constexpr (kAWN[J] == 0) return v;
  else return when_all<J>(a, rawrow, seq<kAWN[J]>{}) ? v : a.ei[J];
}

// ---- accumulators ------------------------------------------------------------
// ops: 0 SUM_F, 1 SUM_I, 2 MIN_F, 3 MAX_F, 4 MIN_I, 5 MAX_I
__device__ __forceinline__ i64 f64_ord(double d) {
  const i64 b = f64_bits(d);
  return b >= 0 ? b : (i64)((u64)b ^ 0x7FFFFFFFFFFFFFFFULL);
}

template <int OP>
__device__ __forceinline__ u64 acc_init() {
  if constexpr (OP == 2) return (u64)f64_ord(1.7976931348623157e308);
  else if constexpr (OP == 3) return (u64)f64_ord(-1.7976931348623157e308);
  else if constexpr (OP == 4) return (u64)kI64Max;
  else if constexpr (OP == 5) return (u64)kI64Min;
  else return 0;
}

// Fold one row's value of aggregate J into register accumulator r (vm: the
// row's slot validity; a NULL argument folds nothing; a count-valid
// accumulator adds one).
template <int J>
__device__ __forceinline__ void fold_reg(u64& r, const Args& a, const i64* rawrow, u32 vm = ~0u) {
  constexpr int OP = kAOP[J];
  if constexpr (acc_mask<J>() != 0)
    if ((vm & acc_mask<J>()) != acc_mask<J>()) return;
  if constexpr (kACN[J]) {
    r += 1;
    return;
  }
  if constexpr (kAISF[J]) {
    const double v = val_f<J>(a, rawrow);
    if constexpr (OP == 0) {
      r = (u64)f64_bits(as_f64((i64)r) + v);
    } else if (v == v) {
      const i64 o = f64_ord(v);
      if constexpr (OP == 2) r = o < (i64)r ? (u64)o : r;
      else r = o > (i64)r ? (u64)o : r;
    }
  } else {
    const i64 v = val_i<J>(a, rawrow);
    if constexpr (OP == 1) r = (u64)((i64)r + v);
    else if constexpr (OP == 4) r = v < (i64)r ? (u64)v : r;
    else r = v > (i64)r ? (u64)v : r;
  }
}

// Same, into this lane's private LDS word (no-return LDS atomic).
template <int J>
__device__ __forceinline__ void fold_lds(u64* w, const Args& a, const i64* rawrow, u32 vm = ~0u) {
  constexpr int OP = kAOP[J];
  if constexpr (acc_mask<J>() != 0)
    if ((vm & acc_mask<J>()) != acc_mask<J>()) return;
  if constexpr (kACN[J]) {
    atomicAdd((unsigned long long*)w, 1ull);
    return;
  }
  if constexpr (kAISF[J]) {
    const double v = val_f<J>(a, rawrow);
    if constexpr (OP == 0) {
      atomicAdd((double*)w, v);
    } else if (v == v) {
      if constexpr (OP == 2) atomicMin((long long*)w, (long long)f64_ord(v));
      else atomicMax((long long*)w, (long long)f64_ord(v));
    }
  } else {
    const i64 v = val_i<J>(a, rawrow);
    if constexpr (OP == 1) atomicAdd((unsigned long long*)w, (unsigned long long)v);
    else if constexpr (OP == 4) atomicMin((long long*)w, (long long)v);
    else atomicMax((long long*)w, (long long)v);
  }
}

// ---- tile iteration (persistent grid) ---------------------------------------
// Tile t holds rows [lo, rend).  A row group can hold any number of rows, so
// a span (and each of its tiles) can start at any row; base, where the
// tile's chunks start, is lo rounded down to a multiple of A, so that
// row-blocked vector loads stay aligned.  Rows below lo (the previous
// tile's) and at or past rend are masked by the caller.
struct TileIt {
  i64 t, lo, base, rend;
  bool have;
};

template <int A>
__device__ __forceinline__ void tile_begin(const Args& a, i64 t, TileIt& it) {
  it.t = t;
  it.have = t < a.ntiles;
  if (!it.have) return;
  int s = 0;
  while (s + 1 < a.nspans && t >= a.sp_tile0[s + 1]) s++;
  it.lo = a.sp_start[s] + (t - a.sp_tile0[s]) * (i64)a.tile_rows;
  it.base = it.lo - (it.lo % A);
  const i64 send = a.sp_start[s] + a.sp_rows[s];
  it.rend = it.lo + a.tile_rows < send ? it.lo + a.tile_rows : send;
}

template <int U>
__device__ __forceinline__ void tile_advance(const Args& a, TileIt& it) {
  it.base += kWave * U;
  if (it.base >= it.rend) tile_begin<U>(a, it.t + gridDim.x, it);
}

__device__ __forceinline__ u64 combine(int op, u64 a, u64 b) {
  switch (op) {
    case 0: return (u64)f64_bits(as_f64((i64)a) + as_f64((i64)b));
    case 1:
    case 100: return a + b;
    case 2:
    case 4: return (i64)b < (i64)a ? b : a;
    case 3:
    case 5: return (i64)b > (i64)a ? b : a;
    default: return b < a ? b : a;  // 101: first row (unsigned min)
  }
}

// The value of lane (this lane's source under DPP control CTRL) in rows
// ROWM; lanes without a source, and rows outside ROWM, take id.
template <int CTRL, int ROWM>
__device__ __forceinline__ u64 dpp_u64(u64 x, u64 id) {
  const u32 lo = (u32)__builtin_amdgcn_update_dpp((int)(u32)id, (int)(u32)x, CTRL, ROWM, 0xf, false);
  const u32 hi =
      (u32)__builtin_amdgcn_update_dpp((int)(u32)(id >> 32), (int)(u32)(x >> 32), CTRL, ROWM, 0xf, false);
  return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 op_identity(int op) {
  switch (op) {
    case 0: return 0x8000000000000000ULL;  // -0.0
    case 2:
    case 4: return (u64)kI64Max;
    case 3:
    case 5: return (u64)kI64Min;
    case 101: return ~0ULL;
    default: return 0;
  }
}

// Fixed-order reduction over the wave: an inclusive scan inside each 16-lane
// row (row_shr 1, 2, 4, 8), rows carried by row_bcast 15 and 31, so that
// lane 63 holds the total, which every lane receives.  Call with the whole
// wave active.
__device__ __forceinline__ u64 wave_reduce(int op, u64 v) {
  const u64 id = op_identity(op);
  v = combine(op, dpp_u64<0x111, 0xf>(v, id), v);
  v = combine(op, dpp_u64<0x112, 0xf>(v, id), v);
  v = combine(op, dpp_u64<0x114, 0xf>(v, id), v);
  v = combine(op, dpp_u64<0x118, 0xf>(v, id), v);
  v = combine(op, dpp_u64<0x142, 0xa>(v, id), v);
  v = combine(op, dpp_u64<0x143, 0xc>(v, id), v);
  const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)v, 63);
  const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(v >> 32), 63);
  return ((u64)hi << 32) | lo;
}

// One accumulator's 64 lane columns folded in lane order from column r0.
template <int OP>
__device__ __forceinline__ u64 fold_col(const u64* c, int r0) {
  u64 v = c[r0 & (kWave - 1)];
  for (int l = 1; l < kWave; l++) v = combine(OP, v, c[(r0 + l) & (kWave - 1)]);
  return v;
}

__device__ __forceinline__ u64 fold_columns(int op, const u64* c, int r0) {
  switch (op) {
    case 0: return fold_col<0>(c, r0);
    case 1: return fold_col<1>(c, r0);
    case 2:
    case 4: return fold_col<2>(c, r0);
    default: return fold_col<3>(c, r0);
  }
}

// The counts (sum) or first rows (unsigned minimum) of one group's columns.
template <bool kMin>
__device__ __forceinline__ u64 fold_counts(const u32* c, int r0) {
  u32 v = c[r0 & (kWave - 1)];
  for (int l = 1; l < kWave; l++) {
    const u32 x = c[(r0 + l) & (kWave - 1)];
    v = kMin ? (x < v ? x : v) : v + x;
  }
  return v;
}

template <int... J>
__device__ __forceinline__ void init_regs(u64* r, Seq<J...>) {
  ((r[J] = acc_init<kAOP[J]>()), ...);
}

template <int... J>
__device__ __forceinline__ void fold_regs(u64* r, const Args& a, const i64* row, Seq<J...>, u32 vm = ~0u) {
  (fold_reg<J>(r[J], a, row, vm), ...);
}

// w points at this lane's word of accumulator 0 of the row's group; the
// accumulator J word is kWave * J further.
template <int... J>
__device__ __forceinline__ void fold_ldss(u64* w, const Args& a, const i64* row, Seq<J...>, u32 vm = ~0u) {
  (fold_lds<J>(w + J * kWave, a, row, vm), ...);
}

}  // namespace vqj

using namespace vqj;

// VQ_KERNELS selects what a module holds: 1 dense aggregate, 2 filter bits,
// 4 join partitioning, 8 GROUP BY through per-workgroup partials, 16 the
// direct join's probe, 32 several one-group aggregates of one shape in one
// pass.
#if VQ_KERNELS & 1
// ---------------------------------------------------------------------------
// Dense GROUP BY (see k_agg_dense in k_scan.hip for the contract).
// ---------------------------------------------------------------------------
extern "C" __global__ __launch_bounds__(64) void vq_dense(const Args a) {
  extern __shared__ u64 lds[];
  constexpr int U = VQ_U, G = VQ_G, NA = VQ_NA;
  const int lane = threadIdx.x;
  u32* cnt = (u32*)(lds + (u64)G * NA * kWave);
  u32* first = cnt + G * kWave;
  u64 racc[NA > 0 ? NA : 1];
  u32 rcnt = 0, rfirst = 0xFFFFFFFFu;
  init_regs(racc, seq<NA>{});
  if constexpr (G > 1) {
#pragma unroll
    for (int j = 0; j < NA; j++)
      for (int gi = 0; gi < G; gi++) lds[(gi * NA + j) * kWave + lane] = racc[j];
    for (int gi = 0; gi < G; gi++) {
      cnt[gi * kWave + lane] = 0;
      first[gi * kWave + lane] = 0xFFFFFFFFu;
    }
  }
  TileIt it;
  tile_begin<U>(a, blockIdx.x, it);
  i64 nraw[VQ_NS][U];
  if (it.have) load_rows_blocked<U>(a, it.base, it.rend, lane, nraw);
  while (it.have) {
    i64 raw[VQ_NS][U];
#pragma unroll
    for (int c = 0; c < VQ_NS; c++)
#pragma unroll
      for (int u = 0; u < U; u++) raw[c][u] = nraw[c][u];
    const i64 base = it.base, lo = it.lo, rend = it.rend;
    tile_advance<U>(a, it);
    if (it.have) load_rows_blocked<U>(a, it.base, it.rend, lane, nraw);
#pragma unroll
    for (int u = 0; u < U; u++) {
      i64 row[VQ_NS];
#pragma unroll
      for (int c = 0; c < VQ_NS; c++) row[c] = raw[c][u];
      const i64 r = base + (i64)lane * U + u;
      if (!(r >= lo && r < rend && pass_all(a, row, seq<VQ_NT>{}))) continue;
#if VQ_NULLS
      // (the lane's U rows share a validity word: base is U-aligned)
      const u32 vm = row_valid(a, r, seq<VQ_NS>{});
      if ((vm & pred_mask()) != pred_mask()) continue;
#else
      const u32 vm = ~0u;
#endif
      if constexpr (G == 1) {
        rfirst = rcnt == 0 ? (u32)r : rfirst;
        rcnt++;
        fold_regs(racc, a, row, seq<NA>{}, vm);
      } else {
#if VQ_NULLS
        const int g = group_of_v(row, vm, seq<VQ_NS>{});
#else
        const int g = group_of(row, seq<VQ_NS>{});
#endif
        atomicAdd(&cnt[g * kWave + lane], 1u);
        atomicMin(&first[g * kWave + lane], (u32)r);
        fold_ldss(&lds[g * NA * kWave + lane], a, row, seq<NA>{}, vm);
      }
    }
  }
  constexpr int NI = NA + 2;
  // partials item-major ([G * NI][grid]): k_agg_final's block for one item
  // then reads the waves' words contiguously
  u64* out = a.part + blockIdx.x;
  const u64 ostride = gridDim.x;
  if constexpr (G > 1) {
    // the partials: lane q folds item q (group q / NI, accumulator, count
    // or first row q % NI) over the 64 lanes' columns, starting at column q
    // (so the lanes read different banks); one fixed order per item
    __syncthreads();
    for (int q = lane; q < G * NI; q += kWave) {
      const int gi = q / NI, item = q - gi * NI;
      u64 v;
      if (item < NA) v = fold_columns(kAOP[item], &lds[(gi * NA + item) * kWave], q);
      else if (item == NA) v = fold_counts<false>(&cnt[gi * kWave], q);
      else v = fold_counts<true>(&first[gi * kWave], q);
      out[q * ostride] = v;
    }
    return;
  }
  for (int gi = 0; gi < G; gi++) {
#pragma unroll
    for (int item = 0; item < NI; item++) {
      u64 v;
      int op;
      if (item < NA) {
        v = G == 1 ? racc[item] : lds[(gi * NA + item) * kWave + lane];
        op = kAOP[item];
      } else if (item == NA) {
        v = G == 1 ? rcnt : cnt[gi * kWave + lane];
        op = 100;
      } else {
        v = G == 1 ? rfirst : first[gi * kWave + lane];
        op = 101;
      }
      v = wave_reduce(op, v);
      if (lane == 0) out[(gi * NI + item) * ostride] = v;
    }
  }
}

#endif

#if VQ_KERNELS & 2
// ---------------------------------------------------------------------------
// Filter pass bits (see k_filter_bits in k_scan.hip): one tile per wave.
// ---------------------------------------------------------------------------
extern "C" __global__ __launch_bounds__(64) void vq_filter_bits(const Args a) {
  constexpr int U = VQ_U;
  const int lane = threadIdx.x;
  TileIt it;
  tile_begin<1>(a, blockIdx.x, it);
  if (!it.have) return;
  u32 count = 0;
  // word k of tile t (rows lo + 64k ..) sits at t * tile_rows / 64 + k, so
  // tiles starting at any row never share a word
  u64* words = a.bits + it.t * (i64)(a.tile_rows >> 6);
  for (i64 base = it.lo; base < it.rend; base += kWave * U) {
    i64 raw[VQ_NS][U];
    load_rows<U>(a, base, it.rend, lane, raw);
    u64 mine = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      i64 row[VQ_NS];
#pragma unroll
      for (int c = 0; c < VQ_NS; c++) row[c] = raw[c][u];
      bool p = base + u * kWave + lane < it.rend && pass_all(a, row, seq<VQ_NT>{});
#if VQ_NULLS
      if (p) {
        const u32 vm = row_valid(a, base + u * kWave + lane, seq<VQ_NS>{});
        p = (vm & pred_mask()) == pred_mask();
      }
#endif
      const u64 m = __ballot(p);
      count += (u32)__popcll(m);
      if (lane == u) mine = m;
    }
    if (lane < U && base + lane * kWave < it.rend) words[((base - it.lo) >> 6) + lane] = mine;
  }
  if (lane == 0) a.tile_counts[it.t] = count;
}
#endif

#if VQ_KERNELS & 28
// ---------------------------------------------------------------------------
// Radix partitioning of one join side (exec/parallel_join.go:322-347,422-550):
// rows that pass the side's predicate go to partition radixHash(key) & (P-1),
// P = a.P a power of two chosen per query (a run-time value: one module serves
// every partition count).  A persistent grid of G 1024-thread workgroups;
// workgroup w takes the contiguous tile range [w*ntiles/G, (w+1)*ntiles/G).
// Both kernels of a side run with the same G, so they see the same rows per
// workgroup.
//   vq_part_count:   counts[p*G + w] = workgroup w's rows of partition p
//                    (an LDS histogram in dynamic LDS of P words when the
//                    launch provides it, else global atomics)
//   vq_part_scatter: workgroup w's run of partition p starts at
//                    pstart[p] + counts[p*G + w] (k_part_runs, k_part_starts);
//                    each row goes there as (key[, row id, values of the
//                    aggregate arguments]).
// The scatter gathers rows in an LDS buffer of VQ_PB entries per partition
// and writes a full buffer as VQ_PB consecutive entries (a 128-byte line of
// keys at VQ_PB = 16), advancing the workgroup's own cursor: no atomics
// outside LDS, and the partial lines are only those at the ends of a
// workgroup's run.  Where a row lands inside its workgroup's run depends on
// LDS atomics, so the order inside a partition can differ from run to run;
// every consumer aggregates with order-independent integer updates.
// VQ_PB = 0 (partition counts too large for LDS buffers) writes each row
// straight to its position, claimed from the workgroup's cursors in LDS (or,
// when the launch gives no LDS for them, in its own column of counts).
// Dynamic LDS layouts (bytes from the host, part_lds_bytes):
//   count:            hist u32[P]                       (or none)
//   scatter, PB > 0:  fill u32[P] | cur u32[P] | key i64[P*PB]
//                     [| row u32[P*PB] | val u64[NV][P*PB]]   (probe side)
//   scatter, PB = 0:  cur u32[P]                        (or none)
// ---------------------------------------------------------------------------
namespace vqj {

constexpr int kPB = 1024;  // threads per partitioning workgroup
constexpr int kBuf = VQ_PB;

__device__ __forceinline__ u64 radix_hash(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

template <int U, int... C>
__device__ __forceinline__ void pload_one(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u,
                                          Seq<C...>) {
  ((raw[C][u] = ld<kST[C]>(a.col[C], r)), ...);
}

// A column the row's liveness depends on: the key or a predicate column.
// With the semi-join filter the others (aggregate arguments) are loaded
// only for the rows that live, after the test: most probe rows of a
// selective join fail it, and a 64-byte line none of whose rows lives is
// never fetched.
template <int C>
__device__ constexpr bool live_slot() {
  if (C == VQ_KEYSLOT) return true;
  for (int k = 0; k < VQ_NKEY; k++)
    if (kKSL[k] == C) return true;
  for (int t = 0; t < VQ_NT; t++)
    if (kTS[t] == C || (kTO[t] >= 16 && kTS2[t] == C)) return true;  // (either side of a compare)
  return false;
}
constexpr int kDjSlots = VQ_DJ_SLOTS;  // the direct join's slots per key (the host's kDjSlots)
constexpr bool kLateVals = ((VQ_BLOOM != 0) || (VQ_KERNELS & 16) != 0) && (VQ_PROBE != 0);

// The row's key: its key column, or (a GROUP BY of several dictionary-code
// columns) their codes packed as the int-key path packs them
// (exec/agg_intkey.go:67-100).
template <int... K>
__device__ __forceinline__ i64 packed_key(const i64* row, Seq<K...>) {
  return (i64)(0ull | ... | ((u64)row[kKSL[K]] << kKSH[K]));
}
__device__ __forceinline__ i64 row_key(const i64* row) {
  if constexpr (VQ_NKEY > 1) return packed_key(row, seq<VQ_NKEY>{});
  else return row[VQ_KEYSLOT];
}

template <int C, int U>
__device__ __forceinline__ void pload_slot(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u) {
  if constexpr (!kLateVals || live_slot<C>()) raw[C][u] = ld<kST[C]>(a.col[C], r);
  else raw[C][u] = 0;
}

template <int U, int... C>
__device__ __forceinline__ void pload_live_one(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u,
                                               Seq<C...>) {
  (pload_slot<C, U>(a, r, raw, u), ...);
}

template <int C>
__device__ __forceinline__ void late_slot(const Args& a, i64 r, i64* row) {
  if constexpr (kLateVals && !live_slot<C>()) row[C] = ld<kST[C]>(a.col[C], r);
}

template <int... C>
__device__ __forceinline__ void late_load(const Args& a, i64 r, i64* row, Seq<C...>) {
  (late_slot<C>(a, r, row), ...);
}

// Rows base + u*kPB + tid, clamped into the tile (with the semi-join
// filter, the probe side's live columns only).
template <int U>
__device__ __forceinline__ void pload_rows(const Args& a, i64 base, i64 rend,
                                           i64 (&raw)[VQ_NS][U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    i64 r = base + u * kPB + threadIdx.x;
    r = r < rend ? r : rend - 1;
    pload_live_one(a, r, raw, u, seq<VQ_NS>{});
  }
}

__device__ __forceinline__ void ptile(const Args& a, i64 t, i64& row0, i64& rend) {
  int s = 0;
  while (s + 1 < a.nspans && t >= a.sp_tile0[s + 1]) s++;
  row0 = a.sp_start[s] + (t - a.sp_tile0[s]) * (i64)a.tile_rows;
  const i64 send = a.sp_start[s] + a.sp_rows[s];
  rend = row0 + a.tile_rows < send ? row0 + a.tile_rows : send;
}

// Value of aggregate argument J for one row, as 64 bits (float bits or int).
template <int J>
__device__ __forceinline__ u64 value_of(const Args& a, const i64* row) {
  if constexpr (kAISF[J]) return (u64)f64_bits(val_f<J>(a, row));
  else return (u64)val_i<J>(a, row);
}

template <int... J>
__device__ __forceinline__ void values_of(const Args& a, const i64* row, u64* v, Seq<J...>) {
  ((v[J] = value_of<J>(a, row)), ...);
}

// The semi-join filter's test of a key with hash h: one 64-byte line of its
// partition's block (hash bits 20..), one 64-bit word of the line (bits
// 40..42), two bits of the word (bits 43..48 and 49..54); the same bits
// k_bloom_build sets.
__device__ __forceinline__ bool bloom_test(const Args& a, u64 h) {
  const u32 p = (u32)h & a.bloom_pmask;
  const u32 lines = 1u << a.bloom_lpp;
  const u32 line = p * lines + ((u32)(h >> 20) & (lines - 1));
  // both bits in one 64-bit word of the line: one load
  const u64 x = ((const u64*)a.bloom)[(u64)line * 8 + ((u32)(h >> 40) & 7u)];
  const u32 b1 = (u32)(h >> 43) & 63u, b2 = (u32)(h >> 49) & 63u;
  return ((x >> b1) & (x >> b2) & 1ull) != 0;
}

constexpr int kNV1 = VQ_NV > 0 ? VQ_NV : 1;

// The rows one thread has in flight: partition, key, row id, values.
template <int U>
struct PRows {
  u32 part[U];
  i64 key[U];
  u32 row[U];
  u64 val[U][kNV1];
  bool live[U];
};

template <int U, bool kVals>
__device__ __forceinline__ void prows_from(const Args& a, i64 base, i64 rend,
                                           const i64 (&raw)[VQ_NS][U], PRows<U>& q) {
  const u64 mask = (u64)a.P - 1;
#pragma unroll
  for (int u = 0; u < U; u++) {
    i64 row[VQ_NS];
#pragma unroll
    for (int c = 0; c < VQ_NS; c++) row[c] = raw[c][u];
    const i64 r = base + u * kPB + threadIdx.x;
    q.live[u] = r < rend && pass_all(a, row, seq<VQ_NT>{});
    q.key[u] = row_key(row);
    const u64 h = radix_hash((u64)q.key[u]);
    q.part[u] = (u32)(h & mask);
#if VQ_BLOOM
    q.live[u] = q.live[u] && bloom_test(a, h);
#endif
    q.row[u] = (u32)r;
    if constexpr (kVals && VQ_PROBE && VQ_NV > 0) {
      if (q.live[u]) late_load(a, r, row, seq<VQ_NS>{});
      values_of(a, row, q.val[u], seq<VQ_NV>{});
    }
  }
}

template <int U, bool kVals>
__device__ __forceinline__ void prows_eval(const Args& a, i64 base, i64 rend, PRows<U>& q) {
  i64 raw[VQ_NS][U];
  pload_rows<U>(a, base, rend, raw);
  prows_from<U, kVals>(a, base, rend, raw, q);
}

#if VQ_LIVEBITS
// The live-bit words of this wave's U groups of 64 rows (rows base + u*kPB +
// wave*64 ..; the host keeps row groups 64-aligned for this).
template <int U>
__device__ __forceinline__ void lw_load(const Args& a, i64 base, i64 rend, u64 (&lw)[U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    const i64 r0 = base + u * kPB + (threadIdx.x & ~63);
    lw[u] = r0 < rend ? a.livebits[r0 >> 6] : 0ULL;
  }
}

template <int... C>
__device__ __forceinline__ void load_all(const Args& a, i64 r, i64* row, Seq<C...>) {
  ((row[C] = ld<kST[C]>(a.col[C], r)), ...);
}

// The scatter's rows from the count pass's live bits: only live rows load
// their columns (key and aggregate arguments).
template <int U>
__device__ __forceinline__ void prows_live(const Args& a, i64 base, i64 rend, const u64 (&lw)[U],
                                           PRows<U>& q) {
  const u64 mask = (u64)a.P - 1;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < U; u++) {
    const i64 r = base + u * kPB + threadIdx.x;
    q.live[u] = r < rend && ((lw[u] >> lane) & 1ULL);
    q.row[u] = (u32)r;
    q.key[u] = 0;
    q.part[u] = 0;
    if (q.live[u]) {
      i64 row[VQ_NS];
      load_all(a, r, row, seq<VQ_NS>{});
      q.key[u] = row_key(row);
      q.part[u] = (u32)(radix_hash((u64)q.key[u]) & mask);
      if constexpr (VQ_NV > 0) values_of(a, row, q.val[u], seq<VQ_NV>{});
    }
  }
}
#endif

// The chunks (kPB * U rows, clamped to their tile) of the tile range
// [t, t1): next() steps to the following chunk, false past the last.
template <int U>
struct Chunks {
  i64 t, t1, base, rend;
  __device__ __forceinline__ bool next(const Args& a) {
    base += kPB * U;
    while (base >= rend) {
      if (++t >= t1) return false;
      ptile(a, t, base, rend);
    }
    return true;
  }
};

template <int... J>
__device__ __forceinline__ void put_vals(const Args& a, u32 pos, const u64* v, Seq<J...>) {
  ((a.vals[(i64)J * a.vstride + pos] = v[J]), ...);
}

}  // namespace vqj
#endif

#if VQ_KERNELS & 4
extern "C" __global__ __launch_bounds__(1024) void vq_part_count(const Args a) {
  extern __shared__ u32 dyn[];
  constexpr int U = VQ_U;
  const int P = a.P;
  const bool lds = a.lds_words >= (u32)P;
  const i64 t0 = (i64)blockIdx.x * a.ntiles / gridDim.x;
  const i64 t1 = (i64)(blockIdx.x + 1) * a.ntiles / gridDim.x;
  const i64 G = gridDim.x;
  for (int p = threadIdx.x; p < P; p += kPB) {
    if (lds) dyn[p] = 0;
    else a.counts[p * G + blockIdx.x] = 0;
  }
  __syncthreads();
  // the next chunk's column loads issued before this chunk's rows are
  // evaluated (and their filter lines read), so the two wait together
  Chunks<U> ck{t0 - 1, t1, 0, 0};
  ck.base = -(i64)kPB * U;  // next() steps onto tile t0
  bool have = t0 < t1 && ck.next(a);
  i64 raw[VQ_NS][U];
  if (have) pload_rows<U>(a, ck.base, ck.rend, raw);
  while (have) {
    const i64 base = ck.base, rend = ck.rend;
    have = ck.next(a);
    i64 nraw[VQ_NS][U];
    if (have) pload_rows<U>(a, ck.base, ck.rend, nraw);
    {
      PRows<U> q;
      prows_from<U, false>(a, base, rend, raw, q);
#if VQ_LIVEBITS
#pragma unroll
      for (int u = 0; u < U; u++) {
        const i64 r0 = base + u * kPB + (threadIdx.x & ~63);
        const u64 m = __ballot(q.live[u]);
        if ((threadIdx.x & 63) == 0 && r0 < rend) a.livebits[r0 >> 6] = m;
      }
#endif
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!q.live[u]) continue;
        if (lds) atomicAdd(&dyn[q.part[u]], 1u);
        else atomicAdd(&a.counts[q.part[u] * G + blockIdx.x], 1u);
      }
    }
    if (have) {
#pragma unroll
      for (int c = 0; c < VQ_NS; c++)
#pragma unroll
        for (int u = 0; u < U; u++) raw[c][u] = nraw[c][u];
    }
  }
  __syncthreads();
  if (lds)
    for (int p = threadIdx.x; p < P; p += kPB) a.counts[p * G + blockIdx.x] = dyn[p];
}

namespace vqj {

#if VQ_PB > 0
constexpr int kGrp = 64 / kBuf;  // partitions one wave flushes per step

// The scatter's LDS buffers, carved from dynamic LDS for P partitions.
struct PBuf {
  u32* fill;
  u32* cur;  // next position of the workgroup's run of each partition
  i64* key;
  u32* row;
  u64* val;  // [NV][P * kBuf]
  int P;
  // word offsets: fill 0, cur P, key 2P (even, so 8-byte aligned), row
  // 2P + 2P*kBuf, val after row rounded up to an even word
  __device__ __forceinline__ void carve(u32* dyn, int p) {
    P = p;
    fill = dyn;
    cur = dyn + P;
    key = (i64*)(dyn + 2 * P);
    row = dyn + 2 * P + 2 * P * kBuf;
    val = (u64*)(dyn + 2 * P + 2 * P * kBuf + ((P * kBuf + 1) & ~1));
  }
};

// Writes the buffers of the partitions that are full (kTail = false) or hold
// anything (kTail = true) at the workgroup's cursors.  Lanes
// [kBuf*g, kBuf*(g+1)) of a wave carry partition p0 + g; each wave takes every
// 16th group of partitions.
template <bool kTail>
__device__ __forceinline__ void flush(const Args& a, PBuf& b) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int e = lane % kBuf;
  const int P = b.P;
  for (int p0 = wv * kGrp; p0 < P; p0 += (kPB / 64) * kGrp) {
    const int p = p0 + lane / kBuf;
    const u32 f = p < P ? b.fill[p] : 0;
    const u32 n = kTail ? (f < (u32)kBuf ? f : (u32)kBuf) : (f >= (u32)kBuf ? (u32)kBuf : 0);
    if (n == 0) continue;
    const u32 g = b.cur[p];
    const bool ok = g + n <= a.pstart[p + 1];  // never past the partition's end
    if ((u32)e < n && ok) {
      const u32 pos = g + e;
      const int i = p * kBuf + e;
      a.keys[pos] = b.key[i];
#if VQ_PROBE
      a.rows[pos] = b.row[i];
#if VQ_NV > 0
      u64 v[VQ_NV];
#pragma unroll
      for (int j = 0; j < VQ_NV; j++) v[j] = b.val[(i64)j * P * kBuf + i];
      put_vals(a, pos, v, seq<VQ_NV>{});
#endif
#endif
    }
    if (e == 0) {
      if (!ok) atomicOr(a.perr, 1u);
      b.cur[p] = g + n;
      b.fill[p] = 0;
    }
  }
}
#endif

}  // namespace vqj

extern "C" __global__ __launch_bounds__(1024) void vq_part_scatter(const Args a) {
  extern __shared__ u32 dyn[];
  constexpr int U = VQ_U;
  const int P = a.P;
  const i64 t0 = (i64)blockIdx.x * a.ntiles / gridDim.x;
  const i64 t1 = (i64)(blockIdx.x + 1) * a.ntiles / gridDim.x;
  const i64 G = gridDim.x;
#if VQ_PB > 0
  PBuf b;
  b.carve(dyn, P);
  for (int p = threadIdx.x; p < P; p += kPB) {
    b.fill[p] = 0;
    b.cur[p] = a.pstart[p] + a.counts[p * G + blockIdx.x];
  }
#else
  // this workgroup's own cursors, in LDS when the launch gives room for them,
  // else advanced in place in its own column of counts
  const bool lds = a.lds_words >= (u32)P;
  for (int p = threadIdx.x; p < P; p += kPB) {
    if (lds) dyn[p] = a.pstart[p] + a.counts[p * G + blockIdx.x];
    else a.counts[p * G + blockIdx.x] += a.pstart[p];
  }
#endif
  __syncthreads();
  // one chunk's column loads in flight while the previous chunk is placed
  Chunks<U> ck{t0 - 1, t1, 0, 0};
  ck.base = -(i64)kPB * U;  // next() steps onto tile t0
  bool have = t0 < t1 && ck.next(a);
#if VQ_LIVEBITS
  u64 lw[U];
  if (have) lw_load<U>(a, ck.base, ck.rend, lw);
#else
  i64 raw[VQ_NS][U];
  if (have) pload_rows<U>(a, ck.base, ck.rend, raw);
#endif
  while (have) {
    const i64 base = ck.base, rend = ck.rend;
    PRows<U> q;
#if VQ_LIVEBITS
    prows_live<U>(a, base, rend, lw, q);
    have = ck.next(a);
    if (have) lw_load<U>(a, ck.base, ck.rend, lw);
#else
    prows_from<U, true>(a, base, rend, raw, q);
    have = ck.next(a);
    if (have) pload_rows<U>(a, ck.base, ck.rend, raw);
#endif
#if VQ_PB > 0
    // insert what fits; flush full buffers; again until every row is in
    for (;;) {
      bool pend = false;
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!q.live[u]) continue;
        const u32 s = atomicAdd(&b.fill[q.part[u]], 1u);
        if (s < (u32)kBuf) {
          const int i = q.part[u] * kBuf + s;
          b.key[i] = q.key[u];
#if VQ_PROBE
          b.row[i] = q.row[u];
#if VQ_NV > 0
#pragma unroll
          for (int j = 0; j < VQ_NV; j++) b.val[(i64)j * P * kBuf + i] = q.val[u][j];
#endif
#endif
          q.live[u] = false;
        } else {
          pend = true;
        }
      }
      __syncthreads();
      flush<false>(a, b);
      if (!__syncthreads_or(pend)) break;
    }
#else
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!q.live[u]) continue;
      const u32 p = q.part[u];
      const u32 pos = lds ? atomicAdd(&dyn[p], 1u) : atomicAdd(&a.counts[p * G + blockIdx.x], 1u);
      if (pos >= a.pstart[p + 1]) {
        atomicOr(a.perr, 1u);
        continue;
      }
      a.keys[pos] = q.key[u];
#if VQ_PROBE
      a.rows[pos] = q.row[u];
#if VQ_NV > 0
      put_vals(a, pos, q.val[u], seq<VQ_NV>{});
#endif
#endif
    }
#endif
  }
#if VQ_PB > 0
  __syncthreads();
  flush<true>(a, b);
#endif
}
#endif

#if VQ_KERNELS & 8
// ---------------------------------------------------------------------------
// GROUP BY of at most a few hundred keys (HashAggregate's int-key path and
// ParallelHashAggregate's partials + merge, exec/agg_intkey.go:206-361,
// exec/parallel.go:18-314): every workgroup takes an even share of the rows
// and aggregates them into its own accumulators in LDS; k_local_merge then
// combines the workgroups' partials per group.  The distinct keys were
// collected beforehand (k_local_keys) into one table that every workgroup
// copies, so a group has the same index in every workgroup.
//
// LDS: keys i64[lcap] | index u32[lcap] | acc u64[groups][lnw][S] |
//      cnt u32[groups][S] | first u32[groups][S]
// S copies of the accumulators, as many as fit (at most one per thread):
// thread t updates copy t mod S, so the lanes of a wave that share a group
// spread their atomics over up to S addresses (a handful of groups would
// otherwise queue every lane of a wave on one word).  Every update is an
// integer add, minimum or maximum, so neither the copy nor the order changes
// the result.  Partials: part[wg][group][lnw + 2] (words, count, first row).
// ---------------------------------------------------------------------------
namespace vqj {

constexpr int kLT = 512;  // threads per workgroup
constexpr u32 kLocalOverflow = 4u;

// The slots the predicate reads.  A join + GROUP BY of the joined rows
// (VQ_LOOKUP) loads only these for every probe row and the others (join
// key, group columns, aggregate arguments) for the rows that pass: a join's
// probe predicate keeps few rows (Q12's about 1 %), so most of those
// columns' lines are never fetched.
__host__ __device__ constexpr u32 lpred_slots() {
  u32 m = 0;
  for (int t = 0; t < VQ_NT; t++) {
    m |= 1u << kTS[t];
    if (kTO[t] >= 16) m |= 1u << kTS2[t];
  }
  return m;
}
constexpr bool kLateLocal = VQ_LOOKUP != 0 && VQ_NT > 0;
template <int C>
__host__ __device__ constexpr bool late_slot() {
  return kLateLocal && !kSB[C] && ((lpred_slots() >> C) & 1u) == 0;
}

// A probe-side slot's value of row r (VQ_LOOKUP: a build-side slot is
// loaded for each matching build row instead; a late slot for a passing row).
template <int C, int U>
__device__ __forceinline__ void lload_slot(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u) {
  if constexpr (kSB[C] || late_slot<C>()) raw[C][u] = 0;
  else raw[C][u] = ld<kST[C]>(a.col[C], r);
}

// The late slots of a passing row r.
template <int... C>
__device__ __forceinline__ void llate_load(const Args& a, i64 r, i64* row, Seq<C...>) {
  ((late_slot<C>() ? (void)(row[C] = ld<kST[C]>(a.col[C], r)) : (void)0), ...);
}

template <int U, int... C>
__device__ __forceinline__ void lload_one(const Args& a, i64 r, i64 (&raw)[VQ_NS][U], int u,
                                          Seq<C...>) {
  (lload_slot<C, U>(a, r, raw, u), ...);
}

#if VQ_LOOKUP
// The first build row with this key (0xFFFFFFFF: none).
__device__ __forceinline__ u32 lookup_head(const Args& a, i64 key) {
  if (a.lk_range) {
    const u64 i = (u64)key - (u64)a.lk_base;
    return i < a.lk_range ? a.lk_head[i] : 0xFFFFFFFFu;
  }
  if (a.lk_packed) {
    if (key < 0 || key >= 0xFFFFFFFFll) return 0xFFFFFFFFu;  // (no build key there)
    const u64 kk = (u64)key << 32;
    u32 i = (u32)(radix_hash((u64)key) >> 32) & a.lk_mask;
    for (u32 pr = 0; pr <= a.lk_mask; pr++) {
      const u64 w = ((const u64*)a.lk_keys)[i];
      if ((w & 0xFFFFFFFF00000000ull) == kk) return (u32)w;
      if (w == ~0ull) break;
      i = (i + 1) & a.lk_mask;
    }
    return 0xFFFFFFFFu;
  }
  if (key == kI64Min) return *a.lk_min;
  u32 i = (u32)(radix_hash((u64)key) >> 32) & a.lk_mask;
  for (u32 pr = 0; pr <= a.lk_mask; pr++) {
    const i64 k = a.lk_keys[i];
    if (k == key) return a.lk_head[i];
    if (k == kI64Min) break;
    i = (i + 1) & a.lk_mask;
  }
  return 0xFFFFFFFFu;
}

template <int C>
__device__ __forceinline__ void build_slot(const Args& a, i64 b, i64* row) {
  if constexpr (kSB[C]) row[C] = ld<kST[C]>(a.col[C], b);
}

template <int... C>
__device__ __forceinline__ void build_load(const Args& a, i64 b, i64* row, Seq<C...>) {
  (build_slot<C>(a, b, row), ...);
}
#endif

// Rows base + u*kLT + tid, clamped into the tile.
template <int U>
__device__ __forceinline__ void lload_rows(const Args& a, i64 base, i64 rend,
                                           i64 (&raw)[VQ_NS][U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    i64 r = base + u * kLT + threadIdx.x;
    r = r < rend ? r : rend - 1;
    lload_one(a, r, raw, u, seq<VQ_NS>{});
  }
}

// Aggregate J's value of one row into copy c of group g (fold kLOP[J]: 0
// float sum, 1 int sum, 2/3 float min/max, 4/5 int min/max, of value
// kLSRC[J]).
template <int J>
__device__ __forceinline__ void local_fold(const Args& a, u64* acc, u32 S, u32 g, u32 c,
                                           const u64* vals) {
  constexpr int op = kLOP[J];
  const u64 v = vals[kLSRC[J]];
  u64* w = acc + ((u64)g * a.lnw + a.vword[J]) * S + c;
  if constexpr (op == 0) {
    vqfx::FixScale fs;
    fs.lsb = a.vlsb[J];
    fs.w = a.vw[J];
    i64 c0, c1, c2;
    if (!vqfx::fix_chunks(as_f64((i64)v), fs, c0, c1, c2)) return;
    if (c0) atomicAdd(w, (u64)c0);
    if (c1) atomicAdd(w + S, (u64)c1);
    if (c2) atomicAdd(w + 2 * S, (u64)c2);
  } else if constexpr (op == 1) {
    atomicAdd(w, v);
  } else if constexpr (op == 2 || op == 3) {
    const double d = as_f64((i64)v);
    if (d != d) return;  // NaN never wins a compare
    if constexpr (op == 2) atomicMin((long long*)w, (long long)f64_ord(d));
    else atomicMax((long long*)w, (long long)f64_ord(d));
  } else if constexpr (op == 4) {
    atomicMin((long long*)w, (long long)v);
  } else {
    atomicMax((long long*)w, (long long)v);
  }
}

template <int... J>
__device__ __forceinline__ void local_folds(const Args& a, u64* acc, u32 S, u32 g, u32 c,
                                            const u64* v, Seq<J...>) {
  (local_fold<J>(a, acc, S, g, c, v), ...);
}

}  // namespace vqj

extern "C" __global__ __launch_bounds__(512) void vq_hash_local(const Args a) {
  extern __shared__ u64 lds[];
  __shared__ u8 s_wop[32];  // per accumulator word: 0 add, 1 min, 2 max
  constexpr int U = VQ_U;
  const int tid = threadIdx.x;
  if (a.lflags[0] & kLocalOverflow) return;  // more keys than this path holds
  // this pass's keys: ranks [k0, k1) of the key list (the regular keys in
  // slot order of the key table, then INT64_MIN when a row has it)
  const int cap = a.lcap;
  const int lg = 31 - __clz(cap);
  const bool has_min = a.lflags[10] != 0;
  const u32 total = a.lflags[2];
  const u32 nreg = total - (has_min ? 1u : 0u);
  const u32 k0 = (u32)a.lk0;
  if (k0 >= total) return;  // (a pass beyond the keys)
  const u32 k1 = total - k0 < (u32)(cap / 2) ? total : k0 + (u32)(cap / 2);
  const u32 nk = k1 - k0;
  const bool whole = k0 == 0 && k1 == total;  // every key in this pass
  const int min_g = has_min && nreg >= k0 && nreg < k1 ? (int)(nreg - k0) : -1;
  const int NW = a.lnw;
  i64* tk = (i64*)lds;
  u32* tix = (u32*)(tk + cap);
  u64* acc = (u64*)(tix + cap + (cap & 1));
  const u64 room = (u64)a.lds_words * 4 - (u64)cap * 12 - (cap & 1) * 4;
  const u64 per_copy = (u64)nk * (NW * 8 + 8);
  u32 S = kLT;
  while (S > 1 && S * per_copy > room) S >>= 1;
  if (S * per_copy > room) {  // (the host sizes LDS for its key limit)
    if (tid == 0) atomicOr(&a.lflags[0], kLocalOverflow);
    return;
  }
  u32* cnt = (u32*)(acc + (u64)nk * NW * S);
  u32* first = cnt + (u64)nk * S;

  for (int i = tid; i < cap; i += kLT) tk[i] = kI64Min;
  if (tid < NW) {
    u8 op = 0;
    for (int j = 0; j < a.nvop; j++) {
      const int nw = a.vop[j] == 0 ? 3 : 1;
      if (tid >= a.vword[j] && tid < a.vword[j] + nw)
        op = (a.vop[j] == 2 || a.vop[j] == 4) ? 1 : (a.vop[j] == 3 || a.vop[j] == 5) ? 2 : 0;
    }
    s_wop[tid] = op;
  }
  __syncthreads();
  // the pass's regular keys into the LDS table, each with its group index
  const u32 kr1 = k1 < nreg ? k1 : nreg;
  // (slots from the hash's bits 20 and up, not its top bits: the key list is
  // in the key table's slot order, by top bits, so a pass's keys share them)
  for (u32 r = k0 + tid; r < kr1; r += kLT) {
    const i64 key = a.lgkey[r];
    u32 i = (u32)(radix_hash((u64)key) >> 20) & (u32)(cap - 1);
    for (int pr = 0; pr < cap; pr++) {
      const u64 old = atomicCAS((unsigned long long*)&tk[i], (unsigned long long)kI64Min,
                                (unsigned long long)key);
      if (old == (u64)kI64Min) {
        tix[i] = r - k0;
        break;
      }
      i = (i + 1) & (u32)(cap - 1);
    }
  }
  for (u64 x = tid; x < (u64)nk * NW * S; x += kLT) {
    const int op = s_wop[(x / S) % NW];
    acc[x] = op == 1 ? (u64)kI64Max : op == 2 ? (u64)kI64Min : 0;
  }
  for (u64 x = tid; x < (u64)nk * S; x += kLT) {
    cnt[x] = 0;
    first[x] = 0xFFFFFFFFu;
  }
  __syncthreads();

  // this workgroup's tiles [t0, t1), kLT * U rows at a time, the next rows'
  // loads in flight while the current ones are folded
  const u32 c = (u32)tid & (S - 1);
  const i64 t0 = (i64)blockIdx.x * a.ntiles / gridDim.x;
  const i64 t1 = (i64)(blockIdx.x + 1) * a.ntiles / gridDim.x;
  i64 t = t0 - 1, base = 0, rend = 0;
  auto step = [&]() -> bool {
    base += (i64)kLT * U;
    while (base >= rend) {
      if (++t >= t1) return false;
      ptile(a, t, base, rend);
    }
    return true;
  };
  base = -(i64)kLT * U;
  bool have = t0 < t1 && step();
  i64 raw[VQ_NS][U];
  if (have) lload_rows<U>(a, base, rend, raw);
  while (have) {
    i64 cur[VQ_NS][U];
#pragma unroll
    for (int q = 0; q < VQ_NS; q++)
#pragma unroll
      for (int u = 0; u < U; u++) cur[q][u] = raw[q][u];
    const i64 cbase = base, cend = rend;
    have = step();
    if (have) lload_rows<U>(a, base, rend, raw);
#pragma unroll
    for (int u = 0; u < U; u++) {
      i64 row[VQ_NS];
#pragma unroll
      for (int q = 0; q < VQ_NS; q++) row[q] = cur[q][u];
      const i64 rr = cbase + u * kLT + tid;
      if (!(rr < cend && pass_all(a, row, seq<VQ_NT>{}))) continue;
      if constexpr (kLateLocal) llate_load(a, rr, row, seq<VQ_NS>{});
      const i64 key = row_key(row);
      u32 g = 0xFFFFFFFFu;
      if (key == kI64Min) {
        if (min_g >= 0) g = (u32)min_g;
      } else {
        u32 i = (u32)(radix_hash((u64)key) >> 20) & (u32)(cap - 1);
        for (int pr = 0; pr < cap; pr++) {
          const i64 k = tk[i];
          if (k == key) {
            g = tix[i];
            break;
          }
          if (k == kI64Min) break;
          i = (i + 1) & (u32)(cap - 1);
        }
      }
      if (g >= nk) {  // another pass's key, or (one pass) one the key pass missed
        if (whole) atomicOr(&a.lflags[0], kLocalOverflow);
        continue;
      }
#if VQ_LOOKUP
      // every build row with the probe row's join key: one joined row each
      for (u32 b = lookup_head(a, row[VQ_JKEY]); b != 0xFFFFFFFFu; b = a.lk_next[b]) {
        build_load(a, (i64)b, row, seq<VQ_NS>{});
#else
      {
#endif
        atomicAdd(&cnt[(u64)g * S + c], 1u);
        atomicMin(&first[(u64)g * S + c], (u32)rr);
#if VQ_NV > 0
        u64 v[VQ_NV];
        values_of(a, row, v, seq<VQ_NV>{});
        local_folds(a, acc, S, g, c, v, seq<VQ_NAG>{});
#endif
      }
    }
  }
  __syncthreads();
  // fold the copies into this workgroup's partials
  const int NI = NW + 2;
  u64* out = a.part + (u64)blockIdx.x * nk * NI;
  for (u64 x = tid; x < (u64)nk * NI; x += kLT) {
    const u32 g = (u32)(x / NI);
    const int w = (int)(x % NI);
    u64 v;
    if (w < NW) {
      const int op = s_wop[w];
      const u64* p = acc + ((u64)g * NW + w) * S;
      v = p[0];
      for (u32 k = 1; k < S; k++) {
        const u64 o = p[k];
        if (op == 0) v += o;
        else if (op == 1) v = (i64)o < (i64)v ? o : v;
        else v = (i64)o > (i64)v ? o : v;
      }
    } else if (w == NW) {
      v = 0;
      for (u32 k = 0; k < S; k++) v += cnt[(u64)g * S + k];
    } else {
      u32 f = 0xFFFFFFFFu;
      for (u32 k = 0; k < S; k++) f = first[(u64)g * S + k] < f ? first[(u64)g * S + k] : f;
      v = f;
    }
    out[x] = v;
  }
}
#endif

#if VQ_KERNELS & 16
// ---------------------------------------------------------------------------
// Probe of the direct join (HashJoin's probe feeding ParallelHashJoinAggregate's
// partial aggregate, exec/join.go:84-155, exec/parallel_join.go:780-988), for
// build keys whose zone maps bound them to a dense range: the build side is
// one bit per key of the range (k_dj_mark, k_dj_pack), so a probe row's
// match is a test of one bit.  Two kernels over the same persistent grid of
// even tile shares:
//   vq_dj_probe  (1024 threads) streams the key and predicate columns (the
//                next chunk's in flight), tests the predicate and the bit,
//                and lists the matching rows in its region (tiles t0.. of
//                tile_rows rows: at most one item per row);
//   vq_dj_match  (256 threads) takes workgroup w's listed rows: loads their
//                columns, evaluates the aggregate arguments, and makes each
//                match entry e (its position in the region) -- probe row,
//                argument values -- one of its key's slots by one atomic add
//                on the key's count (past kDjSlots: chained by an exchange of
//                the key's overflow head); the key's first match lists the
//                key in the workgroup's list, so every met key is listed once.
// Split so the probe's waves never wait on a match's loads or atomic: most
// probe rows of a selective join do not match, and the matches, a few per
// cent of the rows, are then handled all at once.  Global atomics are slow
// next to stores (they are serviced past the caches), so one per match is
// all this costs; the groups are summed exactly, entry by entry, when they are
// collected (k_dj_collect).
// ---------------------------------------------------------------------------
extern "C" __global__ __launch_bounds__(1024, 8) void vq_dj_probe(const Args a) {
  __shared__ u32 s_nm;
  constexpr int U = VQ_U;
  const i64 t0 = (i64)blockIdx.x * a.ntiles / gridDim.x;
  const i64 t1 = (i64)(blockIdx.x + 1) * a.ntiles / gridDim.x;
  const i64 region = t0 * (i64)a.tile_rows;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_nm = 0;
  __syncthreads();
  Chunks<U> ck{t0 - 1, t1, 0, 0};
  ck.base = -(i64)kPB * U;  // next() steps onto tile t0
  bool have = t0 < t1 && ck.next(a);
  i64 raw[VQ_NS][U];
  if (have) pload_rows<U>(a, ck.base, ck.rend, raw);
  while (have) {
    const i64 base = ck.base, rend = ck.rend;
    i64 cur[VQ_NS][U];
#pragma unroll
    for (int c = 0; c < VQ_NS; c++)
#pragma unroll
      for (int u = 0; u < U; u++) cur[c][u] = raw[c][u];
    have = ck.next(a);
    if (have) pload_rows<U>(a, ck.base, ck.rend, raw);
    u64 off[U];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      i64 row[VQ_NS];
#pragma unroll
      for (int c = 0; c < VQ_NS; c++) row[c] = cur[c][u];
      const i64 r = base + u * kPB + threadIdx.x;
      off[u] = (u64)row_key(row) - (u64)a.dj_base;
      live[u] = r < rend && off[u] < a.dj_range && pass_all(a, row, seq<VQ_NT>{});
    }
    u32 pw[U];
#pragma unroll
    for (int u = 0; u < U; u++) pw[u] = live[u] ? a.dj_pres[off[u] >> 5] : 0u;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const bool hit = (pw[u] >> (u32)(off[u] & 31)) & 1u;
      const u64 m = __ballot(hit);
      if (m == 0) continue;
      const u32 below = __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
      const int leader = __ffsll((unsigned long long)m) - 1;
      u32 at = 0;
      if (lane == leader) at = atomicAdd(&s_nm, (u32)__popcll(m));
      at = (u32)__builtin_amdgcn_readlane((int)at, leader) + below;
      if (hit) a.dj_mrow[region + at] = (u32)(base + u * kPB + threadIdx.x);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) a.dj_mcount[blockIdx.x] = s_nm;
}

namespace vqj {
template <int... C>
__device__ __forceinline__ void dj_load(const Args& a, i64 r, i64* row, Seq<C...>) {
  ((row[C] = ld<kST[C]>(a.col[C], r)), ...);
}
}  // namespace vqj

extern "C" __global__ __launch_bounds__(256) void vq_dj_match(const Args a) {
  __shared__ u32 s_nt;
  const i64 t0 = (i64)blockIdx.x * a.ntiles / gridDim.x;
  const i64 region = t0 * (i64)a.tile_rows;
  const int ew = a.dj_ew;
  if (threadIdx.x == 0) s_nt = 0;
  __syncthreads();
  const u32 n = a.dj_mcount[blockIdx.x];
  for (u32 i = threadIdx.x; i < n; i += 256) {
    const u32 e = (u32)region + i;
    const u32 r = a.dj_mrow[e];
    i64 row[VQ_NS];
    dj_load(a, (i64)r, row, seq<VQ_NS>{});
    const u64 off = (u64)row_key(row) - (u64)a.dj_base;
    u32* rec = a.dj_rec + off * 8;
    const u32 s = __hip_atomic_fetch_add(&rec[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    u64* ent = a.dj_ent + (u64)e * ew;
#if VQ_NV > 0
    u64 v[VQ_NV];
    values_of(a, row, v, seq<VQ_NV>{});
#pragma unroll
    for (int j = 0; j < VQ_NV; j++) ent[1 + j] = v[j];
#endif
    u32 prev = 0;
    if (s < (u32)kDjSlots) rec[2 + s] = e;
    else prev = __hip_atomic_exchange(&rec[1], e + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ent[0] = (u64)prev << 32 | (u64)r;
    if (s == 0) a.dj_tlist[region + atomicAdd(&s_nt, 1u)] = (u32)off;
  }
  __syncthreads();
  if (threadIdx.x == 0) a.dj_tcount[blockIdx.x] = s_nt;
}
#endif

#if VQ_KERNELS & 32
// ---------------------------------------------------------------------------
// Several aggregates of one shape in one pass (a served stream's scans that
// differ only in their predicates' literals: continuous batching, no
// counterpart in the reference, whose operators run one query per call).
// The shape is vq_dense's with one group; VQ_NQ queries' literals come from
// a.mlit.  Every row's columns are read once and every query's predicate
// and folds run on it, into that query's own registers; each wave writes one
// partial per query and item, [q][NA + 2][grid] (k_agg_final's layout with a
// group per query), and k_agg_final reduces them in wave order, so each
// query's words are those of the same query run by vq_dense up to the lanes'
// summation order.
// ---------------------------------------------------------------------------
namespace vqj {
template <int T>
__device__ __forceinline__ bool termq(const Args& a, const i64* lit, const i64* rawrow) {
  constexpr int S = kTS[T], OP = kTO[T];
  if constexpr (OP == 6) {
    const i64 v = rawrow[S];
    return (u64)v < 64 && (((u64)lit[T] >> v) & 1ull);
  } else if constexpr (OP >= 16) {
    return term<T>(a, rawrow);  // (a compare of two columns has no literal)
  } else if constexpr (kST[S] == VQ_F64) {
    return cmp<OP>(as_f64(rawrow[S]), as_f64(lit[T]));
  } else {
    return cmp<OP>(rawrow[S], lit[T]);
  }
}

template <int... T>
__device__ __forceinline__ bool pass_q(const Args& a, const i64* lit, const i64* rawrow, Seq<T...>) {
  return (true && ... && termq<T>(a, lit, rawrow));
}
}  // namespace vqj

extern "C" __global__ __launch_bounds__(64) void vq_dense_multi(const Args a) {
  constexpr int U = VQ_U, NA = VQ_NA, Q = VQ_NQ, NT = VQ_NT > 0 ? VQ_NT : 1;
  const int lane = threadIdx.x;
  i64 lit[Q][NT];
#pragma unroll
  for (int q = 0; q < Q; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) lit[q][t] = VQ_NT > 0 ? a.mlit[q * VQ_NT + t] : 0;
  u64 racc[Q][NA > 0 ? NA : 1];
  u32 rcnt[Q], rfirst[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    init_regs(racc[q], seq<NA>{});
    rcnt[q] = 0;
    rfirst[q] = 0xFFFFFFFFu;
  }
  TileIt it;
  tile_begin<U>(a, blockIdx.x, it);
  i64 nraw[VQ_NS][U];
  if (it.have) load_rows_blocked<U>(a, it.base, it.rend, lane, nraw);
  while (it.have) {
    i64 raw[VQ_NS][U];
#pragma unroll
    for (int c = 0; c < VQ_NS; c++)
#pragma unroll
      for (int u = 0; u < U; u++) raw[c][u] = nraw[c][u];
    const i64 base = it.base, lo = it.lo, rend = it.rend;
    tile_advance<U>(a, it);
    if (it.have) load_rows_blocked<U>(a, it.base, it.rend, lane, nraw);
#pragma unroll
    for (int u = 0; u < U; u++) {
      i64 row[VQ_NS];
#pragma unroll
      for (int c = 0; c < VQ_NS; c++) row[c] = raw[c][u];
      const i64 r = base + (i64)lane * U + u;
      if (!(r >= lo && r < rend)) continue;
#pragma unroll
      for (int q = 0; q < Q; q++) {
        if (!pass_q(a, lit[q], row, seq<VQ_NT>{})) continue;
        rfirst[q] = rcnt[q] == 0 ? (u32)r : rfirst[q];
        rcnt[q]++;
        fold_regs(racc[q], a, row, seq<NA>{});
      }
    }
  }
  constexpr int NI = NA + 2;
  u64* out = a.part + blockIdx.x;
  const u64 ostride = gridDim.x;
#pragma unroll
  for (int q = 0; q < Q; q++) {
#pragma unroll
    for (int item = 0; item < NI; item++) {
      u64 v;
      int op;
      if (item < NA) {
        v = racc[q][item];
        op = kAOP[item];
      } else if (item == NA) {
        v = rcnt[q];
        op = 100;
      } else {
        v = rfirst[q];
        op = 101;
      }
      v = wave_reduce(op, v);
      if (lane == 0) out[(q * NI + item) * ostride] = v;
    }
  }
}
#endif
