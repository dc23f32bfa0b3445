// k_scan.hip — gfx950 kernels for the zone-pruned scan, the selection-vector
// filter and the dictionary-code GROUP BY.
//
// Work decomposition: the surviving row groups form contiguous spans; a span
// is cut into tiles of tile_rows rows and each workgroup (one wave64) owns one
// tile.  Lane l reads rows base + u*64 + l, so every load instruction of a
// wave covers 64 consecutive values of one column (fully coalesced) and U
// rows per lane are in flight before the first use.
//
// Determinism: a row's tile, lane and position in the lane's sequence are a
// fixed function of its row index, lanes combine in a fixed butterfly, and the
// cross-tile reduction runs in tile order with double-double compensation.
// Reruns of one query on one table therefore give identical bits.
#include "vq_internal.h"

namespace vq {

constexpr int kWave = 64;
constexpr int kU = 4;  // rows per lane in flight per step

__device__ __forceinline__ int64_t load_slot(const QSlot& s, int64_t r) {
  switch (s.type) {
    case VQ_U8: return ((const uint8_t*)s.p)[r];
    case VQ_U16: return ((const uint16_t*)s.p)[r];
    case VQ_DATE: return ((const int32_t*)s.p)[r];
    default: return ((const int64_t*)s.p)[r];
  }
}

__device__ __forceinline__ bool cmp_i(int op, int64_t a, int64_t b) {
  switch (op) {
    case VQ_EQ: return a == b;
    case VQ_NE: return a != b;
    case VQ_LT: return a < b;
    case VQ_LE: return a <= b;
    case VQ_GT: return a > b;
    default: return a >= b;
  }
}

__device__ __forceinline__ bool cmp_f(int op, double a, double b) {
  switch (op) {
    case VQ_EQ: return a == b;
    case VQ_NE: return a != b;
    case VQ_LT: return a < b;
    case VQ_LE: return a <= b;
    case VQ_GT: return a > b;
    default: return a >= b;
  }
}

// applyArithFloat64 / applyArithInt64 (exec/expr.go:990-1020).
__device__ __forceinline__ double arith_f(int op, double a, double b) {
  switch (op) {
    case VQ_ADD: return a + b;
    case VQ_SUB: return a - b;
    case VQ_MUL: return a * b;
    default: return b == 0.0 ? __longlong_as_double(0x7ff8000000000000LL) : a / b;
  }
}

__device__ __forceinline__ int64_t arith_i(int op, int64_t a, int64_t b) {
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

// Loads every referenced column for U rows of this lane.  Rows past the tile
// end re-read the tile's last row (in bounds) and are masked out afterwards.
template <int U>
__device__ __forceinline__ void load_rows(const QSlot* slot, int nslots, int64_t base,
                                          int64_t rend, int lane, int64_t (&raw)[kMaxSlots][U]) {
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) {
    if (c < nslots) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        int64_t r = base + u * kWave + lane;
        r = r < rend ? r : rend - 1;
        raw[c][u] = load_slot(slot[c], r);
      }
    }
  }
}

template <int U>
__device__ __forceinline__ void eval_pred(const QSlot* slot, int nslots, int64_t base, int64_t rend,
                                          int lane, const int64_t (&raw)[kMaxSlots][U],
                                          bool (&pass)[U], int (&g)[U]) {
#pragma unroll
  for (int u = 0; u < U; u++) {
    pass[u] = base + u * kWave + lane < rend;
    g[u] = 0;
  }
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) {
    if (c < nslots) {
      const QSlot& s = slot[c];
      const bool isf = s.type == VQ_F64;
      for (int j = 0; j < s.nterms; j++) {
        const QTerm t = s.terms[j];
#pragma unroll
        for (int u = 0; u < U; u++) {
          bool ok = isf ? cmp_f(t.op, bits_to_f64(raw[c][u]), t.fval) : cmp_i(t.op, raw[c][u], t.ival);
          pass[u] = pass[u] && ok;
        }
      }
      if (s.key_stride) {
#pragma unroll
        for (int u = 0; u < U; u++) g[u] += s.key_stride * (int)raw[c][u];
      }
    }
  }
}

// Selects column `slot` of the loaded rows (slot is wave-uniform: the branch
// is scalar and the copy is a register move).
template <int U>
__device__ __forceinline__ void pick(const int64_t (&raw)[kMaxSlots][U], int slot, int64_t (&x)[U]) {
  switch (slot) {
#define VQ_PICK(c)                                   \
  case c:                                            \
    _Pragma("unroll") for (int u = 0; u < U; u++) x[u] = raw[c][u]; \
    break;
    VQ_PICK(0)
    VQ_PICK(1)
    VQ_PICK(2)
    VQ_PICK(3)
    VQ_PICK(4)
    VQ_PICK(5)
    VQ_PICK(6)
    default:
      _Pragma("unroll") for (int u = 0; u < U; u++) x[u] = raw[7][u];
#undef VQ_PICK
  }
}

// The left fold of exec/project.go's materialised BinOps, one rounding per
// operation (built with -ffp-contract=off: no fused multiply-add).
template <int U>
__device__ __forceinline__ void eval_f(const QExpr& e, const QSlot* slot,
                                       const int64_t (&raw)[kMaxSlots][U], double (&v)[U]) {
  for (int k = 0; k < e.nfactors; k++) {
    const QFactor f = e.f[k];
    double x[U];
    if (f.kind != VQ_F_LIT) {
      int64_t xi[U];
      pick<U>(raw, f.slot, xi);
      if (slot[f.slot].type == VQ_F64) {
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = bits_to_f64(xi[u]);
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = (double)xi[u];
      }
    }
    switch (f.kind) {
      case VQ_F_LIT:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = f.fval;
        break;
      case VQ_F_LIT_OP_COL:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = arith_f(f.op, f.fval, x[u]);
        break;
      case VQ_F_COL_OP_LIT:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = arith_f(f.op, x[u], f.fval);
        break;
      default:
        break;
    }
    if (k == 0) {
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = x[u];
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = arith_f(f.combine, v[u], x[u]);
    }
  }
}

template <int U>
__device__ __forceinline__ void eval_i(const QExpr& e, const int64_t (&raw)[kMaxSlots][U],
                                       int64_t (&v)[U]) {
  for (int k = 0; k < e.nfactors; k++) {
    const QFactor f = e.f[k];
    int64_t x[U];
    if (f.kind != VQ_F_LIT) pick<U>(raw, f.slot, x);
    switch (f.kind) {
      case VQ_F_LIT:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = f.ival;
        break;
      case VQ_F_LIT_OP_COL:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = arith_i(f.op, f.ival, x[u]);
        break;
      case VQ_F_COL_OP_LIT:
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = arith_i(f.op, x[u], f.ival);
        break;
      default:
        break;
    }
    if (k == 0) {
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = x[u];
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = arith_i(f.combine, v[u], x[u]);
    }
  }
}

__device__ __forceinline__ uint64_t acc_init(int op) {
  switch (op) {
    case ACC_MIN_F: return (uint64_t)f64_to_bits(1.7976931348623157e308);
    case ACC_MAX_F: return (uint64_t)f64_to_bits(-1.7976931348623157e308);
    case ACC_MIN_I: return (uint64_t)INT64_MAX;
    case ACC_MAX_I: return (uint64_t)INT64_MIN;
    default: return 0;
  }
}

// One fold step of foldRow (exec/agg_intkey.go:303-361) on the raw 64-bit
// accumulator word (float64 kept as IEEE bits, like HashAggregate's []int64).
__device__ __forceinline__ uint64_t fold_f(int op, uint64_t acc, double v) {
  double cur = bits_to_f64((int64_t)acc);
  switch (op) {
    case ACC_SUM_F: return (uint64_t)f64_to_bits(cur + v);
    case ACC_MIN_F: return v < cur ? (uint64_t)f64_to_bits(v) : acc;
    default: return v > cur ? (uint64_t)f64_to_bits(v) : acc;
  }
}

__device__ __forceinline__ uint64_t fold_i(int op, uint64_t acc, int64_t v) {
  int64_t cur = (int64_t)acc;
  switch (op) {
    case ACC_SUM_I: return (uint64_t)cur + (uint64_t)v;
    case ACC_MIN_I: return v < cur ? (uint64_t)v : acc;
    default: return v > cur ? (uint64_t)v : acc;
  }
}

// Combine two partial accumulators (mergePartialAgg, exec/parallel.go:237-314).
__device__ __forceinline__ uint64_t combine(int op, uint64_t a, uint64_t b) {
  switch (op) {
    case ACC_SUM_F: return (uint64_t)f64_to_bits(bits_to_f64((int64_t)a) + bits_to_f64((int64_t)b));
    case ACC_SUM_I: return a + b;
    case ACC_MIN_F: return fold_f(ACC_MIN_F, a, bits_to_f64((int64_t)b));
    case ACC_MAX_F: return fold_f(ACC_MAX_F, a, bits_to_f64((int64_t)b));
    case ACC_MIN_I: return fold_i(ACC_MIN_I, a, (int64_t)b);
    default: return fold_i(ACC_MAX_I, a, (int64_t)b);
  }
}

constexpr int ITEM_COUNT = 100;  // pseudo-ops for the per-group count / first row
constexpr int ITEM_FIRST = 101;

__device__ __forceinline__ int item_op(const QDesc& d, int item) {
  return item < d.nacc ? d.acc[item].op : (item == d.nacc ? ITEM_COUNT : ITEM_FIRST);
}

__device__ __forceinline__ uint64_t combine_item(int op, uint64_t a, uint64_t b) {
  if (op == ITEM_COUNT) return a + b;
  if (op == ITEM_FIRST) return a < b ? a : b;
  return combine(op, a, b);
}

// Fixed-order butterfly over the wave; every lane ends with the same value.
__device__ __forceinline__ uint64_t wave_reduce(int op, uint64_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = (uint64_t)__shfl_xor((long long)v, m, kWave);
    // order the pair by lane so both partners compute the same value
    uint64_t lo = (threadIdx.x & m) ? o : v;
    uint64_t hi = (threadIdx.x & m) ? v : o;
    v = combine_item(op, lo, hi);
  }
  return v;
}

__device__ __forceinline__ int tile_span(const QDesc& d, int64_t t) {
  int s = 0;
  while (s + 1 < d.nspans && t >= d.span[s + 1].tile0) s++;
  return s;
}

// ---------------------------------------------------------------------------
// Dense GROUP BY: filter -> arithmetic -> per-lane accumulators.
// SINGLE (no GROUP BY, accumulateDirect): accumulators live in registers.
// Otherwise each lane owns a private column of LDS accumulators
// lds[(g*nacc + j)*64 + lane], so a row updates only its own group with one
// read-modify-write per accumulator, conflict free, whatever the group count.
// Output: per tile and group, nacc accumulators + row count + first row.
// ---------------------------------------------------------------------------
template <bool SINGLE, int U>
__global__ __launch_bounds__(kWave) void k_agg_dense(const QDesc d, uint64_t* __restrict__ part) {
  extern __shared__ uint64_t lds[];
  const int lane = threadIdx.x;
  const int64_t t = blockIdx.x;
  const int s = tile_span(d, t);
  const int64_t row0 = d.span[s].start + (t - d.span[s].tile0) * (int64_t)d.tile_rows;
  const int64_t send = d.span[s].start + d.span[s].rows;
  const int64_t rend = row0 + d.tile_rows < send ? row0 + d.tile_rows : send;
  const int G = d.ngroups, NA = d.nacc;
  uint32_t* cnt = (uint32_t*)(lds + (size_t)G * NA * kWave);
  uint32_t* first = cnt + G * kWave;

  uint64_t racc[kMaxAcc];
  uint32_t rcnt = 0, rfirst = 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < kMaxAcc; j++) racc[j] = j < NA ? acc_init(d.acc[j].op) : 0;
  if (!SINGLE) {
    for (int i = 0; i < G * NA; i++) lds[i * kWave + lane] = acc_init(d.acc[i % NA].op);
    for (int i = 0; i < G; i++) {
      cnt[i * kWave + lane] = 0;
      first[i * kWave + lane] = 0xFFFFFFFFu;
    }
  }

  for (int64_t base = row0; base < rend; base += kWave * U) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(d.slot, d.nslots, base, rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, d.nslots, base, rend, lane, raw, pass, g);
    if (SINGLE) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (pass[u]) {
          if (rcnt == 0) rfirst = (uint32_t)(base + u * kWave + lane);
          rcnt++;
        }
    } else {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (pass[u]) {
          const int ci = g[u] * kWave + lane;
          const uint32_t c = cnt[ci];
          if (c == 0) first[ci] = (uint32_t)(base + u * kWave + lane);
          cnt[ci] = c + 1;
        }
    }
#pragma unroll
    for (int j = 0; j < kMaxAcc; j++) {
      if (j < NA) {
        const QAcc& a = d.acc[j];
        const int op = a.op;
        if (op == ACC_SUM_F || op == ACC_MIN_F || op == ACC_MAX_F) {
          double v[U];
          eval_f<U>(a.e, d.slot, raw, v);
#pragma unroll
          for (int u = 0; u < U; u++) {
            if (SINGLE) {
              if (pass[u]) racc[j] = fold_f(op, racc[j], v[u]);
            } else if (pass[u]) {
              uint64_t* p = &lds[(g[u] * NA + j) * kWave + lane];
              *p = fold_f(op, *p, v[u]);
            }
          }
        } else {
          int64_t v[U];
          eval_i<U>(a.e, raw, v);
#pragma unroll
          for (int u = 0; u < U; u++) {
            if (SINGLE) {
              if (pass[u]) racc[j] = fold_i(op, racc[j], v[u]);
            } else if (pass[u]) {
              uint64_t* p = &lds[(g[u] * NA + j) * kWave + lane];
              *p = fold_i(op, *p, v[u]);
            }
          }
        }
      }
    }
  }

  // Lane combine (fixed butterfly), lane 0 writes the tile partials.
  const int NI = NA + 2;
  uint64_t* out = part + (size_t)t * G * NI;
  for (int gi = 0; gi < G; gi++) {
    for (int item = 0; item < NI; item++) {
      uint64_t v;
      if (SINGLE) {
        v = item < NA ? racc[0] : 0;
#pragma unroll
        for (int j = 1; j < kMaxAcc; j++)
          if (item == j) v = racc[j];
        if (item == NA) v = rcnt;
        if (item == NA + 1) v = rfirst;
      } else {
        if (item < NA) v = lds[(gi * NA + item) * kWave + lane];
        else if (item == NA) v = cnt[gi * kWave + lane];
        else v = first[gi * kWave + lane];
      }
      const int op = item_op(d, item);
      v = wave_reduce(op, v);
      if (lane == 0) out[gi * NI + item] = v;
    }
  }
}

// Double-double helpers (Knuth TwoSum; exact error term, no FMA needed).
struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {
  double s = a + b;
  double bp = s - a;
  double ap = s - bp;
  return {s, (a - ap) + (b - bp)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  double e = s.lo + (a.lo + b.lo);
  return two_sum(s.hi, e);
}

// Cross-tile reduction: block (group, item) folds the tiles in index order;
// float sums are carried in double-double so the result is the correctly
// ordered compensated sum of the tile partials.
__global__ __launch_bounds__(256) void k_agg_final(const QDesc d, const uint64_t* __restrict__ part,
                                                   uint64_t* __restrict__ outv) {
  __shared__ uint64_t s_hi[256];
  __shared__ double s_lo[256];
  const int NI = d.nacc + 2;
  const int gi = blockIdx.x / NI, item = blockIdx.x % NI;
  const int op = item_op(d, item);
  const int tid = threadIdx.x;
  const int G = d.ngroups;
  uint64_t v = op == ITEM_COUNT ? 0 : (op == ITEM_FIRST ? 0xFFFFFFFFu : acc_init(op));
  DD dd = {0.0, 0.0};
  for (int64_t w = tid; w < d.ntiles; w += 256) {
    uint64_t x = part[((size_t)w * G + gi) * NI + item];
    if (op == ACC_SUM_F) dd = dd_add(dd, {bits_to_f64((int64_t)x), 0.0});
    else v = combine_item(op, v, x);
  }
  if (op == ACC_SUM_F) {
    s_hi[tid] = (uint64_t)f64_to_bits(dd.hi);
    s_lo[tid] = dd.lo;
  } else {
    s_hi[tid] = v;
  }
  __syncthreads();
  for (int stride = 128; stride >= 1; stride >>= 1) {
    if (tid < stride) {
      if (op == ACC_SUM_F) {
        DD a = {bits_to_f64((int64_t)s_hi[tid]), s_lo[tid]};
        DD b = {bits_to_f64((int64_t)s_hi[tid + stride]), s_lo[tid + stride]};
        DD c = dd_add(a, b);
        s_hi[tid] = (uint64_t)f64_to_bits(c.hi);
        s_lo[tid] = c.lo;
      } else {
        s_hi[tid] = combine_item(op, s_hi[tid], s_hi[tid + stride]);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (op == ACC_SUM_F) {
      double r = bits_to_f64((int64_t)s_hi[0]) + s_lo[0];
      outv[gi * NI + item] = (uint64_t)f64_to_bits(r);
    } else {
      outv[gi * NI + item] = s_hi[0];
    }
  }
}

// ---------------------------------------------------------------------------
// Selection vector: (1) per tile, evaluate the predicate and store the pass
// bits (one ballot word per 64 rows) and the tile's count; (2) exclusive scan
// of tile counts; (3) per tile, expand the words into ascending row ids with
// coalesced writes.  Tiles start on 64-row boundaries (row groups are 65,536
// rows, tiles a multiple of 64), so a word never straddles two tiles.
// ---------------------------------------------------------------------------
template <int U>
__global__ __launch_bounds__(kWave) void k_filter_bits(const QDesc d, uint64_t* __restrict__ bits,
                                                       uint32_t* __restrict__ tile_counts) {
  const int lane = threadIdx.x;
  const int64_t t = blockIdx.x;
  const int s = tile_span(d, t);
  const int64_t row0 = d.span[s].start + (t - d.span[s].tile0) * (int64_t)d.tile_rows;
  const int64_t send = d.span[s].start + d.span[s].rows;
  const int64_t rend = row0 + d.tile_rows < send ? row0 + d.tile_rows : send;
  uint32_t count = 0;
  for (int64_t base = row0; base < rend; base += kWave * U) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(d.slot, d.nslots, base, rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, d.nslots, base, rend, lane, raw, pass, g);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t wb = base + u * kWave;
      const uint64_t m = __ballot(pass[u]);
      if (wb < rend) {
        if (lane == 0) bits[wb >> 6] = m;
        count += (uint32_t)__popcll(m);
      }
    }
  }
  if (lane == 0) tile_counts[t] = count;
}

// Exclusive scan of the tile counts by one block; total -> *out_count.
__global__ __launch_bounds__(1024) void k_filter_scan(const uint32_t* __restrict__ counts,
                                                      int64_t n, uint32_t* __restrict__ offsets,
                                                      uint64_t* __restrict__ out_count) {
  __shared__ uint64_t s[1024];
  const int tid = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t lo = tid * per, hi = lo + per < n ? lo + per : n;
  uint64_t sum = 0;
  for (int64_t i = lo; i < hi; i++) sum += counts[i];
  s[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint64_t x = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  uint64_t run = s[tid] - sum;  // exclusive
  for (int64_t i = lo; i < hi; i++) {
    offsets[i] = (uint32_t)run;
    run += counts[i];
  }
  if (tid == 1023) *out_count = s[1023];
}

__global__ __launch_bounds__(kWave) void k_filter_write(const QDesc d, const uint64_t* __restrict__ bits,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out_rows) {
  const int lane = threadIdx.x;
  const int64_t t = blockIdx.x;
  const int s = tile_span(d, t);
  const int64_t row0 = d.span[s].start + (t - d.span[s].tile0) * (int64_t)d.tile_rows;
  const int64_t send = d.span[s].start + d.span[s].rows;
  const int64_t rend = row0 + d.tile_rows < send ? row0 + d.tile_rows : send;
  uint32_t pos = offsets[t];
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  for (int64_t wb = row0; wb < rend; wb += kWave) {
    const uint64_t m = bits[wb >> 6];
    if ((m >> lane) & 1ULL) out_rows[pos + (uint32_t)__popcll(m & below)] = (uint32_t)(wb + lane);
    pos += (uint32_t)__popcll(m);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static size_t dense_lds_bytes(const QDesc& d) {
  return (size_t)d.ngroups * d.nacc * kWave * 8 + (size_t)d.ngroups * kWave * 4 * 2;
}

hipError_t launch_agg_dense(const QDesc& d, uint64_t* part, hipStream_t s) {
  if (d.ntiles == 0) return hipSuccess;
  if (d.ngroups == 1) {
    hipLaunchKernelGGL((k_agg_dense<true, kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, part);
  } else {
    hipLaunchKernelGGL((k_agg_dense<false, kU>), dim3((unsigned)d.ntiles), dim3(kWave),
                       dense_lds_bytes(d), s, d, part);
  }
  return hipGetLastError();
}

hipError_t launch_agg_final(const QDesc& d, const uint64_t* part, uint64_t* out, hipStream_t s) {
  const int NI = d.nacc + 2;
  hipLaunchKernelGGL(k_agg_final, dim3((unsigned)(d.ngroups * NI)), dim3(256), 0, s, d, part, out);
  return hipGetLastError();
}

hipError_t launch_filter(const QDesc& d, uint64_t* bits, uint32_t* tile_counts,
                         uint32_t* tile_offsets, uint32_t* out_rows, uint64_t* out_count,
                         hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (d.ntiles == 0) return hipMemsetAsync(out_count, 0, sizeof(uint64_t), s);
  if (ev0) hipEventRecord(ev0, s);
  hipLaunchKernelGGL((k_filter_bits<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, bits,
                     tile_counts);
  if (ev1) hipEventRecord(ev1, s);
  hipLaunchKernelGGL(k_filter_scan, dim3(1), dim3(1024), 0, s, tile_counts, d.ntiles,
                     tile_offsets, out_count);
  hipLaunchKernelGGL(k_filter_write, dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, bits,
                     tile_offsets, out_rows);
  return hipGetLastError();
}

}  // namespace vq
