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
#include <algorithm>
#include <cstdlib>

#include "vq_internal.h"

namespace vq {

constexpr int kWave = 64;
constexpr int kU = 4;  // rows per lane in flight per step

// Branch-free column read: every column type is fetched as the aligned
// 8-byte word holding the element, then shifted, masked and (for DATE)
// sign-extended.  One load instruction shape for all columns keeps all of a
// chunk's loads in flight together (a per-type switch around the loads makes
// the compiler wait at every join).  Column buffers are padded to 256 bytes,
// so the aligned word of the last element is inside the allocation.
struct SlotLd {
  const uint8_t* p;
  uint64_t mask;
  int lw;     // log2(element width)
  bool sext;  // sign-extend a 32-bit value (DATE)
};

__device__ __forceinline__ SlotLd slot_ld(const QSlot& s) {
  SlotLd l;
  l.p = (const uint8_t*)s.p;
  l.lw = s.type == VQ_U8 ? 0 : (s.type == VQ_U16 ? 1 : (s.type == VQ_DATE ? 2 : 3));
  l.mask = l.lw == 3 ? ~0ULL : ((1ULL << (8 << l.lw)) - 1);
  l.sext = s.type == VQ_DATE;
  return l;
}

__device__ __forceinline__ int64_t ld64(const SlotLd& l, int64_t r) {
  const int64_t byte = r << l.lw;
  const uint64_t w = *(const uint64_t*)(l.p + (byte & ~(int64_t)7));
  const uint64_t x = (w >> ((byte & 7) * 8)) & l.mask;
  const int64_t sx = (int64_t)(int32_t)(uint32_t)x;
  return l.sext ? sx : (int64_t)x;
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
    default: return b == 0.0 ? __longlong_as_double(kGoNaNBits) : a / b;
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
__device__ __forceinline__ void load_rows(const SlotLd* ld, int nslots, int64_t base,
                                          int64_t rend, int lane, int64_t (&raw)[kMaxSlots][U]) {
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) {
    if (c < nslots) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        int64_t r = base + u * kWave + lane;
        r = r < rend ? r : rend - 1;
        raw[c][u] = ld64(ld[c], r);
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
        if (t.op == VQ_IN) {  // a mask over [0, 64)
#pragma unroll
          for (int u = 0; u < U; u++) {
            const bool ok = (uint64_t)raw[c][u] < 64 && (((uint64_t)t.ival >> raw[c][u]) & 1ull);
            pass[u] = pass[u] && ok;
          }
        } else if (t.op & VQ_COLCMP) {  // column OP column
          int64_t y[U];
          pick(raw, t.slot2, y);
          const int op = t.op & ~VQ_COLCMP;
#pragma unroll
          for (int u = 0; u < U; u++) {
            const bool ok = isf ? cmp_f(op, bits_to_f64(raw[c][u]), bits_to_f64(y[u])) : cmp_i(op, raw[c][u], y[u]);
            pass[u] = pass[u] && ok;
          }
        } else {
#pragma unroll
          for (int u = 0; u < U; u++) {
            bool ok = isf ? cmp_f(t.op, bits_to_f64(raw[c][u]), t.fval) : cmp_i(t.op, raw[c][u], t.ival);
            pass[u] = pass[u] && ok;
          }
        }
      }
      if (s.key_stride) {
#pragma unroll
        for (int u = 0; u < U; u++) g[u] += s.key_stride * (int)raw[c][u];
      }
    }
  }
}

// A CASE WHEN guard (CaseExpr with one WHEN, exec/expr.go:1386-1435): the
// conjunction of its terms on each row.
template <int U>
__device__ __forceinline__ void eval_when(const QExpr& e, const QSlot* slot, const int64_t (&raw)[kMaxSlots][U],
                                          bool (&ok)[U]) {
#pragma unroll
  for (int u = 0; u < U; u++) ok[u] = true;
  for (int k = 0; k < e.nwhen; k++) {
    const QWhen w = e.when[k];
    const bool isf = slot[w.slot].type == VQ_F64;
    int64_t x[U];
    pick<U>(raw, w.slot, x);
    if (w.op == VQ_IN) {
#pragma unroll
      for (int u = 0; u < U; u++) ok[u] = ok[u] && (uint64_t)x[u] < 64 && (((uint64_t)w.ival >> x[u]) & 1ull);
    } else if (w.op & VQ_COLCMP) {
      int64_t y[U];
      pick<U>(raw, w.slot2, y);
      const int op = w.op & ~VQ_COLCMP;
#pragma unroll
      for (int u = 0; u < U; u++)
        ok[u] = ok[u] && (isf ? cmp_f(op, bits_to_f64(x[u]), bits_to_f64(y[u])) : cmp_i(op, x[u], y[u]));
    } else {
#pragma unroll
      for (int u = 0; u < U; u++)
        ok[u] = ok[u] && (isf ? cmp_f(w.op, bits_to_f64(x[u]), w.fval) : cmp_i(w.op, x[u], w.ival));
    }
  }
}

// The left fold of exec/project.go's materialised BinOps, one rounding per
// operation (built with -ffp-contract=off: no fused multiply-add), under its
// CASE WHEN guard when it has one.
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
  if (e.nwhen) {
    bool ok[U];
    eval_when<U>(e, slot, raw, ok);
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = ok[u] ? v[u] : e.else_f;
  }
}

template <int U>
__device__ __forceinline__ void eval_i(const QExpr& e, const QSlot* slot, const int64_t (&raw)[kMaxSlots][U],
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
  if (e.nwhen) {
    bool ok[U];
    eval_when<U>(e, slot, raw, ok);
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = ok[u] ? v[u] : e.else_i;
  }
}

// float64 <-> order-preserving int64 (MIN/MAX of floats run as integer
// MIN/MAX; NaN rows are skipped, as foldRow's `v < cur` never takes a NaN).
__device__ __forceinline__ int64_t f64_ord(double d) {
  const int64_t b = f64_to_bits(d);
  return b >= 0 ? b : (int64_t)((uint64_t)b ^ 0x7FFFFFFFFFFFFFFFULL);
}

// Accumulator seeds: newAccums (exec/aggregate.go:255-279) puts +-MaxFloat64 /
// +-MaxInt64 sentinels in MIN/MAX slots.
__device__ __forceinline__ uint64_t acc_init(int op) {
  switch (op) {
    case ACC_MIN_F: return (uint64_t)f64_ord(1.7976931348623157e308);
    case ACC_MAX_F: return (uint64_t)f64_ord(-1.7976931348623157e308);
    case ACC_MIN_I: return (uint64_t)INT64_MAX;
    case ACC_MAX_I: return (uint64_t)INT64_MIN;
    default: return 0;
  }
}

constexpr int ITEM_COUNT = 100;  // pseudo-ops for the per-group count / first row
constexpr int ITEM_FIRST = 101;

__device__ __forceinline__ int item_op(const QDesc& d, int item) {
  return item < d.nacc ? d.acc[item].op : (item == d.nacc ? ITEM_COUNT : ITEM_FIRST);
}

// Combine two partials (mergePartialAgg, exec/parallel.go:237-314).  Float
// sums are combined in double-double by the final kernel, not here.
__device__ __forceinline__ uint64_t combine_item(int op, uint64_t a, uint64_t b) {
  switch (op) {
    case ACC_SUM_F: return (uint64_t)f64_to_bits(bits_to_f64((int64_t)a) + bits_to_f64((int64_t)b));
    case ACC_SUM_I:
    case ITEM_COUNT: return a + b;
    case ACC_MIN_F:
    case ACC_MIN_I: return (int64_t)b < (int64_t)a ? b : a;
    case ACC_MAX_F:
    case ACC_MAX_I: return (int64_t)b > (int64_t)a ? b : a;
    default: return b < a ? b : a;  // ITEM_FIRST (unsigned)
  }
}

// The value of lane (this lane's source under DPP control CTRL) in rows
// ROWM; lanes without a source, and rows outside ROWM, take id.
template <int CTRL, int ROWM>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x, uint64_t id) {
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)id, (int)(uint32_t)x, CTRL, ROWM, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(id >> 32),
                                                            (int)(uint32_t)(x >> 32), CTRL, ROWM, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t item_identity(int op) {
  switch (op) {
    case ACC_SUM_F: return 0x8000000000000000ULL;  // -0.0
    case ACC_MIN_F:
    case ACC_MIN_I: return (uint64_t)INT64_MAX;
    case ACC_MAX_F:
    case ACC_MAX_I: return (uint64_t)INT64_MIN;
    case ITEM_FIRST: return ~0ULL;
    default: return 0;
  }
}

// Fixed-order reduction over the wave: an inclusive scan inside each 16-lane
// row (row_shr 1, 2, 4, 8), rows carried by row_bcast 15 and 31, so that
// lane 63 holds the total, which every lane receives.  Call with the whole
// wave active.
__device__ __forceinline__ uint64_t wave_reduce(int op, uint64_t v) {
  const uint64_t id = item_identity(op);
  v = combine_item(op, dpp_u64<0x111, 0xf>(v, id), v);
  v = combine_item(op, dpp_u64<0x112, 0xf>(v, id), v);
  v = combine_item(op, dpp_u64<0x114, 0xf>(v, id), v);
  v = combine_item(op, dpp_u64<0x118, 0xf>(v, id), v);
  v = combine_item(op, dpp_u64<0x142, 0xa>(v, id), v);
  v = combine_item(op, dpp_u64<0x143, 0xc>(v, id), v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

struct TileIt {
  int64_t t, base, rend;
  bool have;
};

__device__ __forceinline__ void tile_begin(const QDesc& d, int64_t t, TileIt& it) {
  it.t = t;
  it.have = t < d.ntiles;
  if (!it.have) return;
  int s = 0;
  while (s + 1 < d.nspans && t >= d.span[s + 1].tile0) s++;
  it.base = d.span[s].start + (t - d.span[s].tile0) * (int64_t)d.tile_rows;
  const int64_t send = d.span[s].start + d.span[s].rows;
  it.rend = it.base + d.tile_rows < send ? it.base + d.tile_rows : send;
}

// Next chunk of kWave*U rows: the rest of this tile, then the wave's next tile
// (tiles t, t + gridDim.x, ... — a fixed assignment, so reruns fold every row
// into the same lane in the same order).
template <int U>
__device__ __forceinline__ void tile_advance(const QDesc& d, TileIt& it) {
  it.base += kWave * U;
  if (it.base >= it.rend) tile_begin(d, it.t + gridDim.x, it);
}

// ---------------------------------------------------------------------------
// Dense GROUP BY: filter -> arithmetic -> per-lane accumulators.
//
// A persistent grid of single-wave workgroups; wave w folds the tiles w,
// w + grid, ... .  Loads for the next chunk are issued before the current
// chunk is folded.
// SINGLE (no GROUP BY, accumulateDirect): accumulators live in registers.
// Otherwise lane l owns LDS column l of every accumulator,
// lds[(g*nacc + j)*64 + l], updated with no-return LDS atomics: conflict free
// (one lane per address), in program order (one lane's updates to an address
// execute in issue order), and with no read-modify-write latency chain.
// Output: per wave and group, nacc accumulators + row count + first row.
// ---------------------------------------------------------------------------
template <bool SINGLE, int U>
__global__ __launch_bounds__(kWave) void k_agg_dense(const QDesc d, uint64_t* __restrict__ part) {
  extern __shared__ uint64_t lds[];
  const int lane = threadIdx.x;
  const int G = d.ngroups, NA = d.nacc, NS = d.nslots;
  uint32_t* cnt = (uint32_t*)(lds + (size_t)G * NA * kWave);
  uint32_t* first = cnt + G * kWave;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);

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

  TileIt it;
  tile_begin(d, blockIdx.x, it);
  int64_t nraw[kMaxSlots][U];
  if (it.have) load_rows<U>(ld, NS, it.base, it.rend, lane, nraw);
  while (it.have) {
    int64_t raw[kMaxSlots][U];
#pragma unroll
    for (int c = 0; c < kMaxSlots; c++)
#pragma unroll
      for (int u = 0; u < U; u++) raw[c][u] = nraw[c][u];
    const int64_t base = it.base, rend = it.rend;
    tile_advance<U>(d, it);
    if (it.have) load_rows<U>(ld, NS, it.base, it.rend, lane, nraw);

    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, base, rend, lane, raw, pass, g);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!pass[u]) continue;
      const uint32_t row = (uint32_t)(base + u * kWave + lane);
      if (SINGLE) {
        rfirst = rcnt == 0 ? row : rfirst;
        rcnt++;
      } else {
        atomicAdd(&cnt[g[u] * kWave + lane], 1u);
        atomicMin(&first[g[u] * kWave + lane], row);
      }
    }
#pragma unroll
    for (int j = 0; j < kMaxAcc; j++) {
      if (j >= NA) continue;
      const QAcc& a = d.acc[j];
      const int op = a.op;
      if (op == ACC_SUM_F || op == ACC_MIN_F || op == ACC_MAX_F) {
        double v[U];
        eval_f<U>(a.e, d.slot, raw, v);
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (!pass[u]) continue;
          if (op == ACC_SUM_F) {
            if (SINGLE) racc[j] = (uint64_t)f64_to_bits(bits_to_f64((int64_t)racc[j]) + v[u]);
            else atomicAdd((double*)&lds[(g[u] * NA + j) * kWave + lane], v[u]);
          } else if (v[u] == v[u]) {  // skip NaN
            const int64_t o = f64_ord(v[u]);
            if (SINGLE) {
              racc[j] = (uint64_t)(op == ACC_MIN_F ? (o < (int64_t)racc[j] ? o : (int64_t)racc[j])
                                                   : (o > (int64_t)racc[j] ? o : (int64_t)racc[j]));
            } else {
              long long* p = (long long*)&lds[(g[u] * NA + j) * kWave + lane];
              if (op == ACC_MIN_F) atomicMin(p, (long long)o);
              else atomicMax(p, (long long)o);
            }
          }
        }
      } else {
        int64_t v[U];
        eval_i<U>(a.e, d.slot, raw, v);
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (!pass[u]) continue;
          if (SINGLE) {
            const int64_t cur = (int64_t)racc[j];
            racc[j] = op == ACC_SUM_I ? (uint64_t)cur + (uint64_t)v[u]
                                      : (uint64_t)(op == ACC_MIN_I ? (v[u] < cur ? v[u] : cur)
                                                                   : (v[u] > cur ? v[u] : cur));
          } else {
            unsigned long long* p = (unsigned long long*)&lds[(g[u] * NA + j) * kWave + lane];
            if (op == ACC_SUM_I) atomicAdd(p, (unsigned long long)v[u]);
            else if (op == ACC_MIN_I) atomicMin((long long*)p, (long long)v[u]);
            else atomicMax((long long*)p, (long long)v[u]);
          }
        }
      }
    }
  }
  if (!SINGLE) __syncthreads();

  // Lane combine (fixed butterfly), lane 0 writes the wave's partials.
  const int NI = NA + 2;
  // partials item-major ([G * NI][grid]), as the run-time compiled kernel
  // writes them
  uint64_t* out = part + blockIdx.x;
  const size_t ostride = gridDim.x;
  for (int gi = 0; gi < G; gi++) {
    for (int item = 0; item < NI; item++) {
      uint64_t v;
      if (SINGLE) {
        v = racc[0];
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
      v = wave_reduce(item_op(d, item), v);
      if (lane == 0) out[(gi * NI + item) * ostride] = v;
    }
  }
}

// Double-double helpers (Knuth TwoSum; exact error term, no FMA needed).
struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {
  const double s = a + b;
  const double bp = s - a;
  const double ap = s - bp;
  return {s, (a - ap) + (b - bp)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  const DD s = two_sum(a.hi, b.hi);
  const double e = s.lo + (a.lo + b.lo);
  return two_sum(s.hi, e);
}

// Cross-wave reduction: block (group, item) folds the wave partials in a
// fixed order; float sums are carried in double-double, so the result is the
// compensated sum of the partials.
__global__ __launch_bounds__(256) void k_agg_final(const QDesc d, int32_t nparts,
                                                   const uint64_t* __restrict__ part,
                                                   uint64_t* __restrict__ outv) {
  __shared__ uint64_t s_hi[256];
  __shared__ double s_lo[256];
  const int NI = d.nacc + 2;
  const int gi = blockIdx.x / NI, item = blockIdx.x % NI;
  const int op = item_op(d, item);
  const int tid = threadIdx.x;
  uint64_t v = op == ITEM_COUNT ? 0 : (op == ITEM_FIRST ? 0xFFFFFFFFu : acc_init(op));
  DD dd = {0.0, 0.0};
  const uint64_t* col = part + (size_t)blockIdx.x * nparts;  // item gi*NI + item, every wave
  for (int w = tid; w < nparts; w += 256) {
    const uint64_t x = col[w];
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
        const DD a = {bits_to_f64((int64_t)s_hi[tid]), s_lo[tid]};
        const DD b = {bits_to_f64((int64_t)s_hi[tid + stride]), s_lo[tid + stride]};
        const DD c = dd_add(a, b);
        s_hi[tid] = (uint64_t)f64_to_bits(c.hi);
        s_lo[tid] = c.lo;
      } else {
        s_hi[tid] = combine_item(op, s_hi[tid], s_hi[tid + stride]);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (op == ACC_SUM_F) outv[gi * NI + item] = (uint64_t)f64_to_bits(bits_to_f64((int64_t)s_hi[0]) + s_lo[0]);
    else outv[gi * NI + item] = s_hi[0];
  }
}

// ---------------------------------------------------------------------------
// Selection vector: (1) per tile, evaluate the predicate and store the pass
// bits (one ballot word per 64 rows) and the tile's count; (2) exclusive scan
// of tile counts; (3) per tile, expand the words into ascending row ids with
// coalesced writes.  A tile can start at any row (row groups hold any number
// of rows); its words sit at tile_index * tile_rows / 64 onward (tile_rows is
// a multiple of 64), word k holding rows lo + 64k .. lo + 64k + 63, so no
// word is shared by two tiles.
// ---------------------------------------------------------------------------
template <int U>
__global__ __launch_bounds__(kWave) void k_filter_bits(const QDesc d, uint64_t* __restrict__ bits,
                                                       uint32_t* __restrict__ tile_counts) {
  const int lane = threadIdx.x;
  const int NS = d.nslots;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  const int64_t t0 = it.t, lo = it.base;
  uint32_t count = 0;
  // word k of tile t (rows lo + 64k ..) at t * tile_rows / 64 + k: tiles
  // starting at any row never share a word
  uint64_t* words = bits + t0 * (int64_t)(d.tile_rows >> 6);
  // this kernel runs one tile per wave (grid = ntiles)
  while (it.have && it.t == t0) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(ld, NS, it.base, it.rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, it.base, it.rend, lane, raw, pass, g);
    uint64_t mine = 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t m = __ballot(pass[u]);
      count += (uint32_t)__popcll(m);
      if (lane == u) mine = m;
    }
    if (lane < U && it.base + lane * kWave < it.rend) words[((it.base - lo) >> 6) + lane] = mine;
    it.base += kWave * U;
    if (it.base >= it.rend) it.have = false;
  }
  if (lane == 0) tile_counts[t0] = count;
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
    const uint64_t x = tid >= off ? s[tid - off] : 0;
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
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  if (!it.have) return;
  uint32_t pos = offsets[it.t];
  const uint64_t below = lane ? (~0ULL >> (64 - lane)) : 0ULL;
  const uint64_t* words = bits + it.t * (int64_t)(d.tile_rows >> 6);
  for (int64_t wb = it.base; wb < it.rend; wb += kWave) {
    const uint64_t m = words[(wb - it.base) >> 6];
    if ((m >> lane) & 1ULL) out_rows[pos + (uint32_t)__popcll(m & below)] = (uint32_t)(wb + lane);
    pos += (uint32_t)__popcll(m);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t dense_lds_bytes(const QDesc& d) {
  if (d.ngroups == 1) return 0;
  return (size_t)d.ngroups * d.nacc * kWave * 8 + (size_t)d.ngroups * kWave * 4 * 2;
}

// Waves the persistent dense kernel runs with: what fits the device at once.
int32_t dense_grid(const QDesc& d) {
  const int cus = device_cus();
  int per_cu = 0;
  if (d.ngroups == 1)
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_agg_dense<true, kU>, kWave, 0);
  else
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_agg_dense<false, kU>, kWave,
                                                 dense_lds_bytes(d));
  if (per_cu <= 0) per_cu = 1;
  // VQ_WAVES_PER_CU overrides the occupancy-derived wave count (tuning sweeps)
  if (const char* e = getenv("VQ_WAVES_PER_CU")) {
    const int v = atoi(e);
    if (v > 0) per_cu = v;
  }
  const int64_t g = (int64_t)cus * per_cu;
  return (int32_t)(d.ntiles < g ? d.ntiles : g);
}

hipError_t launch_agg_dense(const QDesc& d, int32_t grid, uint64_t* part, hipStream_t s) {
  if (d.ntiles == 0 || grid == 0) return hipSuccess;
  if (d.ngroups == 1) {
    hipLaunchKernelGGL((k_agg_dense<true, kU>), dim3((unsigned)grid), dim3(kWave), 0, s, d, part);
  } else {
    hipLaunchKernelGGL((k_agg_dense<false, kU>), dim3((unsigned)grid), dim3(kWave),
                       dense_lds_bytes(d), s, d, part);
  }
  return hipGetLastError();
}

hipError_t launch_agg_final(const QDesc& d, int32_t nparts, const uint64_t* part, uint64_t* out,
                            hipStream_t s) {
  const int NI = d.nacc + 2;
  hipLaunchKernelGGL(k_agg_final, dim3((unsigned)(d.ngroups * NI)), dim3(256), 0, s, d, nparts,
                     part, out);
  return hipGetLastError();
}

// ---- the build side of a join + GROUP BY of the joined rows ---------------
// A device hash table from key to its rows (the chained table of HashJoin's
// build, exec/join_table.go:400-569): rows passing the build predicate are
// chained under their key, head[slot] the last one inserted and next[row]
// the one before.  Chain order is arrival order; the probe side folds every
// match exactly (integer and fixed-point words), so it does not matter.
__global__ __launch_bounds__(256) void k_lookup_init(const LookupBufs lk) {
  if (lk.range) {  // a direct table (cnt is cleared where a key repeats: direct_pass)
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < lk.range; i += gridDim.x * 256)
      lk.head[i] = 0xFFFFFFFFu;
    if (blockIdx.x == 0 && threadIdx.x < 4) lk.min_head[threadIdx.x] = threadIdx.x == 0 ? 0xFFFFFFFFu : 0u;
    return;
  }
  const uint32_t n = lk.mask + 1;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    lk.cnt[i] = 0;
    if (lk.packed) {
      lk.keys[i] = -1;  // all ones
    } else {
      lk.keys[i] = INT64_MIN;
      lk.head[i] = 0xFFFFFFFFu;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) lk.min_head[threadIdx.x] = threadIdx.x == 0 ? 0xFFFFFFFFu : 0u;
}

__device__ __forceinline__ void lookup_insert(const LookupBufs& lk, int64_t key, uint32_t r) {
  if (lk.packed) {
    if (key < 0 || key >= 0xFFFFFFFFll) {  // outside the bounds the zone maps gave
      atomicOr(&lk.min_head[1], 2u);
      return;
    }
    unsigned long long* slot = (unsigned long long*)lk.keys;
    const unsigned long long kk = (unsigned long long)key << 32, empty = ~0ull;
    uint32_t i = (uint32_t)(radix_hash((uint64_t)key) >> 32) & lk.mask;
    lk.next[r] = 0xFFFFFFFFu;
    for (uint32_t pr = 0; pr <= lk.mask; pr++) {
      unsigned long long w = atomicCAS(&slot[i], empty, kk | r);
      if (w == empty) return;  // a new key: r its only row
      if ((w & 0xFFFFFFFF00000000ull) == kk) {
        // a repeated key: r goes in front of its chain
        for (;;) {
          lk.next[r] = (uint32_t)w;
          const unsigned long long prev = atomicCAS(&slot[i], w, kk | r);
          if (prev == w) break;
          w = prev;
        }
        atomicMax(&lk.min_head[2], atomicAdd(&lk.cnt[i], 1u) + 2u);
        return;
      }
      i = (i + 1) & lk.mask;
    }
    atomicOr(&lk.min_head[1], 1u);  // (the host sizes the table above the rows: not reached)
    return;
  }
  if (key == INT64_MIN) {  // (the empty slots' mark: kept apart)
    lk.next[r] = atomicExch(&lk.min_head[0], r);
    atomicAdd(&lk.min_head[3], 1u);
    return;
  }
  uint32_t i = (uint32_t)(radix_hash((uint64_t)key) >> 32) & lk.mask;
  for (uint32_t pr = 0; pr <= lk.mask; pr++) {
    const unsigned long long old = atomicCAS((unsigned long long*)&lk.keys[i],
                                             (unsigned long long)INT64_MIN, (unsigned long long)key);
    if (old == (unsigned long long)INT64_MIN || old == (unsigned long long)key) {
      lk.next[r] = atomicExch(&lk.head[i], r);
      // a repeated key: cnt counts its rows after the first, so this row makes
      // cnt + 2 of them (unique keys take no atomics beyond these two)
      if (old == (unsigned long long)key) atomicMax(&lk.min_head[2], atomicAdd(&lk.cnt[i], 1u) + 2u);
      return;
    }
    i = (i + 1) & lk.mask;
  }
  atomicOr(&lk.min_head[1], 1u);  // (the host sizes the table above the rows: not reached)
}

// A direct table's three passes over the rows it takes: (0) each row stores
// its id at its key (one store survives); (1) a row whose store did not
// marks itself, clears its key's count and raises min_head[3]; (2) the
// marked rows chain in by exchange (a repeated key) and count its rows --
// skipped by every wave when min_head[3] stayed 0 (unique keys).  Keys
// outside [base, base + range) contradict the zone maps the range came
// from: flagged.
constexpr uint32_t kLostRow = 0xFFFFFFFEu;
__device__ __forceinline__ void direct_pass(const LookupBufs& lk, int64_t key, uint32_t r, int pass) {
  const uint64_t i = (uint64_t)key - (uint64_t)lk.base;
  if (i >= lk.range) {
    if (pass == 0) atomicOr(&lk.min_head[1], 2u);
    return;
  }
  if (pass == 0) {
    lk.head[i] = r;
    lk.next[r] = 0xFFFFFFFFu;
  } else if (pass == 1) {
    if (lk.head[i] != r) {
      lk.next[r] = kLostRow;
      lk.cnt[i] = 0;
      lk.min_head[3] = 1u;
    }
  } else if (lk.next[r] == kLostRow) {
    lk.next[r] = atomicExch(&lk.head[i], r);
    atomicMax(&lk.min_head[2], atomicAdd(&lk.cnt[i], 1u) + 2u);
  }
}

// One tile per wave (grid = ntiles), as k_filter_bits; `step` is
// direct_pass's pass for a direct table (a hash table takes step 0 only).
template <int U>
__global__ __launch_bounds__(kWave) void k_lookup_build(const QDesc d, int32_t key_slot, const LookupBufs lk,
                                                        int32_t step) {
  if (step == 2 && lk.min_head[3] == 0) return;  // no key repeated
  const int lane = threadIdx.x;
  const int NS = d.nslots;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  const int64_t t0 = it.t;
  while (it.have && it.t == t0) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(ld, NS, it.base, it.rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, it.base, it.rend, lane, raw, pass, g);
    int64_t key[U];
    pick<U>(raw, key_slot, key);
#pragma unroll
    for (int u = 0; u < U; u++)
      if (pass[u]) {
        const uint32_t r = (uint32_t)(it.base + u * kWave + lane);
        if (lk.range) direct_pass(lk, key[u], r, step);
        else lookup_insert(lk, key[u], r);
      }
    it.base += kWave * U;
    if (it.base >= it.rend) it.have = false;
  }
}

hipError_t launch_lookup_build(const QDesc& d, int32_t key_slot, const LookupBufs& lk, hipStream_t s) {
  const uint32_t n = lk.range ? lk.range : lk.mask + 1;
  const uint32_t blocks = std::min<uint32_t>(2048, (n + 255) / 256);
  hipLaunchKernelGGL(k_lookup_init, dim3(blocks), dim3(256), 0, s, lk);
  for (int pass = 0; d.ntiles > 0 && pass < (lk.range ? 3 : 1); pass++)
    hipLaunchKernelGGL((k_lookup_build<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, key_slot, lk, pass);
  return hipGetLastError();
}

// ---- the direct join's build side ------------------------------------------
// Build keys bounded by their zone maps to [base, base + range).  k_dj_mark:
// every row that passes the build predicate stores a byte at key - base (no
// atomics: stores of one value to one byte agree) and writes its tile's
// passing rows to bcnt.  k_dj_pack: thread w packs bytes [32w, 32w + 32) into
// bit word w and clears them for the next query; a block writes its bits'
// count (distinct keys: fewer than the passing rows means a key repeats) to
// pcnt.
template <int U>
__global__ __launch_bounds__(kWave) void k_dj_mark(const QDesc d, int32_t key_slot, const DirectBufs D) {
  const int lane = threadIdx.x;
  const int NS = d.nslots;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  const int64_t t0 = it.t;
  uint32_t npass = 0, err = 0;
  while (it.have && it.t == t0) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(ld, NS, it.base, it.rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, it.base, it.rend, lane, raw, pass, g);
    int64_t key[U];
    pick<U>(raw, key_slot, key);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!pass[u]) continue;
      const uint64_t off = (uint64_t)key[u] - (uint64_t)D.base;
      if (off >= D.range) {
        err = 1u;
        continue;
      }
      npass++;
      D.mark[off] = 1;
    }
    it.base += kWave * U;
    if (it.base >= it.rend) it.have = false;
  }
  for (int o = 32; o > 0; o >>= 1) {
    npass += (uint32_t)__shfl_xor((int)npass, o, 64);
    err |= (uint32_t)__shfl_xor((int)err, o, 64);
  }
  if (lane == 0) {
    D.bcnt[blockIdx.x] = npass;
    if (err) atomicOr(&D.rflags[0], err);  // (corrupt zone maps only)
  }
}

__global__ __launch_bounds__(256) void k_dj_pack(const DirectBufs D) {
  __shared__ uint32_t s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  uint32_t bits = 0;
  if (w < D.words) {
    uint4* m = (uint4*)(D.mark + w * 32);
    const uint4 a = m[0], b = m[1];
    const uint32_t x[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t v = x[i];
      bits |= ((v & 1u) | ((v >> 7) & 2u) | ((v >> 14) & 4u) | ((v >> 21) & 8u)) << (4 * i);
    }
    D.pres[w] = bits;
    if (bits) {
      m[0] = make_uint4(0, 0, 0, 0);
      m[1] = make_uint4(0, 0, 0, 0);
    }
  }
  uint32_t n = (uint32_t)__popc(bits);
  for (int o = 32; o > 0; o >>= 1) n += (uint32_t)__shfl_xor((int)n, o, 64);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0) D.pcnt[blockIdx.x] = s_n;
}

__global__ __launch_bounds__(256) void k_zero_words(uint64_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0;
}

hipError_t launch_zero_words(uint64_t* p, uint64_t words, hipStream_t s) {
  const uint64_t blocks = std::min<uint64_t>(4096, std::max<uint64_t>(1, (words + 255) / 256));
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, s, p, words);
  return hipGetLastError();
}

// Ranks of a dictionary column's rows: out[r] = rank[code[r]], the position
// of the row's string in byte order (MIN/MAX over strings fold ranks).  The
// codes were checked against the dictionary at upload, so every code
// indexes rank.
template <typename C>
__global__ __launch_bounds__(256) void k_rank_map(const C* __restrict__ codes, const uint16_t* __restrict__ rank,
                                                   uint16_t* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    out[i] = rank[codes[i]];
}

hipError_t launch_rank_map(const void* codes, int32_t type, const uint16_t* rank, uint16_t* out, int64_t n,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>(4096, (n + 255) / 256);
  if (type == VQ_U8)
    hipLaunchKernelGGL(k_rank_map<uint8_t>, dim3((unsigned)blocks), dim3(256), 0, s, (const uint8_t*)codes, rank,
                       out, n);
  else
    hipLaunchKernelGGL(k_rank_map<uint16_t>, dim3((unsigned)blocks), dim3(256), 0, s, (const uint16_t*)codes, rank,
                       out, n);
  return hipGetLastError();
}

hipError_t launch_dj_build(const QDesc& d, int32_t key_slot, const DirectBufs& D, hipStream_t s) {
  if (d.ntiles > 0)
    hipLaunchKernelGGL((k_dj_mark<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, key_slot, D);
  const uint64_t blocks = std::max<uint64_t>(1, (D.words + 255) / 256);
  hipLaunchKernelGGL(k_dj_pack, dim3((unsigned)blocks), dim3(256), 0, s, D);
  return hipGetLastError();
}

// ---- a join whose build side is itself a join (the nested build) ------------
// The inner side's rows are in a lookup table (launch_lookup_build); a build
// row that passes its predicate and meets m inner rows is m rows of the
// nested join's output (HashJoinShared's rows, exec/parallel_join.go:
// 389-552), which become the outer join's build side, kept in row order.

__global__ __launch_bounds__(64) void k_zero_u64(uint64_t* __restrict__ p, int64_t n) {
  for (int64_t i = threadIdx.x; i < n; i += 64) p[i] = 0;
}

__global__ __launch_bounds__(64) void k_zero_u32_1(uint32_t* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0;
}

// The first row chained under key (0xFFFFFFFF: none).
__device__ __forceinline__ uint32_t lookup_first(const LookupBufs& lk, int64_t key) {
  if (lk.range) {
    const uint64_t i = (uint64_t)key - (uint64_t)lk.base;
    return i < lk.range ? lk.head[i] : 0xFFFFFFFFu;
  }
  if (lk.packed) {
    if (key < 0 || key >= 0xFFFFFFFFll) return 0xFFFFFFFFu;
    const uint64_t kk = (uint64_t)key << 32;
    uint32_t i = (uint32_t)(radix_hash((uint64_t)key) >> 32) & lk.mask;
    for (uint32_t pr = 0; pr <= lk.mask; pr++) {
      const uint64_t w = ((const uint64_t*)lk.keys)[i];
      if ((w & 0xFFFFFFFF00000000ull) == kk) return (uint32_t)w;
      if (w == ~0ull) break;
      i = (i + 1) & lk.mask;
    }
    return 0xFFFFFFFFu;
  }
  if (key == INT64_MIN) return lk.min_head[0];
  uint32_t i = (uint32_t)(radix_hash((uint64_t)key) >> 32) & lk.mask;
  for (uint32_t pr = 0; pr <= lk.mask; pr++) {
    const int64_t k = lk.keys[i];
    if (k == key) return lk.head[i];
    if (k == INT64_MIN) break;
    i = (i + 1) & lk.mask;
  }
  return 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t chain_len(const LookupBufs& lk, int64_t key) {
  uint32_t m = 0;
  for (uint32_t r = lookup_first(lk, key); r != 0xFFFFFFFFu; r = lk.next[r]) m++;
  return m;
}

// Rows of each tile (one tile per wave, as k_filter_bits): the sum over its
// passing rows of the inner rows each meets.
template <int U>
__global__ __launch_bounds__(kWave) void k_semi_count(const QDesc d, int32_t fk_slot, const LookupBufs lk,
                                                      uint32_t* __restrict__ tile_counts) {
  const int lane = threadIdx.x;
  const int NS = d.nslots;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  const int64_t t0 = it.t;
  uint32_t count = 0;
  while (it.have && it.t == t0) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(ld, NS, it.base, it.rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, it.base, it.rend, lane, raw, pass, g);
    int64_t fk[U];
    pick<U>(raw, fk_slot, fk);
#pragma unroll
    for (int u = 0; u < U; u++)
      if (pass[u]) count += chain_len(lk, fk[u]);
    it.base += kWave * U;
    if (it.base >= it.rend) it.have = false;
  }
  for (int off = 32; off >= 1; off >>= 1) count += (uint32_t)__shfl_xor((int)count, off);
  if (lane == 0) tile_counts[t0] = count;
}

__device__ __forceinline__ void store_as(void* dst, int32_t type, int64_t i, int64_t v) {
  switch (type) {
    case VQ_U8: ((uint8_t*)dst)[i] = (uint8_t)v; break;
    case VQ_U16: ((uint16_t*)dst)[i] = (uint16_t)v; break;
    case VQ_DATE: ((int32_t*)dst)[i] = (int32_t)v; break;
    default: ((int64_t*)dst)[i] = v; break;
  }
}

// Each passing row's m copies at offsets[tile] + (the copies of the tile's
// rows before it): the output in row order.
template <int U>
__global__ __launch_bounds__(kWave) void k_semi_gather(const QDesc d, int32_t fk_slot, const LookupBufs lk,
                                                       const uint32_t* __restrict__ offsets, const SemiOut o) {
  const int lane = threadIdx.x;
  const int NS = d.nslots;
  SlotLd ld[kMaxSlots];
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) ld[c] = slot_ld(d.slot[c < NS ? c : 0]);
  TileIt it;
  tile_begin(d, blockIdx.x, it);
  const int64_t t0 = it.t;
  int64_t pos = offsets[t0];
  while (it.have && it.t == t0) {
    int64_t raw[kMaxSlots][U];
    load_rows<U>(ld, NS, it.base, it.rend, lane, raw);
    bool pass[U];
    int g[U];
    eval_pred<U>(d.slot, NS, it.base, it.rend, lane, raw, pass, g);
    int64_t fk[U];
    pick<U>(raw, fk_slot, fk);
#pragma unroll
    for (int u = 0; u < U; u++) {  // rows base + u * 64 + lane: row order is (u, lane)
      const uint32_t m = pass[u] ? chain_len(lk, fk[u]) : 0u;
      uint32_t incl = m;  // inclusive scan over the wave's lanes
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += y;
      }
      const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1);
      const int64_t at = pos + incl - m;
      for (uint32_t k = 0; k < m; k++)
        for (int c = 0; c < o.ncols; c++) {
          int64_t v[U];
          pick<U>(raw, o.slot[c], v);
          store_as(o.dst[c], o.type[c], at + k, v[u]);
        }
      pos += total;
    }
    it.base += kWave * U;
    if (it.base >= it.rend) it.have = false;
  }
}

// The output groups' build-side columns: keys[i] -> group i in a lookup
// table, then every row of the nested output writes its columns to its
// group's (pass 0), and checks them (pass 1: rows of one key that disagree
// would be separate groups of GROUP BY key, columns -- flagged).
__global__ __launch_bounds__(256) void k_keys_insert(const int64_t* __restrict__ keys, int64_t n,
                                                     const LookupBufs lk, int32_t pass) {
  if (lk.range && pass == 2 && lk.min_head[3] == 0) return;  // no key repeated
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (lk.range) direct_pass(lk, keys[i], (uint32_t)i, pass);
    else lookup_insert(lk, keys[i], (uint32_t)i);
  }
}

__device__ __forceinline__ int64_t load_as(const void* p, int32_t type, int64_t i) {
  switch (type) {
    case VQ_U8: return ((const uint8_t*)p)[i];
    case VQ_U16: return ((const uint16_t*)p)[i];
    case VQ_DATE: return ((const int32_t*)p)[i];
    default: return ((const int64_t*)p)[i];
  }
}

__global__ __launch_bounds__(256) void k_pay_gather(const PayCols pc, int64_t n, const LookupBufs lk,
                                                    int64_t* __restrict__ pay, uint32_t* __restrict__ flag,
                                                    int32_t check) {
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    const uint32_t g = lookup_first(lk, load_as(pc.key, pc.key_type, r));
    if (g == 0xFFFFFFFFu) continue;
    for (int c = 0; c < pc.npay; c++) {
      const int64_t v = load_as(pc.src[c], pc.type[c], r);
      if (!check) pay[(int64_t)g * pc.npay + c] = v;
      else if (pay[(int64_t)g * pc.npay + c] != v) atomicOr(flag, 1u);
    }
  }
}

hipError_t launch_semi(const QDesc& d, int32_t fk_slot, const LookupBufs& lk, uint32_t* tile_counts,
                       uint32_t* tile_offsets, uint64_t* out_count, hipStream_t s) {
  if (d.ntiles == 0) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out_count, (int64_t)1);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_semi_count<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, fk_slot, lk, tile_counts);
  hipLaunchKernelGGL(k_filter_scan, dim3(1), dim3(1024), 0, s, tile_counts, d.ntiles, tile_offsets, out_count);
  return hipGetLastError();
}

hipError_t launch_semi_gather(const QDesc& d, int32_t fk_slot, const LookupBufs& lk,
                              const uint32_t* tile_offsets, const SemiOut& o, hipStream_t s) {
  if (d.ntiles > 0)
    hipLaunchKernelGGL((k_semi_gather<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, fk_slot, lk,
                       tile_offsets, o);
  return hipGetLastError();
}

hipError_t launch_pay_gather(const int64_t* keys, int64_t nkeys, const LookupBufs& lk, const PayCols& pc,
                             int64_t nrows, int64_t* pay, uint32_t* flag, hipStream_t s) {
  const uint32_t kb = (uint32_t)std::min<int64_t>(1024, (nkeys + 255) / 256 + 1);
  const uint32_t n = lk.range ? lk.range : lk.mask + 1;
  const uint32_t lb = std::min<uint32_t>(2048, (n + 255) / 256);
  hipLaunchKernelGGL(k_lookup_init, dim3(lb), dim3(256), 0, s, lk);
  for (int pass = 0; pass < (lk.range ? 3 : 1); pass++)
    hipLaunchKernelGGL(k_keys_insert, dim3(kb), dim3(256), 0, s, keys, nkeys, lk, pass);
  const uint32_t rb = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(2048, (nrows + 255) / 256));
  hipLaunchKernelGGL(k_zero_u32_1, dim3(1), dim3(64), 0, s, flag);
  hipLaunchKernelGGL(k_pay_gather, dim3(rb), dim3(256), 0, s, pc, nrows, lk, pay, flag, 0);
  hipLaunchKernelGGL(k_pay_gather, dim3(rb), dim3(256), 0, s, pc, nrows, lk, pay, flag, 1);
  return hipGetLastError();
}

hipError_t launch_filter(const QDesc& d, bool jit, uint64_t* bits, uint32_t* tile_counts,
                         uint32_t* tile_offsets, uint32_t* out_rows, uint64_t* out_count,
                         hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, int32_t repeats) {
  if (d.ntiles == 0) return hipMemsetAsync(out_count, 0, sizeof(uint64_t), s);
  if (ev0) hipEventRecord(ev0, s);
  for (int32_t rep = 0; rep < repeats; rep++) {  // identical bits each time
    if (jit) {
      hipError_t e = jit_launch_filter_bits(d, bits, tile_counts, s);
      if (e != hipSuccess) return e;
    } else {
      hipLaunchKernelGGL((k_filter_bits<kU>), dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, bits,
                         tile_counts);
    }
  }
  if (ev1) hipEventRecord(ev1, s);
  hipLaunchKernelGGL(k_filter_scan, dim3(1), dim3(1024), 0, s, tile_counts, d.ntiles,
                     tile_offsets, out_count);
  hipLaunchKernelGGL(k_filter_write, dim3((unsigned)d.ntiles), dim3(kWave), 0, s, d, bits,
                     tile_offsets, out_rows);
  return hipGetLastError();
}

}  // namespace vq
