// k_join.hip — gfx950 kernels for the radix-partitioned hash join feeding a
// GROUP BY on the join key (ParallelHashJoinAggregate, exec/parallel_join.go).
//
// Pipeline, per side (build, then probe):
//   k_part_count   filter + radixHash + LDS histogram per tile -> counts[t][p]
//   k_part_colscan per partition, exclusive prefix over tiles + total
//   k_part_starts  exclusive scan of the partition totals (one block)
//   k_part_scatter filter again, scatter (key[, row, values]) to its partition
// then one workgroup per partition:
//   k_join_agg     build the partition's open-addressing table in LDS from the
//                  build keys, probe it with the partition's probe rows and
//                  aggregate into per-slot accumulators in LDS, then emit the
//                  partition's groups (or its top-k candidates)
//   k_topk_merge   merge the per-partition candidates (top-k queries)
//
// The partition of a key is the low `bits` of radixHash (radixPart,
// parallel_join.go:104-109) and its slot inside the partition table the high
// bits (slotFor, join_table.go:488-492) — the same split as the reference, so
// a partition's keys spread over its table.
//
// Exactness: float sums are accumulated as fixed-point integers (FixScale),
// integer sums as wrapping int64, MIN/MAX as order-preserving integers, first
// row as a minimum: every update is an associative integer atomic, so the
// result does not depend on which lane or workgroup got there first.
#include "vq_internal.h"

namespace vq {

constexpr int kBlk = 256;
constexpr int kJU = 4;
constexpr int kSlotsPerThread = 8;  // table slots + 1 <= kBlk * kSlotsPerThread
constexpr int64_t kEmptyKey = INT64_MIN;

__device__ __forceinline__ int64_t jload(const QSlot& s, int64_t r) {
  switch (s.type) {
    case VQ_U8: return ((const uint8_t*)s.p)[r];
    case VQ_U16: return ((const uint16_t*)s.p)[r];
    case VQ_DATE: return ((const int32_t*)s.p)[r];
    default: return ((const int64_t*)s.p)[r];
  }
}

__device__ __forceinline__ bool jcmp(const QSlot& s, const QTerm& t, int64_t x) {
  if (s.type == VQ_F64) {
    const double a = bits_to_f64(x), b = t.fval;
    switch (t.op) {
      case VQ_EQ: return a == b;
      case VQ_NE: return a != b;
      case VQ_LT: return a < b;
      case VQ_LE: return a <= b;
      case VQ_GT: return a > b;
      default: return a >= b;
    }
  }
  switch (t.op) {
    case VQ_EQ: return x == t.ival;
    case VQ_NE: return x != t.ival;
    case VQ_LT: return x < t.ival;
    case VQ_LE: return x <= t.ival;
    case VQ_GT: return x > t.ival;
    default: return x >= t.ival;
  }
}

__device__ __forceinline__ double jarith_f(int op, double a, double b) {
  switch (op) {
    case VQ_ADD: return a + b;
    case VQ_SUB: return a - b;
    case VQ_MUL: return a * b;
    default: return b == 0.0 ? __longlong_as_double(0x7ff8000000000000LL) : a / b;
  }
}

__device__ __forceinline__ int64_t jarith_i(int op, int64_t a, int64_t b) {
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

// Row-local value of an aggregate argument (one row, already loaded slots).
__device__ __forceinline__ int64_t slot_val(const int64_t (&raw)[kMaxSlots], int slot) {
  int64_t x = raw[0];
#pragma unroll
  for (int c = 1; c < kMaxSlots; c++)
    if (slot == c) x = raw[c];
  return x;
}

__device__ __forceinline__ uint64_t jexpr(const QExpr& e, const QSlot* slot,
                                          const int64_t (&raw)[kMaxSlots]) {
  if (e.is_float) {
    double v = 0.0;
    for (int k = 0; k < e.nfactors; k++) {
      const QFactor f = e.f[k];
      double x = f.fval;
      if (f.kind != VQ_F_LIT) {
        const int64_t xi = slot_val(raw, f.slot);
        const double c = slot[f.slot].type == VQ_F64 ? bits_to_f64(xi) : (double)xi;
        x = f.kind == VQ_F_COL ? c
                               : (f.kind == VQ_F_LIT_OP_COL ? jarith_f(f.op, f.fval, c)
                                                            : jarith_f(f.op, c, f.fval));
      }
      v = k == 0 ? x : jarith_f(f.combine, v, x);
    }
    return (uint64_t)f64_to_bits(v);
  }
  int64_t v = 0;
  for (int k = 0; k < e.nfactors; k++) {
    const QFactor f = e.f[k];
    int64_t x = f.ival;
    if (f.kind != VQ_F_LIT) {
      const int64_t c = slot_val(raw, f.slot);
      x = f.kind == VQ_F_COL ? c
                             : (f.kind == VQ_F_LIT_OP_COL ? jarith_i(f.op, f.ival, c)
                                                          : jarith_i(f.op, c, f.ival));
    }
    v = k == 0 ? x : jarith_i(f.combine, v, x);
  }
  return (uint64_t)v;
}

struct TileRange {
  int64_t row0, rend;
};

__device__ __forceinline__ TileRange jtile(const JSide& sd, int64_t t) {
  int s = 0;
  while (s + 1 < sd.nspans && t >= sd.span[s + 1].tile0) s++;
  TileRange r;
  r.row0 = sd.span[s].start + (t - sd.span[s].tile0) * (int64_t)sd.tile_rows;
  const int64_t send = sd.span[s].start + sd.span[s].rows;
  r.rend = r.row0 + sd.tile_rows < send ? r.row0 + sd.tile_rows : send;
  return r;
}

// Loads one row's slots; returns whether it passes the side's predicate.
__device__ __forceinline__ bool jrow(const JSide& sd, int64_t r, int64_t (&raw)[kMaxSlots]) {
  bool pass = true;
#pragma unroll
  for (int c = 0; c < kMaxSlots; c++) {
    if (c < sd.nslots) {
      raw[c] = jload(sd.slot[c], r);
      for (int j = 0; j < sd.slot[c].nterms; j++) pass = pass && jcmp(sd.slot[c], sd.slot[c].terms[j], raw[c]);
    } else {
      raw[c] = 0;
    }
  }
  return pass;
}

// counts[t * P + p] = rows of tile t that pass and fall in partition p.
__global__ __launch_bounds__(kBlk) void k_part_count(const JSide sd, int32_t P,
                                                     uint32_t* __restrict__ counts) {
  extern __shared__ uint32_t hist[];
  const int64_t t = blockIdx.x;
  for (int p = threadIdx.x; p < P; p += kBlk) hist[p] = 0;
  __syncthreads();
  const TileRange tr = jtile(sd, t);
  const uint64_t mask = (uint64_t)P - 1;
  for (int64_t base = tr.row0; base < tr.rend; base += kBlk * kJU) {
#pragma unroll
    for (int u = 0; u < kJU; u++) {
      const int64_t r = base + u * kBlk + threadIdx.x;
      if (r < tr.rend) {
        int64_t raw[kMaxSlots];
        if (jrow(sd, r, raw)) {
          const uint64_t h = radix_hash((uint64_t)slot_val(raw, sd.key_slot));
          atomicAdd(&hist[h & mask], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += kBlk) counts[t * P + p] = hist[p];
}

// Per partition p: offsets[t*P+p] = sum of counts[t'*P+p] for t' < t; total[p].
__global__ __launch_bounds__(kBlk) void k_part_colscan(const uint32_t* __restrict__ counts,
                                                       int64_t ntiles, int32_t P,
                                                       uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ totals) {
  const int p = blockIdx.x * kBlk + threadIdx.x;
  if (p >= P) return;
  uint32_t run = 0;
  for (int64_t t = 0; t < ntiles; t++) {
    const uint32_t c = counts[t * P + p];
    offsets[t * P + p] = run;
    run += c;
  }
  totals[p] = run;
}

// pstart[p] = exclusive prefix of totals; pstart[P] = grand total.  One block.
__global__ __launch_bounds__(1024) void k_part_starts(const uint32_t* __restrict__ totals,
                                                      int32_t P, uint32_t* __restrict__ pstart) {
  __shared__ uint32_t s[1024];
  const int tid = threadIdx.x;
  const int per = (P + 1023) / 1024;
  const int lo = tid * per, hi = lo + per < P ? lo + per : P;
  uint32_t sum = 0;
  for (int i = lo; i < hi; i++) sum += totals[i];
  s[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t x = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  uint32_t run = s[tid] - sum;
  for (int i = lo; i < hi; i++) {
    pstart[i] = run;
    run += totals[i];
  }
  if (tid == 1023) pstart[P] = s[1023];
}

// Scatter each passing row to its partition: key (both sides), row id and
// argument values (probe side).  Position inside a (tile, partition) segment
// comes from an LDS counter.
__global__ __launch_bounds__(kBlk) void k_part_scatter(const JSide sd, const JDesc jd, int32_t P,
                                                       int32_t with_vals,
                                                       const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ pstart,
                                                       int64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ rows,
                                                       uint64_t* __restrict__ vals, int64_t vstride) {
  extern __shared__ uint32_t lcnt[];
  const int64_t t = blockIdx.x;
  for (int p = threadIdx.x; p < P; p += kBlk) lcnt[p] = pstart[p] + offsets[t * P + p];
  __syncthreads();
  const TileRange tr = jtile(sd, t);
  const uint64_t mask = (uint64_t)P - 1;
  for (int64_t base = tr.row0; base < tr.rend; base += kBlk * kJU) {
#pragma unroll
    for (int u = 0; u < kJU; u++) {
      const int64_t r = base + u * kBlk + threadIdx.x;
      if (r < tr.rend) {
        int64_t raw[kMaxSlots];
        if (jrow(sd, r, raw)) {
          const int64_t key = slot_val(raw, sd.key_slot);
          const uint64_t h = radix_hash((uint64_t)key);
          const uint32_t pos = atomicAdd(&lcnt[h & mask], 1u);
          keys[pos] = key;
          if (with_vals) {
            rows[pos] = (uint32_t)r;
            int vi = 0;
            for (int j = 0; j < jd.naggs; j++) {
              if (jd.agg[j].e.nfactors > 0) {
                vals[vi * vstride + pos] = jexpr(jd.agg[j].e, sd.slot, raw);
                vi++;
              }
            }
          }
        }
      }
    }
  }
}

// ---- per-partition table ----------------------------------------------------

// float64 -> order-preserving int64 (total order on non-NaN values).
__device__ __forceinline__ int64_t f64_ord(double d) {
  const int64_t b = f64_to_bits(d);
  return b >= 0 ? b : (int64_t)((uint64_t)b ^ 0x7FFFFFFFFFFFFFFFULL);
}
__device__ __forceinline__ double ord_f64(int64_t o) {
  return bits_to_f64(o >= 0 ? o : (int64_t)((uint64_t)o ^ 0x7FFFFFFFFFFFFFFFULL));
}

// Add float64 t into three fixed-point limbs (see FixScale).
__device__ __forceinline__ void fix_add(unsigned long long* limb, double t, FixScale fs) {
  const uint64_t b = (uint64_t)f64_to_bits(t);
  const int ex = (int)((b >> 52) & 0x7FF);
  uint64_t m = b & ((1ULL << 52) - 1);
  if (ex == 0 && m == 0) return;
  int e;
  if (ex == 0) {
    e = -1074;
  } else {
    m |= 1ULL << 52;
    e = ex - 1075;
  }
  const int sh = e - fs.lsb;  // T = m * 2^sh
  uint64_t lo, hi;
  if (sh >= 0) {
    if (sh >= 128) return;  // outside the planned range (cannot happen for bounded data)
    lo = sh >= 64 ? 0 : m << sh;
    hi = sh == 0 ? 0 : (sh >= 64 ? m << (sh - 64) : m >> (64 - sh));
  } else {
    if (-sh >= 64) return;
    lo = m >> (-sh);
    hi = 0;
  }
  const int w = fs.w;
  const uint64_t wm = (1ULL << w) - 1;
  const uint64_t c0 = lo & wm;
  const uint64_t c1 = ((lo >> w) | (w == 0 ? 0 : hi << (64 - w))) & wm;
  const uint64_t c2 = ((2 * w >= 64 ? hi >> (2 * w - 64) : (lo >> (2 * w)) | (hi << (64 - 2 * w)))) & wm;
  const bool neg = (b >> 63) != 0;
  atomicAdd(&limb[0], neg ? (unsigned long long)(-(int64_t)c0) : (unsigned long long)c0);
  atomicAdd(&limb[1], neg ? (unsigned long long)(-(int64_t)c1) : (unsigned long long)c1);
  atomicAdd(&limb[2], neg ? (unsigned long long)(-(int64_t)c2) : (unsigned long long)c2);
}

struct Cand {
  int64_t v;       // sort value (aggregate 0, as the sort sees it)
  uint32_t first;  // first-occurrence row (tie break)
  uint32_t slot;   // LDS slot (or global record index)
};

__device__ __forceinline__ bool cand_better(const Cand& a, const Cand& b) {
  return a.v > b.v || (a.v == b.v && a.first < b.first);
}

__device__ __forceinline__ Cand cand_none() {
  Cand c;
  c.v = INT64_MIN;
  c.first = 0xFFFFFFFFu;
  c.slot = 0xFFFFFFFFu;
  return c;
}

__device__ __forceinline__ Cand shfl_cand(const Cand& c, int src) {
  Cand o;
  o.v = __shfl((long long)c.v, src, 64);
  o.first = (uint32_t)__shfl((int)c.first, src, 64);
  o.slot = (uint32_t)__shfl((int)c.slot, src, 64);
  return o;
}

// Best candidate of the block (ties impossible: first rows are distinct).
__device__ Cand block_best(Cand c, Cand* s_c) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    Cand o = shfl_cand(c, lane ^ m);
    if (cand_better(o, c)) c = o;
  }
  if (lane == 0) s_c[wv] = c;
  __syncthreads();
  Cand b = s_c[0];
  for (int i = 1; i < kBlk / 64; i++)
    if (cand_better(s_c[i], b)) b = s_c[i];
  __syncthreads();
  return b;
}

// Record layout (64-bit words): key, count, first row, then one word per
// aggregate (int64 or float64 bits).
__device__ __forceinline__ void emit_record(const JDesc& jd, uint64_t* rec, int64_t key,
                                            uint64_t cnt, uint32_t mult, uint32_t first,
                                            const unsigned long long* acc) {
  rec[0] = (uint64_t)key;
  rec[1] = cnt * mult;
  rec[2] = first;
  for (int j = 0; j < jd.naggs; j++) {
    const JAgg& a = jd.agg[j];
    const unsigned long long* w = acc + a.word0;
    uint64_t out = 0;
    switch (a.kind) {
      case VQ_AGG_COUNT: out = cnt * mult; break;
      case VQ_AGG_SUM:
        if (a.accum == VQ_F64) out = (uint64_t)f64_to_bits(fix_to_f64((const int64_t*)w, 3, a.fs, mult));
        else out = (uint64_t)w[0] * mult;
        break;
      case VQ_AGG_AVG: {
        double s = a.e.is_float ? fix_to_f64((const int64_t*)w, 3, a.fs, mult)
                                : (double)(int64_t)((uint64_t)w[0] * mult);
        out = (uint64_t)f64_to_bits(s / (double)(cnt * mult));
        break;
      }
      default:  // MIN / MAX
        out = a.accum == VQ_F64 ? (uint64_t)f64_to_bits(ord_f64((int64_t)w[0])) : (uint64_t)w[0];
        break;
    }
    rec[3 + j] = out;
  }
}

// ExternalSort's key for aggregate 0 (exec/sort.go:130-163): the int64 the
// row carries, i.e. a float's raw IEEE bits.
__device__ __forceinline__ int64_t sort_key(const JDesc& jd, const uint64_t* rec) {
  return jd.naggs > 0 ? (int64_t)rec[3] : 0;
}

// One workgroup per partition.  LDS: keys[cap+1], mult[cap+1], cnt[cap+1],
// first[cap+1], acc[(cap+1) * nwords]; slot `cap` holds the key INT64_MIN,
// which doubles as the empty marker of the other slots.
__global__ __launch_bounds__(kBlk) void k_join_agg(const JDesc jd, const JoinPlan plan,
                                                   const int64_t* __restrict__ b_keys,
                                                   const uint32_t* __restrict__ b_pstart,
                                                   const int64_t* __restrict__ p_keys,
                                                   const uint32_t* __restrict__ p_rows,
                                                   const uint64_t* __restrict__ p_vals,
                                                   int64_t vstride,
                                                   const uint32_t* __restrict__ p_pstart,
                                                   uint64_t* __restrict__ cand_out,
                                                   uint64_t* __restrict__ full_out,
                                                   uint32_t* __restrict__ full_count,
                                                   uint32_t* __restrict__ flags) {
  extern __shared__ unsigned long long smem[];
  __shared__ Cand s_c[kBlk / 64];
  __shared__ uint32_t s_fail;
  const int p = blockIdx.x;
  const int cap = plan.cap;
  const int ns = cap + 1;
  const int NW = jd.nwords;
  unsigned long long* keys = smem;
  unsigned long long* acc = keys + ns;
  uint32_t* mult = (uint32_t*)(acc + (size_t)ns * NW);
  uint32_t* cnt = mult + ns;
  uint32_t* first = cnt + ns;
  const int lg = __ffs(cap) - 1;  // log2(cap)
  if (threadIdx.x == 0) s_fail = 0;
  for (int i = threadIdx.x; i < ns; i += kBlk) {
    keys[i] = (unsigned long long)kEmptyKey;
    mult[i] = 0;
    cnt[i] = 0;
    first[i] = 0xFFFFFFFFu;
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      for (int k = 0; k < a.nwords; k++) {
        unsigned long long init = 0;
        if (a.kind == VQ_AGG_MIN) init = (unsigned long long)INT64_MAX;
        if (a.kind == VQ_AGG_MAX) init = (unsigned long long)INT64_MIN;
        acc[(size_t)i * NW + a.word0 + k] = init;
      }
    }
  }
  __syncthreads();

  // build: insert the partition's build keys (multiplicity = rows per key)
  const uint32_t b0 = b_pstart[p], b1 = b_pstart[p + 1];
  for (uint32_t e = b0 + threadIdx.x; e < b1; e += kBlk) {
    const int64_t key = b_keys[e];
    if (key == kEmptyKey) {
      atomicAdd(&mult[cap], 1u);
      continue;
    }
    const uint64_t h = radix_hash((uint64_t)key);
    uint32_t s = lg ? (uint32_t)(h >> (64 - lg)) : 0;
    bool done = false;
    for (int probe = 0; probe < cap; probe++) {
      const unsigned long long old =
          atomicCAS(&keys[s], (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (old == (unsigned long long)kEmptyKey || old == (unsigned long long)key) {
        atomicAdd(&mult[s], 1u);
        done = true;
        break;
      }
      s = (s + 1) & (uint32_t)(cap - 1);
    }
    if (!done) s_fail = 1;
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) atomicOr(&flags[0], 1u);
    return;
  }

  // probe + aggregate
  const uint32_t q0 = p_pstart[p], q1 = p_pstart[p + 1];
  for (uint32_t e = q0 + threadIdx.x; e < q1; e += kBlk) {
    const int64_t key = p_keys[e];
    int s = -1;
    if (key == kEmptyKey) {
      s = mult[cap] ? cap : -1;
    } else {
      const uint64_t h = radix_hash((uint64_t)key);
      uint32_t i = lg ? (uint32_t)(h >> (64 - lg)) : 0;
      for (int probe = 0; probe < cap; probe++) {
        const int64_t k = (int64_t)keys[i];
        if (k == kEmptyKey) break;
        if (k == key) {
          s = (int)i;
          break;
        }
        i = (i + 1) & (uint32_t)(cap - 1);
      }
    }
    if (s < 0) continue;
    const uint32_t row = p_rows[e];
    atomicAdd(&cnt[s], 1u);
    atomicMin(&first[s], row);
    int vi = 0;
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      if (a.e.nfactors == 0) continue;
      const uint64_t v = p_vals[vi * vstride + e];
      vi++;
      unsigned long long* w = &acc[(size_t)s * NW + a.word0];
      switch (a.kind) {
        case VQ_AGG_SUM:
        case VQ_AGG_AVG:
          if (a.e.is_float) fix_add(w, bits_to_f64((int64_t)v), a.fs);
          else atomicAdd(w, (unsigned long long)v);
          break;
        case VQ_AGG_MIN: {
          const int64_t o = a.accum == VQ_F64 ? f64_ord(bits_to_f64((int64_t)v)) : (int64_t)v;
          if (!(a.accum == VQ_F64 && bits_to_f64((int64_t)v) != bits_to_f64((int64_t)v)))
            atomicMin((long long*)w, (long long)o);
          break;
        }
        case VQ_AGG_MAX: {
          const int64_t o = a.accum == VQ_F64 ? f64_ord(bits_to_f64((int64_t)v)) : (int64_t)v;
          if (!(a.accum == VQ_F64 && bits_to_f64((int64_t)v) != bits_to_f64((int64_t)v)))
            atomicMax((long long*)w, (long long)o);
          break;
        }
        default:
          break;
      }
    }
  }
  __syncthreads();

  const int RW = 3 + jd.naggs;
  if (plan.topk <= 0) {
    // full output: every matched slot, appended through a global counter
    for (int i = threadIdx.x; i < ns; i += kBlk) {
      if (cnt[i] == 0) continue;
      const int64_t key = i == cap ? kEmptyKey : (int64_t)keys[i];
      const uint32_t at = atomicAdd(full_count, 1u);
      emit_record(jd, full_out + (size_t)at * RW, key, cnt[i], mult[i], first[i],
                  &acc[(size_t)i * NW]);
    }
    return;
  }

  // top-k candidates of this partition: each thread ranks its own slots once
  // (at most kSlotsPerThread of them), then k rounds of a block-wide best pick.
  uint64_t* my_out = cand_out + (size_t)p * plan.topk * RW;
  uint64_t recbuf[3 + VQ_MAX_AGGS];
  Cand mine[kSlotsPerThread];
#pragma unroll
  for (int k = 0; k < kSlotsPerThread; k++) {
    const int i = threadIdx.x + k * kBlk;
    mine[k] = cand_none();
    if (i < ns && cnt[i] != 0) {
      const int64_t key = i == cap ? kEmptyKey : (int64_t)keys[i];
      emit_record(jd, recbuf, key, cnt[i], mult[i], first[i], &acc[(size_t)i * NW]);
      mine[k].v = sort_key(jd, recbuf);
      mine[k].first = first[i];
      mine[k].slot = (uint32_t)i;
    }
  }
  for (int round = 0; round < plan.topk; round++) {
    Cand best = cand_none();
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; k++)
      if (cand_better(mine[k], best)) best = mine[k];
    const Cand b = block_best(best, s_c);
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; k++)
      if (mine[k].slot == b.slot) mine[k] = cand_none();
    if (threadIdx.x == 0) {
      uint64_t* rec = my_out + (size_t)round * RW;
      if (b.slot == 0xFFFFFFFFu) {
        for (int k = 0; k < RW; k++) rec[k] = 0;
        rec[2] = 0xFFFFFFFFFFFFFFFFULL;  // no candidate
      } else {
        const int i = (int)b.slot;
        const int64_t key = i == cap ? kEmptyKey : (int64_t)keys[i];
        emit_record(jd, rec, key, cnt[i], mult[i], first[i], &acc[(size_t)i * NW]);
      }
    }
  }
}

// Merge P * topk candidate records into the global top-k (one block).
__global__ __launch_bounds__(kBlk) void k_topk_merge(const JDesc jd, int32_t P, int32_t topk,
                                                     uint64_t* __restrict__ cand,
                                                     uint64_t* __restrict__ out,
                                                     uint32_t* __restrict__ out_n) {
  __shared__ Cand s_c[kBlk / 64];
  const int RW = 3 + jd.naggs;
  const int64_t n = (int64_t)P * topk;
  int got = 0;
  for (int round = 0; round < topk; round++) {
    Cand best = cand_none();
    for (int64_t i = threadIdx.x; i < n; i += kBlk) {
      const uint64_t* rec = cand + i * RW;
      if (rec[2] == 0xFFFFFFFFFFFFFFFFULL) continue;
      Cand c;
      c.v = sort_key(jd, rec);
      c.first = (uint32_t)rec[2];
      c.slot = (uint32_t)i;
      if (cand_better(c, best)) best = c;
    }
    const Cand b = block_best(best, s_c);
    if (b.slot == 0xFFFFFFFFu) break;
    if (threadIdx.x == 0) {
      uint64_t* rec = cand + (size_t)b.slot * RW;
      for (int k = 0; k < RW; k++) out[(size_t)round * RW + k] = rec[k];
      rec[2] = 0xFFFFFFFFFFFFFFFFULL;
    }
    got++;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_n = (uint32_t)got;
}

static size_t join_lds_bytes(const JDesc& jd, const JoinPlan& plan) {
  const size_t ns = (size_t)plan.cap + 1;
  return ns * 8 + ns * jd.nwords * 8 + ns * 4 * 3;
}

hipError_t launch_join(const JSide& b, const JSide& p, const JDesc& jd, const JoinPlan& plan,
                       const JoinBufs& B, hipStream_t s, hipEvent_t* ev) {
  const int P = plan.P;
  const size_t hist_bytes = (size_t)P * 4;
  const unsigned cs_blocks = (unsigned)((P + kBlk - 1) / kBlk);
  if (ev) hipEventRecord(ev[0], s);
  hipMemsetAsync(B.flags, 0, sizeof(uint32_t) * 2, s);
  // build side
  if (b.ntiles > 0) {
    hipLaunchKernelGGL(k_part_count, dim3((unsigned)b.ntiles), dim3(kBlk), hist_bytes, s, b, P,
                       B.b_counts);
    hipLaunchKernelGGL(k_part_colscan, dim3(cs_blocks), dim3(kBlk), 0, s, B.b_counts, b.ntiles, P,
                       B.b_offsets, B.b_pstart + P + 1);
    hipLaunchKernelGGL(k_part_starts, dim3(1), dim3(1024), 0, s, B.b_pstart + P + 1, P, B.b_pstart);
    hipLaunchKernelGGL(k_part_scatter, dim3((unsigned)b.ntiles), dim3(kBlk), hist_bytes, s, b, jd,
                       P, 0, B.b_offsets, B.b_pstart, B.b_keys, (uint32_t*)nullptr,
                       (uint64_t*)nullptr, (int64_t)0);
  } else {
    hipMemsetAsync(B.b_pstart, 0, sizeof(uint32_t) * (P + 1), s);
  }
  // probe side
  if (p.ntiles > 0) {
    hipLaunchKernelGGL(k_part_count, dim3((unsigned)p.ntiles), dim3(kBlk), hist_bytes, s, p, P,
                       B.p_counts);
    hipLaunchKernelGGL(k_part_colscan, dim3(cs_blocks), dim3(kBlk), 0, s, B.p_counts, p.ntiles, P,
                       B.p_offsets, B.p_pstart + P + 1);
    hipLaunchKernelGGL(k_part_starts, dim3(1), dim3(1024), 0, s, B.p_pstart + P + 1, P, B.p_pstart);
    hipLaunchKernelGGL(k_part_scatter, dim3((unsigned)p.ntiles), dim3(kBlk), hist_bytes, s, p, jd,
                       P, 1, B.p_offsets, B.p_pstart, B.p_keys, B.p_rows, B.p_vals, B.cap_p);
  } else {
    hipMemsetAsync(B.p_pstart, 0, sizeof(uint32_t) * (P + 1), s);
  }
  if (ev) hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_join_agg, dim3((unsigned)P), dim3(kBlk), join_lds_bytes(jd, plan), s, jd,
                     plan, B.b_keys, B.b_pstart, B.p_keys, B.p_rows, B.p_vals, B.cap_p,
                     B.p_pstart, B.cand, B.out, B.flags + 1, B.flags);
  if (ev) hipEventRecord(ev[2], s);
  if (plan.topk > 0)
    hipLaunchKernelGGL(k_topk_merge, dim3(1), dim3(kBlk), 0, s, jd, P, plan.topk, B.cand, B.out,
                       B.flags + 1);
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

}  // namespace vq
