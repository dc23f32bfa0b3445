// k_join.hip — gfx950 kernels for the radix-partitioned hash join feeding a
// GROUP BY on the join key (ParallelHashJoinAggregate, exec/parallel_join.go).
//
// Pipeline, per side (build, then probe):
//   vq_part_count    (run-time compiled, jit_template.hip) filter + radixHash
//                    + LDS histogram per tile -> counts[t][p]
//   k_off_*          counts -> absolute positions (2-D exclusive scan)
//   vq_part_scatter  (run-time compiled) each passing row to its partition:
//                    the key, and on the probe side its row id and the values
//                    of the aggregate arguments
// then:
//   k_join_agg       one workgroup per partition: build the partition's
//                    open-addressing table in LDS from its build keys, probe it
//                    with its probe rows, aggregate into per-slot accumulators
//                    in LDS, emit every matched slot as a group record
//   k_topk           ORDER BY aggregate 0 DESC LIMIT k over the records
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
constexpr int64_t kEmptyKey = INT64_MIN;

// ---- partition offsets: a 2-D exclusive scan over counts[t][p] --------------
// Position of tile t's rows of partition p:
//   pstart[p] + sum_{t' < t} counts[t'][p],  pstart = exclusive scan of totals.
// Tiles are grouped in chunks of kChunk; three coalesced passes:
//   k_off_chunks  csum[c][p] = sum of counts[t][p] over chunk c
//   k_off_starts  cpre[c][p] = pstart[p] + sum of csum[c'][p] for c' < c (one block)
//   k_off_tiles   counts[t][p] <- cpre[c][p] + sum_{t' in c, t' < t} counts[t'][p]
constexpr int kChunk = 32;

__global__ __launch_bounds__(kBlk) void k_off_chunks(const uint32_t* __restrict__ counts,
                                                     int64_t ntiles, int32_t P,
                                                     uint32_t* __restrict__ csum) {
  const int c = blockIdx.x;
  const int p = blockIdx.y * kBlk + threadIdx.x;
  if (p >= P) return;
  const int64_t t0 = (int64_t)c * kChunk;
  const int64_t t1 = t0 + kChunk < ntiles ? t0 + kChunk : ntiles;
  uint32_t s = 0;
  for (int64_t t = t0; t < t1; t++) s += counts[t * P + p];
  csum[(int64_t)c * P + p] = s;
}

// One block: csum[c][p] -> cpre[c][p] (in place) and pstart[0..P].  The sums
// are staged in LDS when they fit, so the per-partition walk over chunks runs
// at LDS latency.
constexpr int kStageMax = 12 * 1024;

__global__ __launch_bounds__(1024) void k_off_starts(uint32_t* __restrict__ csum, int32_t nchunks,
                                                     int32_t P, uint32_t* __restrict__ pstart) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t stage[kStageMax];
  const int tid = threadIdx.x;
  const int64_t n = (int64_t)nchunks * P;
  const bool staged = n <= kStageMax;
  uint32_t* cs = staged ? stage : csum;
  if (staged)
    for (int64_t i = tid; i < n; i += 1024) stage[i] = csum[i];
  __syncthreads();
  const int per = (P + 1023) / 1024;
  const int lo = tid * per, hi = lo + per < P ? lo + per : P;
  uint32_t mine = 0;
  for (int p = lo; p < hi; p++)
    for (int c = 0; c < nchunks; c++) mine += cs[(int64_t)c * P + p];
  s[tid] = mine;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t x = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  uint32_t run = s[tid] - mine;
  for (int p = lo; p < hi; p++) {
    pstart[p] = run;
    for (int c = 0; c < nchunks; c++) {
      const uint32_t v = cs[(int64_t)c * P + p];
      cs[(int64_t)c * P + p] = run;
      run += v;
    }
  }
  if (tid == 1023) pstart[P] = s[1023];
  if (staged) {
    __syncthreads();
    for (int64_t i = tid; i < n; i += 1024) csum[i] = stage[i];
  }
}

__global__ __launch_bounds__(kBlk) void k_off_tiles(uint32_t* __restrict__ counts, int64_t ntiles,
                                                    int32_t P, const uint32_t* __restrict__ cpre) {
  const int c = blockIdx.x;
  const int p = blockIdx.y * kBlk + threadIdx.x;
  if (p >= P) return;
  const int64_t t0 = (int64_t)c * kChunk;
  const int64_t t1 = t0 + kChunk < ntiles ? t0 + kChunk : ntiles;
  uint32_t run = cpre[(int64_t)c * P + p];
  for (int64_t t = t0; t < t1; t++) {
    const uint32_t v = counts[t * P + p];
    counts[t * P + p] = run;
    run += v;
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

// One 1024-thread workgroup per partition.  LDS:
//   table  keys u64[cap+1] | mult u32[cap+1] | pidx u32[cap+1]
//   pool   slot u32[pool] | cnt u32[pool] | first u32[pool] | acc u64[pool*nwords]
// Slot `cap` holds the key INT64_MIN, which doubles as the empty marker of the
// other slots.  Accumulators exist only for build keys a probe row matches: the
// first match of a slot takes the next pool entry (a racing second claimant
// loses the compare-and-swap on pidx, uses the winner's entry and leaves its
// own unused).  flags[0] |= 1 table full, 2 pool full; flags[1] = records;
// flags[2] = distinct build keys.
constexpr int kJoinThreads = 1024;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__global__ __launch_bounds__(kJoinThreads) void k_join_agg(const JDesc jd, const JoinPlan plan,
                                                           const int64_t* __restrict__ b_keys,
                                                           const uint32_t* __restrict__ b_pstart,
                                                           const int64_t* __restrict__ p_keys,
                                                           const uint32_t* __restrict__ p_rows,
                                                           const uint64_t* __restrict__ p_vals,
                                                           int64_t vstride,
                                                           const uint32_t* __restrict__ p_pstart,
                                                           uint64_t* __restrict__ out,
                                                           uint32_t* __restrict__ flags) {
  extern __shared__ unsigned long long smem[];
  __shared__ uint32_t s_fail, s_nrec, s_nkeys, s_base, s_pool;
  const int p = blockIdx.x;
  const int cap = plan.cap, ns = cap + 1, NP = plan.pool;
  const int NW = jd.nwords;
  unsigned long long* keys = smem;
  unsigned long long* acc = keys + ns;                       // [NP * NW]
  uint32_t* mult = (uint32_t*)(acc + (size_t)NP * NW);
  uint32_t* pidx = mult + ns;
  uint32_t* pslot = pidx + ns;
  uint32_t* cnt = pslot + NP;
  uint32_t* first = cnt + NP;
  const int lg = __ffs(cap) - 1;  // log2(cap)
  if (threadIdx.x == 0) {
    s_fail = 0;
    s_nrec = 0;
    s_nkeys = 0;
    s_pool = 0;
  }
  for (int i = threadIdx.x; i < ns; i += kJoinThreads) {
    keys[i] = (unsigned long long)kEmptyKey;
    mult[i] = 0;
    pidx[i] = kNone;
  }
  for (int i = threadIdx.x; i < NP; i += kJoinThreads) {
    cnt[i] = 0;
    first[i] = kNone;
  }
  for (int i = threadIdx.x; i < NP * NW; i += kJoinThreads) {
    const int w = i % NW;
    unsigned long long init = 0;
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      if (w >= a.word0 && w < a.word0 + a.nwords) {
        if (a.kind == VQ_AGG_MIN) init = (unsigned long long)INT64_MAX;
        if (a.kind == VQ_AGG_MAX) init = (unsigned long long)INT64_MIN;
      }
    }
    acc[i] = init;
  }
  __syncthreads();

  // build: the partition's build keys; multiplicity = build rows per key
  const uint32_t b0 = b_pstart[p], b1 = b_pstart[p + 1];
  for (uint32_t e = b0 + threadIdx.x; e < b1; e += kJoinThreads) {
    const int64_t key = b_keys[e];
    if (key == kEmptyKey) {
      if (atomicAdd(&mult[cap], 1u) == 0) atomicAdd(&s_nkeys, 1u);
      continue;
    }
    const uint64_t h = radix_hash((uint64_t)key);
    uint32_t sl = lg ? (uint32_t)(h >> (64 - lg)) : 0;
    bool done = false;
    for (int probe = 0; probe < cap; probe++) {
      const unsigned long long old =
          atomicCAS(&keys[sl], (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (old == (unsigned long long)kEmptyKey || old == (unsigned long long)key) {
        atomicAdd(&mult[sl], 1u);
        if (old == (unsigned long long)kEmptyKey) atomicAdd(&s_nkeys, 1u);
        done = true;
        break;
      }
      sl = (sl + 1) & (uint32_t)(cap - 1);
    }
    if (!done) s_fail = 1;
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) atomicOr(&flags[0], s_fail);
    return;
  }
  if (threadIdx.x == 0 && s_nkeys) atomicAdd(&flags[2], s_nkeys);

  // probe + aggregate; each thread has kPU entries' loads in flight
  const uint32_t q0 = p_pstart[p], q1 = p_pstart[p + 1];
  constexpr int kPU = 4;
  for (uint32_t e0 = q0 + threadIdx.x; e0 < q1; e0 += kJoinThreads * kPU) {
    int64_t k_[kPU];
    uint32_t r_[kPU];
    uint64_t v_[kPU][2];
#pragma unroll
    for (int u = 0; u < kPU; u++) {
      const uint32_t e = e0 + u * kJoinThreads;
      const uint32_t ec = e < q1 ? e : q1 - 1;
      k_[u] = p_keys[ec];
      r_[u] = p_rows[ec];
      v_[u][0] = plan.nvals > 0 ? p_vals[ec] : 0;
      v_[u][1] = plan.nvals > 1 ? p_vals[vstride + ec] : 0;
    }
#pragma unroll
    for (int u = 0; u < kPU; u++) {
      const uint32_t e = e0 + u * kJoinThreads;
      if (e >= q1) continue;
      const int64_t key = k_[u];
      int sl = -1;
      if (key == kEmptyKey) {
        sl = mult[cap] ? cap : -1;
      } else {
        const uint64_t h = radix_hash((uint64_t)key);
        uint32_t i = lg ? (uint32_t)(h >> (64 - lg)) : 0;
        for (int probe = 0; probe < cap; probe++) {
          const int64_t k = (int64_t)keys[i];
          if (k == kEmptyKey) break;
          if (k == key) {
            sl = (int)i;
            break;
          }
          i = (i + 1) & (uint32_t)(cap - 1);
        }
      }
      if (sl < 0) continue;
      uint32_t pi = pidx[sl];
      if (pi == kNone) {
        const uint32_t mine = atomicAdd(&s_pool, 1u);
        if (mine >= (uint32_t)NP) {
          s_fail = 2;
          continue;
        }
        const uint32_t old = atomicCAS(&pidx[sl], kNone, mine);
        pi = old == kNone ? mine : old;
        if (old == kNone) pslot[mine] = (uint32_t)sl;
      }
      if (atomicAdd(&cnt[pi], 1u) == 0) atomicAdd(&s_nrec, 1u);
      atomicMin(&first[pi], r_[u]);
      int vi = 0;
      for (int j = 0; j < jd.naggs; j++) {
        const JAgg& a = jd.agg[j];
        if (a.e.nfactors == 0) continue;
        const uint64_t v = vi < 2 ? v_[u][vi] : p_vals[(int64_t)vi * vstride + e];
        vi++;
        unsigned long long* w = &acc[(size_t)pi * NW + a.word0];
        switch (a.kind) {
          case VQ_AGG_SUM:
          case VQ_AGG_AVG:
            if (a.e.is_float) fix_add(w, bits_to_f64((int64_t)v), a.fs);
            else atomicAdd(w, (unsigned long long)v);
            break;
          case VQ_AGG_MIN:
          case VQ_AGG_MAX: {
            const double dv = bits_to_f64((int64_t)v);
            if (a.accum == VQ_F64 && dv != dv) break;  // NaN never wins a compare
            const int64_t o = a.accum == VQ_F64 ? f64_ord(dv) : (int64_t)v;
            if (a.kind == VQ_AGG_MIN) atomicMin((long long*)w, (long long)o);
            else atomicMax((long long*)w, (long long)o);
            break;
          }
          default:
            break;
        }
      }
    }
  }
  __syncthreads();
  if (s_fail) {
    if (threadIdx.x == 0) atomicOr(&flags[0], s_fail);
    return;
  }

  // emit every matched group
  if (threadIdx.x == 0) s_base = s_nrec ? atomicAdd(&flags[1], s_nrec) : 0;
  __syncthreads();
  if (s_nrec == 0) return;
  const uint32_t used = s_pool < (uint32_t)NP ? s_pool : (uint32_t)NP;
  __syncthreads();
  if (threadIdx.x == 0) s_nrec = 0;
  __syncthreads();
  const int RW = 3 + jd.naggs;
  for (uint32_t i = threadIdx.x; i < used; i += kJoinThreads) {
    if (cnt[i] == 0) continue;
    const uint32_t sl = pslot[i];
    const int64_t key = (int)sl == cap ? kEmptyKey : (int64_t)keys[sl];
    const uint32_t at = s_base + atomicAdd(&s_nrec, 1u);
    emit_record(jd, out + (size_t)at * RW, key, cnt[i], mult[sl], first[i], &acc[(size_t)i * NW]);
  }
}

// ---- top-k over the group records (ExternalSort DESC + Limit) ----------------
// One block.  Each thread keeps its kMaxK best records (sorted insertion),
// each wave merges its 64 lists by k rounds of a wave-wide best pick, and
// wave 0 merges the waves' lists the same way.
constexpr int kMaxK = 16;
constexpr int kTopThreads = 256;

struct KItem {
  int64_t v;
  uint32_t first;
  uint32_t idx;
};

__device__ __forceinline__ bool kbetter(const KItem& a, const KItem& b) {
  return a.v > b.v || (a.v == b.v && a.first < b.first);
}

__device__ __forceinline__ KItem knone() {
  KItem k;
  k.v = INT64_MIN;
  k.first = 0xFFFFFFFFu;
  k.idx = 0xFFFFFFFFu;
  return k;
}

__device__ __forceinline__ KItem kshfl(const KItem& c, int src) {
  KItem o;
  o.v = __shfl((long long)c.v, src, 64);
  o.first = (uint32_t)__shfl((int)c.first, src, 64);
  o.idx = (uint32_t)__shfl((int)c.idx, src, 64);
  return o;
}

// k rounds: the wave's best remaining head wins, its owner advances.
// Lane `r` of the result holds round r's winner.
__device__ __forceinline__ KItem wave_topk(KItem (&list)[kMaxK], int k) {
  const int lane = threadIdx.x & 63;
  int pos = 0;
  KItem mine = knone();
  for (int r = 0; r < k; r++) {
    KItem head = pos < kMaxK ? list[0] : knone();
#pragma unroll
    for (int q = 1; q < kMaxK; q++)
      if (q == pos) head = list[q];
    KItem best = head;
    int owner = lane;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const KItem o = kshfl(best, lane ^ m);
      const int oo = __shfl(owner, lane ^ m, 64);
      if (kbetter(o, best) || (!kbetter(best, o) && oo < owner && o.idx == best.idx)) {
        best = o;
        owner = oo;
      }
    }
    if (lane == owner && best.idx != 0xFFFFFFFFu) pos++;
    if (lane == r) mine = best;
  }
  return mine;
}

__device__ __forceinline__ void kinsert(KItem (&list)[kMaxK], const KItem& x) {
  if (!kbetter(x, list[kMaxK - 1])) return;
  KItem c = x;
#pragma unroll
  for (int q = 0; q < kMaxK; q++) {
    if (kbetter(c, list[q])) {
      const KItem t = list[q];
      list[q] = c;
      c = t;
    }
  }
}

// Block b of the grid ranks records [b*chunk, (b+1)*chunk) of n and writes
// its best topk as full records to out[b*topk ...] (missing ones with first
// row ~0).  Run with a grid of kTopBlocks over the group records, then with
// one block over the kTopBlocks*topk candidates; the single-block run also
// writes the count.
constexpr int kTopBlocks = 128;

__global__ __launch_bounds__(kTopThreads) void k_topk(const JDesc jd, int32_t topk,
                                                       const uint64_t* __restrict__ recs,
                                                       const uint32_t* __restrict__ n_dev,
                                                       uint32_t n_static,
                                                       uint64_t* __restrict__ out,
                                                       uint32_t* __restrict__ out_n) {
  __shared__ KItem s_w[kTopThreads / 64][kMaxK];
  const int RW = 3 + jd.naggs;
  const uint32_t n = n_dev ? *n_dev : n_static;
  const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  KItem list[kMaxK];
#pragma unroll
  for (int q = 0; q < kMaxK; q++) list[q] = knone();
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kTopThreads) {
    const uint64_t* r = recs + (size_t)i * RW;
    if (r[2] == ~0ULL) continue;  // a missing candidate
    KItem x;
    x.v = jd.naggs > 0 ? (int64_t)r[3] : 0;
    x.first = (uint32_t)r[2];
    x.idx = i;
    kinsert(list, x);
  }
  const KItem w = wave_topk(list, topk);
  if (lane < topk) s_w[wv][lane] = w;
  __syncthreads();
  if (wv != 0) return;
#pragma unroll
  for (int q = 0; q < kMaxK; q++) list[q] = knone();
  for (int e = lane; e < (kTopThreads / 64) * topk; e += 64) kinsert(list, s_w[e / topk][e % topk]);
  const KItem f = wave_topk(list, topk);
  uint64_t* o = out + (size_t)blockIdx.x * topk * RW;
  if (lane < topk) {
    if (f.idx != 0xFFFFFFFFu) {
      for (int q = 0; q < RW; q++) o[(size_t)lane * RW + q] = recs[(size_t)f.idx * RW + q];
    } else {
      for (int q = 0; q < RW; q++) o[(size_t)lane * RW + q] = 0;
      o[(size_t)lane * RW + 2] = ~0ULL;
    }
  }
  const uint32_t got = (uint32_t)__popcll(__ballot(lane < topk && f.idx != 0xFFFFFFFFu));
  if (lane == 0 && out_n) *out_n = got;
}

size_t join_lds_for(int32_t cap, int32_t pool, int32_t nwords) {
  const size_t ns = (size_t)cap + 1;
  return ns * (8 + 4 + 4) + (size_t)pool * (4 + 4 + 4 + 8 * (size_t)nwords);
}

// Turns counts[t][p] (written by the partition-count kernel) into absolute
// positions and pstart[0..P].
static void offsets(uint32_t* counts, int64_t ntiles, int32_t P, uint32_t* csum, uint32_t* pstart,
                    hipStream_t s) {
  const unsigned nch = (unsigned)((ntiles + kChunk - 1) / kChunk);
  const dim3 g(nch, (unsigned)((P + kBlk - 1) / kBlk));
  hipLaunchKernelGGL(k_off_chunks, g, dim3(kBlk), 0, s, counts, ntiles, P, csum);
  hipLaunchKernelGGL(k_off_starts, dim3(1), dim3(1024), 0, s, csum, (int32_t)nch, P, pstart);
  hipLaunchKernelGGL(k_off_tiles, g, dim3(kBlk), 0, s, counts, ntiles, P, csum);
}

hipError_t launch_join(const JSide& b, const JSide& p, const JDesc& jd, const JoinPlan& plan,
                       const JoinBufs& B, hipStream_t s, hipEvent_t* ev) {
  const int P = plan.P;
  if (ev) hipEventRecord(ev[0], s);
  hipMemsetAsync(B.flags, 0, sizeof(uint32_t) * 4, s);
  if (b.ntiles > 0) {
    hipError_t e = jit_launch_part(b, nullptr, P, B.b_counts, nullptr, nullptr, nullptr, 0, s, true);
    if (e != hipSuccess) return e;
    offsets(B.b_counts, b.ntiles, P, B.b_csum, B.b_pstart, s);
    e = jit_launch_part(b, nullptr, P, B.b_counts, B.b_keys, nullptr, nullptr, 0, s, false);
    if (e != hipSuccess) return e;
  } else {
    hipMemsetAsync(B.b_pstart, 0, sizeof(uint32_t) * (P + 1), s);
  }
  if (p.ntiles > 0) {
    hipError_t e = jit_launch_part(p, &jd, P, B.p_counts, nullptr, nullptr, nullptr, 0, s, true);
    if (e != hipSuccess) return e;
    offsets(B.p_counts, p.ntiles, P, B.p_csum, B.p_pstart, s);
    e = jit_launch_part(p, &jd, P, B.p_counts, B.p_keys, B.p_rows, B.p_vals, B.cap_p, s, false);
    if (e != hipSuccess) return e;
  } else {
    hipMemsetAsync(B.p_pstart, 0, sizeof(uint32_t) * (P + 1), s);
  }
  if (ev) hipEventRecord(ev[1], s);
  const size_t lds = join_lds_for(plan.cap, plan.pool, jd.nwords);
  if (lds > 64 * 1024)
    hipFuncSetAttribute((const void*)k_join_agg, hipFuncAttributeMaxDynamicSharedMemorySize,
                        plan.lds_max);
  hipLaunchKernelGGL(k_join_agg, dim3((unsigned)P), dim3(kJoinThreads), lds, s, jd,
                     plan, B.b_keys, B.b_pstart, B.p_keys, B.p_rows, B.p_vals, B.cap_p,
                     B.p_pstart, B.recs, B.flags);
  if (ev) hipEventRecord(ev[2], s);
  if (plan.topk > 0) {
    hipLaunchKernelGGL(k_topk, dim3(kTopBlocks), dim3(kTopThreads), 0, s, jd, plan.topk, B.recs,
                       B.flags + 1, 0u, B.cand, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_topk, dim3(1), dim3(kTopThreads), 0, s, jd, plan.topk, B.cand,
                       (const uint32_t*)nullptr, (uint32_t)(kTopBlocks * plan.topk), B.out,
                       B.flags + 3);
  }
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

}  // namespace vq
