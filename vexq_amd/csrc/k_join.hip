// k_join.hip — gfx950 kernels for the radix-partitioned hash join feeding a
// GROUP BY on the join key (ParallelHashJoinAggregate, exec/parallel_join.go).
//
// Pipeline (one stream; eight launches, eleven with the semi-join filter):
//   k_zero_u32       the flags to zero
//   vq_part_count    (run-time compiled, jit_template.hip) per side: filter +
//                    radixHash + LDS histogram per persistent workgroup ->
//                    counts[p][w]
//   k_part_runs,     per side: counts -> partition starts and the start of
//   k_part_starts    each workgroup's run inside each partition
//   vq_part_scatter  (run-time compiled) per side: each passing row into its
//                    workgroup's run through LDS write-combining buffers: the
//                    key, and on the probe side its row id and the values of
//                    the aggregate arguments
//   k_bloom_build    with the filter, between the build side's scatter and
//                    the probe side's count: each partition's build keys as
//                    a blocked Bloom filter, which the probe side's rows must
//                    pass (most probe rows of a selective join match no
//                    passing build key, and are then never copied)
//   k_join_agg       one workgroup per partition: build the partition's
//                    open-addressing table in LDS from its build keys, probe it
//                    with its probe rows, aggregate into per-key accumulators
//                    in LDS, then either emit every matched key as a group
//                    record or (ORDER BY aggregate 0 DESC LIMIT k) keep the
//                    partition's best k; the last workgroup to finish merges
//                    the partitions' candidates into the final top k.
//
// The partition of a key is the low `bits` of radixHash (radixPart,
// parallel_join.go:104-109) and its slot inside the partition table the high
// bits (slotFor, join_table.go:488-492) — the same split as the reference, so
// a partition's keys spread over its table.
//
// Exactness: float sums are accumulated as fixed-point integers (fixpoint.h),
// integer sums as wrapping int64, MIN/MAX as order-preserving integers, first
// row as a minimum: every update is an associative integer atomic, so the
// result does not depend on which lane or workgroup got there first, nor on
// the order of entries inside a partition.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>

#include "vq_internal.h"

namespace vq {

constexpr int64_t kEmptyKey = INT64_MIN;

// ---- zero fill: a kernel rather than a memset node, so that a captured
// graph of the pipeline replays it like any other launch -----------------------
__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = 0;
}

// ---- partition positions ---------------------------------------------------------
// counts[p][w] holds workgroup w's rows of partition p (layout [P][G]).
//   k_part_runs    one block per (partition, side): counts[p][w] <- rows of
//                  partition p in workgroups before w (exclusive scan over w),
//                  tot[p] <- the partition's rows
//   k_part_starts  one block per side: pstart = exclusive scan of tot
// The scatter's workgroup w writes partition p from pstart[p] + counts[p][w].
struct PartScan {
  uint32_t* counts[2];
  uint32_t* tot[2];
  uint32_t* pstart[2];
  uint32_t* total[2];
  int32_t G[2];
  int32_t P;
  int32_t pad_;
};

constexpr int kScanThreads = 256;

// Exclusive scan of v[0..n) in place by one block of kScanThreads; returns
// the total to every thread.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t* v, int64_t n, uint32_t* s) {
  const int tid = threadIdx.x;
  const int64_t per = (n + kScanThreads - 1) / kScanThreads;
  const int64_t lo = tid * per < n ? tid * per : n;
  const int64_t hi = lo + per < n ? lo + per : n;
  uint32_t mine = 0;
  for (int64_t i = lo; i < hi; i++) mine += v[i];
  s[tid] = mine;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t x = tid >= off ? s[tid - off] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  uint32_t run = s[tid] - mine;
  for (int64_t i = lo; i < hi; i++) {
    const uint32_t x = v[i];
    v[i] = run;
    run += x;
  }
  const uint32_t total = s[kScanThreads - 1];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanThreads) void k_part_runs(const PartScan a) {
  __shared__ uint32_t s[kScanThreads];
  const int p = blockIdx.x, side = blockIdx.y;
  const int G = a.G[side];
  const uint32_t t = block_scan_excl(a.counts[side] + (int64_t)p * G, G, s);
  if (threadIdx.x == 0) a.tot[side][p] = t;
}

__global__ __launch_bounds__(kScanThreads) void k_part_starts(const PartScan a) {
  __shared__ uint32_t s[kScanThreads];
  const int side = blockIdx.x, P = a.P;
  uint32_t* ps = a.pstart[side];
  for (int p = threadIdx.x; p < P; p += kScanThreads) ps[p] = a.tot[side][p];
  __syncthreads();
  const uint32_t t = block_scan_excl(ps, P, s);
  if (threadIdx.x == 0) {
    ps[P] = t;
    *a.total[side] = t;
  }
}

// ---- semi-join filter ---------------------------------------------------------
// One block per partition builds the partition's filter block (2^lpp lines
// of 64 bytes) in LDS from its build entries — two bits per key, in one
// 64-bit word of one line, both chosen by the key's hash (the bits
// vq_part_count / vq_part_scatter test,
// jit_template.hip bloom_test) — and writes it whole: no global atomics, no
// zero fill.
__global__ __launch_bounds__(256) void k_bloom_build(const int64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ pstart,
                                                     uint32_t* __restrict__ bloom, int32_t lpp,
                                                     int32_t P1, int32_t P2) {
  extern __shared__ uint32_t seg[];
  // block b holds layout partition b: coarse b / P2, fine b % P2, which is
  // hash partition (b / P2) + (b % P2) * P1 (the filter's block index)
  const uint32_t b = blockIdx.x;
  const uint32_t p = b / (uint32_t)P2 + (b % (uint32_t)P2) * (uint32_t)P1;
  const uint32_t lines = 1u << lpp, words = lines * 16;
  for (uint32_t i = threadIdx.x; i < words; i += 256) seg[i] = 0;
  __syncthreads();
  for (uint32_t e = pstart[b] + threadIdx.x; e < pstart[b + 1]; e += 256) {
    const uint64_t h = radix_hash((uint64_t)keys[e]);
    // both bits in one 64-bit word of the line (one load per test)
    const uint32_t line = (uint32_t)(h >> 20) & (lines - 1);
    const uint32_t w = line * 16 + (((uint32_t)(h >> 40) & 7u) << 1);
    const uint32_t b1 = (uint32_t)(h >> 43) & 63u, b2 = (uint32_t)(h >> 49) & 63u;
    atomicOr(&seg[w + (b1 >> 5)], 1u << (b1 & 31));
    atomicOr(&seg[w + (b2 >> 5)], 1u << (b2 & 31));
  }
  __syncthreads();
  uint32_t* out = bloom + (size_t)p * words;
  for (uint32_t i = threadIdx.x; i < words; i += 256) out[i] = seg[i];
}

// ---- second partitioning pass ------------------------------------------------
// One block per coarse partition c (hash bits [0, bits1)): its entries split
// by the next hash bits into P2 fine partitions, written in fine order to the
// final arrays inside c's range (an LDS histogram, a scan, then LDS
// cursors); fpstart[c * P2 + f] is fine partition f's start.  A coarse
// partition's writes stay within its own range, so they combine in L2.
constexpr int kFineThreads = 1024;

__global__ __launch_bounds__(kFineThreads) void k_part_fine(const PartBufs s, int32_t P1,
                                                            int32_t bits1, int32_t P2,
                                                            int32_t nvals, int32_t probe) {
  __shared__ uint32_t hist[256], cur[256];
  const uint32_t c = blockIdx.x;
  const uint32_t cs = s.pstart[c], ce = s.pstart[c + 1];
  for (int f = threadIdx.x; f < P2; f += kFineThreads) hist[f] = 0;
  __syncthreads();
  for (uint32_t e = cs + threadIdx.x; e < ce; e += kFineThreads) {
    const uint32_t f = (uint32_t)(radix_hash((uint64_t)s.keys[e]) >> bits1) & (uint32_t)(P2 - 1);
    atomicAdd(&hist[f], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t at = cs;
    for (int f = 0; f < P2; f++) {
      cur[f] = at;
      s.fpstart[(size_t)c * P2 + f] = at;
      at += hist[f];
    }
    if ((int)c == P1 - 1) s.fpstart[(size_t)P1 * P2] = ce;
  }
  __syncthreads();
  for (uint32_t e = cs + threadIdx.x; e < ce; e += kFineThreads) {
    const int64_t key = s.keys[e];
    const uint32_t f = (uint32_t)(radix_hash((uint64_t)key) >> bits1) & (uint32_t)(P2 - 1);
    const uint32_t pos = atomicAdd(&cur[f], 1u);
    s.fkeys[pos] = key;
    if (probe) {
      s.frows[pos] = s.rows[e];
      for (int v = 0; v < nvals; v++) s.fvals[(size_t)v * s.vstride + pos] = s.vals[(size_t)v * s.vstride + e];
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

// Add float64 t into three fixed-point limbs (fixpoint.h).
__device__ __forceinline__ void fix_add(unsigned long long* limb, double t, FixScale fs) {
  int64_t c0, c1, c2;
  if (!fix_chunks(t, fs, c0, c1, c2)) return;
  if (c0) atomicAdd(&limb[0], (unsigned long long)c0);
  if (c1) atomicAdd(&limb[1], (unsigned long long)c1);
  if (c2) atomicAdd(&limb[2], (unsigned long long)c2);
}

// The output word of one aggregate of a group: int64, or float64 bits.
// cnt = matched probe rows, mult = build rows of the key (each probe match
// joins every one of them, so counts and sums scale by mult).  Not inlined,
// and given plain values rather than the descriptor: it runs once per group,
// and an inlined copy of the rounding code in every caller would lengthen
// the kernel the instruction cache has to hold.
__device__ __noinline__ uint64_t agg_word(int32_t kind, int32_t accum, int32_t is_float,
                                          FixScale fs, uint64_t cnt, uint32_t mult,
                                          const unsigned long long* w) {
  if (kind == VQ_AGG_COUNT) return cnt * mult;
  if (kind == VQ_AGG_MIN || kind == VQ_AGG_MAX)
    return accum == VQ_F64 ? (uint64_t)f64_to_bits(ord_f64((int64_t)w[0])) : (uint64_t)w[0];
  const bool fixed = kind == VQ_AGG_SUM ? accum == VQ_F64 : is_float != 0;
  const double fsum = fixed ? fix_to_f64((int64_t)w[0], (int64_t)w[1], (int64_t)w[2], fs, mult) : 0.0;
  if (kind == VQ_AGG_SUM) return fixed ? (uint64_t)f64_to_bits(fsum) : (uint64_t)w[0] * mult;
  // AVG
  const double sum = fixed ? fsum : (double)(int64_t)((uint64_t)w[0] * mult);
  return (uint64_t)f64_to_bits(sum / (double)(cnt * mult));
}

__device__ __forceinline__ uint64_t agg_word_of(const JAgg& a, uint64_t cnt, uint32_t mult,
                                                const unsigned long long* acc) {
  return agg_word(a.kind, a.accum, a.e.is_float, a.fs, cnt, mult, acc + a.word0);
}

// Record layout (64-bit words): key, count, first row, then one word per
// aggregate (int64 or float64 bits).  A missing top-k candidate has first
// row ~0.
// Stores and loads at the device's coherence point (past this L2), for data
// one workgroup hands to another inside a kernel: with them the hand-over
// needs no cache write-back, only the writer's stores acknowledged before its
// signal.
__device__ __forceinline__ void st_dev(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_dev(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool kDev = false>
__device__ __forceinline__ void emit_record(const JDesc& jd, uint64_t* rec, int64_t key,
                                            uint64_t cnt, uint32_t mult, uint32_t first,
                                            const unsigned long long* acc) {
  auto put = [&](int i, uint64_t v) {
    if (kDev) st_dev(&rec[i], v);
    else rec[i] = v;
  };
  put(0, (uint64_t)key);
  put(1, cnt * mult);
  put(2, first);
  for (int j = 0; j < jd.naggs; j++) put(3 + j, agg_word_of(jd.agg[j], cnt, mult, acc));
}

// ---- top-k (ExternalSort on aggregate 0 DESC + Limit, exec/sort.go:130-163):
// the sort compares the int64 a row carries (a float's raw bits) and keeps
// first-occurrence order among ties.  The order (value DESC, first row ASC)
// is total over groups, since no two groups share a first row.  The block
// selects the best k by k rounds of a block-wide maximum over the items
// ranked after the previous round's winner.
constexpr int kMaxK = 16;
constexpr int kJoinThreads = 1024;
constexpr int kWaves = kJoinThreads / 64;

struct KItem {
  int64_t v;
  uint32_t first;
  uint32_t idx;
};

__device__ __forceinline__ bool kbetter(const KItem& a, const KItem& b) {
  return a.v > b.v || (a.v == b.v && a.first < b.first);
}

__device__ __forceinline__ KItem knone();
__device__ __forceinline__ void st_item(KItem* p, const KItem& x);
__device__ __forceinline__ KItem ld_item(const KItem* p);

__device__ __forceinline__ KItem knone() {
  KItem k;
  k.v = INT64_MIN;
  k.first = 0xFFFFFFFFu;
  k.idx = 0xFFFFFFFFu;
  return k;
}

__device__ __forceinline__ void st_item(KItem* p, const KItem& x) {
  st_dev((uint64_t*)p, (uint64_t)x.v);
  st_dev((uint64_t*)p + 1, (uint64_t)x.first | ((uint64_t)x.idx << 32));
}
__device__ __forceinline__ KItem ld_item(const KItem* p) {
  KItem x;
  x.v = (int64_t)ld_dev((const uint64_t*)p);
  const uint64_t w = ld_dev((const uint64_t*)p + 1);
  x.first = (uint32_t)w;
  x.idx = (uint32_t)(w >> 32);
  return x;
}

__device__ __forceinline__ KItem kshfl(const KItem& c, int src) {
  KItem o;
  o.v = __shfl((long long)c.v, src, 64);
  o.first = (uint32_t)__shfl((int)c.first, src, 64);
  o.idx = (uint32_t)__shfl((int)c.idx, src, 64);
  return o;
}

// One DPP move of an item across lanes; lanes without a valid source keep
// their own item.
template <int CTRL, int ROWM>
__device__ __forceinline__ KItem kdpp(const KItem& x) {
  const uint64_t u = (uint64_t)x.v;
  const int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
  KItem o;
  const uint32_t lo2 = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWM, 0xf, false);
  const uint32_t hi2 = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWM, 0xf, false);
  o.v = (int64_t)(((uint64_t)hi2 << 32) | lo2);
  o.first = (uint32_t)__builtin_amdgcn_update_dpp((int)x.first, (int)x.first, CTRL, ROWM, 0xf, false);
  o.idx = (uint32_t)__builtin_amdgcn_update_dpp((int)x.idx, (int)x.idx, CTRL, ROWM, 0xf, false);
  return o;
}

__device__ __forceinline__ KItem kmax(const KItem& a, const KItem& b) { return kbetter(b, a) ? b : a; }

// The best item of the wave, returned to every lane: a running maximum over
// row_shr 1, 2, 4, 8 (lane 15 of each 16-lane row holds the row's best), then
// row_bcast 15 and 31 carry it to lane 63.  Call with the whole wave active.
__device__ __forceinline__ KItem wave_best(KItem x) {
  x = kmax(x, kdpp<0x111, 0xf>(x));
  x = kmax(x, kdpp<0x112, 0xf>(x));
  x = kmax(x, kdpp<0x114, 0xf>(x));
  x = kmax(x, kdpp<0x118, 0xf>(x));
  x = kmax(x, kdpp<0x142, 0xa>(x));
  x = kmax(x, kdpp<0x143, 0xc>(x));
  const uint64_t u = (uint64_t)x.v;
  KItem o;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  o.v = (int64_t)(((uint64_t)hi << 32) | lo);
  o.first = (uint32_t)__builtin_amdgcn_readlane((int)x.first, 63);
  o.idx = (uint32_t)__builtin_amdgcn_readlane((int)x.idx, 63);
  return o;
}

// x if it ranks after thr (or thr is none yet), else none.
__device__ __forceinline__ KItem kafter(const KItem& x, const KItem& thr, bool first_round) {
  return (x.idx != 0xFFFFFFFFu && (first_round || kbetter(thr, x))) ? x : knone();
}

// The wave's best k of items striped over its lanes: k rounds of wave_best
// over the items ranked after the last winner.  Lane r returns rank r (none
// past the last item).  Not inlined, and one copy per item source: this is
// cold code, run a few times per workgroup.
//   wave_topk_mem: items a[i], i in [lo, hi) (LDS or global memory), taken
//                  kRegItems per lane at a time into registers, the running
//                  best k riding along as one more item per lane
//   wave_topk_two: each lane's two items x0, x1
constexpr int kRegItems = 4;

__device__ __noinline__ KItem wave_topk_mem(const KItem* a, uint32_t lo, uint32_t hi, int k) {
  const int lane = threadIdx.x & 63;
  KItem best = knone();
  for (uint32_t b = lo; b < hi; b += 64 * kRegItems) {
    KItem x[kRegItems];
#pragma unroll
    for (int j = 0; j < kRegItems; j++) {
      const uint32_t i = b + j * 64 + lane;
      x[j] = i < hi ? a[i] : knone();
    }
    KItem thr = knone(), mine = knone();
    for (int r = 0; r < k; r++) {
      KItem c = kafter(best, thr, r == 0);
#pragma unroll
      for (int j = 0; j < kRegItems; j++) c = kmax(c, kafter(x[j], thr, r == 0));
      const KItem w = wave_best(c);
      if (w.idx == 0xFFFFFFFFu) break;
      if (lane == r) mine = w;
      thr = w;
    }
    best = mine;
  }
  return best;
}

// wave_topk_mem over items other workgroups of this kernel wrote (read at
// the device's coherence point).
__device__ __noinline__ KItem wave_topk_dev(const KItem* a, uint32_t n, int k) {
  const int lane = threadIdx.x & 63;
  KItem best = knone();
  for (uint32_t b = 0; b < n; b += 64 * kRegItems) {
    KItem x[kRegItems];
#pragma unroll
    for (int j = 0; j < kRegItems; j++) {
      const uint32_t i = b + j * 64 + lane;
      x[j] = i < n ? ld_item(&a[i]) : knone();
    }
    KItem thr = knone(), mine = knone();
    for (int r = 0; r < k; r++) {
      KItem c = kafter(best, thr, r == 0);
#pragma unroll
      for (int j = 0; j < kRegItems; j++) c = kmax(c, kafter(x[j], thr, r == 0));
      const KItem w = wave_best(c);
      if (w.idx == 0xFFFFFFFFu) break;
      if (lane == r) mine = w;
      thr = w;
    }
    best = mine;
  }
  return best;
}

__device__ __noinline__ KItem wave_topk_two(KItem x0, KItem x1, int k) {
  const int lane = threadIdx.x & 63;
  KItem thr = knone(), mine = knone();
  for (int r = 0; r < k; r++) {
    const KItem a0 = kafter(x0, thr, r == 0), a1 = kafter(x1, thr, r == 0);
    const KItem w = wave_best(kmax(a0, a1));
    if (w.idx == 0xFFFFFFFFu) break;
    if (lane == r) mine = w;
    thr = w;
  }
  return mine;
}

__device__ __forceinline__ unsigned long long lds_load_key(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The number of items a[0..n) that rank ahead of x (an absent item never
// does), with eight LDS reads in flight at a time.
__device__ __forceinline__ uint32_t rank_in(const KItem* a, uint32_t n, const KItem& x) {
  uint32_t r = 0, b = 0;
  for (; b + 8 <= n; b += 8) {
    KItem y[8];
#pragma unroll
    for (int q = 0; q < 8; q++) y[q] = a[b + q];
#pragma unroll
    for (int q = 0; q < 8; q++) r += kbetter(y[q], x) ? 1u : 0u;
  }
  for (; b < n; b++) r += kbetter(a[b], x) ? 1u : 0u;
  return r;
}

// One 1024-thread workgroup per partition.  LDS:
//   table  keys u64[cap+1] | acc u64[reps][pool][nwords] | tag u32[cap+1]
//   pool   slot u32[pool] | mult u32[pool] | cnt u32[reps][pool] |
//          first u32[reps][pool]
// Slot `cap` holds the key INT64_MIN, which doubles as the empty marker of the
// other slots.  A slot's tag counts its build rows (the key's multiplicity)
// while the table is built.  Accumulators exist only for build keys a probe
// row matches: the first match of a slot marks the tag pending by a
// compare-and-swap, takes the next pool entry, moves the multiplicity there
// and sets the tag to kClaimed | entry; lanes that lose the swap wait for
// the entry.  Each wave updates replica (wave mod reps) of the accumulators;
// the replicas are folded before the results are read.  flags: see JoinBufs.
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kClaimed = 0x80000000u;
constexpr uint32_t kPending = 0xFFFFFFFFu;  // a lane is taking the slot's pool entry

// Position of this lane's entry when the lanes with `pred` append to the
// list *ctr counts: one LDS atomic per wave.  Call from the whole wave.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* ctr) {
  const uint64_t m = __ballot(pred);
  if (m == 0) return 0;
  const uint32_t below =
      __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  return (uint32_t)__builtin_amdgcn_readlane((int)base, leader) + below;
}

// Drops from a[0..n) (n <= kJoinThreads, in LDS) every item that cannot make
// the best k, and absent items; the survivors are compacted to the front of
// a, and their count is returned.  The bound: the k-th best of a sample (the
// first 64, ranked by wave 0 lane against lane), which is at or ahead of the
// whole set's k-th best, so the best k survive.  Call from the whole
// workgroup.
// The k-th best of the wave's lane items (each lane ranked against every
// other), written to *out by the lane that holds it; none when fewer than k
// items are valid.  Call from the whole wave.
__device__ __forceinline__ void wave_kth(const KItem& x, int k, KItem* out) {
  const bool valid = x.idx != kNone;
  const uint32_t xlo = (uint32_t)(uint64_t)x.v, xhi = (uint32_t)((uint64_t)x.v >> 32);
  uint32_t r = 0;
  for (int l = 0; l < 64; l++) {
    KItem y;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)xlo, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)xhi, l);
    y.v = (int64_t)(((uint64_t)hi << 32) | lo);
    y.first = (uint32_t)__builtin_amdgcn_readlane((int)x.first, l);
    y.idx = (uint32_t)__builtin_amdgcn_readlane((int)x.idx, l);
    r += (y.idx != kNone && kbetter(y, x)) ? 1u : 0u;
  }
  const uint64_t m = __ballot(valid && r == (uint32_t)(k - 1));
  if ((threadIdx.x & 63) == 0) *out = knone();
  if (m && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) *out = x;
}

__device__ uint32_t prune_topk(KItem* a, uint32_t n, int k, uint32_t* s_cnt, KItem* s_thr) {
  const bool bound = n > 64;
  if (n > 64 && threadIdx.x < 64) wave_kth(a[threadIdx.x], k, s_thr);
  const KItem x = threadIdx.x < n ? a[threadIdx.x] : knone();
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();
  const KItem t = bound ? *s_thr : knone();
  const bool keep = x.idx != kNone && (t.idx == kNone || !kbetter(t, x));
  const uint32_t at = wave_append(keep, s_cnt);
  if (keep) a[at] = x;
  __syncthreads();
  return *s_cnt;
}

constexpr int kPoolPerThread = 2;  // pool <= kPoolPerThread * kJoinThreads
constexpr int kRankMax = 512;      // top-k by counting ranks up to this many groups

// The merge of P lists of k candidates (citems[p*k + r], rank order, and
// their records cand), by one workgroup of T threads: the best k records
// into out and their count into flags[3].  The lists were
// written by other workgroups of the same kernel (read at the device's
// coherence point).  The bound: each lane of wave 0 takes the best of its
// share of the lists' first items, and the k-th best of those 64 is at or
// ahead of the k-th best of all (k distinct items are at or ahead of it);
// the items at or ahead of it are ranked by counting.  When more survive
// than the ranking room holds, wave 0 selects over all items by rounds of
// wave maxima.  ritems: kRankMax items of LDS; s_bound, s_nkeys: LDS.
// (Plain pointers: a reference to the kernel's argument structs would make
// the compiler copy them into scratch memory.)
template <int T, int kIn = 4>
__device__ __forceinline__ void topk_merge(const KItem* citems, const uint64_t* cand, uint64_t* out, int k,
                                           int RW, uint32_t P, KItem* ritems, KItem* s_bound,
                                           uint32_t* s_nkeys, uint32_t* flags) {
  const uint32_t n = P * (uint32_t)k;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) *s_nkeys = 0;
  // (kIn loads per lane in flight at a time: with thousands of partitions
  // the lists hold tens of thousands of items)
  if (wv == 0) {
    KItem b = knone();
    for (uint32_t l0 = lane; l0 < P; l0 += 64 * kIn) {
      KItem xs[kIn];
#pragma unroll
      for (int u = 0; u < kIn; u++) {
        const uint32_t l = l0 + u * 64;
        xs[u] = l < P ? ld_item(&citems[(size_t)l * k]) : knone();
      }
#pragma unroll
      for (int u = 0; u < kIn; u++) b = kmax(b, xs[u]);
    }
    wave_kth(b, k, s_bound);
  }
  __syncthreads();
  const KItem t = *s_bound;
  for (uint32_t i0 = 0; i0 < n; i0 += T * kIn) {
    KItem xs[kIn];
#pragma unroll
    for (int u = 0; u < kIn; u++) {
      const uint32_t i = i0 + u * T + threadIdx.x;
      xs[u] = i < n ? ld_item(&citems[i]) : knone();
    }
#pragma unroll
    for (int u = 0; u < kIn; u++) {
      const bool keep = xs[u].idx != kNone && (t.idx == kNone || !kbetter(t, xs[u]));
      const uint32_t at = wave_append(keep, s_nkeys);
      if (keep && at < (uint32_t)kRankMax) ritems[at] = xs[u];
    }
  }
  __syncthreads();
  const uint32_t m = *s_nkeys;
  if (m <= (uint32_t)kRankMax) {
    for (uint32_t a = threadIdx.x; a < m; a += T) {
      const KItem x = ritems[a];
      const uint32_t r = rank_in(ritems, m, x);
      if (r >= (uint32_t)k) continue;
      for (int q = 0; q < RW; q++) out[(size_t)r * RW + q] = ld_dev(&cand[(size_t)x.idx * RW + q]);
    }
  } else if (wv == 0) {
    const KItem x = wave_topk_dev(citems, n, k);
    if (lane < k && x.idx != kNone)
      for (int q = 0; q < RW; q++) out[(size_t)lane * RW + q] = ld_dev(&cand[(size_t)x.idx * RW + q]);
  }
  if (threadIdx.x == 0) flags[3] = m < (uint32_t)k ? m : (uint32_t)k;
}

// Eight waves per SIMD (at most 64 VGPRs): two workgroups per CU when the
// tables leave room in LDS.
__global__ __launch_bounds__(kJoinThreads, 8) void k_join_agg(const JDesc jd, const JoinPlan plan,
                                                           const JoinBufs B) {
  extern __shared__ unsigned long long smem[];
  __shared__ uint32_t s_fail, s_nrec, s_nkeys, s_base, s_pool, s_last;
  __shared__ KItem s_bound;
  __shared__ KItem s_cand[kWaves * kMaxK];
  __shared__ uint8_t s_wop[3 * VQ_MAX_AGGS];  // per accumulator word: 0 add, 1 min, 2 max
  const int p = blockIdx.x;
  const int cap = plan.cap, ns = cap + 1, NP = plan.pool, R = plan.reps;
  const int NW = jd.nwords;
  uint32_t* flags = B.flags;
  unsigned long long* keys = smem;
  unsigned long long* acc = keys + ns;                       // [R][NP][NW]
  uint32_t* tag = (uint32_t*)(acc + (size_t)R * NP * NW);
  uint32_t* pslot = tag + ns;
  uint32_t* pmult = pslot + NP;
  uint32_t* cnt = pmult + NP;                                // [R][NP]
  uint32_t* first = cnt + (size_t)R * NP;                    // [R][NP]
  const int lg = __ffs(cap) - 1;  // log2(cap)
  // the top-k's ranking room after the pool, 16-byte aligned (an offset from
  // smem, so the compiler keeps it an LDS address)
  KItem* ritems = (KItem*)(smem + ((((size_t)ns * 12 + (size_t)R * NP * NW * 8 + (size_t)NP * 8 +
                                     (size_t)R * NP * 8) + 15) / 16) * 2);
  // lane l of every wave updates accumulator replica l mod R: with few hot
  // keys the replicas spread what would be one LDS address's atomics (lanes
  // of one instruction on the same address serialize) over R addresses
  const int myrep = (int)threadIdx.x & (R - 1);
  // the partition's bounds, this thread's first build key and first probe
  // entry: loading while the tables are cleared
  const uint32_t b0 = plan.groupby ? 0 : B.b.fpstart[p], b1 = plan.groupby ? 0 : B.b.fpstart[p + 1];
  const uint32_t q0 = B.p.fpstart[p], q1 = B.p.fpstart[p + 1];
  const int64_t vstride = B.p.vstride;
  int64_t bkey0 = 0;
  if (b0 + threadIdx.x < b1) bkey0 = B.b.fkeys[b0 + threadIdx.x];
  uint32_t e = q0 + threadIdx.x;
  int64_t nkey = 0;
  uint32_t nrow = 0;
  uint64_t nva = 0, nvb = 0, nvc = 0;
  auto fetch = [&](uint32_t x) {
    const uint32_t ec = x < q1 ? x : q1 - 1;
    nkey = B.p.fkeys[ec];
    nrow = B.p.frows[ec];
    nva = plan.nvals > 0 ? B.p.fvals[ec] : 0;
    nvb = plan.nvals > 1 ? B.p.fvals[vstride + ec] : 0;
    nvc = plan.nvals > 2 ? B.p.fvals[2 * vstride + ec] : 0;
  };
  if (q1 > q0) fetch(e);
  if (threadIdx.x == 0) {
    s_fail = 0;
    s_nrec = 0;
    s_nkeys = 0;
    s_pool = 0;
  }
  for (int i = threadIdx.x; i < ns; i += kJoinThreads) {
    keys[i] = (unsigned long long)kEmptyKey;
    tag[i] = 0;
  }
  for (int i = threadIdx.x; i < R * NP; i += kJoinThreads) {
    cnt[i] = 0;
    first[i] = kNone;
  }
  if (threadIdx.x < NW) {
    uint8_t op = 0;
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      if ((int)threadIdx.x >= a.word0 && (int)threadIdx.x < a.word0 + a.nwords)
        op = a.kind == VQ_AGG_MIN ? 1 : a.kind == VQ_AGG_MAX ? 2 : 0;
    }
    s_wop[threadIdx.x] = op;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < R * NP * NW; i += kJoinThreads) {
    const int op = s_wop[i % NW];
    acc[i] = op == 1 ? (unsigned long long)INT64_MAX : op == 2 ? (unsigned long long)INT64_MIN : 0ULL;
  }
  __syncthreads();

  // build: the partition's build keys; multiplicity = build rows per key
  // (GROUP BY mode has no build side: keys enter the table as rows arrive)
  for (uint32_t eb = b0 + threadIdx.x; eb < b1; eb += kJoinThreads) {
    const int64_t key = eb == b0 + threadIdx.x ? bkey0 : B.b.fkeys[eb];
    if (key == kEmptyKey) {
      if (atomicAdd(&tag[cap], 1u) == 0) atomicAdd(&s_nkeys, 1u);
      continue;
    }
    const uint64_t h = radix_hash((uint64_t)key);
    uint32_t sl = lg ? (uint32_t)(h >> (64 - lg)) : 0;
    bool done = false;
    for (int probe = 0; probe < cap; probe++) {
      unsigned long long old = lds_load_key(&keys[sl]);
      if (old == (unsigned long long)kEmptyKey)
        old = atomicCAS(&keys[sl], (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (old == (unsigned long long)kEmptyKey || old == (unsigned long long)key) {
        atomicAdd(&tag[sl], 1u);
        if (old == (unsigned long long)kEmptyKey) atomicAdd(&s_nkeys, 1u);
        done = true;
        break;
      }
      sl = (sl + 1) & (uint32_t)(cap - 1);
    }
    if (!done) s_fail = 1;
  }
  __syncthreads();

  // GROUP BY mode: the partition's keys are its rows' keys.  Every key goes
  // into the table first, then each occupied slot gets the pool entry of its
  // rank in slot order (so no two lanes ever race for a new group's entry),
  // and the probe below finds every key claimed.
  if (plan.groupby) {
    for (uint32_t eg = q0 + threadIdx.x; eg < q1; eg += kJoinThreads) {
      const int64_t key = eg == q0 + threadIdx.x ? nkey : B.p.fkeys[eg];
      if (key == kEmptyKey) {
        tag[cap] = 1;
        continue;
      }
      const uint64_t h = radix_hash((uint64_t)key);
      uint32_t sl = lg ? (uint32_t)(h >> (64 - lg)) : 0;
      bool done = false;
      for (int probe = 0; probe < cap; probe++) {
        // (a key already in the table is found without an atomic: a hot
        // key's rows do not serialize on its slot)
        unsigned long long old = lds_load_key(&keys[sl]);
        if (old == (unsigned long long)kEmptyKey)
          old = atomicCAS(&keys[sl], (unsigned long long)kEmptyKey, (unsigned long long)key);
        if (old == (unsigned long long)kEmptyKey || old == (unsigned long long)key) {
          if (old == (unsigned long long)kEmptyKey) tag[sl] = 1;
          done = true;
          break;
        }
        sl = (sl + 1) & (uint32_t)(cap - 1);
      }
      if (!done) s_fail = 1;
    }
    __syncthreads();
    // rank the occupied slots: each thread a contiguous run of slots
    const int per = (ns + kJoinThreads - 1) / kJoinThreads;
    const int lo = threadIdx.x * per < ns ? threadIdx.x * per : ns;
    const int hi = lo + per < ns ? lo + per : ns;
    uint32_t mine = 0;
    for (int i = lo; i < hi; i++) mine += tag[i] ? 1u : 0u;
    uint32_t* sc = (uint32_t*)&s_cand[0];  // kJoinThreads words of scratch
    sc[threadIdx.x] = mine;
    __syncthreads();
    for (int off = 1; off < kJoinThreads; off <<= 1) {
      const uint32_t x = (int)threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
      __syncthreads();
      sc[threadIdx.x] += x;
      __syncthreads();
    }
    uint32_t r = sc[threadIdx.x] - mine;
    const uint32_t total = sc[kJoinThreads - 1];
    for (int i = lo; i < hi; i++) {
      if (!tag[i]) continue;
      if (r < (uint32_t)NP) {
        pslot[r] = (uint32_t)i;
        pmult[r] = 1;
      }
      tag[i] = kClaimed | r;
      r++;
    }
    if (threadIdx.x == 0) {
      s_nkeys = total;
      s_pool = total;
      if (total > (uint32_t)NP) s_fail |= 2;  // more groups than pool entries
    }
    __syncthreads();
  }
  const bool built = s_fail == 0;
  if (built && threadIdx.x == 0 && s_nkeys) atomicAdd(&flags[2], s_nkeys);

  // probe + aggregate: one entry per lane and iteration, the next entry's
  // loads issued before this one's update (an empty partition runs no
  // iteration, and loads nothing)
  for (; built && e < q1; e += kJoinThreads) {
    const int64_t key = nkey;
    const uint32_t row = nrow;
    const uint64_t va = nva, vb = nvb, vc = nvc;
    if (e + kJoinThreads < q1) fetch(e + kJoinThreads);
    int sl = -1;
    if (key == kEmptyKey) {
      sl = tag[cap] ? cap : -1;
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
    uint32_t pi = tag[sl];
    if (!(pi & kClaimed)) {
      // first match of this key: the lane whose compare-and-swap marks the
      // slot pending takes the next pool entry; the others wait for it, so a
      // key claims exactly one entry
      const uint32_t old = atomicCAS(&tag[sl], pi, kPending);
      if (old == pi) {
        const uint32_t mine = atomicAdd(&s_pool, 1u);
        if (mine < (uint32_t)NP) {
          pslot[mine] = (uint32_t)sl;
          pmult[mine] = pi;
        }
        pi = kClaimed | (mine < (uint32_t)NP ? mine : (uint32_t)NP);
        __hip_atomic_store(&tag[sl], pi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        pi = old;
      }
    }
    // (the tag read above may itself be another lane's pending mark)
    while (pi == kPending)
      pi = __hip_atomic_load(&tag[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    pi &= ~kClaimed;
    if (pi >= (uint32_t)NP) {  // the pool is full
      s_fail = 2;
      continue;
    }
    atomicAdd(&cnt[myrep * NP + pi], 1u);
    atomicMin(&first[myrep * NP + pi], row);
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      const int vi = a.vidx;
      if (vi < 0) continue;
      const uint64_t v = vi == 0 ? va : vi == 1 ? vb : vi == 2 ? vc : B.p.fvals[(int64_t)vi * vstride + e];
      unsigned long long* w = &acc[((size_t)myrep * NP + pi) * NW + a.word0];
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
  __syncthreads();
  const uint32_t fail = s_fail;
  if (fail && threadIdx.x == 0) atomicOr(&flags[0], fail);
  const uint32_t used = s_pool < (uint32_t)NP ? s_pool : (uint32_t)NP;
  // fold the replicas into replica 0 (integer adds, minima and maxima: the
  // same result in any order), then count the groups
  if (R > 1 && !fail) {
    for (uint32_t x = threadIdx.x; x < used * (uint32_t)NW; x += kJoinThreads) {
      const uint32_t i = x / NW, w = x % NW;
      const int op = s_wop[w];
      unsigned long long v = acc[(size_t)i * NW + w];
      for (int r = 1; r < R; r++) {
        const unsigned long long o = acc[((size_t)r * NP + i) * NW + w];
        if (op == 0) v += o;
        else if (op == 1) v = (long long)o < (long long)v ? o : v;
        else v = (long long)o > (long long)v ? o : v;
      }
      acc[(size_t)i * NW + w] = v;
    }
    for (uint32_t i = threadIdx.x; i < used; i += kJoinThreads) {
      uint32_t c = cnt[i], f = first[i];
      for (int r = 1; r < R; r++) {
        c += cnt[r * NP + i];
        f = first[r * NP + i] < f ? first[r * NP + i] : f;
      }
      cnt[i] = c;
      first[i] = f;
    }
    __syncthreads();
  }
  for (uint32_t i = threadIdx.x; !fail && i < used; i += kJoinThreads)
    if (cnt[i]) atomicAdd(&s_nrec, 1u);
  __syncthreads();
  const int RW = 3 + jd.naggs;

  if (plan.topk == 0) {
    // every matched key as a group record
    if (fail) return;
    if (threadIdx.x == 0) s_base = s_nrec ? atomicAdd(&flags[1], s_nrec) : 0;
    __syncthreads();
    if (s_nrec == 0) return;
    __syncthreads();
    if (threadIdx.x == 0) s_nrec = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < used; i += kJoinThreads) {
      if (cnt[i] == 0) continue;
      const uint32_t sl = pslot[i];
      const int64_t key = (int)sl == cap ? kEmptyKey : (int64_t)keys[sl];
      const uint32_t at = s_base + atomicAdd(&s_nrec, 1u);
      emit_record(jd, B.recs + (size_t)at * RW, key, cnt[i], pmult[i], first[i],
                  &acc[(size_t)i * NW]);
    }
    return;
  }

  // top-k: the partition's best k as candidates (records, and their sort
  // keys in rank order), then the last workgroup to finish merges them
  const int k = plan.topk;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  if (!fail && threadIdx.x == 0 && s_nrec) atomicAdd(&flags[1], s_nrec);
  if (fail || used <= (uint32_t)kRankMax) {
    // few groups: each one's rank is the number of groups ahead of it
    KItem* items = ritems;
    const uint32_t nu = fail ? 0u : used;
    for (uint32_t i = threadIdx.x; i < nu; i += kJoinThreads) {
      KItem x = knone();
      if (cnt[i] != 0) {
        x.v = jd.naggs > 0 ? (int64_t)agg_word_of(jd.agg[0], cnt[i], pmult[i], &acc[(size_t)i * NW]) : 0;
        x.first = first[i];
        x.idx = i;
      }
      items[i] = x;
    }
    __syncthreads();
    const uint32_t n = prune_topk(items, nu, k, &s_nkeys, &s_cand[0]);
    for (uint32_t a = threadIdx.x; a < n; a += kJoinThreads) {
      const KItem x = items[a];
      const uint32_t r = rank_in(items, n, x);
      if (r >= (uint32_t)k) continue;
      const uint32_t c = (uint32_t)p * k + r;
      const uint32_t i = x.idx, sl = pslot[i];
      const int64_t key = (int)sl == cap ? kEmptyKey : (int64_t)keys[sl];
      emit_record<true>(jd, B.cand + (size_t)c * RW, key, cnt[i], pmult[i], first[i], &acc[(size_t)i * NW]);
      KItem ci = x;
      ci.idx = c;
      st_item(&B.citems[c], ci);
    }
    for (uint32_t r = n + threadIdx.x; r < (uint32_t)k; r += kJoinThreads)
      st_item(&B.citems[(uint32_t)p * k + r], knone());
  } else {
    // many groups: the k-th best of a sample (the first 64 groups) bounds
    // the best k from below; the groups at or ahead of it are ranked by
    // counting, in the ranking room.  When more survive than the room holds:
    // each wave's best k by rounds of wave maxima, then wave 0's best k of
    // those.
    KItem it[kPoolPerThread];
#pragma unroll
    for (int j = 0; j < kPoolPerThread; j++) {
      const uint32_t i = threadIdx.x + j * kJoinThreads;
      it[j] = knone();
      if (!fail && i < used && cnt[i] != 0) {
        it[j].v = jd.naggs > 0 ? (int64_t)agg_word_of(jd.agg[0], cnt[i], pmult[i], &acc[(size_t)i * NW])
                               : 0;
        it[j].first = first[i];
        it[j].idx = i;
      }
    }
    static_assert(kPoolPerThread == 2, "items per lane");
    if (wv == 0) wave_kth(it[0], k, &s_bound);
    if (threadIdx.x == 0) s_nkeys = 0;
    __syncthreads();
    const KItem t = s_bound;
#pragma unroll
    for (int j = 0; j < kPoolPerThread; j++) {
      const bool keep = it[j].idx != kNone && (t.idx == kNone || !kbetter(t, it[j]));
      const uint32_t at = wave_append(keep, &s_nkeys);
      if (keep && at < (uint32_t)kRankMax) ritems[at] = it[j];
    }
    __syncthreads();
    const uint32_t m = s_nkeys;
    if (m <= (uint32_t)kRankMax) {
      for (uint32_t a = threadIdx.x; a < m; a += kJoinThreads) {
        const KItem x = ritems[a];
        const uint32_t r = rank_in(ritems, m, x);
        if (r >= (uint32_t)k) continue;
        const uint32_t c = (uint32_t)p * k + r;
        const uint32_t i = x.idx, sl = pslot[i];
        const int64_t key = (int)sl == cap ? kEmptyKey : (int64_t)keys[sl];
        emit_record<true>(jd, B.cand + (size_t)c * RW, key, cnt[i], pmult[i], first[i], &acc[(size_t)i * NW]);
        KItem ci = x;
        ci.idx = c;
        st_item(&B.citems[c], ci);
      }
      for (uint32_t r = m + threadIdx.x; r < (uint32_t)k; r += kJoinThreads)
        st_item(&B.citems[(uint32_t)p * k + r], knone());
    } else {
    KItem w = wave_topk_two(it[0], it[1], k);
    if (lane < k) s_cand[wv * k + lane] = w;
    __syncthreads();
    if (wv == 0) {
      w = wave_topk_mem(s_cand, 0, (uint32_t)(kWaves * k), k);
      if (lane < k) {
        const uint32_t c = (uint32_t)p * k + lane;
        KItem ci = knone();
        if (w.idx != kNone) {
          const uint32_t i = w.idx, sl = pslot[i];
          const int64_t key = (int)sl == cap ? kEmptyKey : (int64_t)keys[sl];
          emit_record<true>(jd, B.cand + (size_t)c * RW, key, cnt[i], pmult[i], first[i], &acc[(size_t)i * NW]);
          ci = w;
          ci.idx = c;
        }
        st_item(&B.citems[c], ci);
      }
    }
    }
  }
  // The merge: the last workgroup to finish takes every partition's
  // candidates.  Candidates and records go through the device's coherence
  // point (st_dev / ld_dev): every wave waits until its stores are
  // acknowledged, then after a barrier one thread signals.  (A device-scope
  // fence instead writes back the whole L2, which costs microseconds.)  The
  // bound: each lane of wave 0 takes the best of its share of the lists'
  // first items, and the k-th best of those 64 is at or ahead of the k-th
  // best of all (k distinct items are at or ahead of it); the items at or
  // ahead of it are ranked by counting.  When more survive than the ranking
  // room holds, wave 0 selects over all items by rounds of wave maxima.
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (gridDim.x > 1) {
    if (threadIdx.x == 0) s_last = atomicAdd(&B.tctr[0], 1u) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
  }
  topk_merge<kJoinThreads>(B.citems, B.cand, B.out, k, RW, gridDim.x, ritems, &s_bound, &s_nkeys, flags);
}

// ---- the direct join's groups ----------------------------------------------
// (the direct join, DirectBufs: k_dj_mark/k_dj_pack set the passing build
// keys' bits, vq_dj_probe gives each match a slot of its key and lists each
// met key once, in the list of the probe workgroup that met it first.)
// k_dj_collect, workgroup w of the probe's grid: the keys of list w, one per
// thread at a time.  A thread reads its key's matches (the entries of its
// slots, loaded together, then any overflow chain) and sums them exactly in
// its own LDS words (count, first probe row, fixed-point limbs,
// integer sums, order-mapped MIN/MAX: the order of the chain changes no bit),
// then
//   all groups (plan.topk 0): writes the group's record (key, count, first
//     row, the aggregates' words, as k_join_agg writes them) at the list's
//     place in the output (the lists before it, summed: no atomics);
//   top-k: offers it to the workgroup's running best k (value of aggregate 0
//     DESC, first row ASC), ranked by counting, whose records stay in LDS;
//     the best k become the workgroup's candidates and the last workgroup
//     to finish merges every workgroup's (topk_merge).
// Then the workgroup clears the counts of its keys and its share of the bits
// (the region is zero for the next query) and the last workgroup reports
// the flags into B.flags (bit 2 of flags[0] when a passing build key repeats:
// the multiplicities are not counted here, the host runs such a join through
// the partitions) and resets the region's.  Dynamic LDS: the threads' words,
// kDjThreads * (nwords + 1).
constexpr int kDjThreads = 256;
constexpr int kDjRec = 3 + VQ_MAX_AGGS;  // the longest record

// Folds match entry e (its probe row, its values: the first given, v0) into
// the first row and the accumulator words w (k_join_agg's encoding).
__device__ __forceinline__ void dj_fold(const JDesc& jd, const DirectBufs& D, uint32_t e, uint32_t row,
                                        uint64_t v0, unsigned long long* w, uint32_t& first) {
  first = row < first ? row : first;
  for (int j = 0; j < jd.naggs; j++) {
    const JAgg& a = jd.agg[j];
    if (a.vidx < 0) continue;
    const uint64_t v = a.vidx == 0 ? v0 : D.ent[(uint64_t)e * D.ew + 1 + a.vidx];
    unsigned long long* x = w + a.word0;
    switch (a.kind) {
      case VQ_AGG_SUM:
      case VQ_AGG_AVG:
        if (a.e.is_float) {
          int64_t c0, c1, c2;
          if (fix_chunks(bits_to_f64((int64_t)v), a.fs, c0, c1, c2)) {
            x[0] += (unsigned long long)c0;
            x[1] += (unsigned long long)c1;
            x[2] += (unsigned long long)c2;
          }
        } else {
          x[0] += (unsigned long long)v;
        }
        break;
      case VQ_AGG_MIN:
      case VQ_AGG_MAX: {
        const double dv = bits_to_f64((int64_t)v);
        if (a.accum == VQ_F64 && dv != dv) break;  // NaN never wins a compare
        const int64_t ov = a.accum == VQ_F64 ? f64_ord(dv) : (int64_t)v;
        if (a.kind == VQ_AGG_MIN ? ov < (int64_t)x[0] : ov > (int64_t)x[0]) x[0] = (unsigned long long)ov;
        break;
      }
      default:
        break;
    }
  }
}

// A key's group: first row and the accumulator words in w[0, NW)
// (k_join_agg's encoding) from its c matches: the entries of its slots (sl,
// read with its record), each entry's row and first value in one load, then,
// past kDjSlots matches, its overflow chain from ovh.
__device__ __forceinline__ void dj_walk(const JDesc& jd, const DirectBufs& D, const uint8_t* wop, uint32_t c,
                                        uint32_t ovh, const uint32_t (&sl)[kDjSlots], unsigned long long* w,
                                        uint32_t& first) {
  const int NW = jd.nwords;
  for (int i = 0; i < NW; i++)
    w[i] = wop[i] == 1 ? (unsigned long long)INT64_MAX : wop[i] == 2 ? (unsigned long long)INT64_MIN : 0ull;
  first = 0xFFFFFFFFu;
  const uint32_t ns = c < (uint32_t)kDjSlots ? c : (uint32_t)kDjSlots;
  const bool two = D.ew == 2;  // entries of 16 bytes: row and value in one load
  uint64_t lk[kDjSlots], va[kDjSlots];
#pragma unroll
  for (int q = 0; q < kDjSlots; q++) {
    lk[q] = 0;
    va[q] = 0;
    if ((uint32_t)q >= ns) continue;
    const uint64_t* ep = D.ent + (uint64_t)sl[q] * D.ew;
    if (two) {
      const ulonglong2 x = *(const ulonglong2*)ep;
      lk[q] = x.x;
      va[q] = x.y;
    } else {
      lk[q] = ep[0];
      va[q] = D.nvals > 0 ? ep[1] : 0ull;
    }
  }
#pragma unroll
  for (int q = 0; q < kDjSlots; q++)
    if ((uint32_t)q < ns) dj_fold(jd, D, sl[q], (uint32_t)lk[q], va[q], w, first);
  for (uint32_t e1 = c > (uint32_t)kDjSlots ? ovh : 0u; e1 != 0;) {
    const uint32_t e = e1 - 1;
    const uint64_t* ep = D.ent + (uint64_t)e * D.ew;
    const uint64_t link = ep[0];
    e1 = (uint32_t)(link >> 32);
    dj_fold(jd, D, e, (uint32_t)link, D.nvals > 0 ? ep[1] : 0ull, w, first);
  }
}

// The sum of a[0, n) by the workgroup (kDjThreads threads) into *acc (zero),
// eight loads per thread in flight at a time.
__device__ __forceinline__ uint32_t dj_sum(const uint32_t* a, int64_t n, uint32_t* acc) {
  uint32_t x = 0;
  for (int64_t i0 = threadIdx.x; i0 < n; i0 += 8 * kDjThreads) {
    uint32_t y[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int64_t i = i0 + u * kDjThreads;
      y[u] = i < n ? a[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) x += y[u];
  }
  for (int o = 32; o > 0; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
  if ((threadIdx.x & 63) == 0 && x) atomicAdd(acc, x);
  __syncthreads();
  return *acc;
}

__global__ __launch_bounds__(kDjThreads, 4) void k_dj_collect(const JDesc jd, const JoinPlan plan,
                                                              const JoinBufs B) {
  extern __shared__ unsigned long long s_w[];
  __shared__ uint32_t s_last, s_nbest, s_nkeys, s_sum[3];
  __shared__ KItem s_bound;
  __shared__ KItem s_best[kMaxK];
  __shared__ KItem s_cand[kDjThreads + kMaxK];
  __shared__ KItem ritems[kRankMax];
  __shared__ uint32_t s_key[kDjThreads], s_cnt[kDjThreads], s_first[kDjThreads];
  __shared__ unsigned long long s_brec[2][kMaxK * kDjRec];  // the best's records (double-buffered)
  __shared__ uint8_t s_wop[3 * VQ_MAX_AGGS];
  const int k = plan.topk;
  const int NW = jd.nwords, RW = 3 + jd.naggs, WS = NW + 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t base = B.d.base;
  const uint32_t* tlist = B.d.tlist;
  const int64_t region = (int64_t)blockIdx.x * B.d.ntiles / gridDim.x * B.d.tile_rows;
  unsigned long long* w = s_w + (size_t)threadIdx.x * WS;
  if (threadIdx.x == 0) {
    s_nbest = 0;
    s_sum[0] = s_sum[1] = s_sum[2] = 0;
  }
  if ((int)threadIdx.x < NW) {
    uint8_t op = 0;
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      if ((int)threadIdx.x >= a.word0 && (int)threadIdx.x < a.word0 + a.nwords)
        op = a.kind == VQ_AGG_MIN ? 1 : a.kind == VQ_AGG_MAX ? 2 : 0;
    }
    s_wop[threadIdx.x] = op;
  }
  // (the list's first keys loaded with its length: the region is in bounds)
  const uint32_t n = B.d.tcount[blockIdx.x];
  const uint32_t o_first = tlist[region + threadIdx.x];
  __syncthreads();
  // workgroup 0 sums the counters while the others collect: the groups, the
  // passing build rows and the distinct passing keys, for the last workgroup
  if (blockIdx.x == 0) {
    const uint32_t g = dj_sum(B.d.tcount, gridDim.x, &s_sum[0]);
    const uint32_t b = dj_sum(B.d.bcnt, B.d.nbcnt, &s_sum[1]);
    const uint32_t d = dj_sum(B.d.pcnt, B.d.npcnt, &s_sum[2]);
    if (threadIdx.x == 0) {
      st_dev((uint64_t*)&B.d.rflags[4], (uint64_t)g);
      st_dev((uint64_t*)&B.d.rflags[6], (uint64_t)b | (uint64_t)d << 32);
    }
    if (threadIdx.x == 0) s_sum[0] = 0;
  }
  // all groups: this list's records follow the lists before it
  const uint32_t out0 = k == 0 ? dj_sum(B.d.tcount, blockIdx.x, &s_sum[0]) : 0;
  int cur = 0;
  for (uint32_t i0 = 0; i0 < n; i0 += kDjThreads) {
    const uint32_t i = i0 + threadIdx.x;
    const bool mine = i < n;
    uint32_t c = 0, first = 0, o = 0;
    if (mine) {
      o = i0 == 0 ? o_first : tlist[region + i];
      // the key's record (count, overflow head, slots) in one load (slots
      // past the count are not read as entries)
      uint4* rp = (uint4*)(B.d.rec + (uint64_t)o * 8);
      const uint4 r0 = rp[0], r1 = rp[1];
      c = r0.x;
      const uint32_t sl[kDjSlots] = {r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
      dj_walk(jd, B.d, s_wop, c, r0.y, sl, w, first);
      // this key is done: its count and overflow head cleared now
      *(uint64_t*)rp = 0;
    }
    if (k == 0) {
      if (mine) emit_record(jd, B.recs + (size_t)(out0 + i) * RW, base + (int64_t)o, c, 1, first, w);
      continue;
    }
    // top-k: this round's items (ritems[t], t a thread) and the running best
    // (ritems[kDjThreads + r]); the k-th best of the first 64 bounds the best
    // k from below (k valid items are at or ahead of it), the items at or
    // ahead of it are ranked by counting.  idx names the item's record: a
    // thread of this round (< kDjThreads), or kDjThreads + its rank in the
    // running best.
    KItem x = knone();
    if (mine) {
      x.v = jd.naggs > 0 ? (int64_t)agg_word_of(jd.agg[0], c, 1, w) : 0;
      x.first = first;
      x.idx = threadIdx.x;
      s_key[threadIdx.x] = o;
      s_cnt[threadIdx.x] = c;
      s_first[threadIdx.x] = first;
    }
    const uint32_t nb = s_nbest;
    ritems[threadIdx.x] = x;
    if (threadIdx.x < nb) {
      KItem y = s_best[threadIdx.x];
      y.idx = kDjThreads + threadIdx.x;
      ritems[kDjThreads + threadIdx.x] = y;
    }
    if (threadIdx.x == 0) s_nkeys = 0;
    __syncthreads();
    if (wv == 0) wave_kth(ritems[lane], k, &s_bound);
    __syncthreads();
    const KItem t = s_bound;
    KItem* sv = s_cand;  // the survivors (at most kDjThreads + kMaxK)
    {
      const bool keep = x.idx != kNone && (t.idx == kNone || !kbetter(t, x));
      const uint32_t at = wave_append(keep, &s_nkeys);
      if (keep) sv[at] = x;
      KItem y = threadIdx.x < nb ? ritems[kDjThreads + threadIdx.x] : knone();
      const bool keepb = y.idx != kNone && (t.idx == kNone || !kbetter(t, y));
      const uint32_t atb = wave_append(keepb, &s_nkeys);
      if (keepb) sv[atb] = y;
    }
    __syncthreads();
    const uint32_t m = s_nkeys;
    const unsigned long long* orec = s_brec[cur];
    unsigned long long* nrec = s_brec[cur ^ 1];
    for (uint32_t a = threadIdx.x; a < m; a += kDjThreads) {
      const KItem y = sv[a];
      const uint32_t r = rank_in(sv, m, y);
      if (r >= (uint32_t)k) continue;
      s_best[r] = y;
      unsigned long long* d = nrec + r * kDjRec;
      if (y.idx >= (uint32_t)kDjThreads) {
        for (int q = 0; q < RW; q++) d[q] = orec[(y.idx - kDjThreads) * kDjRec + q];
      } else {
        const uint32_t tt = y.idx;
        d[0] = (unsigned long long)(base + (int64_t)s_key[tt]);
        d[1] = s_cnt[tt];
        d[2] = s_first[tt];
        for (int j = 0; j < jd.naggs; j++) d[3 + j] = agg_word_of(jd.agg[j], s_cnt[tt], 1, s_w + (size_t)tt * WS);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_nbest = m < (uint32_t)k ? m : (uint32_t)k;
    cur ^= 1;
    __syncthreads();
  }
  if (k > 0 && threadIdx.x < (uint32_t)k) {
    // this workgroup's candidates: its best k and their records
    const uint32_t cc = blockIdx.x * (uint32_t)k + threadIdx.x;
    KItem ci = knone();
    if (threadIdx.x < s_nbest) {
      for (int q = 0; q < RW; q++) st_dev(&B.cand[(size_t)cc * RW + q], s_brec[cur][threadIdx.x * kDjRec + q]);
      ci = s_best[threadIdx.x];
      ci.idx = cc;
    }
    st_item(&B.citems[cc], ci);
  }
  // this workgroup's share of the bits, cleared (the region is zero for the
  // next query)
  {
    const uint64_t w0 = (uint64_t)blockIdx.x * B.d.words / gridDim.x;
    const uint64_t w1 = (uint64_t)(blockIdx.x + 1) * B.d.words / gridDim.x;
    for (uint64_t q = w0 + threadIdx.x; q < w1; q += kDjThreads) B.d.pres[q] = 0;
  }
  // the last workgroup: the merge, the flags
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&B.d.rflags[2], 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  if (k > 0)
    topk_merge<kDjThreads, 12>(B.citems, B.cand, B.out, k, RW, gridDim.x, ritems, &s_bound, &s_nkeys,
                               B.flags);
  if (threadIdx.x == 0) {
    const uint32_t groups = (uint32_t)ld_dev((const uint64_t*)&B.d.rflags[4]);
    const uint64_t bd = ld_dev((const uint64_t*)&B.d.rflags[6]);
    const uint32_t built = (uint32_t)bd, distinct = (uint32_t)(bd >> 32);
    B.flags[0] = __hip_atomic_load(&B.d.rflags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
                 (distinct != built ? 2u : 0u);
    B.flags[1] = groups;
    B.flags[2] = distinct;
    if (k == 0) B.flags[3] = groups;
    B.flags[4] = 0;
    B.flags[5] = 0;
    B.flags[6] = built;
    B.flags[7] = 0;
  }
  __syncthreads();
  if (threadIdx.x < 16) B.d.rflags[threadIdx.x] = 0;
}

hipError_t launch_direct(const QDesc& bq, int32_t bkey_slot, const DirectLaunch& dl, const JDesc& jd,
                         const JoinPlan& plan, const JoinBufs& B, hipStream_t s, hipEvent_t* ev) {
  if (ev) hipEventRecord(ev[0], s);
  hipError_t e = launch_dj_build(bq, bkey_slot, B.d, s);
  if (e != hipSuccess) return e;
  if (ev) hipEventRecord(ev[1], s);
  e = jit_direct_launch(dl, B.d, s);
  if (e != hipSuccess) return e;
  if (ev) hipEventRecord(ev[2], s);
  // (one workgroup per probe workgroup: list w is workgroup w's)
  const size_t lds = (size_t)kDjThreads * (jd.nwords + 1) * 8;
  hipLaunchKernelGGL(k_dj_collect, dim3((unsigned)B.d.grid), dim3(kDjThreads), lds, s, jd, plan, B);
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

size_t join_lds_for(int32_t cap, int32_t pool, int32_t nwords, int32_t reps, bool topk) {
  const size_t ns = (size_t)cap + 1;
  return ns * (8 + 4) + (size_t)pool * (4 + 4) + (size_t)reps * pool * (4 + 4 + 8 * (size_t)nwords) +
         (topk ? 16 * (size_t)kRankMax + 16 : 0);
}

int32_t topk_tree_nodes(int32_t P) { return P > 1 ? 1 : 0; }

uint32_t part_scatter_lds(int32_t P, int32_t pbuf, int32_t entry) {
  if (pbuf <= 0) return P <= 16384 ? (uint32_t)P * 4 : 0;
  // fill + cur words and the keys; on the probe side (entries beyond the
  // 8-byte key: a 4-byte row id and 8-byte values) the row ids, rounded up to
  // an even word count, and the values
  const int64_t n = (int64_t)P * pbuf;
  int64_t words = 2 * (int64_t)P + 2 * n;
  if (entry > 8) words += ((n + 1) & ~(int64_t)1) + n * ((entry - 12) / 4);
  return (uint32_t)(words * 4);
}

int32_t part_buffer_entries(int32_t P, int32_t entry, int32_t lds_max) {
  int32_t want = 16;
  if (const char* e = getenv("VQ_PART_BUF")) want = atoi(e);
  const int64_t limit = (int64_t)lds_max - 1024;
  for (int32_t b = 16; b >= 1; b >>= 1)
    if (b <= want && (int64_t)part_scatter_lds(P, b, entry) <= limit) return b;
  return 0;
}

void join_allow_lds(int32_t bytes) {
  static std::atomic<int32_t> allowed{0};
  if (allowed.load() >= bytes) return;
  hipFuncSetAttribute((const void*)k_join_agg, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  allowed.store(bytes);
}

// ---- GROUP BY of few keys through per-workgroup partials ---------------------
// (ParallelHashAggregate's per-worker partials and mergePartialAgg,
// exec/parallel.go:18-314.)  Four launches:
//   k_local_init   the key table to empty, the flags to zero
//   k_local_keys   every distinct key of the scanned rows (predicates
//                  aside: a key whose rows all fail them makes a group of no
//                  rows, which emits nothing) into one table of lcap slots;
//                  more than `limit` keys sets flags[0] bit 4, and the query
//                  runs through the partitions instead
//   vq_hash_local  (run-time compiled) each workgroup aggregates an even
//                  share of the rows into LDS accumulators per group index
//                  and writes its partials
//   k_local_merge  one block per group: the workgroups' partials combined,
//                  the group's record emitted (the same record as k_join_agg)
constexpr uint32_t kLocalOverflow = 4u;
constexpr int kLocalThreads = 512;

struct LocalKeys {
  const void* col;
  int32_t type;
  // several dictionary-code columns packed into one key (JSide.nkeys > 1)
  int32_t nkeys;
  const void* kcol[8];
  int32_t ktype[8];
  int32_t kshift[8];
  int32_t lcap;
  int32_t limit;
  int32_t nspans;
  int64_t tile_rows;
  int64_t ntiles;
  int64_t sp_start[kMaxSpans];
  int64_t sp_rows[kMaxSpans];
  int64_t sp_tile0[kMaxSpans];
  int64_t* keys;
  uint32_t* flags;
  // dictionary-code keys: seed the table with every code combination
  // (card[c] codes of column c) instead of reading the rows
  int64_t seed;
  int32_t card[8];
};

__global__ __launch_bounds__(256) void k_local_init(int64_t* __restrict__ keys, int32_t lcap,
                                                    uint32_t* __restrict__ flags) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < lcap; i += gridDim.x * 256) keys[i] = kEmptyKey;
  if (blockIdx.x == 0 && threadIdx.x < 16) flags[threadIdx.x] = 0;
}

// A key column's value as the run-time kernels read it: codes zero-extended,
// dates sign-extended, 64-bit columns as their bits.
__device__ __forceinline__ int64_t load_key(const void* col, int32_t type, int64_t r) {
  switch (type) {
    case VQ_U8: return ((const uint8_t*)col)[r];
    case VQ_U16: return ((const uint16_t*)col)[r];
    case VQ_DATE: return ((const int32_t*)col)[r];
    default: return ((const int64_t*)col)[r];
  }
}

// Inserts key into an open-addressing table of cap slots; returns 1 if it
// was new, 0 if present, -1 if it found no place within kMaxProbes slots (a
// table filled past its key limit: a full one would cost a walk over every
// slot for each key).  At the tables' load (at most one half while within
// the limit) a probe sequence that long does not occur in practice; when it
// does, the query runs through the partitions.
constexpr int kMaxProbes = 128;
template <int SCOPE>
__device__ __forceinline__ int table_insert(unsigned long long* t, int cap, int lg, int64_t key) {
  uint32_t i = (uint32_t)(radix_hash((uint64_t)key) >> (64 - lg));
  const int np = cap < kMaxProbes ? cap : kMaxProbes;
  for (int pr = 0; pr < np; pr++) {
    unsigned long long cur = __hip_atomic_load(&t[i], __ATOMIC_RELAXED, SCOPE);
    if (cur == (unsigned long long)key) return 0;
    if (cur == (unsigned long long)kEmptyKey) {
      cur = atomicCAS(&t[i], (unsigned long long)kEmptyKey, (unsigned long long)key);
      if (cur == (unsigned long long)kEmptyKey) return 1;
      if (cur == (unsigned long long)key) return 0;
    }
    i = (i + 1) & (uint32_t)(cap - 1);
  }
  return -1;
}

// Each workgroup collects the keys of its tiles in LDS (a key repeated over
// millions of rows costs a read of its slot, not an atomic), then adds the
// ones it found to the global table.
__global__ __launch_bounds__(kLocalThreads) void k_local_keys(const LocalKeys k) {
  extern __shared__ unsigned long long lt[];
  __shared__ uint32_t s_n, s_min, s_stop;
  const int cap = k.lcap, lg = __ffs(cap) - 1;
  for (int i = threadIdx.x; i < cap; i += kLocalThreads) lt[i] = (unsigned long long)kEmptyKey;
  if (threadIdx.x == 0) {
    s_n = 0;
    s_min = 0;
    s_stop = 0;
  }
  __syncthreads();
  for (int64_t t = blockIdx.x; t < k.ntiles; t += gridDim.x) {
    int s = 0;
    while (s + 1 < k.nspans && t >= k.sp_tile0[s + 1]) s++;
    const int64_t row0 = k.sp_start[s] + (t - k.sp_tile0[s]) * k.tile_rows;
    const int64_t send = k.sp_start[s] + k.sp_rows[s];
    const int64_t rend = row0 + k.tile_rows < send ? row0 + k.tile_rows : send;
    bool over = false;
    for (int64_t r = row0 + threadIdx.x; r < rend && !over; r += kLocalThreads) {
      int64_t key;
      if (k.nkeys > 1) {
        uint64_t pk = 0;
        for (int c = 0; c < k.nkeys; c++) pk |= (uint64_t)load_key(k.kcol[c], k.ktype[c], r) << k.kshift[c];
        key = (int64_t)pk;
      } else {
        key = load_key(k.col, k.type, r);
      }
      if (key == kEmptyKey) {
        s_min = 1;
        continue;
      }
      // (past the limit every thread stops: a table that fills makes each
      // further insert probe every slot)
      if (__hip_atomic_load(&s_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > (uint32_t)k.limit) {
        over = true;
        break;
      }
      const int got = table_insert<__HIP_MEMORY_SCOPE_WORKGROUP>(lt, cap, lg, key);
      if (got < 0 || (got > 0 && atomicAdd(&s_n, 1u) + 1 > (uint32_t)k.limit)) over = true;
    }
    if (over) s_stop = 1;
    __syncthreads();
    // stop early once this workgroup or another has seen too many keys
    if (threadIdx.x == 0 && (__hip_atomic_load(&k.flags[0], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT) & kLocalOverflow))
      s_stop = 1;
    __syncthreads();
    if (s_stop) break;
  }
  if (s_stop) {
    if (threadIdx.x == 0) atomicOr(&k.flags[0], kLocalOverflow);
    return;
  }
  for (int i = threadIdx.x; i < cap; i += kLocalThreads) {
    const int64_t key = (int64_t)lt[i];
    if (key == kEmptyKey) continue;
    // (past the limit, stop: another workgroup's keys already overflowed)
    if (__hip_atomic_load(&k.flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kLocalOverflow) break;
    const int got = table_insert<__HIP_MEMORY_SCOPE_AGENT>((unsigned long long*)k.keys, cap, lg, key);
    if (got < 0 || (got > 0 && atomicAdd(&k.flags[2], 1u) + 1 > (uint32_t)k.limit))
      atomicOr(&k.flags[0], kLocalOverflow);
  }
  if (threadIdx.x == 0 && s_min && atomicCAS(&k.flags[10], 0u, 1u) == 0u &&
      atomicAdd(&k.flags[2], 1u) + 1 > (uint32_t)k.limit)
    atomicOr(&k.flags[0], kLocalOverflow);
}

// Every combination of the key columns' codes into the key table (a group
// whose rows all fail the predicate, or meet no build row, has a count of
// zero and leaves no record): for dictionary-code keys whose code space is
// within the per-workgroup path's limit, no pass over the rows.
__global__ __launch_bounds__(256) void k_local_seed(const LocalKeys k) {
  const int cap = k.lcap, lg = __ffs(cap) - 1;
  for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < k.seed; x += (int64_t)gridDim.x * 256) {
    int64_t key = x;
    if (k.nkeys > 1) {
      uint64_t pk = 0;
      int64_t rem = x;
      for (int c = 0; c < k.nkeys; c++) {
        pk |= (uint64_t)(rem % k.card[c]) << k.kshift[c];
        rem /= k.card[c];
      }
      key = (int64_t)pk;
    }
    const int got = table_insert<__HIP_MEMORY_SCOPE_AGENT>((unsigned long long*)k.keys, cap, lg, key);
    if (got > 0) atomicAdd(&k.flags[2], 1u);
    else if (got < 0) atomicOr(&k.flags[0], kLocalOverflow);
  }
}

// The key list: the table's regular keys in slot order (a key's rank is its
// group index), then INT64_MIN when a row has it.  One workgroup.
constexpr int kRankThreads = 1024;
__global__ __launch_bounds__(kRankThreads) void k_local_rank(const int64_t* __restrict__ keys,
                                                             int32_t gcap, uint32_t* __restrict__ flags,
                                                             int64_t* __restrict__ gkey) {
  __shared__ uint32_t s_scan[kRankThreads];
  if (flags[0] & kLocalOverflow) return;
  const int per = (gcap + kRankThreads - 1) / kRankThreads;
  const int s0 = (int)threadIdx.x * per < gcap ? (int)threadIdx.x * per : gcap;
  const int s1 = s0 + per < gcap ? s0 + per : gcap;
  uint32_t mine = 0;
  for (int i = s0; i < s1; i++) mine += keys[i] != kEmptyKey ? 1u : 0u;
  s_scan[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < kRankThreads; off <<= 1) {
    const uint32_t x = (int)threadIdx.x >= off ? s_scan[threadIdx.x - off] : 0;
    __syncthreads();
    s_scan[threadIdx.x] += x;
    __syncthreads();
  }
  uint32_t r = s_scan[threadIdx.x] - mine;
  for (int i = s0; i < s1; i++)
    if (keys[i] != kEmptyKey) gkey[r++] = keys[i];
  if (threadIdx.x == kRankThreads - 1 && flags[10]) gkey[s_scan[kRankThreads - 1]] = kEmptyKey;
}

struct LocalMerge {
  const uint64_t* part;
  const int64_t* gkey;  // the key list
  uint32_t* flags;
  uint64_t* recs;
  int32_t grid;   // workgroups that wrote partials
  int32_t k0;     // this pass's first key
  int32_t limit;  // keys a pass holds
  int32_t pad_;
};

constexpr int kMergeThreads = 1024;

__device__ __forceinline__ unsigned long long merge_op(int op, unsigned long long a,
                                                       unsigned long long b) {
  if (op == 0) return a + b;
  if (op == 1) return (long long)b < (long long)a ? b : a;
  if (op == 2) return (long long)b > (long long)a ? b : a;
  return b < a ? b : a;
}

// One block per group index: each word of the group's partials combined over
// the workgroups (integer adds, minima and maxima) by one wave — its lanes'
// loads all in flight, then a butterfly over the lanes — and then the
// group's record.
__global__ __launch_bounds__(kMergeThreads) void k_local_merge(const JDesc jd, const LocalMerge m) {
  __shared__ unsigned long long s_words[3 * VQ_MAX_AGGS + 2];
  __shared__ uint8_t s_wop[3 * VQ_MAX_AGGS + 2];
  const uint32_t g = blockIdx.x;
  if (m.flags[0] & kLocalOverflow) return;
  const uint32_t total = m.flags[2], k0 = (uint32_t)m.k0;
  if (k0 >= total) return;
  const uint32_t nk = total - k0 < (uint32_t)m.limit ? total - k0 : (uint32_t)m.limit;
  if (g >= nk) return;
  const int NW = jd.nwords, NI = NW + 2;
  if ((int)threadIdx.x < NI) {
    uint8_t op = 0;  // add
    if ((int)threadIdx.x == NI - 1) op = 3;  // first row: unsigned minimum
    for (int j = 0; j < jd.naggs; j++) {
      const JAgg& a = jd.agg[j];
      if ((int)threadIdx.x >= a.word0 && (int)threadIdx.x < a.word0 + a.nwords)
        op = a.kind == VQ_AGG_MIN ? 1 : a.kind == VQ_AGG_MAX ? 2 : 0;
    }
    s_wop[threadIdx.x] = op;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i < NI; i += kMergeThreads / 64) {
    const int op = s_wop[i];
    const unsigned long long id = op == 1 ? (unsigned long long)INT64_MAX
                                  : op == 2 ? (unsigned long long)INT64_MIN
                                  : op == 3 ? 0xFFFFFFFFull : 0ull;
    unsigned long long v = id;
    // eight partials per lane in flight at a time (integer adds, minima and
    // maxima: any order gives the same words)
    for (int w0 = lane; w0 < m.grid; w0 += 64 * 8) {
      unsigned long long x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int w = w0 + u * 64;
        x[u] = w < m.grid ? m.part[((size_t)w * nk + g) * NI + i] : id;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) v = merge_op(op, v, x[u]);
    }
    for (int off = 32; off > 0; off >>= 1)
      v = merge_op(op, v, (unsigned long long)__shfl_xor((long long)v, off, 64));
    if (lane == 0) s_words[i] = v;
  }
  __syncthreads();
  // the record, by wave 0: lane j the word of aggregate j (their roundings
  // run side by side), lane 0 the key, count and first row
  const unsigned long long cnt = s_words[NW];
  if (cnt == 0 || wv != 0) return;  // (a key whose rows all failed the predicate)
  uint32_t at = 0;
  if (lane == 0) at = atomicAdd(&m.flags[1], 1u);
  at = (uint32_t)__shfl((int)at, 0, 64);
  uint64_t* rec = m.recs + (size_t)at * (3 + jd.naggs);
  if (lane == 0) {
    rec[0] = (uint64_t)m.gkey[k0 + g];
    rec[1] = cnt;
    rec[2] = (uint32_t)s_words[NW + 1];
  }
  if (lane < jd.naggs) rec[3 + lane] = agg_word_of(jd.agg[lane], cnt, 1, s_words);
}

size_t local_lds_for(int32_t lcap, int32_t limit, int32_t nwords) {
  return (size_t)lcap * 12 + (size_t)limit * (8 * (size_t)nwords + 8);
}

hipError_t launch_local(const LocalLaunch& ll, const JSide& ps, const JDesc& jd,
                        const JoinPlan& plan, const JoinBufs& B, hipStream_t s, hipEvent_t* ev,
                        bool ev0_done) {
  if (ev && !ev0_done) hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_local_init, dim3((unsigned)std::max(1, plan.gcap / 256)), dim3(256), 0, s,
                     B.l.keys, plan.gcap, B.flags);
  LocalKeys k;
  memset(&k, 0, sizeof k);
  k.col = ps.slot[ps.key_slot].p;
  k.type = ps.slot[ps.key_slot].type;
  k.nkeys = ps.nkeys;
  for (int c = 0; c < ps.nkeys && c < 8; c++) {
    k.kcol[c] = ps.slot[ps.key_slots[c]].p;
    k.ktype[c] = ps.slot[ps.key_slots[c]].type;
    k.kshift[c] = ps.key_shift[c];
  }
  k.lcap = plan.gcap;
  k.limit = plan.limit * plan.passes;
  k.nspans = ps.nspans;
  k.tile_rows = ps.tile_rows;
  k.ntiles = ps.ntiles;
  for (int i = 0; i < ps.nspans; i++) {
    k.sp_start[i] = ps.span[i].start;
    k.sp_rows[i] = ps.span[i].rows;
    k.sp_tile0[i] = ps.span[i].tile0;
  }
  k.keys = B.l.keys;
  k.flags = B.flags;
  {
    // the key table's LDS copy: up to 64 KB
    static std::atomic<int32_t> allowed{0};
    const int32_t bytes = plan.gcap * 8;
    if (allowed.load() < bytes) {
      hipFuncSetAttribute((const void*)k_local_keys, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
      allowed.store(bytes);
    }
  }
  k.seed = plan.seed_groups;
  for (int c = 0; c < 8; c++) k.card[c] = plan.seed_card[c];
  if (plan.seed_groups > 0)
    hipLaunchKernelGGL(k_local_seed, dim3((unsigned)std::min<int64_t>(64, (plan.seed_groups + 255) / 256)),
                       dim3(256), 0, s, k);
  else
    hipLaunchKernelGGL(k_local_keys, dim3((unsigned)plan.kgrid), dim3(kLocalThreads),
                       (size_t)plan.gcap * 8, s, k);
  hipLaunchKernelGGL(k_local_rank, dim3(1), dim3(kRankThreads), 0, s, (const int64_t*)B.l.keys,
                     plan.gcap, B.flags, B.l.gkey);
  if (ev) hipEventRecord(ev[1], s);
  // one pass over the rows per `limit` keys: partials, then their merge
  LocalMerge m;
  m.part = B.l.part;
  m.gkey = B.l.gkey;
  m.flags = B.flags;
  m.recs = B.recs;
  m.grid = ll.grid;
  m.limit = plan.limit;
  m.pad_ = 0;
  for (int c = 0; c < plan.passes; c++) {
    hipError_t e = jit_local_launch(ll, B, s, c * plan.limit);
    if (e != hipSuccess) return e;
    if (ev && c == plan.passes - 1) hipEventRecord(ev[2], s);
    m.k0 = c * plan.limit;
    hipLaunchKernelGGL(k_local_merge, dim3((unsigned)plan.limit), dim3(kMergeThreads), 0, s, jd, m);
  }
  if (ev) hipEventRecord(ev[3], s);
  return hipGetLastError();
}

hipError_t launch_join(const PartLaunch& b, const PartLaunch& p, const JDesc& jd,
                       const JoinPlan& plan, const JoinBufs& B, hipStream_t s, hipEvent_t* ev) {
  const int P = plan.P1;  // the partition kernels' (coarse) partitions
  const int P2 = plan.P / plan.P1;
  auto fine = [&](const PartBufs& side, bool probe) {
    if (P2 > 1)
      hipLaunchKernelGGL(k_part_fine, dim3((unsigned)P), dim3(kFineThreads), 0, s, side, P,
                         plan.bits1, P2, probe ? plan.nvals : 0, probe ? 1 : 0);
  };
  if (ev) hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, B.flags, (int64_t)16 + plan.tnodes);
  const bool has_build = !plan.groupby;
  // the run and start scans of one or two sides (build, then probe)
  auto scans = [&](const PartBufs* s0, uint32_t* t0, const PartBufs* s1, uint32_t* t1) {
    PartScan ps;
    memset(&ps, 0, sizeof ps);
    const PartBufs* sd[2] = {s0, s1};
    uint32_t* tt[2] = {t0, t1};
    const int ns = s1 ? 2 : 1;
    for (int i = 0; i < ns; i++) {
      ps.counts[i] = sd[i]->counts;
      ps.tot[i] = sd[i]->tot;
      ps.pstart[i] = sd[i]->pstart;
      ps.total[i] = tt[i];
      ps.G[i] = sd[i]->G;
    }
    ps.P = P;
    hipLaunchKernelGGL(k_part_runs, dim3((unsigned)P, (unsigned)ns), dim3(kScanThreads), 0, s, ps);
    hipLaunchKernelGGL(k_part_starts, dim3((unsigned)ns), dim3(kScanThreads), 0, s, ps);
  };
  hipError_t e;
  if (has_build && B.bloom) {
    // build side partitioned first, its filter, then the probe side
    if ((e = jit_part_launch(b, B.b, P, s, true)) != hipSuccess) return e;
    scans(&B.b, B.flags + 6, nullptr, nullptr);
    if ((e = jit_part_launch(b, B.b, P, s, false)) != hipSuccess) return e;
    fine(B.b, false);
    hipLaunchKernelGGL(k_bloom_build, dim3((unsigned)plan.P), dim3(256), (size_t)64 << plan.bloom_lpp, s,
                       B.b.fkeys, B.b.fpstart, B.bloom, plan.bloom_lpp, P, P2);
    if ((e = jit_part_launch(p, B.p, P, s, true)) != hipSuccess) return e;
    scans(&B.p, B.flags + 7, nullptr, nullptr);
    if ((e = jit_part_launch(p, B.p, P, s, false)) != hipSuccess) return e;
    fine(B.p, true);
  } else {
    e = has_build ? jit_part_launch(b, B.b, P, s, true) : hipSuccess;
    if (e != hipSuccess) return e;
    if ((e = jit_part_launch(p, B.p, P, s, true)) != hipSuccess) return e;
    if (has_build) scans(&B.b, B.flags + 6, &B.p, B.flags + 7);
    else scans(&B.p, B.flags + 7, nullptr, nullptr);
    e = has_build ? jit_part_launch(b, B.b, P, s, false) : hipSuccess;
    if (e != hipSuccess) return e;
    if ((e = jit_part_launch(p, B.p, P, s, false)) != hipSuccess) return e;
    if (has_build) fine(B.b, false);
    fine(B.p, true);
  }
  if (ev) hipEventRecord(ev[1], s);
  const size_t lds = plan.lds_bytes;
  if (lds > 64 * 1024) join_allow_lds(plan.lds_max);
  hipLaunchKernelGGL(k_join_agg, dim3((unsigned)plan.P), dim3(kJoinThreads), lds, s, jd, plan, B);
  if (ev) {
    hipEventRecord(ev[2], s);
    hipEventRecord(ev[3], s);
  }
  return hipGetLastError();
}

}  // namespace vq
