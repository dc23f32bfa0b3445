// vq_internal.h — descriptors shared by the host dispatch (vq_api.hip) and the
// gfx950 kernels (k_scan.hip, k_join.hip).  Everything here is plain data
// passed to kernels by value, so every field a kernel reads sits in the
// kernarg segment and is fetched with scalar loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/vexq_gpu.h"

namespace vq {

constexpr int kMaxSlots = 8;      // distinct columns one query touches
constexpr int kMaxSlotTerms = 4;  // predicate terms on one column
constexpr int kMaxAcc = 8;        // accumulators after de-duplication
constexpr int kMaxSpans = 24;     // contiguous runs of surviving row groups

// Accumulator operations of the dense (dictionary-code) path.
enum AccOp : int32_t {
  ACC_SUM_F = 0,  // float64 sum
  ACC_SUM_I = 1,  // int64 sum (wrapping, order free)
  ACC_MIN_F = 2,
  ACC_MAX_F = 3,
  ACC_MIN_I = 4,
  ACC_MAX_I = 5,
};

struct QTerm {
  int32_t op;
  int32_t pad_;
  int64_t ival;
  double fval;
};

// One referenced column.
struct QSlot {
  const void* p;
  int32_t type;
  int32_t nterms;
  int32_t key_stride;  // dense group slot contribution (0 = not a key)
  int32_t pad_;
  QTerm terms[kMaxSlotTerms];
};

struct QFactor {
  int32_t kind;
  int32_t slot;
  int32_t op;
  int32_t combine;
  double fval;
  int64_t ival;
};

struct QExpr {
  int32_t nfactors;
  int32_t is_float;
  QFactor f[VQ_MAX_FACTORS];
};

struct QAcc {
  int32_t op;
  int32_t pad_;
  QExpr e;
};

// Rows [start, start+rows) of the table; tiles [tile0, tile0 + ceil(rows/T)).
struct QSpan {
  int64_t start;
  int64_t rows;
  int64_t tile0;
};

struct QDesc {
  int32_t nslots;
  int32_t nacc;
  int32_t ngroups;
  int32_t nspans;
  int32_t tile_rows;
  int32_t pad_;
  int64_t ntiles;
  QSlot slot[kMaxSlots];
  QAcc acc[kMaxAcc];
  QSpan span[kMaxSpans];
};

static_assert(sizeof(QDesc) < 4096, "QDesc must fit the kernarg segment");

// ---- exact fixed-point accumulation (join / high-cardinality path) --------
// A float64 term t is added as the integer T = trunc(t / 2^lsb) split into
// three signed chunks of w bits (limb i carries bits [i w, (i+1) w)).  Each
// int64 limb therefore has 63 - w bits of headroom, enough for 2^(62-w)
// additions.  Integer addition is associative, so any order of atomics gives
// the same bits.
struct FixScale {
  int32_t lsb;  // exponent of the least significant bit
  int32_t w;    // chunk width in bits
};

// Join + aggregate descriptors.
struct JAgg {
  int32_t kind;     // VQ_AGG_*
  int32_t accum;    // VQ_I64 / VQ_F64
  int32_t nwords;   // 64-bit accumulator words this aggregate holds per slot
  int32_t word0;    // first word index within a slot's accumulator block
  FixScale fs;      // for float sums
  QExpr e;
};

struct JSide {
  int32_t nslots;
  int32_t key_slot;
  int32_t nspans;
  int32_t tile_rows;
  int64_t ntiles;
  QSlot slot[kMaxSlots];
  QSpan span[kMaxSpans];
};

struct JDesc {
  int32_t naggs;
  int32_t nwords;  // accumulator words per table slot
  JAgg agg[VQ_MAX_AGGS];
};

// ---- small host/device helpers ---------------------------------------------
__host__ __device__ inline uint64_t radix_hash(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

__host__ __device__ inline double bits_to_f64(int64_t b) {
  union {
    int64_t i;
    double d;
  } u;
  u.i = b;
  return u.d;
}

__host__ __device__ inline int64_t f64_to_bits(double d) {
  union {
    int64_t i;
    double d;
  } u;
  u.d = d;
  return u.i;
}

// Round a signed multi-limb fixed-point value  sum_i limb[i] * 2^(lsb + i w)
// times the non-negative integer mult to the nearest float64 (ties to even).
// Exact: the value is first assembled as a 256-bit two's complement integer.
__host__ __device__ inline double fix_to_f64(const int64_t* limb, int nlimbs, FixScale fs,
                                             uint64_t mult) {
  // acc = sum limb[i] << (i w), in 4 x 64-bit words, two's complement
  uint64_t a[4] = {0, 0, 0, 0};
  for (int i = 0; i < nlimbs; i++) {
    int64_t v = limb[i];
    int sh = i * fs.w;
    uint64_t ext = v < 0 ? ~0ULL : 0ULL;
    // x = sign-extended v shifted left by sh into 256 bits
    uint64_t x[4] = {(uint64_t)v, ext, ext, ext};
    int ws = sh >> 6, bs = sh & 63;
    uint64_t y[4] = {0, 0, 0, 0};
    for (int k = 3; k >= 0; k--) {
      int src = k - ws;
      uint64_t lo = src >= 0 ? x[src] : 0;
      uint64_t lo1 = src - 1 >= 0 ? x[src - 1] : 0;
      y[k] = bs ? (lo << bs) | (lo1 >> (64 - bs)) : lo;
    }
    uint64_t c = 0;
    for (int k = 0; k < 4; k++) {
      uint64_t s = a[k] + y[k];
      uint64_t c1 = s < a[k];
      uint64_t s2 = s + c;
      uint64_t c2 = s2 < s;
      a[k] = s2;
      c = c1 | c2;
    }
  }
  bool neg = (int64_t)a[3] < 0;
  if (neg) {  // negate
    uint64_t c = 1;
    for (int k = 0; k < 4; k++) {
      uint64_t s = ~a[k] + c;
      c = (c && s == 0) ? 1 : 0;
      a[k] = s;
    }
  }
  // multiply magnitude by mult (fits: |value| < 2^(3w+62) and mult < 2^32)
  if (mult != 1) {
    uint64_t m_lo = mult & 0xFFFFFFFFULL, m_hi = mult >> 32;
    uint64_t r[4] = {0, 0, 0, 0};
    // schoolbook over 32-bit halves of a
    uint32_t h[8];
    for (int k = 0; k < 4; k++) {
      h[2 * k] = (uint32_t)a[k];
      h[2 * k + 1] = (uint32_t)(a[k] >> 32);
    }
    uint32_t out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t m[2] = {(uint32_t)m_lo, (uint32_t)m_hi};
    for (int j = 0; j < 2; j++) {
      uint64_t carry = 0;
      for (int k = 0; k + j < 8; k++) {
        uint64_t t = (uint64_t)h[k] * m[j] + out[k + j] + carry;
        out[k + j] = (uint32_t)t;
        carry = t >> 32;
      }
    }
    for (int k = 0; k < 4; k++) r[k] = (uint64_t)out[2 * k] | ((uint64_t)out[2 * k + 1] << 32);
    for (int k = 0; k < 4; k++) a[k] = r[k];
  }
  int top = -1;
  for (int k = 3; k >= 0; k--)
    if (a[k]) {
      top = k * 64 + 63 - __builtin_clzll(a[k]);
      break;
    }
  if (top < 0) return 0.0;
  auto bit = [&](int p) -> uint64_t { return p < 0 ? 0 : (a[p >> 6] >> (p & 63)) & 1ULL; };
  double r;
  if (top < 53) {
    r = (double)a[0];
    r = ldexp(r, fs.lsb);
  } else {
    uint64_t mant = 0;
    for (int p = top; p >= top - 52; p--) mant = (mant << 1) | bit(p);
    uint64_t rb = bit(top - 53);
    uint64_t sticky = 0;
    int lim = top - 54;
    for (int k = 0; k < 4 && !sticky; k++) {
      int lo = k * 64, hi = lo + 63;
      if (lo > lim) break;
      if (hi <= lim) sticky = a[k] != 0;
      else {
        uint64_t mask = (lim - lo + 1) >= 64 ? ~0ULL : ((1ULL << (lim - lo + 1)) - 1);
        sticky = (a[k] & mask) != 0;
      }
    }
    if (rb && (sticky || (mant & 1))) mant++;
    r = ldexp((double)mant, top - 52 + fs.lsb);
  }
  return neg ? -r : r;
}

}  // namespace vq

// Launchers (defined in the kernel translation units).
namespace vq {
int32_t dense_grid(const QDesc& d);
size_t dense_lds_bytes(const QDesc& d);
hipError_t launch_agg_dense(const QDesc& d, int32_t grid, uint64_t* part, hipStream_t s);
hipError_t launch_agg_final(const QDesc& d, int32_t nparts, const uint64_t* part, uint64_t* out,
                            hipStream_t s);
hipError_t launch_filter(const QDesc& d, bool jit, uint64_t* bits, uint32_t* tile_counts,
                         uint32_t* tile_offsets, uint32_t* out_rows, uint64_t* out_count,
                         hipStream_t s, hipEvent_t ev_main0, hipEvent_t ev_main1);

// Shape-specialized kernels compiled at run time (vq_jit.hip).  jit_dense_grid
// returns the persistent grid, or -1 (with *err) when no kernel is available.
int jit_dense_grid(const QDesc& d, std::string* err);
hipError_t jit_launch_dense(const QDesc& d, int32_t grid, uint64_t* part, hipStream_t s);
bool jit_available_filter(const QDesc& d, std::string* err);
hipError_t jit_launch_filter_bits(const QDesc& d, uint64_t* bits, uint32_t* tile_counts,
                                  hipStream_t s);

struct JoinBufs {
  // partitioned build side
  int64_t* b_keys;
  uint32_t* b_counts;   // [ntiles_b][P] counts, then absolute positions
  uint32_t* b_csum;     // [nchunks_b][P]
  uint32_t* b_pstart;   // [P + 1]
  // partitioned probe side
  int64_t* p_keys;
  uint32_t* p_rows;
  uint64_t* p_vals;     // [nvals][cap_p]
  uint32_t* p_counts;
  uint32_t* p_csum;
  uint32_t* p_pstart;
  // group records and the top-k
  uint64_t* recs;       // one record per matched group
  uint64_t* cand;       // top-k candidates, kTopBlocks * topk records
  uint64_t* out;        // top-k records
  uint32_t* flags;      // [0] overflow, [1] records, [2] build keys, [3] top-k count
  int64_t cap_p;        // probe entries capacity (value stride)
  int64_t cap_b;
};

struct JoinPlan {
  int32_t bits;
  int32_t P;
  int32_t cap;          // table slots per partition (power of two)
  int32_t pool;         // accumulator entries per partition (matched keys)
  int32_t topk;
  int32_t nvals;        // aggregates that carry a per-row value
  int32_t lds_max;      // per-workgroup LDS the device allows (bytes)
  int32_t pad_;
};

// LDS bytes of one partition's table and accumulator pool.
size_t join_lds_for(int32_t cap, int32_t pool, int32_t nwords);

// Run-time compiled partition kernels of one join side (vq_jit.hip).  With
// count = true, writes counts[t][p]; otherwise scatters to the positions in
// counts.  jd (probe side) supplies the value expressions.
hipError_t jit_launch_part(const JSide& sd, const JDesc* jd, int32_t P, uint32_t* counts,
                           int64_t* keys, uint32_t* rows, uint64_t* vals, int64_t vstride,
                           hipStream_t s, bool count);
bool jit_part_available(const JSide& sd, const JDesc* jd, std::string* err);

hipError_t launch_join(const JSide& b, const JSide& p, const JDesc& jd, const JoinPlan& plan,
                       const JoinBufs& bufs, hipStream_t s, hipEvent_t* ev);
}  // namespace vq
