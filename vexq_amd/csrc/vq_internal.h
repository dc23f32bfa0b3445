// vq_internal.h — descriptors shared by the host dispatch (vq_api.hip) and the
// gfx950 kernels (k_scan.hip, k_join.hip).  Everything here is plain data
// passed to kernels by value, so every field a kernel reads sits in the
// kernarg segment and is fetched with scalar loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/vexq_gpu.h"
#include "fixpoint.h"

namespace vq {

constexpr int kMaxSlots = 8;      // distinct columns one query touches
constexpr int kMaxSlotTerms = 4;  // predicate terms on one column
constexpr int kMaxAcc = 8;        // accumulators after de-duplication
constexpr int kMaxSpans = 24;     // contiguous runs of surviving row groups
// The bits of Go's math.NaN() (uvnan), which applyArithFloat64 returns for
// x/0 (exec/expr.go:1016-1018).
constexpr long long kGoNaNBits = 0x7FF8000000000001LL;

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
  int32_t op;     // VQ_EQ..VQ_GE, VQ_IN (ival: mask), or op | VQ_COLCMP
  int32_t slot2;  // VQ_COLCMP: the other column's slot
  int64_t ival;
  double fval;
};

// One referenced column.
struct QSlot {
  const void* p;
  int32_t type;
  int32_t nterms;
  int32_t key_stride;  // dense group slot contribution (0 = not a key)
  int32_t null_code;   // a key column with NULLs: the code a NULL key takes (its dictionary size)
  QTerm terms[kMaxSlotTerms];
  const uint64_t* nul; // validity words (bit r of word r / 64 set: row r holds a value); null: no NULLs
};

struct QFactor {
  int32_t kind;
  int32_t slot;
  int32_t op;
  int32_t combine;
  double fval;
  int64_t ival;
};

// A CASE WHEN guard's term: as QTerm, on slot `slot`.
struct QWhen {
  int32_t slot;
  int32_t op;
  int32_t slot2;
  int32_t pad_;
  int64_t ival;
  double fval;
};

struct QExpr {
  int32_t nfactors;
  int32_t is_float;
  QFactor f[VQ_MAX_FACTORS];
  int32_t nwhen;  // CASE WHEN (conjunction of when[0, nwhen)) THEN fold ELSE else_*
  int32_t pad_;
  QWhen when[VQ_MAX_WHEN];
  double else_f;
  int64_t else_i;
};

struct QAcc {
  int32_t op;
  int32_t count_valid;  // 1: counts the rows whose e holds a value (SUM_I of 1s), for a nullable e
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
// Join + aggregate descriptors.
struct JAgg {
  int32_t kind;     // VQ_AGG_*
  int32_t accum;    // VQ_I64 / VQ_F64
  int32_t nwords;   // 64-bit accumulator words this aggregate holds per slot
  int32_t word0;    // first word index within a slot's accumulator block
  FixScale fs;      // for float sums
  int32_t vidx;     // its argument's value index (aggregates with the same
                    // argument share one); -1 for COUNT(*)
  int32_t pad_;
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
  // a GROUP BY of several dictionary-code columns: the key is their codes
  // packed as the int-key path packs them (exec/agg_intkey.go:67-100), code i
  // shifted left by key_shift[i]; nkeys 1 is the one key column key_slot
  int32_t nkeys;
  int32_t key_slots[8];
  int32_t key_shift[8];
  // a join + GROUP BY of the joined rows (vq_join_group): the probe side's
  // join key slot, and the slots (bit mask) that hold build-side columns,
  // loaded for each matching build row instead of the probe row
  int32_t lookup;
  int32_t jkey_slot;
  int32_t build_slots;
  int32_t pad_;
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

}  // namespace vq

// Shared by the library's translation units (defined in vq_api.hip).
namespace vq {
// Sets this thread's vq_last_error text; returns code.
int set_error(int code, const char* msg);
bool ctx_reserve(vq_ctx* c, uint64_t bytes);
void ctx_release(vq_ctx* c, uint64_t bytes);
int ctx_device(const vq_ctx* c);
int ctx_kernel_repeats(const vq_ctx* c);
// A table whose columns are already on the device (hipMalloc'd, padded to a
// 256-byte multiple plus 256 bytes, `bytes` reserved on the context).
struct TableParts {
  int64_t nrows = 0;
  std::vector<int32_t> types, dict_sizes;
  std::vector<void*> dev;
  std::vector<int64_t> rg_rows;
  std::vector<vq_zone> zones;  // nrg * ncols, row-group-major
  uint64_t bytes = 0;
  std::vector<std::vector<std::string>> dicts;
};
// The table takes the columns and the reservation; never fails after
// allocating it.
int table_adopt(vq_ctx* c, TableParts&& p, vq_table** out);
}  // namespace vq

// Launchers (defined in the kernel translation units).
namespace vq {
int32_t dense_grid(const QDesc& d);
size_t dense_lds_bytes(const QDesc& d);
// Rows one partitioning workgroup (1024 threads) loads per step, for a side
// of `nslots` columns; the device's compute units.
int64_t part_chunk_rows(int nslots);
// Rows one wave of the dense kernel loads per step.
int64_t dense_chunk_rows(int nslots);
int device_cus();
hipError_t launch_agg_dense(const QDesc& d, int32_t grid, uint64_t* part, hipStream_t s);
hipError_t launch_agg_final(const QDesc& d, int32_t nparts, const uint64_t* part, uint64_t* out,
                            hipStream_t s);
hipError_t launch_filter(const QDesc& d, bool jit, uint64_t* bits, uint32_t* tile_counts,
                         uint32_t* tile_offsets, uint32_t* out_rows, uint64_t* out_count,
                         hipStream_t s, hipEvent_t ev_main0, hipEvent_t ev_main1,
                         int32_t repeats);

// Shape-specialized kernels compiled at run time (vq_jit.hip).
// jit_dense_prepare resolves d's dense kernel once per query: its function,
// its LDS and the waves the device holds at once (slots); -1 (with *err)
// when no kernel is available.
struct DenseJit {
  hipFunction_t fn = nullptr;
  size_t lds = 0;
  int slots = 0;      // the waves a large query runs as (one per SIMD)
  int slots_max = 0;  // the most the kernel's occupancy allows (small queries)
  // the grid's cap for a query of kr kept rows: below 2^20 rows, short waves
  // and as many as fit (their ramp and tail dominate)
  int cap(int64_t kr) const { return kr >= ((int64_t)1 << 20) ? slots : slots_max; }
};
int jit_dense_prepare(const QDesc& d, DenseJit* out, std::string* err);
hipError_t jit_dense_launch(const DenseJit& h, const QDesc& d, int32_t grid, uint64_t* part,
                            hipStream_t s);
bool jit_available_filter(const QDesc& d, std::string* err);
// Several one-group aggregates of d's shape in one pass (vq_dense_multi):
// nq queries (a compiled kernel per shape and nq); -1 (with *err) when none.
int jit_multi_prepare(const QDesc& d, int32_t nq, DenseJit* out, std::string* err);
hipError_t jit_multi_launch(const DenseJit& h, const QDesc& d, int32_t grid, uint64_t* part,
                            const int64_t* mlit, hipStream_t s);
// The literals of d's predicate terms in the run-time kernels' order (floats
// as bits; 0 for a compare of two columns) into lit; returns their count.
int jit_term_literals(const QDesc& d, int64_t* lit);
hipError_t jit_launch_filter_bits(const QDesc& d, uint64_t* bits, uint32_t* tile_counts,
                                  hipStream_t s);

// One partitioned join side.  Partition p's entries are
// [pstart[p], pstart[p+1]).  A persistent grid of G workgroups partitions the
// side; counts[p][w] holds workgroup w's rows of partition p after the count
// pass, and after k_part_runs the rows of partition p in workgroups before w
// (workgroup w's run starts at pstart[p] + counts[p][w]).
// With many partitions (P > 512) a side is partitioned in two passes: the
// partition kernels write P1 = 256 coarse partitions (keys .. pstart), and
// k_part_fine splits each into P / P1 fine ones (fkeys .. fpstart, in
// coarse-major order: coarse c's fine partition f is layout index
// c * P2 + f).  With one pass the f* arrays are the same arrays.  The join
// kernel reads the f* arrays.
struct PartBufs {
  int64_t* keys;
  uint32_t* rows;       // probe side: row id of each entry
  uint64_t* vals;       // probe side: [nvals][vstride] aggregate argument values
  int64_t vstride;      // entries capacity
  uint32_t* counts;     // [P1][G]
  uint32_t* tot;        // [P1] rows per partition
  uint32_t* pstart;     // [P1 + 1]
  uint32_t* perr;       // a write past its partition's end (must stay 0)
  const uint32_t* bloom; // the join's semi-join filter (null: none)
  uint32_t bloom_lpp;    // log2 of its 64-byte lines per partition
  uint32_t bloom_pmask;  // P - 1 (the filter is kept per final partition)
  int32_t G;
  int32_t pad_;
  uint64_t* livebits;    // the count pass's live rows (with the filter; null: none)
  int64_t* fkeys;       // the final layout
  uint32_t* frows;
  uint64_t* fvals;
  uint32_t* fpstart;    // [P + 1]
};

// GROUP BY through per-workgroup partials: the distinct keys in one table
// (slots of kEmptyKey are free; key INT64_MIN is flagged in flags[10]), the
// key of each group index, and the workgroups' partial accumulators.
struct LocalBufs {
  int64_t* keys;        // [gcap] the key table
  int64_t* gkey;        // [limit * passes] the keys in rank order
  uint64_t* part;       // [grid][limit][nwords + 2] one pass's partials
};

// The build side's lookup table (vq_join_group): open addressing on
// radixHash(key) >> 32, linear probing; each slot's key (INT64_MIN: empty)
// and the head of its chain of build rows (next[row], 0xFFFFFFFF ends it);
// rows whose key is INT64_MIN are chained from *min_head.  packed: every
// build key lies in [0, 2^32 - 1), and a slot is one word, key << 32 | head
// (all ones: empty; head unused), so a new key takes one compare-and-swap.
// direct: every key lies in [base, base + range) with range at most a few
// times the rows (the zone maps say so): head[key - base] starts its chain,
// and the table is built with plain stores (a row whose store another row's
// overwrote chains itself in afterwards), no hashing and, for unique keys,
// no atomics.
struct LookupBufs {
  int64_t* keys;
  uint32_t* head;
  uint32_t* next;
  uint32_t* cnt;       // rows per slot after the first
  uint32_t* min_head;  // [0] the INT64_MIN chain, [1] set when the table filled,
                       // [2] the most rows one key has, [3] rows of key INT64_MIN
  uint32_t mask;       // slots - 1
  uint32_t packed;
  int64_t base;        // direct: the smallest key
  uint32_t range;      // direct: head's entries (0: a hash table)
  uint32_t pad_;
};

// The direct join (build keys whose zone maps bound them to a dense range
// [base, base + range)).  A persistent region per slot, zero between queries
// where noted: rflags (zero); rec[o][8], key base + o's record -- [0] its
// matches so far (zero), [1] 1 + its latest match past kDjSlots, chained
// through the entries (zero), [2, 8) the entries of its first kDjSlots
// matches (not cleared: read only below the count) -- one 32-byte sector per
// key; pres, bit o: a passing build row has key base + o (zero: cleared by
// k_dj_collect); mark, one byte per key, the build's marks (zero: packed into
// pres and cleared by k_dj_pack).  Per query (the arena), in regions of the
// probe workgroups (tile t0 * tile_rows onwards: at most one item per row):
// the matched probe rows (mrow) and their count per workgroup (mcount), the
// match entries ent[e][1 + nvals] -- word 0 (1 + the key's previous further
// match) << 32 | probe row, then the argument values -- and the keys each
// workgroup met first (tlist) with their count (tcount); the passing build
// rows per build tile (bcnt) and the distinct passing keys per pack block
// (pcnt): counters are summed from arrays, never added into one word by many
// waves (atomics on one address serialize).  rflags: [0] error bits (1 a
// build key outside the zone range), [2] collect workgroups finished, [4]
// groups, [6] passing build rows and [7] distinct passing keys (summed by the
// collect kernel's workgroup 0); reset by the collect kernel's last
// workgroup.
constexpr int kDjSlots = 6;
struct DirectBufs {
  uint32_t* rflags;
  uint32_t* rec;
  uint32_t* pres;
  uint8_t* mark;
  uint32_t* mrow;
  uint32_t* mcount;
  uint64_t* ent;
  uint32_t* tlist;
  uint32_t* tcount;
  uint32_t* bcnt;
  uint32_t* pcnt;
  int64_t nbcnt;     // build tiles
  int64_t npcnt;     // pack blocks
  int64_t base;
  uint64_t range;
  uint64_t words;    // (range + 31) / 32
  int64_t cap;       // items capacity: probe tiles * tile_rows
  int64_t ntiles;    // the probe's tiles and their rows (the workgroup regions)
  int32_t tile_rows;
  int32_t grid;      // probe workgroups
  int32_t nvals;
  int32_t ew;        // entry words: 1 + nvals
};

struct JoinBufs {
  PartBufs b, p;
  LocalBufs l;
  LookupBufs lk;
  DirectBufs d;
  uint64_t* recs;       // one record per matched group (full-output mode)
  uint64_t* cand;       // [P][topk] per-partition top-k candidates (top-k mode)
  struct KItem* citems; // [P][topk] their sort keys (value, first row, index)
  uint64_t* out;        // top-k records
  uint32_t* bloom;      // semi-join filter of the build keys (null: none)
  uint32_t* tctr;       // top-k merge: workgroups finished (zeroed with the flags)
  // flags[0] overflow (1 a table, 2 a pool, 4 more keys than the
  // per-workgroup path holds), [1] groups, [2] distinct build keys, [3] top-k
  // count, [4] partitions finished, [5] partition write error, [6] build
  // entries, [7] probe entries, [10] key INT64_MIN seen; 16 words (then
  // the top-k merge's counter), zeroed by the pipeline's first launch
  uint32_t* flags;
};

struct JoinPlan {
  int32_t bits;
  int32_t P;
  int32_t cap;          // table slots per partition (power of two)
  int32_t pool;         // accumulator entries per partition (matched keys)
  int32_t topk;
  int32_t nvals;        // aggregates that carry a per-row value
  int32_t lds_max;      // per-workgroup LDS the device allows (bytes)
  int32_t bbuf;         // build scatter: LDS buffer entries per partition (0: none)
  int32_t pbuf;         // probe scatter: the same
  int32_t lds_bytes;    // the join kernel's dynamic LDS: join_lds_for(cap, pool, nwords)
  int32_t groupby;      // no build side: GROUP BY the probe side's key
  int32_t reps;         // accumulator replicas (power of two, 1..16)
  // GROUP BY of few keys through per-workgroup partials instead of partitions
  int32_t local;
  int32_t lcap;         // a pass's LDS key table slots (a power of two)
  int32_t limit;        // the most keys a pass holds (lcap / 2)
  int32_t passes;       // passes over the rows, `limit` keys each
  int32_t gcap;         // the key table's slots (2 * limit * passes, rounded up)
  int32_t lgrid;        // workgroups that aggregate (partials)
  int32_t kgrid;        // workgroups that collect the keys
  int32_t tnodes;       // top-k merge counters (1 with more than one partition)
  int64_t seed_groups;  // > 0: the key table seeded with every code combination
  int32_t seed_card[8]; // (codes of each key column)
  int32_t bloom_lpp;    // semi-join filter: log2 of its 64-byte lines per
                        // partition (-1: no filter)
  int32_t P1;           // coarse partitions of the partition kernels (P: one pass)
  int32_t bits1;        // log2(P1)
  int32_t direct;       // the direct join (DirectBufs) instead of partitions
};

// LDS bytes of one partition's table and accumulator pool (with `reps`
// replicas of the accumulators, and the top-k's ranking room).
size_t join_lds_for(int32_t cap, int32_t pool, int32_t nwords, int32_t reps = 1, bool topk = false);
// Counters of the top-k merge over P partitions' candidates.
int32_t topk_tree_nodes(int32_t P);

// Arguments of every run-time compiled kernel; must match vqj::Args in
// jit_template.hip field for field.
constexpr int kJitTerms = 8 + 8 * VQ_MAX_WHEN;  // the predicate's, then each aggregate's CASE WHEN terms
struct JitArgs {
  const void* col[8];
  int64_t ti[kJitTerms];
  double tf[kJitTerms];
  int64_t fi[8][4];
  double ff[8][4];
  int64_t ei[8];  // ELSE literals
  double ef[8];
  int64_t sp_start[kMaxSpans];
  int64_t sp_rows[kMaxSpans];
  int64_t sp_tile0[kMaxSpans];
  int32_t nspans;
  int32_t tile_rows;
  int64_t ntiles;
  uint64_t* part;
  uint64_t* bits;
  uint32_t* tile_counts;
  uint32_t* counts;
  const uint32_t* pstart;
  uint32_t* perr;
  int64_t* keys;
  uint32_t* rows;
  uint64_t* vals;
  int64_t vstride;
  int32_t P;
  uint32_t lds_words;  // dynamic LDS of the launch, in 4-byte words
  // GROUP BY through per-workgroup partials (vq_hash_local)
  const int64_t* lkeys;
  uint32_t* lflags;
  int64_t* lgkey;
  int32_t lcap;
  int32_t lnw;
  int32_t vop[8];   // per aggregate with an argument: 0 float sum, 1 int sum, 2/3 float min/max, 4/5 int min/max
  int32_t vword[8];
  int32_t vlsb[8];
  int32_t vw[8];
  int32_t vsrc[8];  // its argument's value index
  int32_t nvop;     // aggregates with an argument
  // semi-join filter of the build keys (the probe side's partition kernels)
  const uint32_t* bloom;
  uint32_t bloom_lpp;
  uint32_t bloom_pmask;
  uint64_t* livebits;  // count pass -> scatter: one bit per live row
  int32_t lk0;         // vq_hash_local: the first key (rank) of this pass
  int32_t lpad_;
  // vq_join_group: the build side's lookup table (LookupBufs)
  const int64_t* lk_keys;
  const uint32_t* lk_head;
  const uint32_t* lk_next;
  const uint32_t* lk_min;
  uint32_t lk_mask;
  uint32_t lk_packed;
  int64_t lk_base;     // direct tables (lk_range > 0): head[key - lk_base]
  uint32_t lk_range;
  uint32_t lk_pad;
  // the direct join's probe (vq_dj_probe, vq_dj_match): DirectBufs
  const uint32_t* dj_pres;
  uint32_t* dj_rec;
  uint32_t* dj_mrow;
  uint32_t* dj_mcount;
  uint64_t* dj_ent;
  uint32_t* dj_tlist;
  uint32_t* dj_tcount;
  int64_t dj_base;
  uint64_t dj_range;
  int32_t dj_ew;
  int32_t dj_pad;
  // vq_dense_multi: each query's predicate literals, [q][terms]
  const int64_t* mlit;
  // validity words of each slot's column (null: no NULLs)
  const uint64_t* nul[8];
};

// A join side's partition kernels, resolved once per query (vq_jit.hip):
// the compiled count and scatter kernels of the side's shape, their
// persistent grid (one workgroup per scatter slot the device holds at once,
// at most one per tile) and the arguments that do not depend on buffers.
struct PartLaunch {
  hipFunction_t count = nullptr;
  hipFunction_t scatter = nullptr;
  int32_t grid = 0;
  uint32_t count_lds = 0;    // dynamic LDS bytes of each kernel
  uint32_t scatter_lds = 0;
  JitArgs args;
};
// bloom: this (probe) side keeps only the rows that pass the join's
// semi-join filter (and livebits: its count pass records them for the
// scatter, which then loads only their columns).
int jit_part_prepare(const JSide& sd, const JDesc* jd, int32_t P, int32_t pbuf, PartLaunch* out,
                     std::string* err, bool bloom = false, bool livebits = false);
// count = true: adds the side's rows per partition and workgroup into counts;
// otherwise scatters them into their runs.
hipError_t jit_part_launch(const PartLaunch& pl, const PartBufs& pb, int32_t P, hipStream_t s,
                           bool count);
// LDS buffer entries per partition for a scatter with entries of `entry`
// bytes: the largest of 16, 8, 4, 2, 1 that fits, else 0.
int32_t part_buffer_entries(int32_t P, int32_t entry, int32_t lds_max);
// Dynamic LDS bytes of the scatter: buffers of pbuf entries of `entry` bytes
// per partition, or the cursors alone when pbuf is 0 (0 when they do not fit).
uint32_t part_scatter_lds(int32_t P, int32_t pbuf, int32_t entry);

// Lets the join kernel launch with up to `bytes` of dynamic LDS (once per
// process and size; outside any graph capture).
void join_allow_lds(int32_t bytes);

hipError_t launch_join(const PartLaunch& b, const PartLaunch& p, const JDesc& jd,
                       const JoinPlan& plan, const JoinBufs& bufs, hipStream_t s, hipEvent_t* ev);

// The run-time compiled per-workgroup GROUP BY of one probe side
// (vq_hash_local): its kernel, grid and dynamic LDS, and the arguments that
// do not depend on buffers.
struct LocalLaunch {
  hipFunction_t fn = nullptr;
  int32_t grid = 0;
  uint32_t lds = 0;
  JitArgs args;
};
// The run-time compiled probe of the direct join (vq_dj_probe).
struct DirectLaunch {
  hipFunction_t fn = nullptr;     // vq_dj_probe
  hipFunction_t match = nullptr;  // vq_dj_match
  int32_t grid = 0;
  JitArgs args;
};
int jit_direct_prepare(const JSide& sd, const JDesc& jd, DirectLaunch* out, std::string* err);
hipError_t jit_direct_launch(const DirectLaunch& dl, const DirectBufs& D, hipStream_t s);
// The direct join: build (k_dj_build), probe, collect (k_dj_collect: groups,
// top-k, the region cleared for the next query); events ev[0..3] around it.
hipError_t launch_direct(const QDesc& bq, int32_t bkey_slot, const DirectLaunch& dl, const JDesc& jd,
                         const JoinPlan& plan, const JoinBufs& B, hipStream_t s, hipEvent_t* ev);
// The direct join's build side (k_scan.hip): marks, then the packed bits.
hipError_t launch_dj_build(const QDesc& d, int32_t key_slot, const DirectBufs& D, hipStream_t s);
// Clears a direct region of `words` 64-bit words.
hipError_t launch_zero_words(uint64_t* p, uint64_t words, hipStream_t s);
hipError_t launch_rank_map(const void* codes, int32_t type, const uint16_t* rank, uint16_t* out, int64_t n,
                           hipStream_t s);
constexpr int32_t kLocalLds = 60 * 1024;  // vq_hash_local's dynamic LDS
int jit_local_prepare(const JSide& sd, const JDesc& jd, int32_t lcap, LocalLaunch* out,
                      std::string* err);
hipError_t jit_local_launch(const LocalLaunch& ll, const JoinBufs& B, hipStream_t s, int32_t k0);
// LDS the per-workgroup path needs for `limit` keys of `nwords` words.
size_t local_lds_for(int32_t lcap, int32_t limit, int32_t nwords);
// The nested build (vq_join_agg_nested, k_scan.hip): the columns of the
// nested join's output that are kept (slots of the build side's QDesc) and
// where they go; the output groups' build-side columns.
struct SemiOut {
  int32_t ncols;
  int32_t slot[8];
  int32_t type[8];
  void* dst[8];
};
struct PayCols {
  int32_t npay;
  int32_t key_type;
  const void* key;
  const void* src[8];
  int32_t type[8];
};
// Per tile of d: the inner rows its passing rows meet (lookup table lk, key
// slot fk_slot), scanned into tile_offsets and *out_count.
hipError_t launch_semi(const QDesc& d, int32_t fk_slot, const LookupBufs& lk, uint32_t* tile_counts,
                       uint32_t* tile_offsets, uint64_t* out_count, hipStream_t s);
// Each passing row's copies (one per inner row it meets), in row order.
hipError_t launch_semi_gather(const QDesc& d, int32_t fk_slot, const LookupBufs& lk,
                              const uint32_t* tile_offsets, const SemiOut& o, hipStream_t s);
// keys[0, nkeys) into lk (group i), then the pc columns of every row whose key
// is one of them into pay[group][npay]; *flag set when rows of one key disagree.
hipError_t launch_pay_gather(const int64_t* keys, int64_t nkeys, const LookupBufs& lk, const PayCols& pc,
                             int64_t nrows, int64_t* pay, uint32_t* flag, hipStream_t s);
// vq_join_group's build side: the lookup table of the rows of d's spans that
// pass its predicate, keyed by slot key_slot (k_scan.hip).
hipError_t launch_lookup_build(const QDesc& d, int32_t key_slot, const LookupBufs& lk, hipStream_t s);
hipError_t launch_local(const LocalLaunch& ll, const JSide& ps, const JDesc& jd,
                        const JoinPlan& plan, const JoinBufs& bufs, hipStream_t s,
                        hipEvent_t* ev, bool ev0_done = false);
}  // namespace vq
