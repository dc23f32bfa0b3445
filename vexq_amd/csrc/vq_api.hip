// vq_api.hip — the C ABI (include/vexq_gpu.h): contexts, resident tables,
// zone-map pruning, query planning onto the kernels and result assembly.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vq_internal.h"

using namespace vq;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(VQ_E_DEVICE, "exec: gpu: %s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

int width_of(int type) {
  switch (type) {
    case VQ_U8: return 1;
    case VQ_U16: return 2;
    case VQ_DATE: return 4;
    case VQ_I64:
    case VQ_F64: return 8;
    default: return 0;
  }
}

// Grow-only device scratch owned by a context.
struct Arena {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

// One query slot: a stream and the per-query resources that go with it.  A
// query holds one slot from its first launch to its result; queries on
// different slots run concurrently on the device.
struct Slot {
  hipStream_t stream = nullptr;
  Arena arena;
  void* pinned = nullptr;
  void* pinned_dev = nullptr;  // the device's address of `pinned` (null: copy instead)
  size_t pinned_bytes = 0;
  // full-output GROUP BY and join records on their way to the host (grown on
  // demand; a pageable copy of megabytes costs milliseconds): one for a query
  // run alone, one per result-assembling thread of a batch
  struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
  };
  static constexpr int kAssemblers = 8;
  Pinned rec_pin[kAssemblers];
  hipEvent_t ev[6] = {};
  // batches: branch streams forked from `stream` while a graph is captured,
  // their join events, and the instantiated graphs of recent batches
  static constexpr int kBranches = 8;
  hipStream_t side[kBranches] = {};
  hipEvent_t fork = nullptr, join[kBranches] = {};
  struct CachedGraph {
    uint64_t h1, h2;
    hipGraphExec_t exec;
  };
  std::vector<CachedGraph> graphs;
  // the direct join's region (all zero between queries; a query that may
  // have left it otherwise marks it dirty, and the next one clears it first)
  void* dreg = nullptr;
  size_t dreg_bytes = 0;
  uint64_t dreg_range = 0;  // the keys its layout holds (every query lays it out alike)
  bool dreg_dirty = true;
};

struct Ticket;

struct vq_ctx {
  int device = 0;
  uint64_t budget = 0;
  std::atomic<uint64_t> in_use{0};
  std::mutex mu;  // the slot pool and the caches below
  std::condition_variable slot_cv;
  std::vector<Slot*> slots, idle;
  // build rows that passed, by (build table, key column, predicate): the
  // partition count of the next join over the same build side
  std::unordered_map<std::string, int64_t> build_est;
  // launches of each query's main kernel inside its event pair (timing aid)
  int32_t kernel_repeats = 1;
  // cached selection-vector buffer
  void* sv_cache = nullptr;
  size_t sv_cache_bytes = 0;
  // the autotuner's choices per instance class (under mu)
  std::unordered_map<std::string, std::string> tune;
  // asynchronous queries (vq_submit): FIFO admission against serve_budget
  std::mutex qmu;
  std::condition_variable qcv;
  std::deque<Ticket*> queue;
  std::vector<std::thread> workers;
  uint64_t serve_budget = 0;
  uint64_t serve_reserved = 0;
  int running = 0;
  bool stopping = false;
  std::vector<Ticket*> active;  // the tickets the workers are running (under qmu)
};

struct ColDev {
  int32_t type;
  int32_t dict_size;
  void* d;
  uint64_t* nul = nullptr;  // validity words (bit r of word r / 64: row r holds a value); null: no NULLs
};

struct vq_table {
  vq_ctx* ctx;
  int64_t nrows;
  int32_t ncols;
  std::vector<ColDev> cols;
  int32_t nrg;
  std::vector<int64_t> rg_rows, rg_start;
  std::vector<vq_zone> zones;  // nrg * ncols
  uint64_t bytes;
  // the strings of each dictionary column's global codes (tables ingested
  // from a .vxq image; empty for uploads of decoded columns)
  std::vector<std::vector<std::string>> dicts;
  // per dictionary column, its strings' byte order for MIN/MAX over strings:
  // a device column of ranks (u16 a row; d null until the order is known)
  // and the code of each rank
  std::vector<ColDev> rank;
  std::vector<std::vector<uint32_t>> code_of_rank;
  uint64_t rank_bytes = 0;
  std::mutex rank_mu;
};

struct vq_selvec {
  vq_ctx* ctx;
  uint32_t* rows;
  size_t bytes;
  int64_t count;
};

namespace {

bool reserve(vq_ctx* c, uint64_t bytes) {
  uint64_t cur = c->in_use.load();
  while (true) {
    if (cur + bytes > c->budget) return false;
    if (c->in_use.compare_exchange_weak(cur, cur + bytes)) return true;
  }
}

void release(vq_ctx* c, uint64_t bytes) { c->in_use.fetch_sub(bytes); }

}  // namespace

namespace vq {
// Shared with the other translation units of the library (vq_internal.h).
int set_error(int code, const char* msg) { return fail(code, "%s", msg); }
bool ctx_reserve(vq_ctx* c, uint64_t bytes) { return reserve(c, bytes); }
void ctx_release(vq_ctx* c, uint64_t bytes) { release(c, bytes); }
int ctx_device(const vq_ctx* c) { return c->device; }
int ctx_kernel_repeats(const vq_ctx* c) { return c->kernel_repeats; }

int table_adopt(vq_ctx* c, TableParts&& p, vq_table** out) {
  vq_table* t = new (std::nothrow) vq_table();
  if (!t) return fail(VQ_E_NOMEM, "exec: gpu: table allocation");
  t->ctx = c;
  t->nrows = p.nrows;
  t->ncols = (int32_t)p.dev.size();
  t->nrg = (int32_t)p.rg_rows.size();
  t->bytes = p.bytes;
  for (size_t i = 0; i < p.dev.size(); i++) t->cols.push_back(ColDev{p.types[i], p.dict_sizes[i], p.dev[i]});
  t->rg_rows = std::move(p.rg_rows);
  t->rg_start.resize(t->nrg);
  int64_t st = 0;
  for (int32_t g = 0; g < t->nrg; g++) {
    t->rg_start[g] = st;
    st += t->rg_rows[g];
  }
  t->zones = std::move(p.zones);
  t->dicts = std::move(p.dicts);
  t->rank.assign(t->ncols, ColDev{VQ_U16, 0, nullptr});
  t->code_of_rank.resize(t->ncols);
  *out = t;
  return VQ_OK;
}
}  // namespace vq

namespace {

int arena_get(vq_ctx* c, Slot* S, size_t bytes, void** out) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (S->arena.bytes < bytes) {
    size_t want = std::max(bytes, S->arena.bytes * 3 / 2);
    if (S->arena.p) {
      HIP_TRY(hipStreamSynchronize(S->stream));
      HIP_TRY(hipFree(S->arena.p));
      release(c, S->arena.bytes);
      S->arena.p = nullptr;
      S->arena.bytes = 0;
    }
    if (!reserve(c, want))
      return fail(VQ_E_BUDGET, "exec: gpu: query arena of %zu bytes exceeds the device budget", want);
    hipError_t e = hipMalloc(&S->arena.p, want);
    if (e != hipSuccess) {
      release(c, want);
      return fail(VQ_E_DEVICE, "exec: gpu: hipMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    S->arena.bytes = want;
  }
  *out = S->arena.p;
  return VQ_OK;
}

int pinned_get(Slot* S, size_t bytes, void** out) {
  if (S->pinned_bytes < bytes) {
    if (S->pinned) HIP_TRY(hipHostFree(S->pinned));
    S->pinned = nullptr;
    S->pinned_dev = nullptr;
    size_t want = std::max(bytes, (size_t)1 << 16);
    HIP_TRY(hipHostMalloc(&S->pinned, want, hipHostMallocMapped));
    S->pinned_bytes = want;
    // kernels may write small results here directly (VQ_NO_ZERO_COPY=1:
    // always through a copy)
    const char* e = getenv("VQ_NO_ZERO_COPY");
    if (!(e && *e == '1') && hipHostGetDevicePointer(&S->pinned_dev, S->pinned, 0) != hipSuccess) {
      (void)hipGetLastError();
      S->pinned_dev = nullptr;
    }
  }
  *out = S->pinned;
  return VQ_OK;
}

int slot_create(Slot** out) {
  Slot* S = new (std::nothrow) Slot();
  if (!S) return fail(VQ_E_NOMEM, "exec: gpu: slot allocation");
  hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
  for (auto& ev : S->ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  if (e != hipSuccess) {
    for (auto& ev : S->ev)
      if (ev) hipEventDestroy(ev);
    if (S->stream) hipStreamDestroy(S->stream);
    delete S;
    return fail(VQ_E_DEVICE, "exec: gpu: stream/event creation: %s", hipGetErrorString(e));
  }
  *out = S;
  return VQ_OK;
}

void slot_destroy(vq_ctx* c, Slot* S) {
  hipStreamSynchronize(S->stream);
  for (auto& g : S->graphs) hipGraphExecDestroy(g.exec);
  for (int k = 0; k < Slot::kBranches; k++) {
    if (S->side[k]) hipStreamDestroy(S->side[k]);
    if (S->join[k]) hipEventDestroy(S->join[k]);
  }
  if (S->fork) hipEventDestroy(S->fork);
  if (S->arena.p) {
    hipFree(S->arena.p);
    release(c, S->arena.bytes);
  }
  if (S->dreg) {
    hipFree(S->dreg);
    release(c, S->dreg_bytes);
  }
  if (S->pinned) hipHostFree(S->pinned);
  for (auto& r : S->rec_pin)
    if (r.p) hipHostFree(r.p);
  for (auto& ev : S->ev)
    if (ev) hipEventDestroy(ev);
  hipStreamDestroy(S->stream);
  delete S;
}

void stop_workers(vq_ctx* c);
void tune_load(vq_ctx* c);
int scratch_setup(int device);

// A query's hold on one slot: waits for an idle slot, gives it back at the
// end of the scope.
class SlotLease {
 public:
  explicit SlotLease(vq_ctx* c) : c_(c) {
    std::unique_lock<std::mutex> lk(c->mu);
    c->slot_cv.wait(lk, [&] { return !c->idle.empty(); });
    s_ = c->idle.back();
    c->idle.pop_back();
  }
  ~SlotLease() {
    {
      std::lock_guard<std::mutex> lk(c_->mu);
      c_->idle.push_back(s_);
    }
    c_->slot_cv.notify_one();
  }
  Slot* operator->() const { return s_; }
  Slot* get() const { return s_; }

 private:
  vq_ctx* c_;
  Slot* s_;
};

// Bump allocator over the arena.
struct Carve {
  size_t off = 0;
  template <class T>
  T* take(char* base, size_t n) {
    off = (off + 255) & ~(size_t)255;
    T* p = (T*)(base + off);
    off += n * sizeof(T);
    return p;
  }
};

// ---- zone maps: storage/writer.go:205-348 -----------------------------------
// Zone maps as storage/writer.go:205-348 computes them: per row group and
// column the minimum and maximum of the rows that hold a value (raw bits),
// and the NULL count (no value at all: has_minmax 0).
void compute_zones(const vq_col_desc* cols, int32_t ncols, int32_t nrg,
                   const std::vector<int64_t>& rg_start, const std::vector<int64_t>& rg_rows,
                   std::vector<vq_zone>& out) {
  out.assign((size_t)nrg * ncols, vq_zone{});
  for (int32_t g = 0; g < nrg; g++) {
    const int64_t b = rg_start[g], n = rg_rows[g];
    for (int32_t c = 0; c < ncols; c++) {
      vq_zone& z = out[(size_t)g * ncols + c];
      if (n <= 0) continue;
      const vq_col_desc& cd = cols[c];
      auto has = [&](int64_t r) { return !cd.valid || ((cd.valid[r >> 3] >> (r & 7)) & 1); };
      int64_t nn = 0;
      for (int64_t r = b; r < b + n; r++) nn += has(r) ? 0 : 1;
      z.null_count = nn;
      if (nn == n) continue;  // no value in the row group
      z.has_minmax = 1;
      if ((cd.type == VQ_U8 || cd.type == VQ_U16) && cd.dict_size > 0) {
        z.min_bits = 0;
        z.max_bits = (uint64_t)(cd.dict_size - 1);
        continue;
      }
      if (cd.type == VQ_F64) {
        const double* d = (const double*)cd.data;
        bool any = false;
        double mn = 0, mx = 0;
        for (int64_t r = b; r < b + n; r++) {
          if (!has(r)) continue;
          if (!any || d[r] < mn) mn = d[r];
          if (!any || d[r] > mx) mx = d[r];
          any = true;
        }
        memcpy(&z.min_bits, &mn, 8);
        memcpy(&z.max_bits, &mx, 8);
        continue;
      }
      auto val = [&](int64_t r) -> int64_t {
        switch (cd.type) {
          case VQ_U8: return ((const uint8_t*)cd.data)[r];
          case VQ_U16: return ((const uint16_t*)cd.data)[r];
          case VQ_DATE: return ((const int32_t*)cd.data)[r];
          default: return ((const int64_t*)cd.data)[r];
        }
      };
      bool any = false;
      int64_t mn = 0, mx = 0;
      for (int64_t r = b; r < b + n; r++) {
        if (!has(r)) continue;
        const int64_t v = val(r);
        if (!any || v < mn) mn = v;
        if (!any || v > mx) mx = v;
        any = true;
      }
      z.min_bits = cd.type == VQ_DATE ? (uint64_t)(uint32_t)(int32_t)mn : (uint64_t)mn;
      z.max_bits = cd.type == VQ_DATE ? (uint64_t)(uint32_t)(int32_t)mx : (uint64_t)mx;
    }
  }
}

// zoneRangePred (planner/physical.go:667-750) on one conjunct.
bool zone_term_keep(const ColDev& c, const vq_zone& z, const vq_term& t) {
  if (!z.has_minmax) return true;
  if (t.op == VQ_IN || (t.op & VQ_COLCMP)) return true;  // (never prunes)
  if ((c.type == VQ_U8 || c.type == VQ_U16) && c.dict_size > 0) return true;
  int64_t v;
  if (c.type == VQ_F64) memcpy(&v, &t.fval, 8);
  else v = t.ival;
  const int64_t mn = (int64_t)z.min_bits, mx = (int64_t)z.max_bits;
  switch (t.op) {
    case VQ_EQ: return mn <= v && v <= mx;
    case VQ_NE: return true;
    case VQ_LT: return mn < v;
    case VQ_LE: return mn <= v;
    case VQ_GT: return mx > v;
    default: return mx >= v;
  }
}

void prune(const vq_table* t, const vq_pred* pred, std::vector<uint8_t>& keep) {
  keep.assign(t->nrg, 1);
  if (!pred) return;
  for (int32_t g = 0; g < t->nrg; g++)
    for (int k = 0; k < pred->nterms && keep[g]; k++) {
      const vq_term& tm = pred->terms[k];
      keep[g] = zone_term_keep(t->cols[tm.col], t->zones[(size_t)g * t->ncols + tm.col], tm);
    }
}

// Surviving row groups -> contiguous spans -> tiles.  Returns rows scanned.
int make_spans(const vq_table* t, const std::vector<uint8_t>& keep, int32_t tile_rows,
               QSpan* spans, int32_t* nspans, int64_t* ntiles, int64_t* rows, int64_t* rg_kept) {
  int ns = 0;
  int64_t tiles = 0, nrows = 0, kept = 0;
  for (int32_t g = 0; g < t->nrg; g++) {
    if (!keep[g] || t->rg_rows[g] == 0) continue;
    kept++;
    const int64_t st = t->rg_start[g], n = t->rg_rows[g];
    if (ns > 0 && spans[ns - 1].start + spans[ns - 1].rows == st) {
      spans[ns - 1].rows += n;
    } else {
      if (ns == kMaxSpans)
        return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d disjoint row-group runs survive pruning", kMaxSpans);
      spans[ns].start = st;
      spans[ns].rows = n;
      ns++;
    }
    nrows += n;
  }
  for (int i = 0; i < ns; i++) {
    spans[i].tile0 = tiles;
    tiles += (spans[i].rows + tile_rows - 1) / tile_rows;
  }
  *nspans = ns;
  *ntiles = tiles;
  *rows = nrows;
  *rg_kept = kept;
  return VQ_OK;
}

// Maps table columns to the query's slots (first use order).
struct SlotMap {
  int32_t col_of[kMaxSlots];
  int32_t n = 0;
  int get(int32_t col) {
    for (int i = 0; i < n; i++)
      if (col_of[i] == col) return i;
    if (n == kMaxSlots) return -1;
    col_of[n] = col;
    return n++;
  }
};

int check_col(const vq_table* t, int32_t c) {
  if (c < 0 || c >= t->ncols) return fail(VQ_E_ARG, "exec: gpu: column %d out of range", c);
  return VQ_OK;
}

// A term's operator and operands: a comparison with a literal; IN (a mask
// over [0, 64)) on an integer-typed column; or a comparison of two columns
// of the same kind (both float64, or both integer-typed).
int check_term(const vq_table* t, const vq_term& tm) {
  const bool isf = t->cols[tm.col].type == VQ_F64;
  if (tm.op >= VQ_EQ && tm.op <= VQ_GE) return VQ_OK;
  if (tm.op == VQ_IN) {
    if (isf) return fail(VQ_E_UNSUPPORTED, "exec: gpu: IN on a float64 column");
    return VQ_OK;
  }
  const int base = tm.op & ~VQ_COLCMP;
  if (!(tm.op & VQ_COLCMP) || base < VQ_EQ || base > VQ_GE)
    return fail(VQ_E_ARG, "exec: gpu: bad comparison op %d", tm.op);
  if (tm.ival < 0 || tm.ival >= t->ncols) return fail(VQ_E_ARG, "exec: gpu: column %lld out of range", (long long)tm.ival);
  if ((t->cols[tm.ival].type == VQ_F64) != isf)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: a comparison of a float64 column with an integer one");
  // two dictionary columns compare as strings (exec/expr.go:320-896), which
  // their codes do not order, and two dictionaries need not agree on a code
  if (t->cols[tm.col].dict_size > 0 || t->cols[tm.ival].dict_size > 0)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: a column comparison of dictionary-coded (string) columns");
  return VQ_OK;
}

int add_pred(const vq_table* t, const vq_pred* pred, SlotMap& sm, QSlot* slots) {
  if (!pred) return VQ_OK;
  if (pred->nterms < 0 || pred->nterms > VQ_MAX_TERMS)
    return fail(VQ_E_ARG, "exec: gpu: predicate has %d terms", pred->nterms);
  for (int k = 0; k < pred->nterms; k++) {
    const vq_term& tm = pred->terms[k];
    if (int rc = check_col(t, tm.col)) return rc;
    if (int rc = check_term(t, tm)) return rc;
    int s = sm.get(tm.col);
    if (s < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
    int s2 = 0;
    if (tm.op & VQ_COLCMP) {
      s2 = sm.get((int32_t)tm.ival);
      if (s2 < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
    }
    QSlot& sl = slots[s];
    if (sl.nterms == kMaxSlotTerms)
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d terms on one column", kMaxSlotTerms);
    QTerm& q = sl.terms[sl.nterms++];
    q.op = tm.op;
    q.slot2 = s2;
    q.ival = tm.ival;
    q.fval = tm.fval;
  }
  return VQ_OK;
}

int add_expr(const vq_table* t, const vq_expr& e, SlotMap& sm, QExpr& out) {
  memset(&out, 0, sizeof out);
  if (e.nfactors < 0 || e.nfactors > VQ_MAX_FACTORS)
    return fail(VQ_E_ARG, "exec: gpu: expression has %d factors", e.nfactors);
  out.nfactors = e.nfactors;
  out.is_float = e.is_float ? 1 : 0;
  for (int k = 0; k < e.nfactors; k++) {
    const vq_factor& f = e.f[k];
    QFactor& q = out.f[k];
    if (f.kind < VQ_F_COL || f.kind > VQ_F_COL_OP_LIT) return fail(VQ_E_ARG, "exec: gpu: bad factor kind");
    if (f.op < VQ_ADD || f.op > VQ_DIV || f.combine < VQ_ADD || f.combine > VQ_DIV)
      return fail(VQ_E_ARG, "exec: gpu: bad arithmetic op");
    q.kind = f.kind;
    q.op = f.op;
    q.combine = f.combine;
    q.fval = f.fval;
    q.ival = f.ival;
    if (f.kind != VQ_F_LIT) {
      if (int rc = check_col(t, f.col)) return rc;
      if (!e.is_float && t->cols[f.col].type == VQ_F64)
        return fail(VQ_E_ARG, "exec: gpu: float column in an integer expression");
      int s = sm.get(f.col);
      if (s < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
      q.slot = s;
    }
  }
  // CASE WHEN: predicate terms on the row, the ELSE literal
  if (e.nwhen < 0 || e.nwhen > VQ_MAX_WHEN) return fail(VQ_E_ARG, "exec: gpu: CASE with %d WHEN terms", e.nwhen);
  if (e.nwhen > 0 && e.nfactors == 0) return fail(VQ_E_ARG, "exec: gpu: CASE without a THEN expression");
  out.nwhen = e.nwhen;
  for (int k = 0; k < e.nwhen; k++) {
    const vq_term& tm = e.when[k];
    if (int rc = check_col(t, tm.col)) return rc;
    if (int rc = check_term(t, tm)) return rc;
    QWhen& w = out.when[k];
    w.slot = sm.get(tm.col);
    w.slot2 = (tm.op & VQ_COLCMP) ? sm.get((int32_t)tm.ival) : 0;
    if (w.slot < 0 || w.slot2 < 0)
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
    w.op = tm.op;
    w.ival = tm.ival;
    w.fval = tm.fval;
  }
  if (e.nwhen > 0) {
    if (e.is_float) out.else_f = e.else_f;
    else out.else_i = e.else_i;
  }
  return VQ_OK;
}

// A slot's column: a column of the table, or (kRankCol + c) the device
// column of ranks of dictionary column c (MIN/MAX over strings).
constexpr int32_t kRankCol = 1 << 24;
const ColDev& col_dev(const vq_table* t, int32_t c) {
  return c >= kRankCol ? t->rank[c - kRankCol] : t->cols[c];
}

void fill_slot_cols(const vq_table* t, const SlotMap& sm, QSlot* slots) {
  for (int i = 0; i < sm.n; i++) {
    const ColDev& cd = col_dev(t, sm.col_of[i]);
    slots[i].p = cd.d;
    slots[i].type = cd.type;
    slots[i].nul = cd.nul;
  }
}

// A query's columns, none with NULLs: the operators other than the
// dictionary GROUP BY and the filter (the reference's eligibility check
// falls back to the Go operators when ZoneMap.NullCount > 0).
int refuse_nulls(const vq_table* t, const SlotMap& sm) {
  for (int i = 0; i < sm.n; i++)
    if (col_dev(t, sm.col_of[i]).nul)
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: column %d has NULLs (this operator takes none)", sm.col_of[i]);
  return VQ_OK;
}

bool any_nulls(const vq_table* t, const SlotMap& sm) {
  for (int i = 0; i < sm.n; i++)
    if (col_dev(t, sm.col_of[i]).nul) return true;
  return false;
}

int64_t slot_bytes_per_row(const vq_table* t, const SlotMap& sm) {
  int64_t w = 0;
  for (int i = 0; i < sm.n; i++) w += width_of(col_dev(t, sm.col_of[i]).type);
  return w;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
  return ms;
}

// AccumTypeFor (exec/aggregate.go:51-72) agreement check.
// strings: the operator folds MIN/MAX over strings (VQ_STR; its planner
// checks the argument).
int check_agg(const vq_agg& a, bool strings = false) {
  if (a.kind < VQ_AGG_COUNT || a.kind > VQ_AGG_MAX) return fail(VQ_E_ARG, "exec: gpu: bad aggregate kind %d", a.kind);
  if (a.kind == VQ_AGG_COUNT) return VQ_OK;
  if (a.arg.nfactors <= 0) return fail(VQ_E_ARG, "exec: gpu: aggregate %d needs an argument", a.kind);
  if (a.accum == VQ_STR) {
    if (a.kind != VQ_AGG_MIN && a.kind != VQ_AGG_MAX)
      return fail(VQ_E_ARG, "exec: gpu: a string accumulator for aggregate kind %d", a.kind);
    if (!strings) return fail(VQ_E_UNSUPPORTED, "exec: gpu: MIN/MAX over strings on this operator");
    return VQ_OK;
  }
  const int want = a.kind == VQ_AGG_AVG ? VQ_F64 : (a.arg.is_float ? VQ_F64 : VQ_I64);
  if (a.accum != want)
    return fail(VQ_E_ARG, "exec: gpu: accumulator type %d disagrees with AccumTypeFor (%d)", a.accum, want);
  return VQ_OK;
}

}  // namespace

// ===========================================================================
// context
// ===========================================================================
// The cancel request of the query this thread runs for a serving worker
// (null: a direct call, never canceled).
thread_local const std::atomic<bool>* g_cancel = nullptr;
bool canceled() { return g_cancel && g_cancel->load(std::memory_order_relaxed); }
int fail_canceled() { return fail(VQ_E_CANCELED, "exec: gpu: query canceled"); }

extern "C" const char* vq_last_error(void) { return g_err.c_str(); }

extern "C" int vq_ctx_create(int device, uint64_t budget_bytes, vq_ctx** out) {
  if (!out) return fail(VQ_E_ARG, "exec: gpu: out is NULL");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(VQ_E_ARG, "exec: gpu: device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device));
  vq_ctx* c = new (std::nothrow) vq_ctx();
  if (!c) return fail(VQ_E_NOMEM, "exec: gpu: context allocation");
  c->device = device;
  size_t fr = 0, tot = 0;
  hipError_t e = hipMemGetInfo(&fr, &tot);
  if (e != hipSuccess) {
    delete c;
    return fail(VQ_E_DEVICE, "exec: gpu: hipMemGetInfo: %s", hipGetErrorString(e));
  }
  c->budget = budget_bytes ? budget_bytes : (uint64_t)(fr * 0.9);
  Slot* S = nullptr;
  if (int rc = slot_create(&S)) {
    delete c;
    return rc;
  }
  c->slots.push_back(S);
  c->idle.push_back(S);
  if (int rc = scratch_setup(device)) {
    slot_destroy(c, S);
    delete c;
    return rc;
  }
  tune_load(c);
  *out = c;
  return VQ_OK;
}

extern "C" int vq_ctx_destroy(vq_ctx* c) {
  if (!c) return VQ_OK;
  hipSetDevice(c->device);
  stop_workers(c);
  for (Slot* S : c->slots) slot_destroy(c, S);
  if (c->sv_cache) hipFree(c->sv_cache);
  delete c;
  return VQ_OK;
}

extern "C" int vq_ctx_set_kernel_repeats(vq_ctx* c, int32_t repeats) {
  if (!c || repeats < 1 || repeats > 1000) return fail(VQ_E_ARG, "exec: gpu: kernel repeats %d", repeats);
  std::lock_guard<std::mutex> lk(c->mu);
  c->kernel_repeats = repeats;
  return VQ_OK;
}

extern "C" uint64_t vq_ctx_bytes_in_use(const vq_ctx* c) { return c ? c->in_use.load() : 0; }

// ===========================================================================
// tables
// ===========================================================================
extern "C" int vq_table_upload(vq_ctx* c, int64_t nrows, int32_t ncols, const vq_col_desc* cols,
                               int32_t nrg, const int64_t* rg_rows, const vq_zone* zones,
                               vq_table** out) {
  if (!c || !out || (ncols > 0 && !cols) || ncols < 0 || nrows < 0 || nrows > 0xFFFFFFFELL)
    return fail(VQ_E_ARG, "exec: gpu: bad table upload arguments");
  *out = nullptr;
  if (nrg < 0 || (nrg > 0 && !rg_rows)) return fail(VQ_E_ARG, "exec: gpu: bad row-group list");
  int64_t tot = 0;
  for (int32_t g = 0; g < nrg; g++) {
    if (rg_rows[g] < 0) return fail(VQ_E_ARG, "exec: gpu: negative row-group size");
    tot += rg_rows[g];
  }
  if (tot != nrows) return fail(VQ_E_ARG, "exec: gpu: row groups hold %lld rows, table %lld",
                                (long long)tot, (long long)nrows);
  uint64_t bytes = 0;
  for (int32_t i = 0; i < ncols; i++) {
    int w = width_of(cols[i].type);
    if (!w) return fail(VQ_E_ARG, "exec: gpu: column %d has unknown type %d", i, cols[i].type);
    if (nrows > 0 && !cols[i].data) return fail(VQ_E_ARG, "exec: gpu: column %d has no data", i);
    if (cols[i].dict_size < 0 || (cols[i].type == VQ_U8 && cols[i].dict_size > 256) ||
        (cols[i].type == VQ_U16 && cols[i].dict_size > 65536) ||
        (cols[i].dict_size > 0 && cols[i].type != VQ_U8 && cols[i].type != VQ_U16))
      return fail(VQ_E_ARG, "exec: gpu: column %d: bad dictionary size %d", i, cols[i].dict_size);
    bytes += (((uint64_t)nrows * w + 255) & ~255ULL) + 256;
  }
  // columns with NULLs: their validity words (padded: kernels read the word
  // of any row of a chunk)
  std::vector<uint8_t> nullable((size_t)ncols, 0);
  const size_t vwords = (size_t)(nrows + 63) / 64 + 8;
  for (int32_t i = 0; i < ncols; i++) {
    if (!cols[i].valid) continue;
    for (int64_t r = 0; r < nrows && !nullable[i]; r++)
      if (!((cols[i].valid[r >> 3] >> (r & 7)) & 1)) nullable[i] = 1;
    if (nullable[i]) bytes += vwords * 8;
  }
  // dictionary codes must index their dictionary
  for (int32_t i = 0; i < ncols; i++) {
    if (cols[i].dict_size <= 0) continue;
    for (int64_t r = 0; r < nrows; r++) {
      int v = cols[i].type == VQ_U8 ? ((const uint8_t*)cols[i].data)[r] : ((const uint16_t*)cols[i].data)[r];
      if (v >= cols[i].dict_size)
        return fail(VQ_E_ARG, "exec: gpu: column %d row %lld: code %d outside dictionary of %d", i,
                    (long long)r, v, cols[i].dict_size);
    }
  }
  if (!reserve(c, bytes)) return fail(VQ_E_BUDGET, "exec: gpu: table of %llu bytes exceeds the device budget",
                                      (unsigned long long)bytes);
  vq_table* t = new (std::nothrow) vq_table();
  if (!t) {
    release(c, bytes);
    return fail(VQ_E_NOMEM, "exec: gpu: table allocation");
  }
  t->ctx = c;
  t->nrows = nrows;
  t->ncols = ncols;
  t->nrg = nrg;
  t->bytes = bytes;
  t->rg_rows.assign(rg_rows, rg_rows + nrg);
  t->rg_start.resize(nrg);
  int64_t st = 0;
  for (int32_t g = 0; g < nrg; g++) {
    t->rg_start[g] = st;
    st += rg_rows[g];
  }
  hipSetDevice(c->device);
  for (int32_t i = 0; i < ncols; i++) {
    ColDev cd{cols[i].type, cols[i].dict_size, nullptr};
    size_t nb = (size_t)nrows * width_of(cols[i].type);
    // padded: kernels read the aligned 8-byte word holding an element
    hipError_t e = hipMalloc(&cd.d, ((nb + 255) & ~(size_t)255) + 256);
    if (e == hipSuccess && nb) e = hipMemcpy(cd.d, cols[i].data, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess && nullable[i]) {
      e = hipMalloc((void**)&cd.nul, vwords * 8);
      if (e == hipSuccess) e = hipMemset(cd.nul, 0, vwords * 8);
      if (e == hipSuccess) e = hipMemcpy(cd.nul, cols[i].valid, (size_t)(nrows + 7) / 8, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      if (cd.d) hipFree(cd.d);
      if (cd.nul) hipFree(cd.nul);
      for (auto& x : t->cols) {
        hipFree(x.d);
        if (x.nul) hipFree(x.nul);
      }
      release(c, bytes);
      delete t;
      return fail(VQ_E_DEVICE, "exec: gpu: upload of column %d: %s", i, hipGetErrorString(e));
    }
    t->cols.push_back(cd);
  }
  if (zones) t->zones.assign(zones, zones + (size_t)nrg * ncols);
  else compute_zones(cols, ncols, nrg, t->rg_start, t->rg_rows, t->zones);
  t->rank.assign(ncols, ColDev{VQ_U16, 0, nullptr});
  t->code_of_rank.resize(ncols);
  *out = t;
  return VQ_OK;
}

extern "C" int vq_table_destroy(vq_table* t) {
  if (!t) return VQ_OK;
  hipSetDevice(t->ctx->device);
  // the caller ends every query on the table first; finish any in flight
  for (Slot* S : t->ctx->slots) hipStreamSynchronize(S->stream);
  for (auto& x : t->cols) {
    hipFree(x.d);
    if (x.nul) hipFree(x.nul);
  }
  for (auto& x : t->rank) hipFree(x.d);  // (their validity is the column's)
  release(t->ctx, t->bytes + t->rank_bytes);
  {
    std::lock_guard<std::mutex> lk(t->ctx->mu);
    char pre[32];
    snprintf(pre, sizeof pre, "%p/", (const void*)t);
    for (auto it = t->ctx->build_est.begin(); it != t->ctx->build_est.end();)
      it = it->first.find(pre) != std::string::npos ? t->ctx->build_est.erase(it) : std::next(it);
  }
  delete t;
  return VQ_OK;
}

extern "C" int64_t vq_table_rows(const vq_table* t) { return t ? t->nrows : -1; }

extern "C" int vq_table_zones(const vq_table* t, vq_zone* out, int64_t cap) {
  if (!t || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (cap < (int64_t)t->zones.size()) return fail(VQ_E_ARG, "exec: gpu: zone buffer too small");
  memcpy(out, t->zones.data(), t->zones.size() * sizeof(vq_zone));
  return VQ_OK;
}

extern "C" int vq_table_column(const vq_table* t, int32_t col, int32_t* type, int32_t* dict_size) {
  if (!t || !type || !dict_size) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (col < 0 || col >= t->ncols) return fail(VQ_E_ARG, "exec: gpu: column %d out of range", col);
  *type = t->cols[col].type;
  *dict_size = t->cols[col].dict_size;
  return VQ_OK;
}

extern "C" int vq_table_dict_entry(const vq_table* t, int32_t col, int32_t code, const char** str,
                                   int64_t* len) {
  if (!t || !str || !len) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (col < 0 || col >= t->ncols) return fail(VQ_E_ARG, "exec: gpu: column %d out of range", col);
  if ((size_t)col >= t->dicts.size() || t->dicts[col].empty())
    return fail(VQ_E_ARG, "exec: gpu: column %d carries no dictionary strings", col);
  if (code < 0 || (size_t)code >= t->dicts[col].size())
    return fail(VQ_E_ARG, "exec: gpu: code %d outside column %d's dictionary", code, col);
  *str = t->dicts[col][code].data();
  *len = (int64_t)t->dicts[col][code].size();
  return VQ_OK;
}

extern "C" int vq_zone_prune(const vq_table* t, const vq_pred* pred, uint8_t* keep) {
  if (!t || !keep) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (pred) {
    if (pred->nterms < 0 || pred->nterms > VQ_MAX_TERMS) return fail(VQ_E_ARG, "exec: gpu: bad predicate");
    for (int k = 0; k < pred->nterms; k++) {
      if (int rc = check_col(t, pred->terms[k].col)) return rc;
      if (int rc = check_term(t, pred->terms[k])) return rc;
    }
  }
  std::vector<uint8_t> k;
  prune(t, pred, k);
  memcpy(keep, k.data(), k.size());
  return VQ_OK;
}

extern "C" int32_t vq_radix_bits_for(int64_t est) {
  if (est < 65536) return 0;
  int32_t bits = 4;
  while (bits < 6 && ((int64_t)1 << bits) * 8192 < est) bits++;
  return bits;
}

extern "C" uint64_t vq_radix_hash(int64_t key) { return radix_hash((uint64_t)key); }

extern "C" int64_t vq_join_table_capacity(int64_t keys) {
  int64_t slots = 8;
  if (keys <= 0) return slots;
  int64_t need = (keys * 10 + 6) / 7;
  while (slots < need) slots *= 2;
  return slots;
}

// ===========================================================================
// dictionary-code GROUP BY
// ===========================================================================
namespace {

constexpr size_t kDenseLdsMax = 64 * 1024;
// Dense group slots: one slot when there is no GROUP BY; with one, every
// slot holds at least a count and a first row per lane (8 bytes x 64 lanes)
// inside kDenseLdsMax, so at most 128.
constexpr int kDenseGroupsMax = 128;

// Tile rows for a scan of `rows` rows: a power of two between 256 and 16384
// that cuts the scan into about 2048 tiles, so small (zone-pruned) scans still
// spread over every CU and large ones keep per-tile overhead low.
// Rows per tile (one wave each) of the kernels that insert into or look up
// a device hash table row by row.
constexpr int32_t kLookupTileRows = 256;

int32_t scan_tile_rows(int64_t rows) {
  int64_t t = 256;
  while (t < 16384 && t * 2048 < rows) t *= 2;
  return (int32_t)t;
}

// Tiles of a join side: at least 4096 rows, and about 2048 of them, so a
// persistent grid of one or two workgroups per CU takes near-equal shares.
// Given the partition kernels' chunk (rows one workgroup loads per step, a
// step never crossing a tile), the tile that needs the fewest steps of the
// busiest of `grid` workgroups (a smaller tile on a tie).
int32_t join_tile_rows(int64_t rows, int64_t chunk = 0, int64_t grid = 256) {
  if (chunk <= 0) {
    int64_t t = 4096;
    while (t < (1 << 20) && t * 2048 < rows) t *= 2;
    return (int32_t)t;
  }
  int64_t best = 4096, best_cost = INT64_MAX;
  for (int64_t t = 1024; t <= (1 << 20); t *= 2) {
    const int64_t tiles = (rows + t - 1) / t;
    const int64_t cost = ((tiles + grid - 1) / grid) * ((t + chunk - 1) / chunk);
    if (cost < best_cost) {
      best_cost = cost;
      best = t;
    }
  }
  return (int32_t)best;
}

int64_t kept_rows(const vq_table* t, const std::vector<uint8_t>& keep) {
  int64_t n = 0;
  for (int32_t g = 0; g < t->nrg; g++)
    if (keep[g]) n += t->rg_rows[g];
  return n;
}

// The device column of ranks of dictionary column `col` (its strings' byte
// order), built from rank (rank[code], a permutation of the codes) on the
// first call, or again when rank is given.  (Under the table's rank lock.)
int build_rank_column(vq_table* t, int32_t col, const uint32_t* rank) {
  const ColDev& c = t->cols[col];
  const int32_t n = c.dict_size;
  std::vector<uint32_t> inv((size_t)n, UINT32_MAX);
  std::vector<uint16_t> r16((size_t)n);
  for (int32_t k = 0; k < n; k++) {
    if (rank[k] >= (uint32_t)n || inv[rank[k]] != UINT32_MAX)
      return fail(VQ_E_ARG, "exec: gpu: column %d: the ranks are not a permutation of its %d codes", col, n);
    inv[rank[k]] = (uint32_t)k;
    r16[k] = (uint16_t)rank[k];
  }
  hipSetDevice(t->ctx->device);
  ColDev& rc = t->rank[col];
  if (!rc.d) {
    const size_t bytes = (((size_t)t->nrows * 2 + 255) & ~(size_t)255) + 256;
    if (!reserve(t->ctx, bytes))
      return fail(VQ_E_BUDGET, "exec: gpu: the string order of column %d exceeds the device budget", col);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      release(t->ctx, bytes);
      return fail(VQ_E_DEVICE, "exec: gpu: allocation of the string order of column %d", col);
    }
    if (hipMemset(p, 0, bytes) != hipSuccess) {
      hipFree(p);
      release(t->ctx, bytes);
      return fail(VQ_E_DEVICE, "exec: gpu: clearing the string order of column %d", col);
    }
    t->rank_bytes += bytes;
    rc.d = p;
  }
  rc.type = VQ_U16;
  rc.dict_size = 0;
  rc.nul = c.nul;
  uint16_t* dr = nullptr;
  hipError_t e = hipMalloc((void**)&dr, std::max<size_t>(2, (size_t)n * 2));
  if (e == hipSuccess) e = hipMemcpy(dr, r16.data(), (size_t)n * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_rank_map(c.d, c.type, dr, (uint16_t*)rc.d, t->nrows, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (dr) hipFree(dr);
  if (e != hipSuccess) return fail(VQ_E_DEVICE, "exec: gpu: string order of column %d: %s", col, hipGetErrorString(e));
  t->code_of_rank[col] = std::move(inv);
  return VQ_OK;
}

// Rank column of dictionary column `col` for MIN/MAX over strings: built
// from the dictionary a .vxq table carries if not yet known; `build` false
// (a caller that cannot allocate): only an order already set.
int rank_column(const vq_table* tc, int32_t col, bool build) {
  vq_table* t = const_cast<vq_table*>(tc);
  std::lock_guard<std::mutex> lk(t->rank_mu);
  if (t->rank[col].d && !t->code_of_rank[col].empty()) return VQ_OK;
  const bool have_dict = (size_t)col < t->dicts.size() && (int64_t)t->dicts[col].size() == t->cols[col].dict_size;
  if (!build || !have_dict)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: MIN/MAX over the strings of column %d: their order is not set "
                "(vq_table_dict_order)", col);
  const std::vector<std::string>& dict = t->dicts[col];
  std::vector<uint32_t> idx(dict.size()), rank(dict.size());
  for (size_t k = 0; k < idx.size(); k++) idx[k] = (uint32_t)k;
  // std::string's order compares bytes as unsigned, as Go's string compare
  std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return dict[a] < dict[b]; });
  for (size_t k = 0; k < idx.size(); k++) rank[idx[k]] = (uint32_t)k;
  return build_rank_column(t, col, rank.data());
}

struct AccMap {
  int32_t acc[VQ_MAX_AGGS];  // accumulator index per aggregate (-1 = COUNT)
  // MIN/MAX over strings: the dictionary code of each rank (null: not one)
  const uint32_t* str_code[VQ_MAX_AGGS];
  // an aggregate whose argument can be NULL: the accumulator counting its
  // rows that hold a value (-1: every row does)
  int32_t nn[VQ_MAX_AGGS];
  // a key column with NULLs: the code its NULL takes (-1: none)
  int32_t null_code[4];
};

// *over is set when the query is well formed but its group slots exceed what
// the dense kernel's per-lane accumulators hold (more than 4 key columns,
// more than kDenseGroupsMax slots, or more LDS than kDenseLdsMax).
int plan_dense(const vq_table* t, const vq_pred* pred, int32_t nkeys, const int32_t* key_cols,
               int32_t naggs, const vq_agg* aggs, QDesc& d, AccMap& am, SlotMap& sm,
               int32_t card[4], bool* over = nullptr, bool build_order = true) {
  memset(&d, 0, sizeof d);
  if (over) *over = false;
  if (nkeys < 0 || nkeys > 8 || (nkeys > 0 && !key_cols))
    return fail(VQ_E_ARG, "exec: gpu: %d group-by columns", nkeys);
  if (nkeys > 4) {
    if (over) *over = true;
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: %d group-by columns for the dense slots", nkeys);
  }
  if (naggs < 0 || naggs > VQ_MAX_AGGS || (naggs > 0 && !aggs))
    return fail(VQ_E_ARG, "exec: gpu: %d aggregates", naggs);
  if (naggs == 0 && nkeys == 0)
    return fail(VQ_E_ARG, "exec: hash aggregate: no group-by columns or aggregate expressions");
  if (int rc = add_pred(t, pred, sm, d.slot)) return rc;
  for (int j = 0; j < VQ_MAX_AGGS; j++) {
    am.nn[j] = -1;
    am.str_code[j] = nullptr;
  }
  for (int i = 0; i < 4; i++) am.null_code[i] = -1;
  int64_t G = 1;
  for (int i = 0; i < nkeys; i++) {
    if (int rc = check_col(t, key_cols[i])) return rc;
    const ColDev& c = t->cols[key_cols[i]];
    if ((c.type != VQ_U8 && c.type != VQ_U16) || c.dict_size <= 0) {
      // an integer key column (a composite key with codes, or one integer
      // key): packed into one key by agg_packed
      if (over && c.type != VQ_F64) *over = true;
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: group-by column %d is not a dictionary code", key_cols[i]);
    }
    // a NULL key is a group of its own (exec/aggregate.go:424-427): the code
    // past the dictionary
    card[i] = c.dict_size + (c.nul ? 1 : 0);
    if (c.nul) am.null_code[i] = c.dict_size;
    G = std::min<int64_t>(G * card[i], (int64_t)1 << 40);  // (no overflow past 4 columns)
  }
  if (G > kDenseGroupsMax) {
    if (over) *over = true;
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: %lld dense group slots", (long long)G);
  }
  int64_t stride = 1;
  for (int i = nkeys - 1; i >= 0; i--) {
    int s = sm.get(key_cols[i]);
    if (s < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
    d.slot[s].key_stride += (int32_t)stride;
    d.slot[s].null_code = am.null_code[i] >= 0 ? am.null_code[i] : 0;
    stride *= card[i];
  }
  // does an expression read a column with NULLs (its value can be NULL)?
  auto nullable_expr = [&](const vq_expr& e) {
    for (int k = 0; k < e.nfactors; k++)
      if (e.f[k].kind != VQ_F_LIT && e.f[k].col >= 0 && e.f[k].col < t->ncols && t->cols[e.f[k].col].nul)
        return true;
    return false;
  };
  auto nullable_when = [&](const vq_expr& e) {
    for (int k = 0; k < e.nwhen; k++) {
      const vq_term& w = e.when[k];
      if (w.col >= 0 && w.col < t->ncols && t->cols[w.col].nul) return true;
      if ((w.op & VQ_COLCMP) && w.ival >= 0 && w.ival < t->ncols && t->cols[(int32_t)w.ival].nul) return true;
    }
    return false;
  };
  auto add_acc = [&](const QAcc& qa) -> int {
    for (int k = 0; k < d.nacc; k++)
      if (memcmp(&d.acc[k], &qa, sizeof qa) == 0) return k;
    if (d.nacc == kMaxAcc) return -1;
    d.acc[d.nacc] = qa;
    return d.nacc++;
  };
  d.ngroups = (int32_t)G;
  // accumulators, de-duplicated (SUM(price) and AVG(price) share one); an
  // argument that can be NULL also gets a count of the rows where it holds
  // a value (the reference's nonNull count, exec/aggregate.go:458-546)
  for (int j = 0; j < naggs; j++) {
    if (int rc = check_agg(aggs[j], true)) return rc;
    am.acc[j] = -1;
    const vq_expr& ex = aggs[j].arg;
    if (aggs[j].accum == VQ_STR) {
      // MIN/MAX over strings (minMaxString, exec/aggregate.go:670-730): the
      // integer MIN/MAX of the rows' ranks in byte order, mapped back to a
      // code; a NULL string is skipped like any NULL argument
      const vq_factor& f = ex.f[0];
      if (ex.nfactors != 1 || f.kind != VQ_F_COL || ex.nwhen != 0)
        return fail(VQ_E_ARG, "exec: gpu: MIN/MAX over strings takes one column");
      if (int rc = check_col(t, f.col)) return rc;
      const ColDev& sc = t->cols[f.col];
      if ((sc.type != VQ_U8 && sc.type != VQ_U16) || sc.dict_size <= 0)
        return fail(VQ_E_ARG, "exec: gpu: MIN/MAX over strings of column %d, not a dictionary column", f.col);
      if (int rc = rank_column(t, f.col, build_order)) return rc;
      QAcc qa;
      memset(&qa, 0, sizeof qa);
      qa.e.nfactors = 1;
      qa.e.f[0].kind = VQ_F_COL;
      qa.e.f[0].op = VQ_ADD;
      qa.e.f[0].combine = VQ_ADD;
      qa.e.f[0].slot = sm.get(kRankCol + f.col);
      if (qa.e.f[0].slot < 0)
        return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
      if (sc.nul) {
        QAcc qn = qa;
        qn.op = ACC_SUM_I;
        qn.count_valid = 1;
        am.nn[j] = add_acc(qn);
        if (am.nn[j] < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d accumulators", kMaxAcc);
      }
      qa.op = aggs[j].kind == VQ_AGG_MIN ? ACC_MIN_I : ACC_MAX_I;
      am.acc[j] = add_acc(qa);
      if (am.acc[j] < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d accumulators", kMaxAcc);
      am.str_code[j] = t->code_of_rank[f.col].data();
      continue;
    }
    const bool nullable = nullable_expr(ex);
    // (a CASE whose WHEN fails takes its ELSE literal even where the THEN
    // value is NULL: the kernels' per-argument validity mask does not model
    // that, so CASE over columns with NULLs is left to the fallback)
    if (nullable_when(ex) || (ex.nwhen > 0 && nullable))
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: CASE WHEN over a column with NULLs");
    if (aggs[j].kind == VQ_AGG_COUNT && (ex.nfactors == 0 || !nullable)) continue;  // COUNT(*)
    QAcc qa;
    memset(&qa, 0, sizeof qa);
    if (int rc = add_expr(t, ex, sm, qa.e)) return rc;
    if (nullable) {
      QAcc qn = qa;
      qn.op = ACC_SUM_I;
      qn.count_valid = 1;
      am.nn[j] = add_acc(qn);
      if (am.nn[j] < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d accumulators", kMaxAcc);
    }
    if (aggs[j].kind == VQ_AGG_COUNT) continue;  // COUNT(x): the count of values alone
    const bool f = ex.is_float;
    switch (aggs[j].kind) {
      case VQ_AGG_SUM:
      case VQ_AGG_AVG: qa.op = f ? ACC_SUM_F : ACC_SUM_I; break;
      case VQ_AGG_MIN: qa.op = f ? ACC_MIN_F : ACC_MIN_I; break;
      default: qa.op = f ? ACC_MAX_F : ACC_MAX_I; break;
    }
    am.acc[j] = add_acc(qa);
    if (am.acc[j] < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: more than %d accumulators", kMaxAcc);
  }
  fill_slot_cols(t, sm, d.slot);
  d.nslots = sm.n;
  if (d.nslots == 0) {  // COUNT(*) with no predicate still needs a row source
    if (t->ncols == 0) return fail(VQ_E_ARG, "exec: gpu: table has no columns");
    int s = sm.get(0);
    d.slot[s].p = t->cols[0].d;
    d.slot[s].type = t->cols[0].type;
    d.nslots = 1;
  }
  if (G > 1 && dense_lds_bytes(d) > kDenseLdsMax) {
    if (over) *over = true;
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: %lld groups x %d accumulators exceed LDS", (long long)G, d.nacc);
  }
  return VQ_OK;
}

// The final words of a dense GROUP BY ([G][nacc + 2]: accumulators, count,
// first row) -> the caller's result: groups in first-occurrence order (the
// serial HashAggregate's key order), AVG divided on the host, float MIN/MAX
// mapped back from their order-preserving integers, the empty global
// aggregate as buildEmptyGlobalResult makes it.
int dense_result(const uint64_t* hfin, const QDesc& d, const AccMap& am, const int32_t card[4],
                 int32_t nkeys, int32_t naggs, const vq_agg* aggs, vq_agg_result* out) {
  const int NI = d.nacc + 2;
  // groups in first-occurrence order (the serial HashAggregate's key order)
  // (at most kDenseGroupsMax groups: no heap allocation here)
  std::pair<uint64_t, int> order[kDenseGroupsMax];
  int64_t norder = 0;
  for (int g = 0; g < d.ngroups; g++) {
    const uint64_t cnt = hfin[g * NI + d.nacc];
    if (cnt) order[norder++] = {hfin[g * NI + d.nacc + 1], g};
  }
  std::sort(order, order + norder);
  const bool empty_global = norder == 0 && nkeys == 0;
  const int64_t ng = empty_global ? 1 : norder;
  if (ng > out->capacity) {
    out->ngroups = ng;  // the size query: groups needed, the arrays hold no result
    return fail(VQ_E_ARG, "exec: gpu: %lld groups exceed result capacity %lld", (long long)ng,
                (long long)out->capacity);
  }
  out->ngroups = ng;
  if (empty_global) {  // buildEmptyGlobalResult (exec/aggregate.go:961-982)
    out->counts[0] = 0;
    out->first_row[0] = -1;
    for (int j = 0; j < naggs; j++) {
      out->ivals[j] = 0;
      out->fvals[j] = 0;
      out->valid[j] = aggs[j].kind == VQ_AGG_COUNT;
    }
    return VQ_OK;
  }
  for (int64_t i = 0; i < ng; i++) {
    const int g = order[i].second;
    const uint64_t* r = &hfin[g * NI];
    const int64_t cnt = (int64_t)r[d.nacc];
    int64_t rem = g;
    for (int k = nkeys - 1; k >= 0; k--) {
      const int64_t code = rem % card[k];
      out->keys[i * nkeys + k] = am.null_code[k] >= 0 && code == am.null_code[k] ? -1 : code;  // -1: NULL
      rem /= card[k];
    }
    out->counts[i] = cnt;
    out->first_row[i] = (int64_t)r[d.nacc + 1];
    for (int j = 0; j < naggs; j++) {
      const int64_t o = i * naggs + j;
      out->ivals[o] = 0;
      out->fvals[o] = 0;
      out->valid[o] = 1;
      const vq_agg& a = aggs[j];
      // rows whose argument holds a value: SUM/AVG/MIN/MAX of none are NULL
      // (exec/aggregate.go:780-876); COUNT(x) counts them
      const int64_t vals = am.nn[j] >= 0 ? (int64_t)r[am.nn[j]] : cnt;
      if (a.kind == VQ_AGG_COUNT) {
        out->ivals[o] = vals;
        continue;
      }
      if (vals == 0) {
        out->valid[o] = 0;
        continue;
      }
      uint64_t w = r[am.acc[j]];
      if (am.str_code[j]) {  // a rank in byte order -> its string's code
        out->ivals[o] = (int64_t)am.str_code[j][w];
        continue;
      }
      if ((a.kind == VQ_AGG_MIN || a.kind == VQ_AGG_MAX) && a.accum == VQ_F64) {
        const int64_t o = (int64_t)w;  // order-preserving int64 -> IEEE bits
        w = (uint64_t)(o >= 0 ? o : (int64_t)((uint64_t)o ^ 0x7FFFFFFFFFFFFFFFULL));
      }
      if (a.kind == VQ_AGG_AVG) {
        const double s = a.arg.is_float ? bits_to_f64((int64_t)w) : (double)(int64_t)w;
        out->fvals[o] = s / (double)vals;
      } else if (a.accum == VQ_F64) {
        out->fvals[o] = bits_to_f64((int64_t)w);
      } else {
        out->ivals[o] = (int64_t)w;
      }
    }
  }
  return VQ_OK;
}

int agg_packed(vq_ctx* c, const vq_table* t, const vq_pred* pred, int32_t nkeys,
               const int32_t* key_cols, int32_t naggs, const vq_agg* aggs, vq_agg_result* out,
               vq_stats* stats);

}  // namespace

extern "C" int vq_agg_dict(vq_ctx* c, const vq_table* t, const vq_pred* pred, int32_t nkeys,
                           const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                           vq_agg_result* out, vq_stats* stats) {
  if (!c || !t || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  out->ngroups = 0;  // no result unless this call succeeds
  if (t->ctx != c) return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  QDesc d;
  AccMap am;
  SlotMap sm;
  int32_t card[4] = {1, 1, 1, 1};
  bool over = false;
  if (int rc = plan_dense(t, pred, nkeys, key_cols, naggs, aggs, d, am, sm, card, &over)) {
    // more dense group slots than the per-lane accumulators hold: the codes
    // packed into one key, through the hash GROUP BY's kernels
    if (over) return agg_packed(c, t, pred, nkeys, key_cols, naggs, aggs, out, stats);
    return rc;
  }
  std::vector<uint8_t> keep;
  prune(t, pred, keep);
  // tiles: an even share of the kept rows for each wave the run-time
  // compiled kernel keeps on the device (a step of 64 lanes x U rows never
  // crosses a tile), at least 1024 rows
  DenseJit dj;
  std::string jerr;
  const bool jit = jit_dense_prepare(d, &dj, &jerr) == 0;
  if (!jit && any_nulls(t, sm))  // (the generic kernels take no NULLs)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: NULLs without the run-time compiled kernel: %s", jerr.c_str());
  {
    const int64_t kr = kept_rows(t, keep);
    if (jit && kr >= (int64_t)1 << 20) {
      const int64_t chunk = dense_chunk_rows(d.nslots);
      const int64_t share = (kr + dj.slots - 1) / dj.slots;
      d.tile_rows = (int32_t)std::max<int64_t>(1024, (share + chunk - 1) / chunk * chunk);
    } else {
      d.tile_rows = scan_tile_rows(kr);
    }
  }
  int64_t rows = 0, kept = 0;
  if (int rc = make_spans(t, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, &rows, &kept)) return rc;

  SlotLease S(c);
  if (canceled()) return fail_canceled();
  hipSetDevice(c->device);
  const int NI = d.nacc + 2;
  int32_t grid = 0;
  if (d.ntiles > 0) grid = jit ? (int32_t)std::min<int64_t>(d.ntiles, dj.cap(kept)) : dense_grid(d);
  const size_t part_words = (size_t)std::max<int32_t>(grid, 1) * d.ngroups * NI;
  const size_t out_words = (size_t)d.ngroups * NI;
  void* ar;
  if (int rc = arena_get(c, S.get(), (part_words + out_words) * 8 + 1024, &ar)) return rc;
  Carve cv;
  uint64_t* part = cv.take<uint64_t>((char*)ar, part_words);
  uint64_t* fin = cv.take<uint64_t>((char*)ar, out_words);
  void* hp;
  if (int rc = pinned_get(S.get(), out_words * 8, &hp)) return rc;
  uint64_t* hfin = (uint64_t*)hp;

  // events: around the whole query when stats are asked for, and between
  // the two kernels only when the main kernel is timed on its own (repeats
  // above 1): an event between two launches delays the second by microseconds
  const bool timed = stats != nullptr;
  const bool split = timed && c->kernel_repeats > 1;
  if (d.ntiles > 0) {
    if (timed) HIP_TRY(hipEventRecord(S->ev[0], S->stream));
    for (int32_t rep = 0; rep < c->kernel_repeats; rep++) {  // identical partials each time
      if (jit) HIP_TRY(jit_dense_launch(dj, d, grid, part, S->stream));
      else HIP_TRY(launch_agg_dense(d, grid, part, S->stream));
    }
    if (split) HIP_TRY(hipEventRecord(S->ev[1], S->stream));
    // the final words straight into the slot's pinned host buffer when the
    // device can address it (no copy command), else through a copy
    if (S->pinned_dev) {
      HIP_TRY(launch_agg_final(d, grid, part, (uint64_t*)S->pinned_dev, S->stream));
      if (timed) HIP_TRY(hipEventRecord(S->ev[2], S->stream));
    } else {
      HIP_TRY(launch_agg_final(d, grid, part, fin, S->stream));
      if (timed) HIP_TRY(hipEventRecord(S->ev[2], S->stream));
      HIP_TRY(hipMemcpyAsync(hfin, fin, out_words * 8, hipMemcpyDeviceToHost, S->stream));
    }
    HIP_TRY(hipStreamSynchronize(S->stream));
  } else {
    for (int g = 0; g < d.ngroups; g++)
      for (int i = 0; i < NI; i++) hfin[g * NI + i] = (i == d.nacc + 1) ? 0xFFFFFFFFULL : 0;
  }
  if (stats) {
    memset(stats, 0, sizeof *stats);
    if (d.ntiles > 0 && split) {
      stats->main_kernel_ms = elapsed(S->ev[0], S->ev[1]) / c->kernel_repeats;
      stats->device_ms = stats->main_kernel_ms + elapsed(S->ev[1], S->ev[2]);
    } else if (d.ntiles > 0) {  // one launch: the query's two kernels together
      stats->device_ms = elapsed(S->ev[0], S->ev[2]);
      stats->main_kernel_ms = stats->device_ms;
    }
    stats->rows_scanned = rows;
    stats->rg_total = t->nrg;
    stats->rg_scanned = kept;
    stats->bytes_alg = rows * slot_bytes_per_row(t, sm);
    stats->grid = grid;
    stats->tile_rows = d.tile_rows;
  }

  return dense_result(hfin, d, am, card, nkeys, naggs, aggs, out);
}

// ===========================================================================
// several one-group aggregates of one shape in one pass (continuous batching)
// ===========================================================================
namespace {

constexpr int32_t kMultiMax = 16;  // queries in one vq_agg_dict_multi pass

// Same shape: the same columns, predicate terms (columns, comparisons) and
// accumulators (expressions with their literals); only the predicate
// literals may differ.
bool same_shape(const QDesc& a, const QDesc& b) {
  if (a.nslots != b.nslots || a.nacc != b.nacc || a.ngroups != 1 || b.ngroups != 1) return false;
  for (int c = 0; c < a.nslots; c++) {
    const QSlot &x = a.slot[c], &y = b.slot[c];
    if (x.p != y.p || x.type != y.type || x.nterms != y.nterms || x.key_stride != y.key_stride) return false;
    for (int t = 0; t < x.nterms; t++)
      if (x.terms[t].op != y.terms[t].op || x.terms[t].slot2 != y.terms[t].slot2) return false;
  }
  return memcmp(a.acc, b.acc, sizeof(QAcc) * (size_t)a.nacc) == 0;
}

}  // namespace

extern "C" int vq_agg_dict_multi(vq_ctx* c, const vq_table* t, int32_t n, const vq_pred* preds,
                                 int32_t naggs, const vq_agg* aggs, vq_agg_result* outs,
                                 vq_stats* stats) {
  if (!c || !t || !outs || n < 1) return fail(VQ_E_ARG, "exec: gpu: NULL argument or %d queries", n);
  for (int32_t q = 0; q < n; q++) outs[q].ngroups = 0;
  if (n > kMultiMax) return fail(VQ_E_ARG, "exec: gpu: %d queries in one pass (at most %d)", n, kMultiMax);
  if (t->ctx != c) return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  auto pred_of = [&](int32_t q) -> const vq_pred* { return preds && preds[q].nterms > 0 ? &preds[q] : nullptr; };
  // each query planned as vq_agg_dict plans it; a batch that is not of one
  // shape (or one the run-time kernels cannot take) runs query by query
  std::vector<QDesc> dq((size_t)n);
  std::vector<AccMap> am((size_t)n);
  std::vector<SlotMap> smv((size_t)n);
  int32_t card[4] = {1, 1, 1, 1};
  bool one_shape = true;
  for (int32_t q = 0; q < n; q++) {
    if (int rc = plan_dense(t, pred_of(q), 0, nullptr, naggs, aggs, dq[q], am[q], smv[q], card)) return rc;
    one_shape = one_shape && same_shape(dq[0], dq[q]) && !any_nulls(t, smv[q]);
  }
  int32_t nq = 1;
  while (nq < n) nq *= 2;  // (one compiled kernel per power of two)
  DenseJit dj;
  std::string jerr;
  if (!one_shape || jit_multi_prepare(dq[0], nq, &dj, &jerr)) {
    for (int32_t q = 0; q < n; q++)
      if (int rc = vq_agg_dict(c, t, pred_of(q), 0, nullptr, naggs, aggs, &outs[q], stats ? &stats[q] : nullptr))
        return rc;
    return VQ_OK;
  }
  // the rows any query keeps (pruning never changes a result)
  std::vector<uint8_t> keep, kq;
  prune(t, pred_of(0), keep);
  for (int32_t q = 1; q < n; q++) {
    prune(t, pred_of(q), kq);
    for (size_t g = 0; g < keep.size(); g++) keep[g] |= kq[g];
  }
  // tiles and grid as vq_agg_dict sizes them for one query of this shape, so
  // each wave folds the same rows in the same order as there: over the same
  // kept row groups a query's words are the same bits alone and in the pass
  QDesc d = dq[0];
  DenseJit d1;
  if (jit_dense_prepare(d, &d1, &jerr)) {
    d1.slots = dj.slots;
    d1.slots_max = dj.slots_max;
  }
  {
    const int64_t kr = kept_rows(t, keep);
    if (kr >= (int64_t)1 << 20) {
      const int64_t chunk = dense_chunk_rows(d.nslots);
      const int64_t share = (kr + d1.slots - 1) / d1.slots;
      d.tile_rows = (int32_t)std::max<int64_t>(1024, (share + chunk - 1) / chunk * chunk);
    } else {
      d.tile_rows = scan_tile_rows(kr);
    }
  }
  int64_t rows = 0, kept = 0;
  if (int rc = make_spans(t, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, &rows, &kept)) return rc;
  // the literals, one row per query (the last repeated up to nq)
  int64_t lit0[kJitTerms];
  const int nt = jit_term_literals(d, lit0);
  const int ntw = std::max(nt, 1);

  SlotLease S(c);
  if (canceled()) return fail_canceled();
  hipSetDevice(c->device);
  const int NI = d.nacc + 2;
  const int32_t grid = d.ntiles > 0 ? (int32_t)std::min<int64_t>(d.ntiles, d1.cap(kept)) : 0;
  const size_t part_words = (size_t)std::max<int32_t>(grid, 1) * nq * NI;
  const size_t out_words = (size_t)nq * NI;
  const size_t lit_words = (size_t)nq * ntw;
  void* ar;
  if (int rc = arena_get(c, S.get(), (part_words + out_words + lit_words) * 8 + 1024, &ar)) return rc;
  Carve cv;
  uint64_t* part = cv.take<uint64_t>((char*)ar, part_words);
  uint64_t* fin = cv.take<uint64_t>((char*)ar, out_words);
  int64_t* mlit = (int64_t*)cv.take<uint64_t>((char*)ar, lit_words);
  void* hp;
  if (int rc = pinned_get(S.get(), (out_words + lit_words) * 8, &hp)) return rc;
  uint64_t* hfin = (uint64_t*)hp;
  int64_t* hlit = (int64_t*)(hfin + out_words);
  for (int32_t q = 0; q < nq; q++) {
    int64_t l[kJitTerms];
    jit_term_literals(dq[std::min(q, n - 1)], l);
    for (int i = 0; i < ntw; i++) hlit[(size_t)q * ntw + i] = i < nt ? l[i] : 0;
  }
  // the queries as the groups of one reduction
  QDesc dm = d;
  dm.ngroups = nq;
  if (d.ntiles > 0) {
    HIP_TRY(hipMemcpyAsync(mlit, hlit, lit_words * 8, hipMemcpyHostToDevice, S->stream));
    if (stats) HIP_TRY(hipEventRecord(S->ev[0], S->stream));
    HIP_TRY(jit_multi_launch(dj, d, grid, part, mlit, S->stream));
    HIP_TRY(launch_agg_final(dm, grid, part, fin, S->stream));
    if (stats) HIP_TRY(hipEventRecord(S->ev[2], S->stream));
    HIP_TRY(hipMemcpyAsync(hfin, fin, out_words * 8, hipMemcpyDeviceToHost, S->stream));
    HIP_TRY(hipStreamSynchronize(S->stream));
  } else {
    for (int32_t q = 0; q < nq; q++)
      for (int i = 0; i < NI; i++) hfin[q * NI + i] = (i == d.nacc + 1) ? 0xFFFFFFFFULL : 0;
  }
  for (int32_t q = 0; q < n; q++) {
    if (stats) {
      vq_stats& st = stats[q];
      memset(&st, 0, sizeof st);
      if (d.ntiles > 0) st.device_ms = st.main_kernel_ms = elapsed(S->ev[0], S->ev[2]);
      st.rows_scanned = rows;
      st.rg_total = t->nrg;
      st.rg_scanned = kept;
      st.bytes_alg = rows * slot_bytes_per_row(t, smv[q]);
      st.grid = grid;
      st.tile_rows = d.tile_rows;
    }
    if (int rc = dense_result(hfin + (size_t)q * NI, dq[q], am[q], card, 0, naggs, aggs, &outs[q])) return rc;
  }
  return VQ_OK;
}

// ===========================================================================
// dictionary-code GROUP BY that allocates nothing
// ===========================================================================
// The partials and final words of vq_agg_dict_noalloc live in this file-scope
// device array (each device holds its own copy for the life of the process).
// In host code `scratch` names a host-side stand-in, never the device memory:
// its device address is looked up once per device (scratch_setup, from
// vq_ctx_create with that device current) and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)2 << 20;
constexpr size_t kNoAllocScratchWords = kNoAllocScratchBytes / 8;
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchWords];

namespace {

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch[kNoAllocDevices];  // scratch's device address, per device
// calls on one device take turns on its scratch and host staging
std::mutex g_noalloc_mu[kNoAllocDevices];
uint64_t g_noalloc_host[kNoAllocDevices][kDenseGroupsMax * (kMaxAcc + 2)];

int scratch_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return VQ_OK;  // (the entry point then refuses)
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return VQ_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(VQ_E_DEVICE, "exec: gpu: device %d: scratch address: %s", device, hipGetErrorString(e));
  g_scratch[device] = (uint64_t*)p;
  return VQ_OK;
}

}  // namespace

extern "C" int vq_table_dict_order(vq_table* t, int32_t col, const uint32_t* rank) {
  if (!t || !rank) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (int rc = check_col(t, col)) return rc;
  const ColDev& c = t->cols[col];
  if ((c.type != VQ_U8 && c.type != VQ_U16) || c.dict_size <= 0)
    return fail(VQ_E_ARG, "exec: gpu: column %d is not a dictionary column", col);
  std::lock_guard<std::mutex> lk(t->rank_mu);
  return build_rank_column(t, col, rank);
}

extern "C" int vq_agg_dict_noalloc(vq_ctx* c, const vq_table* t, const vq_pred* pred, int32_t nkeys,
                                   const int32_t* key_cols, int32_t naggs, const vq_agg* aggs,
                                   vq_agg_result* out, vq_stats* stats) {
  if (!c || !t || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  out->ngroups = 0;  // no result unless this call succeeds
  if (t->ctx != c) return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  if (c->device < 0 || c->device >= kNoAllocDevices || !g_scratch[c->device])
    return fail(VQ_E_DEVICE, "exec: gpu: no scratch for device %d", c->device);
  QDesc d;
  AccMap am;
  SlotMap sm;
  int32_t card[4] = {1, 1, 1, 1};
  // (no allocation here: MIN/MAX over strings only with their order already set)
  if (int rc = plan_dense(t, pred, nkeys, key_cols, naggs, aggs, d, am, sm, card, nullptr, false)) return rc;
  if (int rc = refuse_nulls(t, sm)) return rc;  // (the generic kernels take no NULLs)
  const int NI = d.nacc + 2;
  // final words [G][NI] first (256-byte aligned), then partials [grid][G][NI]
  const size_t fin_words = (size_t)d.ngroups * NI;
  const size_t part_off = (fin_words + 31) & ~(size_t)31;
  if (part_off + fin_words > kNoAllocScratchWords)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: %d groups x %d words exceed the %zu-byte scratch", d.ngroups,
                NI, kNoAllocScratchBytes);
  std::vector<uint8_t> keep;
  prune(t, pred, keep);
  d.tile_rows = scan_tile_rows(kept_rows(t, keep));
  int64_t rows = 0, kept = 0;
  if (int rc = make_spans(t, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, &rows, &kept)) return rc;
  // the persistent kernel's waves: as many as the device holds, fewer when
  // their partials would not fit the scratch (each wave then takes more tiles)
  int64_t grid = d.ntiles > 0 ? dense_grid(d) : 0;
  grid = std::min<int64_t>(grid, (int64_t)((kNoAllocScratchWords - part_off) / fin_words));

  std::lock_guard<std::mutex> lk(g_noalloc_mu[c->device]);
  SlotLease S(c);
  hipSetDevice(c->device);
  uint64_t* fin = g_scratch[c->device];
  uint64_t* part = fin + part_off;
  uint64_t* hfin = g_noalloc_host[c->device];
  if (d.ntiles > 0) {
    // (events only when stats are asked for: one between two launches
    // delays the second by microseconds)
    if (stats) HIP_TRY(hipEventRecord(S->ev[0], S->stream));
    HIP_TRY(launch_agg_dense(d, (int32_t)grid, part, S->stream));
    if (stats) HIP_TRY(hipEventRecord(S->ev[1], S->stream));
    HIP_TRY(launch_agg_final(d, (int32_t)grid, part, fin, S->stream));
    if (stats) HIP_TRY(hipEventRecord(S->ev[2], S->stream));
    HIP_TRY(hipMemcpyAsync(hfin, fin, fin_words * 8, hipMemcpyDeviceToHost, S->stream));
    HIP_TRY(hipStreamSynchronize(S->stream));
  } else {
    for (int g = 0; g < d.ngroups; g++)
      for (int i = 0; i < NI; i++) hfin[g * NI + i] = (i == d.nacc + 1) ? 0xFFFFFFFFULL : 0;
  }
  if (stats) {
    memset(stats, 0, sizeof *stats);
    if (d.ntiles > 0) {
      stats->main_kernel_ms = elapsed(S->ev[0], S->ev[1]);
      stats->device_ms = elapsed(S->ev[0], S->ev[2]);
    }
    stats->rows_scanned = rows;
    stats->rg_total = t->nrg;
    stats->rg_scanned = kept;
    stats->bytes_alg = rows * slot_bytes_per_row(t, sm);
    stats->grid = (int32_t)grid;
    stats->tile_rows = d.tile_rows;
  }
  return dense_result(hfin, d, am, card, nkeys, naggs, aggs, out);
}

// ===========================================================================
// Filter -> selection vector
// ===========================================================================
extern "C" int vq_filter(vq_ctx* c, const vq_table* t, const vq_pred* pred, vq_selvec** out,
                         vq_stats* stats) {
  if (!c || !t || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if (t->ctx != c) return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  *out = nullptr;
  QDesc d;
  memset(&d, 0, sizeof d);
  SlotMap sm;
  if (int rc = add_pred(t, pred, sm, d.slot)) return rc;
  fill_slot_cols(t, sm, d.slot);
  d.nslots = sm.n;
  if (d.nslots == 0 && t->ncols > 0) {
    int s = sm.get(0);
    d.slot[s].p = t->cols[0].d;
    d.slot[s].type = t->cols[0].type;
    d.nslots = 1;
  }
  d.ngroups = 1;
  std::vector<uint8_t> keep;
  prune(t, pred, keep);
  d.tile_rows = scan_tile_rows(kept_rows(t, keep));
  int64_t rows = 0, kept = 0;
  if (int rc = make_spans(t, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, &rows, &kept)) return rc;
  // (columns with NULLs: the run-time compiled kernel only)
  if (any_nulls(t, sm) && !jit_available_filter(d, nullptr))
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: NULLs without the run-time compiled filter");

  SlotLease S(c);
  hipSetDevice(c->device);
  const size_t nt = (size_t)std::max<int64_t>(d.ntiles, 1);
  // pass words: tile_rows / 64 per tile (a tile's words start at its index)
  const size_t nwords = nt * (size_t)(d.tile_rows / 64) + 2;
  void* ar;
  if (int rc = arena_get(c, S.get(), nwords * 8 + nt * 8 + 64 + 2048, &ar)) return rc;
  Carve cv;
  uint64_t* bits = cv.take<uint64_t>((char*)ar, nwords);
  uint32_t* tcount = cv.take<uint32_t>((char*)ar, nt);
  uint32_t* toff = cv.take<uint32_t>((char*)ar, nt);
  uint64_t* dcount = cv.take<uint64_t>((char*)ar, 1);
  // selection-vector storage: reuse the cached buffer when it is large enough
  const size_t need = (size_t)std::max<int64_t>(rows, 1) * 4;
  vq_selvec* sv = new (std::nothrow) vq_selvec();
  if (!sv) return fail(VQ_E_NOMEM, "exec: gpu: selvec allocation");
  sv->ctx = c;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->sv_cache && c->sv_cache_bytes >= need) {
      sv->rows = (uint32_t*)c->sv_cache;
      sv->bytes = c->sv_cache_bytes;
      c->sv_cache = nullptr;
      c->sv_cache_bytes = 0;
    }
  }
  if (!sv->rows) {
    if (!reserve(c, need)) {
      delete sv;
      return fail(VQ_E_BUDGET, "exec: gpu: selection vector of %zu bytes exceeds the budget", need);
    }
    hipError_t e = hipMalloc((void**)&sv->rows, need);
    if (e != hipSuccess) {
      release(c, need);
      delete sv;
      return fail(VQ_E_DEVICE, "exec: gpu: hipMalloc: %s", hipGetErrorString(e));
    }
    sv->bytes = need;
  }
  void* hp;
  if (int rc = pinned_get(S.get(), 64, &hp)) {
    vq_selvec_destroy(sv);
    return rc;
  }
  // events around the filter when stats are asked for, and around the bits
  // kernel only when it is timed on its own (repeats above 1): an event
  // between two launches delays the second by microseconds
  const bool timed = stats != nullptr, split = timed && d.ntiles > 0 && c->kernel_repeats > 1;
  if (timed) HIP_TRY(hipEventRecord(S->ev[0], S->stream));
  const bool jit = d.ntiles > 0 && jit_available_filter(d, nullptr);
  HIP_TRY(launch_filter(d, jit, bits, tcount, toff, sv->rows, dcount, S->stream, split ? S->ev[1] : nullptr,
                        split ? S->ev[2] : nullptr, c->kernel_repeats));
  if (timed) HIP_TRY(hipEventRecord(S->ev[3], S->stream));
  HIP_TRY(hipMemcpyAsync(hp, dcount, 8, hipMemcpyDeviceToHost, S->stream));
  HIP_TRY(hipStreamSynchronize(S->stream));
  sv->count = (int64_t)*(uint64_t*)hp;
  if (stats) {
    memset(stats, 0, sizeof *stats);
    const double whole = elapsed(S->ev[0], S->ev[3]);
    // (not timed on its own: the bits kernel's share is not separated)
    stats->main_kernel_ms = split ? elapsed(S->ev[1], S->ev[2]) / c->kernel_repeats : d.ntiles ? whole : 0;
    stats->device_ms = whole - (split ? (c->kernel_repeats - 1) * stats->main_kernel_ms : 0);
    stats->rows_scanned = rows;
    stats->rg_total = t->nrg;
    stats->rg_scanned = kept;
    stats->bytes_alg = rows * slot_bytes_per_row(t, sm) + sv->count * 4;
  }
  *out = sv;
  return VQ_OK;
}

extern "C" int64_t vq_selvec_count(const vq_selvec* sv) { return sv ? sv->count : -1; }

extern "C" int vq_selvec_copy(const vq_selvec* sv, uint32_t* host_rows, int64_t cap) {
  if (!sv || (!host_rows && cap > 0)) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  const int64_t n = std::min(cap, sv->count);
  if (n <= 0) return VQ_OK;
  hipSetDevice(sv->ctx->device);
  HIP_TRY(hipMemcpy(host_rows, sv->rows, (size_t)n * 4, hipMemcpyDeviceToHost));
  return VQ_OK;
}

extern "C" int vq_selvec_destroy(vq_selvec* sv) {
  if (!sv) return VQ_OK;
  vq_ctx* c = sv->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->sv_cache) {
    c->sv_cache = sv->rows;
    c->sv_cache_bytes = sv->bytes;
  } else if (c->sv_cache_bytes < sv->bytes) {
    hipFree(c->sv_cache);
    release(c, c->sv_cache_bytes);
    c->sv_cache = sv->rows;
    c->sv_cache_bytes = sv->bytes;
  } else {
    hipFree(sv->rows);
    release(c, sv->bytes);
  }
  delete sv;
  return VQ_OK;
}

// ===========================================================================
// radix-partitioned join + aggregate
// ===========================================================================
namespace {

struct Interval {
  double lo, hi;
  bool ok;
};

Interval col_interval(const vq_table* t, int32_t col) {
  const ColDev& c = t->cols[col];
  Interval iv{0, 0, true};
  bool any = false;
  for (int32_t g = 0; g < t->nrg; g++) {
    const vq_zone& z = t->zones[(size_t)g * t->ncols + col];
    if (!z.has_minmax) continue;
    double lo, hi;
    switch (c.type) {
      case VQ_F64: lo = bits_to_f64((int64_t)z.min_bits); hi = bits_to_f64((int64_t)z.max_bits); break;
      case VQ_DATE: lo = (double)(int32_t)(uint32_t)z.min_bits; hi = (double)(int32_t)(uint32_t)z.max_bits; break;
      default: lo = (double)(int64_t)z.min_bits; hi = (double)(int64_t)z.max_bits; break;
    }
    if (!std::isfinite(lo) || !std::isfinite(hi)) iv.ok = false;
    if (!any) {
      iv.lo = lo;
      iv.hi = hi;
      any = true;
    } else {
      iv.lo = std::min(iv.lo, lo);
      iv.hi = std::max(iv.hi, hi);
    }
  }
  return iv;
}

Interval iv_op(int op, Interval a, Interval b) {
  Interval r{0, 0, a.ok && b.ok};
  switch (op) {
    case VQ_ADD: r.lo = a.lo + b.lo; r.hi = a.hi + b.hi; break;
    case VQ_SUB: r.lo = a.lo - b.hi; r.hi = a.hi - b.lo; break;
    case VQ_MUL: {
      double p[4] = {a.lo * b.lo, a.lo * b.hi, a.hi * b.lo, a.hi * b.hi};
      r.lo = *std::min_element(p, p + 4);
      r.hi = *std::max_element(p, p + 4);
      break;
    }
    default:
      if (b.lo <= 0 && b.hi >= 0) {
        r.ok = false;
      } else {
        double p[4] = {a.lo / b.lo, a.lo / b.hi, a.hi / b.lo, a.hi / b.hi};
        r.lo = *std::min_element(p, p + 4);
        r.hi = *std::max_element(p, p + 4);
      }
  }
  if (!std::isfinite(r.lo) || !std::isfinite(r.hi)) r.ok = false;
  return r;
}

// Bound of |expr| over the table from its zone maps (interval arithmetic).
Interval expr_interval(const vq_table* t, const vq_expr& e) {
  Interval v{0, 0, true};
  for (int k = 0; k < e.nfactors; k++) {
    const vq_factor& f = e.f[k];
    Interval x;
    const Interval lit{f.fval, f.fval, std::isfinite(f.fval)};
    switch (f.kind) {
      case VQ_F_COL: x = col_interval(t, f.col); break;
      case VQ_F_LIT: x = lit; break;
      case VQ_F_LIT_OP_COL: x = iv_op(f.op, lit, col_interval(t, f.col)); break;
      default: x = iv_op(f.op, col_interval(t, f.col), lit); break;
    }
    v = k == 0 ? x : iv_op(f.combine, v, x);
  }
  if (e.nwhen > 0) {  // CASE WHEN: the ELSE literal is a value too
    const double el = e.is_float ? e.else_f : (double)e.else_i;
    v = Interval{std::min(v.lo, el), std::max(v.hi, el), v.ok && std::isfinite(el)};
  }
  return v;
}

int ilog2_ceil(uint64_t x) {
  int l = 0;
  while (l < 64 && ((uint64_t)1 << l) < x) l++;
  return l;
}

// Identity of a join side for the partition-count estimate: the table, the
// key column and the predicate's columns and comparisons (not its literals,
// which a served stream varies from query to query: the last query of the
// same shape is the estimate, and a partition that overflows is retried).
std::string build_est_key(const vq_table* t, int32_t key, const vq_pred* pred) {
  char buf[96];
  snprintf(buf, sizeof buf, "%p/%d", (const void*)t, key);
  std::string k = buf;
  if (pred)
    for (int i = 0; i < pred->nterms && i < VQ_MAX_TERMS; i++) {
      const vq_term& m = pred->terms[i];
      snprintf(buf, sizeof buf, "|%d,%d", m.col, m.op);
      k += buf;
    }
  return k;
}

int plan_side(const vq_table* t, int32_t key, const vq_pred* pred, const vq_agg* aggs,
              int32_t naggs, JSide& sd, JDesc* jd, SlotMap& sm, int32_t npack = 0,
              const int32_t* pack_cols = nullptr, const int32_t* pack_shift = nullptr,
              int32_t jkey = -1) {
  memset(&sd, 0, sizeof sd);
  if (int rc = check_col(t, key)) return rc;
  if (t->cols[key].type == VQ_F64) return fail(VQ_E_UNSUPPORTED, "exec: gpu: float join key");
  if (int rc = add_pred(t, pred, sm, sd.slot)) return rc;
  int ks = sm.get(key);
  if (ks < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
  sd.key_slot = ks;
  sd.nkeys = 1;
  sd.key_slots[0] = ks;
  if (npack > 1) {
    sd.nkeys = npack;
    for (int i = 0; i < npack; i++) {
      const int s2 = sm.get(pack_cols[i]);
      if (s2 < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
      sd.key_slots[i] = s2;
      sd.key_shift[i] = pack_shift[i];
    }
  }
  if (jd) {
    for (int j = 0; j < naggs; j++)
      if (aggs[j].kind != VQ_AGG_COUNT)
        if (int rc = add_expr(t, aggs[j].arg, sm, jd->agg[j].e)) return rc;
  }
  if (jkey >= 0) {  // a join + GROUP BY of the joined rows: the probe's join key
    if (int rc = check_col(t, jkey)) return rc;
    const int s = sm.get(jkey);
    if (s < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
    sd.lookup = 1;
    sd.jkey_slot = s;
  }
  fill_slot_cols(t, sm, sd.slot);
  sd.nslots = sm.n;
  return refuse_nulls(t, sm);
}

}  // namespace

namespace {

// Tuning of one join / hash GROUP BY (the autotuner's choices; the results do
// not depend on them): radix bits relative to the planner's choice, the
// probe scatter's LDS buffer entries per partition (-1: the largest that
// fits) and the tile size relative to the default (a power of two).
struct JoinTune {
  int32_t bits_delta = 0;
  int32_t pbuf = -1;
  int32_t tile_shift = 0;
  int32_t reps = 0;  // accumulator replicas (0: the planner's choice)
};

constexpr int kRetry = -1;  // internal: repartition and run again

// A whole column's zone over all row groups (has_minmax: any group has
// rows), compared as col_interval reads the bits.
vq_zone merged_zone(const vq_table* t, int32_t col) {
  vq_zone m;
  memset(&m, 0, sizeof m);
  const int type = t->cols[col].type;
  auto val = [&](uint64_t b) {
    return type == VQ_F64 ? bits_to_f64((int64_t)b)
                          : type == VQ_DATE ? (double)(int32_t)(uint32_t)b : (double)(int64_t)b;
  };
  for (int32_t g = 0; g < t->nrg; g++) {
    const vq_zone& z = t->zones[(size_t)g * t->ncols + col];
    m.null_count += z.null_count;
    if (!z.has_minmax) continue;
    if (!m.has_minmax || val(z.min_bits) < val(m.min_bits)) m.min_bits = z.min_bits;
    if (!m.has_minmax || val(z.max_bits) > val(m.max_bits)) m.max_bits = z.max_bits;
    m.has_minmax = 1;
  }
  return m;
}

// A device lookup table's shape for keys in [lo, hi] over `rows` rows: a
// direct table (head[key - lo]) when the range is at most four times the
// rows, else a hash table of a power of two of slots, at least twice the
// rows, one-word slots when every key lies in [0, 2^32 - 1).
struct LookupPlan {
  uint32_t cap = 16;
  uint32_t range = 0;
  int64_t base = 0;
  bool packed = false;
};

LookupPlan plan_lookup(int64_t rows, bool bounded, int64_t lo, int64_t hi) {
  LookupPlan p;
  if (bounded && hi >= lo && (uint64_t)(hi - lo) < (uint64_t)std::min<int64_t>(INT32_MAX, std::max<int64_t>(4 * rows, 4096))) {
    p.range = (uint32_t)(hi - lo + 1);
    p.base = lo;
    return p;
  }
  while ((int64_t)p.cap < 2 * rows) p.cap *= 2;
  p.packed = !bounded || (lo >= 0 && hi < (int64_t)0xFFFFFFFFll);
  return p;
}

// The same from a key column's zone maps.
LookupPlan plan_lookup_col(const vq_table* t, int32_t col, int64_t rows) {
  const vq_zone z = merged_zone(t, col);
  const int type = t->cols[col].type;
  const int64_t lo = type == VQ_DATE ? (int64_t)(int32_t)(uint32_t)z.min_bits : (int64_t)z.min_bits;
  const int64_t hi = type == VQ_DATE ? (int64_t)(int32_t)(uint32_t)z.max_bits : (int64_t)z.max_bits;
  return plan_lookup(rows, z.has_minmax != 0, lo, hi);
}

size_t lookup_bytes(const LookupPlan& p, int64_t rows) {
  const size_t slots = p.range ? p.range : p.cap;
  return (p.range ? 0 : (size_t)p.cap * 8) + slots * 8 + (size_t)std::max<int64_t>(rows, 1) * 4 + 16 + 6 * 256;
}

void carve_lookup(Carve& cv, char* dev, const LookupPlan& p, int64_t rows, LookupBufs& lk) {
  const size_t slots = p.range ? p.range : p.cap;
  lk.keys = cv.take<int64_t>(dev, p.range ? 1 : p.cap);
  lk.head = cv.take<uint32_t>(dev, slots);
  lk.next = cv.take<uint32_t>(dev, (size_t)std::max<int64_t>(rows, 1));
  lk.cnt = cv.take<uint32_t>(dev, slots);
  lk.min_head = cv.take<uint32_t>(dev, 4);
  lk.mask = p.cap - 1;
  lk.packed = p.packed ? 1u : 0u;
  lk.base = p.base;
  lk.range = p.range;
  lk.pad_ = 0;
}

// One join + aggregate (or, with groupby, one GROUP BY of the probe side's
// key with no build side): planned on the host, bound to device and pinned
// memory, enqueued on a stream (or captured into a graph), then turned into
// a result.  A partition table or accumulator pool that overflows makes the
// result step ask for another attempt with a larger pool or one more bit.
struct JoinJob {
  vq_ctx* c = nullptr;
  const vq_table* build = nullptr;
  const vq_table* probe = nullptr;
  int32_t build_key = 0, probe_key = 0;
  vq_pred bpred, ppred;
  const vq_pred* bp = nullptr;
  const vq_pred* pp = nullptr;
  int32_t naggs = 0;
  vq_agg aggs[VQ_MAX_AGGS];
  int32_t radix_bits = -1, topk = 0;
  bool groupby = false;
  // GROUP BY several dictionary-code columns: their codes packed into one
  // key (exec/agg_intkey.go:67-100), unpacked again into the result
  int32_t npack = 0;
  int32_t pack_cols[8] = {0}, pack_shift[8] = {0};
  JoinTune tune;
  // fixed by setup
  JDesc jd;
  JSide bs, ps;
  SlotMap bsm, psm;
  int nw = 0, nvals = 0;
  std::string est_key, grp_key;
  int64_t est = 1, groups_seen = -1;
  int lds_dev = 0, lds_max = 0;
  int32_t cap_max = 4096;
  bool dev_topk = false;
  int bits = 0;
  int32_t pool_req = 0;
  Slot::Pinned* pin = nullptr;  // pinned staging for the records (null: pageable)
  bool no_local = false;   // GROUP BY: too many keys for per-workgroup partials
  bool livebits = false;   // the probe side's count pass records its live rows
  LocalLaunch ll;
  int attempts = 0;        // plans run (more than one: a table or pool overflowed)
  // per attempt
  JoinPlan plan;
  int32_t pool_max = 0;
  PartLaunch bl, pl;
  JoinBufs B;
  int64_t brows = 0, bkept = 0, prows = 0, pkept = 0, cap_b = 1, cap_p = 1, rec_cap = 1;
  int RW = 3;
  size_t bytes = 0;        // device scratch
  size_t host_bytes = 0;   // pinned result staging
  uint32_t* hflags = nullptr;
  uint64_t* hout = nullptr;
  // vq_join_group: the GROUP BY runs over `view` (the probe's columns, then
  // the build's, over the probe's row groups) with the build side in a
  // lookup table: lq its predicate's and key's slots over its kept spans
  bool lookup = false;
  vq_table view;
  const vq_table* lprobe = nullptr;
  const vq_table* lbuild = nullptr;
  int32_t lbuild_key = 0, ljkey = 0, lsplit = 0, lkey_slot = 0;
  vq_pred lbpred;
  const vq_pred* lbp = nullptr;
  QDesc lq;
  LookupPlan lplan;
  int64_t lrows = 0, lkept = 0, lrow_bytes = 0;
  int64_t lmult = 1;  // the most build rows one key has, as planned
  // the direct join (build keys in a dense range): the build side's scan,
  // the compiled probe, the region's size and, once bound, its address
  bool no_direct = false;  // batches (instances share a slot), VQ_NO_DIRECT
  QDesc dq;
  int32_t dkey_slot = 0;
  DirectLaunch dl;
  int64_t dbase = 0;
  uint64_t drange = 0;
  char* dreg = nullptr;
  uint64_t dreg_range = 0;  // the region's layout (its key capacity, >= drange)
};

// Table fill: a GROUP BY's rows all find their keys, but fill the table by
// compare-and-swap as they arrive; a join's build keys go in first, and the
// semi-join filter leaves few probe rows whose key is missing (a miss walks
// to an empty slot), so a join's tables fill further.
constexpr double kLoad = 0.4;
constexpr double kLoadJoin = 0.55;

// Partition tables sized for the keys a partition may get, not just the mean:
// `per` keys on average, with Poisson spread over the partitions.
double keys_high(double per) { return per + 4 * std::sqrt(per) + 8; }
int32_t table_cap_for(double per, int32_t cap_max, double load = kLoad) {
  int32_t cap = 8;
  while (cap < cap_max && (cap * load < per || cap < keys_high(per))) cap *= 2;
  return cap;
}
constexpr int32_t kPoolMax = 2048;  // the join kernel's top-k holds 2 entries per thread
// A GROUP BY of few keys runs through per-workgroup partials, in passes over
// the rows of local_limit(nwords) keys each: 1024, 512 or 256 keys, as many
// as one copy of their accumulators leaves room for in the kernel's LDS.  Up
// to kLocalPasses passes (a pass reads every row, so more keys go through the
// partitions, which read them about three times); a GROUP BY of more keys,
// or whose key count is not known yet and turns out larger, through the
// partitions.
// The key table (every pass's keys at load <= 0.5) is at most kLocalSlots
// slots: the key-collecting kernel holds a copy in LDS.
constexpr int32_t kLocalPasses = 8;
constexpr int32_t kLocalSlots = 8192;
int32_t local_limit(int32_t nwords) {
  for (int32_t limit = 1024; limit >= 256; limit /= 2)
    if (local_lds_for(2 * limit, limit, nwords) <= (size_t)kLocalLds) return limit;
  return 0;
}
int32_t local_passes_max(int32_t limit) {
  return limit > 0 ? std::max(1, std::min(kLocalPasses, kLocalSlots / (2 * limit))) : 0;
}
int64_t local_max_keys(int32_t nwords) {
  const int32_t limit = local_limit(nwords);
  return (int64_t)limit * local_passes_max(limit);
}

// Accumulator words and fixed-point scales: a float sum's three limbs of w
// bits leave 63 - w bits of headroom each, so w follows the additions a
// group can take (the probe's rows).
int join_scales(JoinJob& j) {
  // (a join + GROUP BY of joined rows: each probe row may meet lmult build
  // rows, each one addition)
  const int L = ilog2_ceil((uint64_t)j.probe->nrows * (uint64_t)std::max<int64_t>(j.lmult, 1) + 1);
  int nw = 0, nvals = 0;
  for (int k = 0; k < j.naggs; k++) {
    if (int rc = check_agg(j.aggs[k])) return rc;
    JAgg& a = j.jd.agg[k];
    a.kind = j.aggs[k].kind;
    a.accum = j.aggs[k].accum;
    a.word0 = nw;
    a.vidx = -1;
    if (a.kind == VQ_AGG_COUNT) {
      a.nwords = 0;
      continue;
    }
    // one value per distinct argument expression (partition entries carry
    // each once)
    for (int k2 = 0; k2 < k && a.vidx < 0; k2++)
      if (j.jd.agg[k2].vidx >= 0 && memcmp(&j.jd.agg[k2].e, &a.e, sizeof a.e) == 0)
        a.vidx = j.jd.agg[k2].vidx;
    if (a.vidx < 0) a.vidx = nvals++;
    if ((a.kind == VQ_AGG_SUM || a.kind == VQ_AGG_AVG) && j.aggs[k].arg.is_float) {
      Interval iv = expr_interval(j.probe, j.aggs[k].arg);
      if (!iv.ok)
        return fail(VQ_E_UNSUPPORTED, "exec: gpu: aggregate %d has no finite bound for exact summation", k);
      double m = std::max(std::fabs(iv.lo), std::fabs(iv.hi)) * (1 + 1e-9);
      int E = 0;
      if (m > 0) std::frexp(m, &E);  // m < 2^E
      const int w = std::min(40, 62 - L);
      if (w < 20) return fail(VQ_E_UNSUPPORTED, "exec: gpu: probe side too large for exact sums");
      a.fs.w = w;
      a.fs.lsb = E - 3 * w;
      a.nwords = 3;
    } else {
      a.nwords = 1;
    }
    nw += a.nwords;
  }
  j.jd.nwords = nw;
  j.nw = nw;
  j.nvals = nvals;
  return VQ_OK;
}

int join_setup(JoinJob& j) {
  vq_ctx* c = j.c;
  if (!c || (!j.build && !j.groupby) || !j.probe) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  if ((!j.groupby && j.build->ctx != c) || j.probe->ctx != c)
    return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  if (j.naggs < 0 || j.naggs > VQ_MAX_AGGS) return fail(VQ_E_ARG, "exec: gpu: %d aggregates", j.naggs);
  if (j.topk < 0) return fail(VQ_E_ARG, "exec: gpu: top-k of %d", j.topk);
  if (j.radix_bits > 16) return fail(VQ_E_UNSUPPORTED, "exec: gpu: %d radix bits", j.radix_bits);
  memset(&j.jd, 0, sizeof j.jd);
  j.jd.naggs = j.naggs;
  memset(&j.bs, 0, sizeof j.bs);
  if (!j.groupby)
    if (int rc = plan_side(j.build, j.build_key, j.bp, nullptr, 0, j.bs, nullptr, j.bsm)) return rc;
  if (int rc = plan_side(j.probe, j.probe_key, j.pp, j.aggs, j.naggs, j.ps, &j.jd, j.psm, j.npack,
                         j.pack_cols, j.pack_shift, j.lookup ? j.ljkey : -1))
    return rc;
  if (j.lookup)  // the slots that hold build-side columns
    for (int i = 0; i < j.psm.n; i++)
      if (j.psm.col_of[i] >= j.lsplit) j.ps.build_slots |= 1 << i;
  if (int rc = join_scales(j)) return rc;
  const int nw = j.nw;
  // Partitioning plan.  Each partition's table lives in one workgroup's LDS:
  // cap slots of key and tag (multiplicity, then pool index), plus a pool of
  // accumulator entries (count, first row, multiplicity, accumulator words)
  // for the keys a probe row matches.  The build-side estimate is the scan's
  // row count before filtering, as buildSideRadixOptions uses
  // (planner/parallel_join.go:327-339), until a query of the same shape over
  // the same build side has run: then the rows that passed last time (for a
  // GROUP BY, the distinct keys).  A table at most 4096 slots filled to at
  // most kLoad (a join: kLoadJoin) keeps linear probing short for every lane
  // of a wave (a wave waits for its longest probe).  The pool starts at half the table, or at
  // twice the matched keys per partition of the last query of the same shape,
  // and doubles when it fills; a table that fills adds one bit.
  j.est_key = j.groupby ? "G|" + build_est_key(j.lookup ? j.lprobe : j.probe, j.probe_key, j.pp)
                        : build_est_key(j.build, j.build_key, j.bp);
  if (j.lookup) j.est_key += "|J" + build_est_key(j.lbuild, j.lbuild_key, j.lbp) + "/" + std::to_string(j.ljkey);
  for (int i = 0; j.npack > 1 && i < j.npack; i++) j.est_key += "#k" + std::to_string(j.pack_cols[i]);
  j.grp_key = j.groupby ? j.est_key : j.est_key + "#" + build_est_key(j.probe, j.probe_key, j.pp);
  j.est = std::max<int64_t>(j.groupby ? j.probe->nrows : j.build->nrows, 1);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->build_est.find(j.est_key);
    if (it != c->build_est.end()) j.est = std::max<int64_t>(it->second, 1);
    it = c->build_est.find(j.grp_key);
    if (it != c->build_est.end()) j.groups_seen = it->second;
    if (j.groupby && (c->build_est.count("L|" + j.est_key) ||
                      (j.groups_seen >= 0 && j.est > local_max_keys(nw))))
      j.no_local = true;
  }
  hipDeviceGetAttribute(&j.lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
  if (j.lds_dev <= 0) j.lds_dev = 64 * 1024;
  j.lds_max = j.lds_dev - 8192;  // the join kernel's own static LDS (about 4 KB)
  j.dev_topk = j.topk > 0 && j.topk <= 16;
  j.cap_max = 4096;
  while (j.cap_max > 8 && join_lds_for(j.cap_max, 64, nw, 1, j.dev_topk) > (size_t)j.lds_max)
    j.cap_max /= 2;
  j.bits = j.radix_bits;
  if (j.bits < 0) {
    // enough partitions for the keys to fit their tables, and enough for the
    // work: one workgroup per partition takes at most ~16K probe rows
    // (a handful of keys over millions of rows still spreads over the CUs)
    // (a GROUP BY needs a pool entry for every key of its partition, so
    // the partition's keys must fit the pool next to the table as well)
    auto fits = [&](int b) {
      const double per = (double)j.est / (1 << b);
      if (per > (j.groupby ? kLoad : kLoadJoin) * j.cap_max) return false;
      if (!j.groupby) return true;
      const double need = keys_high(per) + 56;
      return need <= kPoolMax &&
             join_lds_for(table_cap_for(per, j.cap_max), (int32_t)need, nw) <= (size_t)j.lds_max;
    };
    j.bits = 0;
    while (j.bits < 16 && !fits(j.bits)) j.bits++;
    const int64_t prows = j.probe->nrows;
    int pbits = 0;
    while (pbits < 10 && (prows >> pbits) > 16384) pbits++;
    j.bits = std::max(j.bits, pbits);
    j.bits = std::max(0, std::min(16, j.bits + j.tune.bits_delta));
  }
  return VQ_OK;
}

int join_plan_local(JoinJob& j) {
  std::string jerr;
  JoinPlan& plan = j.plan;
  plan.local = 1;
  plan.groupby = 1;
  plan.P = 1;
  plan.limit = local_limit(j.nw);
  plan.lcap = 2 * plan.limit;
  // passes: as many as the keys need (all of them when the key count is not
  // known yet); the key table holds every pass's keys at load <= 0.5
  plan.passes = local_passes_max(plan.limit);
  if (plan.limit > 0 && j.groups_seen >= 0)
    plan.passes = (int32_t)std::max<int64_t>(
        1, std::min<int64_t>(plan.passes, (j.est + plan.limit - 1) / plan.limit));
  // dictionary-code keys whose every combination fits the passes: the key
  // table is seeded with all of them (no pass over the rows to find them)
  plan.seed_groups = 0;
  if (plan.limit > 0) {
    const int nk = j.npack > 1 ? j.npack : 1;
    int64_t G = 1;
    bool codes = true;
    for (int i = 0; i < nk && codes; i++) {
      const ColDev& cd = j.probe->cols[j.npack > 1 ? j.pack_cols[i] : j.probe_key];
      codes = (cd.type == VQ_U8 || cd.type == VQ_U16) && cd.dict_size > 0;
      plan.seed_card[i] = codes ? cd.dict_size : 0;
      G = std::min<int64_t>(G * std::max(cd.dict_size, 1), (int64_t)1 << 40);
    }
    if (codes && G <= (int64_t)plan.limit * local_passes_max(plan.limit)) {
      plan.seed_groups = G;
      plan.passes = (int32_t)((G + plan.limit - 1) / plan.limit);
    }
  }
  plan.gcap = plan.lcap;
  while (plan.gcap < 2 * plan.limit * plan.passes) plan.gcap *= 2;
  plan.lds_max = j.lds_max;
  if (plan.limit == 0) {
    j.no_local = true;
    return kRetry;
  }
  std::vector<uint8_t> pkeep;
  j.brows = j.bkept = j.prows = j.pkept = 0;
  const int ts = std::max(0, std::min(4, j.tune.tile_shift));
  prune(j.probe, j.pp, pkeep);
  j.ps.tile_rows = join_tile_rows(kept_rows(j.probe, pkeep)) << ts;
  if (int rc = make_spans(j.probe, pkeep, j.ps.tile_rows, j.ps.span, &j.ps.nspans, &j.ps.ntiles,
                         &j.prows, &j.pkept))
    return rc;
  if (j.prows >= (int64_t)UINT32_MAX)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: GROUP BY above 2^32 rows");
  j.RW = 3 + j.naggs;
  j.rec_cap = (int64_t)plan.limit * plan.passes;
  memset(&j.B, 0, sizeof j.B);
  if (jit_local_prepare(j.ps, j.jd, plan.lcap, &j.ll, &jerr))
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: GROUP BY kernels unavailable: %s", jerr.c_str());
  plan.lgrid = j.ll.grid;
  plan.kgrid = (int32_t)std::max<int64_t>(1, std::min<int64_t>(j.ps.ntiles, 2 * (int64_t)j.ll.grid));
  size_t bytes = 16 * 4 + (size_t)plan.gcap * 8 + (size_t)plan.limit * plan.passes * 8 +
                 (size_t)plan.lgrid * plan.limit * (j.nw + 2) * 8 + (size_t)j.rec_cap * j.RW * 8;
  if (j.lookup)  // the build side's lookup table: keys, heads, counts, a next row per build row
    bytes += lookup_bytes(j.lplan, j.lbuild->nrows);
  bytes += 12 * 256;  // alignment of the carved pieces
  j.bytes = (bytes + 255) & ~(size_t)255;
  j.host_bytes = 64 + (size_t)16 * j.RW * 8;
  return VQ_OK;
}

int scan_desc(const vq_table* t, const vq_pred* pred, int nkeys, const int32_t* keys, QDesc& d, int* key_slots,
              int64_t* rows, int64_t* kept, int32_t tile_rows = kLookupTileRows);

// The direct join: when the build key's zone maps bound it to a range of at
// most kDirectDensity times the build rows (primary keys, surrogate keys)
// and the range's accumulators fit kDirectMaxBytes, the build side is one
// bit per key of the range and the groups are indexed by key (no
// partitions).  Returns kRetry when the join does not qualify.
constexpr uint64_t kDirectDensity = 16;
constexpr uint64_t kDirectMaxBytes = (uint64_t)8 << 30;

bool direct_disabled() {
  const char* e = getenv("VQ_NO_DIRECT");
  return e && *e && *e != '0';
}

// The region of a direct join over `range` keys: flags, the keys' 32-byte
// records, bits, marks.
uint64_t direct_region_bytes(uint64_t range) {
  auto up = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
  const uint64_t words = (range + 31) / 32;
  return 256 + up(range * 32) + up(words * 4) + up(words * 32);
}

int join_plan_direct(JoinJob& j) {
  // (radix bits given by the caller ask for the partitioned join)
  if (j.groupby || j.lookup || j.no_direct || j.radix_bits >= 0 || direct_disabled()) return kRetry;
  const int bt = j.build->cols[j.build_key].type, pt = j.probe->cols[j.probe_key].type;
  auto keytype = [](int t) { return t == VQ_I64 || t == VQ_DATE || t == VQ_U8 || t == VQ_U16; };
  if (!keytype(bt) || !keytype(pt)) return kRetry;
  const vq_zone z = merged_zone(j.build, j.build_key);
  if (!z.has_minmax) return kRetry;
  const int64_t lo = bt == VQ_DATE ? (int64_t)(int32_t)(uint32_t)z.min_bits : (int64_t)z.min_bits;
  const int64_t hi = bt == VQ_DATE ? (int64_t)(int32_t)(uint32_t)z.max_bits : (int64_t)z.max_bits;
  if (hi < lo) return kRetry;
  const uint64_t span = (uint64_t)hi - (uint64_t)lo;  // range - 1
  if (span >= (uint64_t)0xFFFFFFFEull ||
      span >= kDirectDensity * (uint64_t)std::max<int64_t>(j.build->nrows, 4096))
    return kRetry;
  const uint64_t range = span + 1;
  const uint64_t rbytes = direct_region_bytes(range);
  if (rbytes > kDirectMaxBytes) return kRetry;
  {
    // a build side whose passing keys repeated last time: the partitions
    // count multiplicities, this path does not
    std::lock_guard<std::mutex> lk(j.c->mu);
    if (j.c->build_est.count("D|" + build_est_key(j.build, j.build_key, j.bp))) return kRetry;
  }
  std::string jerr;
  JoinPlan& plan = j.plan;
  const int ts = std::max(0, std::min(4, j.tune.tile_shift));
  std::vector<uint8_t> pkeep;
  prune(j.probe, j.pp, pkeep);
  j.ps.tile_rows = join_tile_rows(kept_rows(j.probe, pkeep), part_chunk_rows(j.ps.nslots), device_cus()) << ts;
  j.brows = j.bkept = j.prows = j.pkept = 0;
  if (int rc = make_spans(j.probe, pkeep, j.ps.tile_rows, j.ps.span, &j.ps.nspans, &j.ps.ntiles,
                         &j.prows, &j.pkept))
    return rc;
  const int64_t cap = std::max<int64_t>(1, j.ps.ntiles * (int64_t)j.ps.tile_rows);
  if (cap >= (int64_t)UINT32_MAX) return kRetry;
  // (256-row tiles: one step of a wave each; 1024 took 18.1 us over the
  // 1.5M-row build against 16.0, 4096 45.9)
  if (int rc = scan_desc(j.build, j.bp, 1, &j.build_key, j.dq, &j.dkey_slot, &j.brows, &j.bkept)) return rc;
  if (jit_direct_prepare(j.ps, j.jd, &j.dl, &jerr)) return kRetry;
  j.dbase = lo;
  j.drange = range;
  plan.direct = 1;
  plan.bits = 0;
  plan.topk = j.dev_topk ? j.topk : 0;
  plan.nvals = j.nvals;
  plan.P = j.dl.grid;  // the collect kernel's workgroups: one per probe workgroup
  j.RW = 3 + j.naggs;
  j.rec_cap = std::max<int64_t>(1, std::min<int64_t>(j.brows, j.prows));
  memset(&j.B, 0, sizeof j.B);
  size_t bytes = 64 * 4;
  bytes += (size_t)cap * (4 + 4 + 8 * (1 + (size_t)j.nvals)) + (size_t)plan.P * 8;
  bytes += (size_t)std::max<int64_t>(j.dq.ntiles, 1) * 4 + (size_t)((range + 31) / 32 + 255) / 256 * 4;
  bytes += j.dev_topk ? (size_t)plan.P * j.topk * (j.RW * 8 + 16) + (size_t)16 * j.RW * 8
                      : (size_t)j.rec_cap * j.RW * 8;
  bytes += 14 * 256;
  j.bytes = (bytes + 255) & ~(size_t)255;
  j.host_bytes = 64 + (size_t)16 * j.RW * 8;
  return VQ_OK;
}

int join_plan(JoinJob& j) {
  std::string jerr;
  JoinPlan& plan = j.plan;
  memset(&plan, 0, sizeof plan);
  {
    const int rc = join_plan_direct(j);
    if (rc != kRetry) return rc;
    memset(&plan, 0, sizeof plan);
  }
  if (j.groupby && !j.no_local) {
    const int rc = join_plan_local(j);
    if (rc != kRetry) return rc;
    memset(&plan, 0, sizeof plan);
  }
  if (j.lookup)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: join GROUP BY with more group keys than the per-workgroup partials hold");
  plan.bits = j.bits;
  plan.P = 1 << j.bits;
  const double per = (double)j.est / plan.P;
  const int32_t cap = table_cap_for(per, j.cap_max, j.groupby ? kLoad : kLoadJoin);
  plan.cap = cap;
  // the largest pool the LDS and the kernel's top-k allow
  j.pool_max = std::min<int32_t>(kPoolMax, cap);
  while (j.pool_max > 64 && join_lds_for(cap, j.pool_max, j.nw, 1, j.dev_topk) > (size_t)j.lds_max)
    j.pool_max -= 64;
  // twice the matched keys per partition of the last query of this shape
  // (a GROUP BY matches every key: at least their high estimate), else half
  // the table
  const double gper = (double)j.groups_seen / plan.P;
  int32_t pool = j.groups_seen >= 0
                     ? (int32_t)std::min<double>(kPoolMax, std::max(2 * gper, j.groupby ? keys_high(gper) : 0) + 64)
                     : cap / 2;
  if (j.pool_req) pool = j.pool_req;
  plan.pool = std::max<int32_t>(64, std::min(pool, j.pool_max));
  // accumulator replicas when a partition's keys take many rows each (their
  // atomics would queue on a few LDS addresses), as many as fit without
  // costing the kernel its second workgroup per CU (or the LDS limit when it
  // has none)
  plan.reps = 1;
  {
    const double keys_pp = std::max(1.0, (double)(j.groups_seen >= 0 ? j.groups_seen : j.est) / plan.P);
    const double rows_pk = (double)j.probe->nrows / plan.P / keys_pp;
    const size_t one = join_lds_for(cap, plan.pool, j.nw, 1, j.dev_topk);
    const size_t half = (size_t)j.lds_dev / 2 - 8192;
    const size_t budget = one <= half ? half : (size_t)j.lds_max;
    int reps = j.tune.reps > 0 ? j.tune.reps : rows_pk >= 64 ? 16 : 1;
    while (reps > 1 && join_lds_for(cap, plan.pool, j.nw, reps, j.dev_topk) > budget) reps /= 2;
    plan.reps = reps;
  }
  plan.lds_bytes = (int32_t)join_lds_for(cap, plan.pool, j.nw, plan.reps, j.dev_topk);
  plan.lds_max = j.lds_max;
  plan.groupby = j.groupby ? 1 : 0;
  plan.topk = j.dev_topk ? j.topk : 0;
  plan.tnodes = j.dev_topk ? topk_tree_nodes(plan.P) : 0;
  // the semi-join filter: about 16 bits per expected build key (two set per
  // key), in 64-byte lines, a power of two of them per partition; at most
  // 2^25 bits (4 MB, so it stays in one XCD's L2) and 64 KB per partition
  plan.bloom_lpp = -1;
  if (!j.groupby) {
    const double want = 16.0 * (double)std::max<int64_t>(j.est, 4096) / 512.0 / plan.P;
    int lpp = 0;
    while (lpp < 10 && (double)(1u << lpp) < want && plan.bits + lpp + 9 < 25) lpp++;
    plan.bloom_lpp = lpp;
  }
  plan.nvals = j.nvals;
  const int P = plan.P;
  // more than 512 partitions: two passes, the first into 256 coarse ones
  // (so its LDS buffers stay deep enough to write whole lines)
  plan.P1 = P > 512 ? 256 : P;
  plan.bits1 = P > 512 ? 8 : plan.bits;
  const int P1 = plan.P1;
  const bool two = P1 < P;
  plan.bbuf = part_buffer_entries(P1, 8, j.lds_dev);
  plan.pbuf = part_buffer_entries(P1, 8 + 4 + 8 * j.nvals, j.lds_dev);
  if (j.tune.pbuf >= 0) plan.pbuf = std::min(plan.pbuf, j.tune.pbuf);

  std::vector<uint8_t> bkeep, pkeep;
  j.brows = j.bkept = j.prows = j.pkept = 0;
  const int ts = std::max(0, std::min(4, j.tune.tile_shift));
  if (!j.groupby) {
    prune(j.build, j.bp, bkeep);
    j.bs.tile_rows = join_tile_rows(kept_rows(j.build, bkeep), part_chunk_rows(j.bs.nslots), device_cus()) << ts;
    if (int rc = make_spans(j.build, bkeep, j.bs.tile_rows, j.bs.span, &j.bs.nspans, &j.bs.ntiles,
                           &j.brows, &j.bkept))
      return rc;
  }
  prune(j.probe, j.pp, pkeep);
  j.ps.tile_rows = join_tile_rows(kept_rows(j.probe, pkeep), part_chunk_rows(j.ps.nslots), device_cus()) << ts;
  if (int rc = make_spans(j.probe, pkeep, j.ps.tile_rows, j.ps.span, &j.ps.nspans, &j.ps.ntiles,
                         &j.prows, &j.pkept))
    return rc;
  j.RW = 3 + j.naggs;
  j.rec_cap = std::max<int64_t>(1, j.groupby ? j.prows : std::min(j.brows, j.prows));
  memset(&j.B, 0, sizeof j.B);
  j.cap_b = std::max<int64_t>(j.brows, 1);
  j.cap_p = std::max<int64_t>(j.prows, 1);
  if (j.cap_b >= (int64_t)INT32_MAX || j.cap_p >= (int64_t)UINT32_MAX)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: join side above 2^31 build / 2^32 probe rows");
  const bool bloom = plan.bloom_lpp >= 0;
  // with the filter, the probe side's count pass hands its live rows to the
  // scatter as bits (row groups starting at multiples of 64 rows)
  j.livebits = bloom;
  for (int i = 0; i < j.ps.nspans; i++)
    if (j.ps.span[i].start % 64) j.livebits = false;
  if ((!j.groupby && jit_part_prepare(j.bs, nullptr, P1, plan.bbuf, &j.bl, &jerr)) ||
      jit_part_prepare(j.ps, &j.jd, P1, plan.pbuf, &j.pl, &jerr, bloom, j.livebits))
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: join kernels unavailable: %s", jerr.c_str());
  j.B.b.G = j.groupby ? 1 : j.bl.grid;
  j.B.p.G = j.pl.grid;
  size_t bytes = 0;
  bytes += (size_t)j.cap_b * 8 + (size_t)(P1 + 1) * 4 + (size_t)j.B.b.G * P1 * 4;
  bytes += (size_t)j.cap_p * 12 + (size_t)std::max(j.nvals, 1) * j.cap_p * 8 + (size_t)(P1 + 1) * 4 +
           (size_t)j.B.p.G * P1 * 4;
  if (two)  // the final layout of both sides
    bytes += (size_t)j.cap_b * 8 + (size_t)j.cap_p * 12 + (size_t)std::max(j.nvals, 1) * j.cap_p * 8 +
             (size_t)2 * (P + 1) * 4;
  bytes += (16 + (size_t)plan.tnodes) * 4 + (size_t)2 * P * 4;
  bytes += j.dev_topk ? (size_t)P * j.topk * (j.RW * 8 + 16) + (size_t)16 * j.RW * 8
                      : (size_t)j.rec_cap * j.RW * 8;
  bytes += bloom ? (size_t)P * 64 << plan.bloom_lpp : 0;
  bytes += j.livebits ? ((size_t)j.probe->nrows / 64 + 2) * 8 : 0;
  bytes += 16 * 256;  // alignment of the carved pieces
  j.bytes = (bytes + 255) & ~(size_t)255;
  j.host_bytes = 64 + (size_t)16 * j.RW * 8;
  return VQ_OK;
}

void join_bind(JoinJob& j, char* dev, char* host) {
  const int P = j.plan.P;
  Carve cv;
  JoinBufs& B = j.B;
  B.flags = cv.take<uint32_t>(dev, 16 + (size_t)j.plan.tnodes);
  B.tctr = B.flags + 16;
  j.hflags = (uint32_t*)host;
  j.hout = (uint64_t*)(host + 64);
  if (j.plan.direct) {
    // the region (laid out for its key capacity, the same for every query:
    // what one query leaves uncleared, its slots, never lies where another's
    // cleared arrays do) and this query's entries
    DirectBufs& D = B.d;
    auto up = [](uint64_t x) { return (size_t)((x + 255) & ~(uint64_t)255); };
    const uint64_t words = (j.drange + 31) / 32, rc = j.dreg_range, rwords = (rc + 31) / 32;
    char* r = j.dreg;
    D.rflags = (uint32_t*)r;
    r += 256;
    D.rec = (uint32_t*)r;
    r += up(rc * 32);
    D.pres = (uint32_t*)r;
    r += up(rwords * 4);
    D.mark = (uint8_t*)r;
    D.base = j.dbase;
    D.range = j.drange;
    D.words = words;
    D.ntiles = j.ps.ntiles;
    D.tile_rows = j.ps.tile_rows;
    D.grid = j.dl.grid;
    D.nvals = j.nvals;
    const int64_t cap = std::max<int64_t>(1, j.ps.ntiles * (int64_t)j.ps.tile_rows);
    D.cap = cap;
    D.ew = 1 + j.nvals;
    D.mrow = cv.take<uint32_t>(dev, (size_t)cap);
    D.mcount = cv.take<uint32_t>(dev, (size_t)P);
    D.ent = cv.take<uint64_t>(dev, (size_t)cap * D.ew);
    D.tlist = cv.take<uint32_t>(dev, (size_t)cap);
    D.tcount = cv.take<uint32_t>(dev, (size_t)P);
    D.nbcnt = j.dq.ntiles;  // (written by the mark kernel, one per tile)
    D.npcnt = (int64_t)std::max<uint64_t>(1, (words + 255) / 256);
    D.bcnt = cv.take<uint32_t>(dev, (size_t)std::max<int64_t>(D.nbcnt, 1));
    D.pcnt = cv.take<uint32_t>(dev, (size_t)D.npcnt);
    if (j.dev_topk) {
      B.cand = cv.take<uint64_t>(dev, (size_t)P * j.topk * j.RW);
      B.citems = (KItem*)cv.take<uint64_t>(dev, (size_t)P * j.topk * 2);
      B.out = cv.take<uint64_t>(dev, (size_t)16 * j.RW);
    } else {
      B.recs = cv.take<uint64_t>(dev, (size_t)j.rec_cap * j.RW);
    }
    return;
  }
  if (j.plan.local) {
    B.l.keys = cv.take<int64_t>(dev, j.plan.gcap);
    B.l.gkey = cv.take<int64_t>(dev, (size_t)j.plan.limit * j.plan.passes);
    B.l.part = cv.take<uint64_t>(dev, (size_t)j.plan.lgrid * j.plan.limit * (j.nw + 2));
    B.recs = cv.take<uint64_t>(dev, (size_t)j.rec_cap * j.RW);
    if (j.lookup) carve_lookup(cv, dev, j.lplan, j.lbuild->nrows, B.lk);
    return;
  }
  B.b.perr = B.p.perr = B.flags + 5;
  if (j.plan.bloom_lpp >= 0) {
    B.bloom = cv.take<uint32_t>(dev, (size_t)P * 16 << j.plan.bloom_lpp);
    B.p.bloom = B.bloom;
    B.p.bloom_lpp = (uint32_t)j.plan.bloom_lpp;
  }
  const int P1 = j.plan.P1;
  B.b.keys = cv.take<int64_t>(dev, j.cap_b);
  B.b.vstride = j.cap_b;
  B.b.pstart = cv.take<uint32_t>(dev, P1 + 1);
  B.b.counts = cv.take<uint32_t>(dev, (size_t)B.b.G * P1);
  B.b.tot = cv.take<uint32_t>(dev, P1);
  B.p.keys = cv.take<int64_t>(dev, j.cap_p);
  B.p.rows = cv.take<uint32_t>(dev, j.cap_p);
  B.p.vals = cv.take<uint64_t>(dev, (size_t)std::max(j.nvals, 1) * j.cap_p);
  B.p.vstride = j.cap_p;
  B.p.pstart = cv.take<uint32_t>(dev, P1 + 1);
  B.p.counts = cv.take<uint32_t>(dev, (size_t)B.p.G * P1);
  B.p.tot = cv.take<uint32_t>(dev, P1);
  if (P1 < P) {
    B.b.fkeys = cv.take<int64_t>(dev, j.cap_b);
    B.b.fpstart = cv.take<uint32_t>(dev, P + 1);
    B.p.fkeys = cv.take<int64_t>(dev, j.cap_p);
    B.p.frows = cv.take<uint32_t>(dev, j.cap_p);
    B.p.fvals = cv.take<uint64_t>(dev, (size_t)std::max(j.nvals, 1) * j.cap_p);
    B.p.fpstart = cv.take<uint32_t>(dev, P + 1);
  } else {
    B.b.fkeys = B.b.keys;
    B.b.fpstart = B.b.pstart;
    B.p.fkeys = B.p.keys;
    B.p.frows = B.p.rows;
    B.p.fvals = B.p.vals;
    B.p.fpstart = B.p.pstart;
  }
  B.p.bloom_pmask = (uint32_t)(P - 1);
  if (j.livebits) B.p.livebits = cv.take<uint64_t>(dev, (size_t)j.probe->nrows / 64 + 2);
  if (j.dev_topk) {
    B.cand = cv.take<uint64_t>(dev, (size_t)P * j.topk * j.RW);
    B.citems = (KItem*)cv.take<uint64_t>(dev, (size_t)P * j.topk * 2);
    B.out = cv.take<uint64_t>(dev, (size_t)16 * j.RW);
  } else {
    B.recs = cv.take<uint64_t>(dev, (size_t)j.rec_cap * j.RW);
  }
}

// Every launch and copy of one attempt, in stream order (capturable: no
// memset nodes, no host synchronisation).
hipError_t join_enqueue(JoinJob& j, hipStream_t s, hipEvent_t* ev) {
  if (j.lookup) {  // the build side's table first, inside the timed span
    if (ev) hipEventRecord(ev[0], s);
    const hipError_t e = launch_lookup_build(j.lq, j.lkey_slot, j.B.lk, s);
    if (e != hipSuccess) return e;
  }
  hipError_t e = j.plan.direct ? launch_direct(j.dq, j.dkey_slot, j.dl, j.jd, j.plan, j.B, s, ev)
                 : j.plan.local ? launch_local(j.ll, j.ps, j.jd, j.plan, j.B, s, ev, j.lookup)
                              : launch_join(j.bl, j.pl, j.jd, j.plan, j.B, s, ev);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(j.hflags, j.B.flags, 32, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && j.lookup)  // the lookup table's words after the flags
    e = hipMemcpyAsync(j.hflags + 8, j.B.lk.min_head, 16, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && j.dev_topk)
    e = hipMemcpyAsync(j.hout, j.B.out, (size_t)16 * j.RW * 8, hipMemcpyDeviceToHost, s);
  return e;
}

// After the stream (or graph) has finished: the result, or kRetry with the
// next attempt's pool / bits set.
int join_result(JoinJob& j, vq_agg_result* out, vq_stats* stats, hipEvent_t* ev) {
  vq_ctx* c = j.c;
  const uint32_t* hflags = j.hflags;
  if (j.plan.local && (hflags[0] & 4)) {
    // more keys than the per-workgroup path holds: through the partitions,
    // now and for later queries of this shape
    std::lock_guard<std::mutex> lk(c->mu);
    c->build_est["L|" + j.est_key] = 1;
    j.no_local = true;
    return kRetry;
  }
  if (j.plan.local && hflags[0]) return fail(VQ_E_DEVICE, "exec: gpu: GROUP BY flags %#x", hflags[0]);
  if (j.plan.direct) {
    if (hflags[0] & 1) return fail(VQ_E_CORRUPT, "exec: gpu: a build key outside its column's zone maps");
    if (hflags[0] & 2) {
      // a passing build key repeats: through the partitions, which count
      // multiplicities (now and for later joins over this build side)
      std::lock_guard<std::mutex> lk(c->mu);
      c->build_est["D|" + build_est_key(j.build, j.build_key, j.bp)] = 1;
      j.no_direct = true;
      return kRetry;
    }
  }
  if (j.lookup) {
    if (hflags[9] & 2) return fail(VQ_E_CORRUPT, "exec: gpu: a build key outside its column's zone maps");
    if (hflags[9]) return fail(VQ_E_DEVICE, "exec: gpu: join lookup table filled");
    // a key with more build rows than planned: the exact sums' limbs need
    // headroom for that many additions per probe row; plan again
    const int64_t mult = std::max<int64_t>(1, std::max<int64_t>(hflags[10], hflags[11]));
    if (mult > j.lmult) {
      j.lmult = mult;
      if (int rc = join_scales(j)) return rc;
      return kRetry;
    }
  }
  if (hflags[5]) return fail(VQ_E_DEVICE, "exec: gpu: partition scatter wrote past a partition's end");
  if (hflags[0] && !j.plan.direct) {
    // a partition's accumulator pool filled: a larger pool if the LDS has
    // room, else (and when a table filled) one bit more of partitioning
    if (!(hflags[0] & 1) && j.plan.pool < j.pool_max) {
      j.pool_req = std::min(j.pool_max, j.plan.pool * 2);
      return kRetry;
    }
    if (j.bits >= 16) return fail(VQ_E_UNSUPPORTED, "exec: gpu: join partitions overflow at 2^16");
    j.bits++;
    j.pool_req = 0;
    return kRetry;
  }
  {
    std::lock_guard<std::mutex> lk(c->mu);
    c->build_est[j.est_key] = j.groupby ? hflags[2] : hflags[6];  // keys / passing build rows
    c->build_est[j.grp_key] = hflags[1];
  }
  const int RW = j.RW, naggs = j.naggs;
  const bool dev_topk = j.dev_topk;
  const uint32_t ngroups_all = hflags[1];
  const uint32_t nrec = dev_topk ? hflags[3] : ngroups_all;
  if (nrec > (uint64_t)j.rec_cap)
    return fail(VQ_E_DEVICE, "exec: gpu: join produced %u records > %lld", nrec, (long long)j.rec_cap);
  // the records, copied into this thread's staging vector (kept between
  // calls: its pages are already mapped, and it is never zero-filled again)
  thread_local std::vector<uint64_t> hrec;
  const uint64_t* recs = j.hout;
  if (!dev_topk && j.pin) {
    // through pinned staging (the slot's)
    Slot::Pinned& pin = *j.pin;
    const size_t bytes = (size_t)nrec * RW * 8 + 8;
    if (pin.bytes < bytes) {
      if (pin.p) HIP_TRY(hipHostFree(pin.p));
      pin.p = nullptr;
      pin.bytes = 0;
      size_t want = (size_t)1 << 20;
      while (want < bytes) want *= 2;
      HIP_TRY(hipHostMalloc(&pin.p, want, hipHostMallocDefault));
      pin.bytes = want;
    }
    if (nrec) HIP_TRY(hipMemcpy(pin.p, j.B.recs, (size_t)nrec * RW * 8, hipMemcpyDeviceToHost));
    recs = (const uint64_t*)pin.p;
  } else if (!dev_topk) {
    const size_t words = (size_t)nrec * RW + 1;
    if (hrec.size() < words) {
      try {
        hrec.resize(words);
      } catch (const std::bad_alloc&) {
        return fail(VQ_E_NOMEM, "exec: gpu: %u result records", nrec);
      }
    }
    if (nrec) HIP_TRY(hipMemcpy(hrec.data(), j.B.recs, (size_t)nrec * RW * 8, hipMemcpyDeviceToHost));
    recs = hrec.data();
  }
  if (stats) {
    memset(stats, 0, sizeof *stats);
    if (ev) {
      stats->device_ms = elapsed(ev[0], ev[3]);
      stats->main_kernel_ms = elapsed(ev[1], ev[2]);
    }
    stats->rows_scanned = j.brows + j.prows;
    stats->rg_total = (j.groupby ? 0 : j.build->nrg) + j.probe->nrg;
    stats->rg_scanned = j.bkept + j.pkept;
    stats->bytes_alg = (j.groupby ? 0 : j.brows * slot_bytes_per_row(j.build, j.bsm)) +
                       j.prows * slot_bytes_per_row(j.probe, j.psm);
    if (j.lookup) {
      // probe rows read their own columns; each build row its predicate,
      // key and the build-side columns the aggregates read
      int64_t bw = 0;
      for (int i = 0; i < j.psm.n; i++)
        if (j.psm.col_of[i] >= j.lsplit) bw += width_of(j.probe->cols[j.psm.col_of[i]].type);
      stats->bytes_alg = j.prows * (slot_bytes_per_row(j.probe, j.psm) - bw) + j.lrows * (j.lrow_bytes + bw);
      stats->rows_scanned += j.lrows;
      stats->rg_total += j.lbuild->nrg;
      stats->rg_scanned += j.lkept;
      stats->build_rows = j.lrows;
    }
    stats->build_rows = j.groupby ? 0 : hflags[6];
    stats->build_keys = hflags[2];
    stats->partitions = j.plan.local || j.plan.direct ? 0 : j.plan.P;
    stats->radix_bits = j.plan.local || j.plan.direct ? 0 : j.bits;
    stats->grid = j.plan.local ? j.plan.lgrid : j.plan.P;
    stats->tile_rows = j.ps.tile_rows;
  }
  // the records' order, as (first row << 32 | record index): rank order for
  // a device top-k, first-occurrence order otherwise (then the host top-k,
  // for k above the device limit, keeps the stable order).  The buffers are
  // this thread's, kept between calls.
  thread_local std::vector<uint64_t> ord, tmp;
  thread_local std::vector<uint32_t> pos;
  try {
    if (ord.size() < nrec) ord.resize(nrec);
  } catch (const std::bad_alloc&) {
    return fail(VQ_E_NOMEM, "exec: gpu: %u result records", nrec);
  }
  uint64_t* o = ord.data();
  uint64_t fmax = 0;
  for (uint32_t i = 0; i < nrec; i++) {
    const uint64_t f = dev_topk ? 0 : (recs[(size_t)i * RW + 2] & 0xFFFFFFFFull);
    fmax = f > fmax ? f : fmax;
    o[i] = f << 32 | i;
  }
  if (!dev_topk && nrec > 4096) {
    // first rows are distinct 32-bit row ids: a least-significant-digit radix
    // sort on their 11-bit digits, as many as the largest one has
    try {
      if (tmp.size() < nrec) tmp.resize(nrec);
      pos.resize(1 << 11);
    } catch (const std::bad_alloc&) {
      return fail(VQ_E_NOMEM, "exec: gpu: %u result records", nrec);
    }
    uint64_t* a = o;
    uint64_t* b = tmp.data();
    for (int sh = 32; sh < 64 && (fmax >> (sh - 32)) != 0; sh += 11) {
      std::fill(pos.begin(), pos.end(), 0u);
      for (uint32_t i = 0; i < nrec; i++) pos[(a[i] >> sh) & 0x7FF]++;
      uint32_t sum = 0;
      for (auto& x : pos) {
        const uint32_t c = x;
        x = sum;
        sum += c;
      }
      for (uint32_t i = 0; i < nrec; i++) b[pos[(a[i] >> sh) & 0x7FF]++] = a[i];
      std::swap(a, b);
    }
    o = a;
  } else if (!dev_topk) {
    std::sort(o, o + nrec);
  }
  size_t nord = nrec;
  if (j.topk > 0 && !dev_topk) {
    auto key = [&](uint64_t x) {
      return naggs > 0 ? (int64_t)recs[(size_t)(uint32_t)x * RW + 3] : (int64_t)0;
    };
    std::stable_sort(o, o + nrec, [&](uint64_t x, uint64_t y) { return key(x) > key(y); });
    if ((int64_t)nord > j.topk) nord = (size_t)j.topk;
  }
  const int64_t nout = (int64_t)nord;
  if (nout > out->capacity) {
    out->ngroups = nout;  // the size query: groups needed, the arrays hold no result
    return fail(VQ_E_ARG, "exec: gpu: %lld groups exceed result capacity %lld", (long long)nout,
                (long long)out->capacity);
  }
  out->ngroups = nout;
  bool isf[VQ_MAX_AGGS];
  for (int k = 0; k < naggs; k++) {
    const vq_agg& a = j.aggs[k];
    isf[k] = a.kind == VQ_AGG_AVG || (a.kind != VQ_AGG_COUNT && a.accum == VQ_F64);
  }
  auto fill = [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; i++) {
      if (i + 8 < i1) __builtin_prefetch(&recs[(size_t)(uint32_t)o[i + 8] * RW]);
      const uint64_t* r = &recs[(size_t)(uint32_t)o[i] * RW];
      if (j.npack > 1) {  // the codes of the packed key, in GROUP BY order
        const int w = j.npack <= 2 ? 32 : j.npack <= 4 ? 16 : 8;
        const uint64_t mask = (1ULL << w) - 1;
        for (int k = 0; k < j.npack; k++)
          out->keys[i * j.npack + k] = (int64_t)(((uint64_t)r[0] >> j.pack_shift[k]) & mask);
      } else {
        out->keys[i] = (int64_t)r[0];
      }
      out->counts[i] = (int64_t)r[1];
      out->first_row[i] = (int64_t)r[2];
      for (int k = 0; k < naggs; k++) {
        const int64_t q = i * naggs + k;
        out->valid[q] = 1;
        out->ivals[q] = isf[k] ? 0 : (int64_t)r[3 + k];
        out->fvals[q] = isf[k] ? bits_to_f64((int64_t)r[3 + k]) : 0.0;
      }
    }
  };
  // many groups: the rows filled by several threads (most of the time goes
  // to first touches of the caller's fresh result pages)
  const int64_t kRowsPerThread = 32768;
  const int nt = (int)std::min<int64_t>(8, nout / kRowsPerThread);
  if (nt >= 2) {
    std::vector<std::thread> th;
    const int64_t step = (nout + nt - 1) / nt;
    for (int t = 1; t < nt; t++) {
      const int64_t i0 = std::min(nout, t * step), i1 = std::min(nout, (t + 1) * step);
      try {
        th.emplace_back(fill, i0, i1);
      } catch (const std::system_error&) {
        fill(i0, i1);  // (no thread to be had: this one fills the range)
      }
    }
    fill(0, std::min(nout, step));
    for (auto& x : th) x.join();
  } else {
    fill(0, nout);
  }
  return VQ_OK;
}

void join_job_init(JoinJob& j, vq_ctx* c, const vq_table* build, int32_t build_key,
                   const vq_pred* build_pred, const vq_table* probe, int32_t probe_key,
                   const vq_pred* probe_pred, int32_t naggs, const vq_agg* aggs,
                   int32_t radix_bits, int32_t topk, bool groupby) {
  j.c = c;
  j.build = build;
  j.probe = probe;
  j.build_key = build_key;
  j.probe_key = probe_key;
  if (build_pred) {
    j.bpred = *build_pred;
    j.bp = &j.bpred;
  }
  if (probe_pred) {
    j.ppred = *probe_pred;
    j.pp = &j.ppred;
  }
  j.naggs = naggs;
  for (int k = 0; k < naggs && k < VQ_MAX_AGGS; k++) j.aggs[k] = aggs[k];
  j.radix_bits = radix_bits;
  j.topk = topk;
  j.groupby = groupby;
}

// The slot's direct-join region of at least `bytes` (grown on demand), clear:
// a new or dirty region is cleared on the slot's stream first.
int direct_region(vq_ctx* c, Slot* S, uint64_t range, char** out, uint64_t* cap_range) {
  if (S->dreg_range < range) {
    // grown by half again at least, so a stream of growing ranges
    // reallocates rarely
    range = std::max<uint64_t>(range, S->dreg_range + S->dreg_range / 2);
    const size_t bytes = (size_t)direct_region_bytes(range);
    if (S->dreg) {
      HIP_TRY(hipStreamSynchronize(S->stream));
      HIP_TRY(hipFree(S->dreg));
      release(c, S->dreg_bytes);
      S->dreg = nullptr;
      S->dreg_bytes = 0;
      S->dreg_range = 0;
    }
    if (!reserve(c, bytes))
      return fail(VQ_E_BUDGET, "exec: gpu: direct join region of %zu bytes exceeds the device budget", bytes);
    hipError_t e = hipMalloc(&S->dreg, bytes);
    if (e != hipSuccess) {
      release(c, bytes);
      S->dreg = nullptr;
      return fail(VQ_E_DEVICE, "exec: gpu: hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    S->dreg_bytes = bytes;
    S->dreg_range = range;
    S->dreg_dirty = true;
  }
  if (S->dreg_dirty) {
    HIP_TRY(launch_zero_words((uint64_t*)S->dreg, S->dreg_bytes / 8, S->stream));
    S->dreg_dirty = false;
  }
  *out = (char*)S->dreg;
  *cap_range = S->dreg_range;
  return VQ_OK;
}

// One join (or hash GROUP BY) on one slot, start to finish.
int join_run(JoinJob& j, vq_agg_result* out, vq_stats* stats) {
  if (int rc = join_setup(j)) return rc;
  vq_ctx* c = j.c;
  SlotLease S(c);
  hipSetDevice(c->device);
  // Each retry doubles the accumulator pool (up to its LDS bound) or adds a
  // radix bit (up to 16), so a query converges within kMaxAttempts plans; a
  // query that has not is reported instead of planning ever more partitions.
  constexpr int kMaxAttempts = 40;
  for (;;) {
    if (j.attempts >= kMaxAttempts)
      return fail(VQ_E_DEVICE, "exec: gpu: join/GROUP BY plan did not converge after %d attempts (bits %d)",
                  j.attempts, j.bits);
    if (int rc = join_plan(j)) return rc;
    j.attempts++;
    if (j.plan.direct) {
      const int rc = direct_region(c, S.get(), j.drange, &j.dreg, &j.dreg_range);
      if (rc == VQ_E_BUDGET) {  // no room for the region: the partitioned join
        j.no_direct = true;
        continue;
      }
      if (rc) return rc;
    }
    void* ar;
    void* hp;
    if (int rc = arena_get(c, S.get(), j.bytes, &ar)) return rc;
    if (int rc = pinned_get(S.get(), j.host_bytes, &hp)) return rc;
    join_bind(j, (char*)ar, (char*)hp);
    j.pin = &S->rec_pin[0];
    if (canceled()) return fail_canceled();
    if (j.plan.direct) S->dreg_dirty = true;  // until the query has run through
    HIP_TRY(join_enqueue(j, S->stream, S->ev));
    HIP_TRY(hipStreamSynchronize(S->stream));
    if (j.plan.direct) S->dreg_dirty = false;
    if (canceled()) return fail_canceled();
    const int rc = join_result(j, out, stats, S->ev);
    if (rc != kRetry) return rc;
  }
}

int join_impl(vq_ctx* c, const vq_table* build, int32_t build_key, const vq_pred* build_pred,
              const vq_table* probe, int32_t probe_key, const vq_pred* probe_pred, int32_t naggs,
              const vq_agg* aggs, int32_t radix_bits, int32_t topk, vq_agg_result* out,
              vq_stats* stats, bool groupby, const JoinTune* tune = nullptr) {
  if (!out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  out->ngroups = 0;  // no result unless this call succeeds
  if (naggs < 0 || naggs > VQ_MAX_AGGS || (naggs > 0 && !aggs))
    return fail(VQ_E_ARG, "exec: gpu: %d aggregates", naggs);
  std::unique_ptr<JoinJob> j(new (std::nothrow) JoinJob());
  if (!j) return fail(VQ_E_NOMEM, "exec: gpu: join job allocation");
  join_job_init(*j, c, build, build_key, build_pred, probe, probe_key, probe_pred, naggs, aggs,
                radix_bits, topk, groupby);
  if (tune) j->tune = *tune;
  return join_run(*j, out, stats);
}

// A dictionary-code GROUP BY whose code space exceeds the dense slots: the
// int-key path's packing (exec/agg_intkey.go:67-100: 32 bits per column for
// up to 2 columns, 16 for up to 4, 8 for up to 8) makes one 64-bit key per
// row, grouped by the hash GROUP BY's kernels and unpacked into the result.
// Codes wider than their packed width would collide: the reference then
// leaves the int-key path for its string keys, and this returns
// VQ_E_UNSUPPORTED (fall back).
int agg_packed(vq_ctx* c, const vq_table* t, const vq_pred* pred, int32_t nkeys,
               const int32_t* key_cols, int32_t naggs, const vq_agg* aggs, vq_agg_result* out,
               vq_stats* stats) {
  if (nkeys < 1 || nkeys > 8 || !key_cols) return fail(VQ_E_ARG, "exec: gpu: %d group-by columns", nkeys);
  if (naggs < 0 || naggs > VQ_MAX_AGGS || (naggs > 0 && !aggs))
    return fail(VQ_E_ARG, "exec: gpu: %d aggregates", naggs);
  const int w = nkeys <= 2 ? 32 : nkeys <= 4 ? 16 : 8;
  for (int i = 0; i < nkeys; i++) {
    if (int rc = check_col(t, key_cols[i])) return rc;
    const ColDev& cd = t->cols[key_cols[i]];
    if (cd.type == VQ_F64) return fail(VQ_E_UNSUPPORTED, "exec: gpu: float64 group-by column %d", key_cols[i]);
    if ((cd.type == VQ_U8 || cd.type == VQ_U16) && cd.dict_size > 0) {
      if (w < 32 && cd.dict_size > (1 << w))
        return fail(VQ_E_UNSUPPORTED, "exec: gpu: %d codes of column %d exceed the %d-bit packed key", cd.dict_size,
                    key_cols[i], w);
      continue;
    }
    if (nkeys == 1) continue;  // one integer key: the hash GROUP BY's own key
    // an integer column of a composite key (the reference's string-key path,
    // exec/aggregate.go:409-557) packs its values as they are: every value
    // must lie in [0, 2^w) (its zone maps bound them)
    const vq_zone z = merged_zone(t, key_cols[i]);
    if (!z.has_minmax) continue;  // (no value: nothing to pack)
    auto as_i64 = [&](uint64_t b) { return cd.type == VQ_DATE ? (int64_t)(int32_t)(uint32_t)b : (int64_t)b; };
    const int64_t lo = as_i64(z.min_bits), hi = as_i64(z.max_bits);
    if (lo < 0 || (uint64_t)hi >= (1ULL << w))
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: values of column %d in [%lld, %lld] outside the %d-bit packed key",
                  key_cols[i], (long long)lo, (long long)hi, w);
  }
  std::unique_ptr<JoinJob> j(new (std::nothrow) JoinJob());
  if (!j) return fail(VQ_E_NOMEM, "exec: gpu: GROUP BY job allocation");
  join_job_init(*j, c, nullptr, 0, nullptr, t, key_cols[0], pred, naggs, aggs, -1, 0, true);
  j->npack = nkeys;
  for (int i = 0; i < nkeys; i++) {
    j->pack_cols[i] = key_cols[i];
    j->pack_shift[i] = w * i;
  }
  return join_run(*j, out, stats);
}

// The column space of a GROUP BY over joined rows: the probe's columns, then
// the build's (build column c is column probe->ncols + c), over the probe's
// row groups.  A build column's zone map in every row group is its whole
// table's (the bound of its values, for exact sums); no predicate on this
// view touches a build column.
void joined_view(const vq_table* probe, const vq_table* build, vq_table& v) {
  v.ctx = probe->ctx;
  v.nrows = probe->nrows;
  v.ncols = probe->ncols + build->ncols;
  v.cols = probe->cols;
  v.cols.insert(v.cols.end(), build->cols.begin(), build->cols.end());
  v.nrg = probe->nrg;
  v.rg_rows = probe->rg_rows;
  v.rg_start = probe->rg_start;
  std::vector<vq_zone> bz(build->ncols);
  for (int32_t c = 0; c < build->ncols; c++) bz[c] = merged_zone(build, c);
  v.zones.resize((size_t)v.nrg * v.ncols);
  for (int32_t g = 0; g < v.nrg; g++) {
    for (int32_t c = 0; c < probe->ncols; c++)
      v.zones[(size_t)g * v.ncols + c] = probe->zones[(size_t)g * probe->ncols + c];
    for (int32_t c = 0; c < build->ncols; c++) v.zones[(size_t)g * v.ncols + probe->ncols + c] = bz[c];
  }
  v.bytes = 0;
}

// The build side of a join + GROUP BY: its predicate's and key's slots over
// the row groups its predicate keeps, and the lookup table's size (a power
// of two at least twice its rows).
int lookup_setup(JoinJob& j) {
  const vq_table* b = j.lbuild;
  QDesc& d = j.lq;
  memset(&d, 0, sizeof d);
  SlotMap sm;
  if (int rc = add_pred(b, j.lbp, sm, d.slot)) return rc;
  j.lkey_slot = sm.get(j.lbuild_key);
  if (j.lkey_slot < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: build side touches more than %d columns", kMaxSlots);
  if (int rc = refuse_nulls(b, sm)) return rc;
  fill_slot_cols(b, sm, d.slot);
  d.nslots = sm.n;
  d.ngroups = 1;
  std::vector<uint8_t> keep;
  prune(b, j.lbp, keep);
  // one tile per wave of 256 rows: each lane's inserts wait on their
  // compare-and-swaps, so the pass wants many waves in flight
  d.tile_rows = kLookupTileRows;
  if (int rc = make_spans(b, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, &j.lrows, &j.lkept)) return rc;
  if (b->nrows >= (int64_t)INT32_MAX) return fail(VQ_E_UNSUPPORTED, "exec: gpu: build side above 2^31 rows");
  j.lrow_bytes = slot_bytes_per_row(b, sm);
  j.lplan = plan_lookup_col(b, j.lbuild_key, j.lrows);
  return VQ_OK;
}

}  // namespace

extern "C" int vq_join_agg(vq_ctx* c, const vq_table* build, int32_t build_key,
                           const vq_pred* build_pred, const vq_table* probe, int32_t probe_key,
                           const vq_pred* probe_pred, int32_t naggs, const vq_agg* aggs,
                           int32_t radix_bits, int32_t topk, vq_agg_result* out, vq_stats* stats) {
  return join_impl(c, build, build_key, build_pred, probe, probe_key, probe_pred, naggs, aggs,
                   radix_bits, topk, out, stats, false);
}

extern "C" int vq_join_group(vq_ctx* c, const vq_table* build, int32_t build_key,
                             const vq_pred* build_pred, const vq_table* probe, int32_t probe_key,
                             const vq_pred* probe_pred, int32_t nkeys, const int32_t* key_cols,
                             int32_t naggs, const vq_agg* aggs, vq_agg_result* out, vq_stats* stats) {
  if (!c || !build || !probe || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  out->ngroups = 0;  // no result unless this call succeeds
  if (build->ctx != c || probe->ctx != c) return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  if (nkeys < 1 || nkeys > 8 || !key_cols) return fail(VQ_E_ARG, "exec: gpu: %d group-by columns", nkeys);
  if (naggs < 0 || naggs > VQ_MAX_AGGS || (naggs > 0 && !aggs))
    return fail(VQ_E_ARG, "exec: gpu: %d aggregates", naggs);
  if (int rc = check_col(build, build_key)) return rc;
  if (int rc = check_col(probe, probe_key)) return rc;
  if (build->cols[build_key].type == VQ_F64 || probe->cols[probe_key].type == VQ_F64)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: float join key");
  // group keys: probe columns (codes packed as the int-key path packs them)
  const int w = nkeys <= 2 ? 32 : nkeys <= 4 ? 16 : 8;
  for (int i = 0; i < nkeys; i++) {
    if (int rc = check_col(probe, key_cols[i])) return rc;
    const ColDev& cd = probe->cols[key_cols[i]];
    if (cd.type == VQ_F64) return fail(VQ_E_UNSUPPORTED, "exec: gpu: float group-by column");
    if (nkeys > 1 && ((cd.type != VQ_U8 && cd.type != VQ_U16) || cd.dict_size <= 0))
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: group-by column %d is not a dictionary code", key_cols[i]);
    if (nkeys > 1 && w < 32 && cd.dict_size > (1 << w))
      return fail(VQ_E_UNSUPPORTED, "exec: gpu: %d codes of column %d exceed the %d-bit packed key", cd.dict_size,
                  key_cols[i], w);
  }
  // the probe predicate is on probe columns only
  if (probe_pred)
    for (int k = 0; k < probe_pred->nterms && k < VQ_MAX_TERMS; k++) {
      const vq_term& tm = probe_pred->terms[k];
      if (tm.col < 0 || tm.col >= probe->ncols ||
          ((tm.op & VQ_COLCMP) && (tm.ival < 0 || tm.ival >= probe->ncols)))
        return fail(VQ_E_ARG, "exec: gpu: probe predicate column out of range");
    }
  std::unique_ptr<JoinJob> j(new (std::nothrow) JoinJob());
  if (!j) return fail(VQ_E_NOMEM, "exec: gpu: join job allocation");
  joined_view(probe, build, j->view);
  // aggregates' column references in the view's column space
  vq_agg va[VQ_MAX_AGGS];
  auto map_col = [&](int64_t col, int64_t* outc) -> int {
    if (col & VQ_BUILD_COL) {
      const int64_t b = col & ~(int64_t)VQ_BUILD_COL;
      if (b < 0 || b >= build->ncols) return fail(VQ_E_ARG, "exec: gpu: build column %lld out of range", (long long)b);
      *outc = probe->ncols + b;
    } else {
      if (col < 0 || col >= probe->ncols) return fail(VQ_E_ARG, "exec: gpu: column %lld out of range", (long long)col);
      *outc = col;
    }
    return VQ_OK;
  };
  for (int k = 0; k < naggs; k++) {
    va[k] = aggs[k];
    vq_expr& e = va[k].arg;
    for (int f = 0; f < e.nfactors && f < VQ_MAX_FACTORS; f++) {
      if (e.f[f].kind == VQ_F_LIT) continue;
      int64_t m;
      if (int rc = map_col(e.f[f].col, &m)) return rc;
      e.f[f].col = (int32_t)m;
    }
    for (int t = 0; t < e.nwhen && t < VQ_MAX_WHEN; t++) {
      int64_t m;
      if (int rc = map_col(e.when[t].col, &m)) return rc;
      e.when[t].col = (int32_t)m;
      if (e.when[t].op & VQ_COLCMP) {
        if (int rc = map_col(e.when[t].ival, &m)) return rc;
        e.when[t].ival = m;
      }
    }
  }
  join_job_init(*j, c, nullptr, 0, nullptr, &j->view, key_cols[0], probe_pred, naggs, va, -1, 0, true);
  j->lookup = true;
  j->lprobe = probe;
  j->lbuild = build;
  j->lbuild_key = build_key;
  j->ljkey = probe_key;
  j->lsplit = probe->ncols;
  if (build_pred) {
    j->lbpred = *build_pred;
    j->lbp = &j->lbpred;
  }
  if (nkeys > 1) {
    j->npack = nkeys;
    for (int i = 0; i < nkeys; i++) {
      j->pack_cols[i] = key_cols[i];
      j->pack_shift[i] = w * i;
    }
  }
  if (int rc = lookup_setup(*j)) return rc;
  return join_run(*j, out, stats);
}

namespace {

// One side of the nested build as a scan: its predicate's slots and the
// given key columns' (slots returned in key_slots), over the row groups its
// predicate keeps.
int scan_desc(const vq_table* t, const vq_pred* pred, int nkeys, const int32_t* keys, QDesc& d, int* key_slots,
              int64_t* rows, int64_t* kept, int32_t tile_rows) {
  memset(&d, 0, sizeof d);
  SlotMap sm;
  if (int rc = add_pred(t, pred, sm, d.slot)) return rc;
  for (int i = 0; i < nkeys; i++) {
    if (int rc = check_col(t, keys[i])) return rc;
    key_slots[i] = sm.get(keys[i]);
    if (key_slots[i] < 0) return fail(VQ_E_UNSUPPORTED, "exec: gpu: query touches more than %d columns", kMaxSlots);
  }
  if (int rc = refuse_nulls(t, sm)) return rc;
  fill_slot_cols(t, sm, d.slot);
  d.nslots = sm.n;
  d.ngroups = 1;
  std::vector<uint8_t> keep;
  prune(t, pred, keep);
  d.tile_rows = tile_rows;  // (lookups and inserts: latency-bound, many waves)
  return make_spans(t, keep, d.tile_rows, d.span, &d.nspans, &d.ntiles, rows, kept);
}

}  // namespace

extern "C" int vq_join_agg_nested(vq_ctx* c, const vq_table* inner, int32_t inner_key, const vq_pred* inner_pred,
                                  const vq_table* build, int32_t build_fk, int32_t build_key,
                                  const vq_pred* build_pred, const vq_table* probe, int32_t probe_key,
                                  const vq_pred* probe_pred, int32_t npay, const int32_t* pay_cols, int32_t naggs,
                                  const vq_agg* aggs, int32_t radix_bits, int32_t topk, vq_agg_result* out,
                                  int64_t* pay_out, vq_stats* stats) {
  if (!c || !inner || !build || !probe || !out) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  out->ngroups = 0;  // no result unless this call succeeds
  if (inner->ctx != c || build->ctx != c || probe->ctx != c)
    return fail(VQ_E_ARG, "exec: gpu: table belongs to another context");
  if (npay < 0 || npay > 7 || (npay > 0 && (!pay_cols || !pay_out)))
    return fail(VQ_E_ARG, "exec: gpu: %d columns with the groups", npay);
  if (int rc = check_col(inner, inner_key)) return rc;
  if (int rc = check_col(build, build_fk)) return rc;
  if (int rc = check_col(build, build_key)) return rc;
  if (inner->cols[inner_key].type == VQ_F64 || build->cols[build_fk].type == VQ_F64 ||
      build->cols[build_key].type == VQ_F64)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: float join key");
  // the inner side (into a lookup table) and the build side (scanned against it)
  QDesc iq, bq;
  int ikey = 0;
  int64_t irows = 0, ikept = 0, brows = 0, bkept = 0;
  if (int rc = scan_desc(inner, inner_pred, 1, &inner_key, iq, &ikey, &irows, &ikept)) return rc;
  int32_t bkeys[9] = {build_fk, build_key};
  for (int i = 0; i < npay; i++) bkeys[2 + i] = pay_cols[i];
  int bslots[9] = {0};
  if (int rc = scan_desc(build, build_pred, 2 + npay, bkeys, bq, bslots, &brows, &bkept)) return rc;
  if (build->nrows >= (int64_t)INT32_MAX || inner->nrows >= (int64_t)INT32_MAX)
    return fail(VQ_E_UNSUPPORTED, "exec: gpu: join side above 2^31 rows");
  const LookupPlan ilp = plan_lookup_col(inner, inner_key, irows);
  float dev_ms = 0;
  // phase 1: the nested join's output rows (key and the columns that ride
  // with the groups) as a table of their own
  vq_table* nested = nullptr;
  int64_t total = 0;
  {
    SlotLease S(c);
    hipSetDevice(c->device);
    const size_t nt = (size_t)std::max<int64_t>(bq.ntiles, 1);
    void* ar;
    if (int rc = arena_get(c, S.get(), lookup_bytes(ilp, inner->nrows) + nt * 8 + 64 + 2048, &ar)) return rc;
    Carve cv;
    LookupBufs lk;
    carve_lookup(cv, (char*)ar, ilp, inner->nrows, lk);
    uint32_t* tcount = cv.take<uint32_t>((char*)ar, nt);
    uint32_t* toff = cv.take<uint32_t>((char*)ar, nt);
    uint64_t* dcount = cv.take<uint64_t>((char*)ar, 1);
    void* hp;
    if (int rc = pinned_get(S.get(), 64, &hp)) return rc;
    uint64_t* hcount = (uint64_t*)hp;
    uint32_t* hflag = (uint32_t*)((char*)hp + 16);
    hipStream_t s = S->stream;
    HIP_TRY(hipEventRecord(S->ev[0], s));
    HIP_TRY(launch_lookup_build(iq, ikey, lk, s));
    HIP_TRY(launch_semi(bq, bslots[0], lk, tcount, toff, dcount, s));
    HIP_TRY(hipMemcpyAsync(hcount, dcount, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hflag, lk.min_head + 1, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(S->ev[1], s));
    HIP_TRY(hipStreamSynchronize(s));
    dev_ms += elapsed(S->ev[0], S->ev[1]);
    if (*hflag & 2) return fail(VQ_E_CORRUPT, "exec: gpu: an inner key outside its column's zone maps");
    if (*hflag) return fail(VQ_E_DEVICE, "exec: gpu: inner lookup table filled");
    total = (int64_t)*hcount;
    if (total >= (int64_t)INT32_MAX) return fail(VQ_E_UNSUPPORTED, "exec: gpu: nested join above 2^31 rows");
    TableParts tp;
    tp.nrows = total;
    const int ncol = 1 + npay;
    SemiOut o;
    memset(&o, 0, sizeof o);
    o.ncols = ncol;
    uint64_t col_bytes = 0;
    for (int i = 0; i < ncol; i++) {
      const ColDev& cd = build->cols[bkeys[1 + i]];
      tp.types.push_back(cd.type);
      tp.dict_sizes.push_back(cd.dict_size);
      col_bytes += (((uint64_t)total * width_of(cd.type) + 255) & ~255ULL) + 256;
      o.slot[i] = bslots[1 + i];
      o.type[i] = cd.type;
    }
    if (!ctx_reserve(c, col_bytes))
      return fail(VQ_E_BUDGET, "exec: gpu: the nested join's %lld rows exceed the device budget", (long long)total);
    hipError_t e = hipSuccess;
    tp.dev.assign(ncol, nullptr);
    for (int i = 0; i < ncol && e == hipSuccess; i++) {
      e = hipMalloc(&tp.dev[i], (((uint64_t)total * width_of(tp.types[i]) + 255) & ~255ULL) + 256);
      o.dst[i] = tp.dev[i];
    }
    if (e == hipSuccess) e = hipEventRecord(S->ev[0], s);
    if (e == hipSuccess) e = launch_semi_gather(bq, bslots[0], lk, toff, o, s);
    if (e == hipSuccess) e = hipEventRecord(S->ev[1], s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      for (void* p : tp.dev)
        if (p) hipFree(p);
      ctx_release(c, col_bytes);
      return fail(VQ_E_DEVICE, "exec: gpu: nested build: %s", hipGetErrorString(e));
    }
    dev_ms += elapsed(S->ev[0], S->ev[1]);
    if (total > 0) tp.rg_rows.push_back(total);
    tp.zones.assign(tp.rg_rows.size() * ncol, vq_zone{});  // (no bounds: nothing prunes it)
    tp.bytes = col_bytes;
    std::vector<void*> owned = tp.dev;
    if (int rc = table_adopt(c, std::move(tp), &nested)) {
      for (void* p : owned)
        if (p) hipFree(p);
      ctx_release(c, col_bytes);
      return rc;
    }
  }
  // phase 2: the outer join, grouped by its key, top-k
  int rc = join_impl(c, nested, 0, nullptr, probe, probe_key, probe_pred, naggs, aggs, radix_bits, topk, out,
                     stats, false);
  // phase 3: the columns that ride with each output group
  if (rc == VQ_OK && npay > 0 && out->ngroups > 0) {
    const int64_t N = out->ngroups;
    int64_t klo = out->keys[0], khi = out->keys[0];
    for (int64_t i = 1; i < N; i++) {
      klo = std::min(klo, out->keys[i]);
      khi = std::max(khi, out->keys[i]);
    }
    const LookupPlan klp = plan_lookup(N, true, klo, khi);
    SlotLease S(c);
    hipSetDevice(c->device);
    void* ar;
    void* hp;
    if ((rc = arena_get(c, S.get(), lookup_bytes(klp, N) + (size_t)N * 8 * (1 + npay) + 3 * 256, &ar)) == VQ_OK &&
        (rc = pinned_get(S.get(), 64, &hp)) == VQ_OK) {
      Carve cv;
      LookupBufs lk;
      carve_lookup(cv, (char*)ar, klp, N, lk);
      int64_t* dkeys = cv.take<int64_t>((char*)ar, N);
      int64_t* dpay = cv.take<int64_t>((char*)ar, (size_t)N * npay);
      uint32_t* dflag = cv.take<uint32_t>((char*)ar, 1);
      PayCols pc;
      memset(&pc, 0, sizeof pc);
      pc.npay = npay;
      pc.key_type = nested->cols[0].type;
      pc.key = nested->cols[0].d;
      for (int i = 0; i < npay; i++) {
        pc.src[i] = nested->cols[1 + i].d;
        pc.type[i] = nested->cols[1 + i].type;
      }
      hipStream_t s = S->stream;
      hipError_t e = hipMemcpyAsync(dkeys, out->keys, (size_t)N * 8, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipEventRecord(S->ev[0], s);
      if (e == hipSuccess) e = launch_pay_gather(dkeys, N, lk, pc, total, dpay, dflag, s);
      if (e == hipSuccess) e = hipEventRecord(S->ev[1], s);
      if (e == hipSuccess) e = hipMemcpyAsync(pay_out, dpay, (size_t)N * npay * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipMemcpyAsync(hp, dflag, 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = fail(VQ_E_DEVICE, "exec: gpu: nested join columns: %s", hipGetErrorString(e));
      } else {
        dev_ms += elapsed(S->ev[0], S->ev[1]);
        if (*(uint32_t*)hp) rc = fail(VQ_E_UNSUPPORTED, "exec: gpu: rows of one join key disagree on a column grouped with it");
      }
    }
    if (rc != VQ_OK) out->ngroups = 0;
  }
  vq_table_destroy(nested);
  if (rc == VQ_OK && stats) {
    stats->device_ms += dev_ms;
    stats->rows_scanned += irows + brows;
    stats->rg_total += inner->nrg + build->nrg;
    stats->rg_scanned += ikept + bkept;
    int64_t iw = 0, bw = 0;
    for (int i = 0; i < iq.nslots; i++) iw += width_of(iq.slot[i].type);
    for (int i = 0; i < bq.nslots; i++) bw += width_of(bq.slot[i].type);
    stats->bytes_alg += irows * iw + brows * bw;
    stats->build_rows = total;
  }
  return rc;
}

extern "C" int vq_agg_hash(vq_ctx* c, const vq_table* t, const vq_pred* pred, int32_t key_col,
                           int32_t naggs, const vq_agg* aggs, vq_agg_result* out, vq_stats* stats) {
  return join_impl(c, nullptr, 0, nullptr, t, key_col, pred, naggs, aggs, -1, 0, out, stats, true);
}

// ===========================================================================
// concurrent and asynchronous queries
// ===========================================================================
struct Ticket {
  vq_ctx* c = nullptr;
  vq_query q;
  vq_pred pred, build_pred;
  bool has_pred = false, has_build_pred = false;
  std::vector<int32_t> key_cols;
  std::vector<vq_agg> aggs;
  vq_agg_result* out = nullptr;
  vq_stats* stats = nullptr;
  uint64_t est = 0;  // device bytes the query is estimated to need
  std::mutex mu;
  std::condition_variable cv;
  bool done = false;
  int status = VQ_OK;
  std::string err;
  std::atomic<bool> cancel{false};  // asked to stop (vq_ticket_cancel)
};

struct vq_ticket {
  Ticket* t;
};

namespace {

// Device memory of one query, from its tables' sizes: the join's partitioned
// copies of both sides (keys, probe row ids and values) and its per-partition
// bookkeeping; the aggregate's per-workgroup partials.
uint64_t estimate_bytes(const Ticket& k) {
  const vq_query& q = k.q;
  if (q.kind == VQ_Q_JOIN_AGG) {
    const uint64_t b = (uint64_t)std::max<int64_t>(q.build->nrows, 1);
    const uint64_t p = (uint64_t)std::max<int64_t>(q.table->nrows, 1);
    const uint64_t nv = (uint64_t)std::max(q.naggs, 1);
    return b * 8 + p * (12 + 8 * nv) + (16ull << 20);
  }
  if (q.kind == VQ_Q_AGG_HASH) {  // the partitioned rows, and the groups' records
    const uint64_t p = (uint64_t)std::max<int64_t>(q.table->nrows, 1);
    const uint64_t nv = (uint64_t)std::max(q.naggs, 1);
    return p * (12 + 8 * nv) + p * 8 * (3 + nv) + (16ull << 20);
  }
  return 8ull << 20;
}

void run_ticket(Ticket* k) {
  const vq_query& q = k->q;
  int rc;
  g_cancel = &k->cancel;
  if (k->cancel.load()) rc = fail_canceled();
  else if (q.kind == VQ_Q_AGG_DICT)
    rc = vq_agg_dict(k->c, q.table, k->has_pred ? &k->pred : nullptr, q.nkeys,
                     k->key_cols.empty() ? nullptr : k->key_cols.data(), q.naggs,
                     k->aggs.empty() ? nullptr : k->aggs.data(), k->out, k->stats);
  else if (q.kind == VQ_Q_AGG_HASH)
    rc = vq_agg_hash(k->c, q.table, k->has_pred ? &k->pred : nullptr, k->key_cols[0], q.naggs,
                     k->aggs.empty() ? nullptr : k->aggs.data(), k->out, k->stats);
  else
    rc = vq_join_agg(k->c, q.build, q.build_key, k->has_build_pred ? &k->build_pred : nullptr,
                     q.table, q.probe_key, k->has_pred ? &k->pred : nullptr, q.naggs,
                     k->aggs.empty() ? nullptr : k->aggs.data(), q.radix_bits, q.topk, k->out,
                     k->stats);
  g_cancel = nullptr;
  std::lock_guard<std::mutex> lk(k->mu);
  k->status = rc;
  if (rc) k->err = g_err;
  k->done = true;
  k->cv.notify_all();
}

// A global aggregate (no GROUP BY) another can share a pass with: the same
// table and aggregates, a predicate on the same columns with the same
// comparisons (vq_agg_dict_multi checks the planned shapes again).
bool shareable(const Ticket& a, const Ticket& b) {
  const vq_query &x = a.q, &y = b.q;
  if (x.kind != VQ_Q_AGG_DICT || y.kind != VQ_Q_AGG_DICT || x.nkeys != 0 || y.nkeys != 0) return false;
  if (x.table != y.table || a.aggs.size() != b.aggs.size() || a.has_pred != b.has_pred) return false;
  if (!a.aggs.empty() && memcmp(a.aggs.data(), b.aggs.data(), a.aggs.size() * sizeof(vq_agg)) != 0) return false;
  if (!a.has_pred) return true;
  if (a.pred.nterms != b.pred.nterms) return false;
  for (int i = 0; i < a.pred.nterms; i++)
    if (a.pred.terms[i].col != b.pred.terms[i].col || a.pred.terms[i].op != b.pred.terms[i].op) return false;
  return true;
}

bool share_disabled() {
  const char* e = getenv("VQ_SERVE_NO_SHARE");
  return e && *e && *e != '0';
}

// Queued global aggregates of one shape in one pass (vq_agg_dict_multi).
void run_shared(std::vector<Ticket*>& ks) {
  const size_t n = ks.size();
  const Ticket& k0 = *ks[0];
  std::vector<vq_pred> preds(n);
  std::vector<vq_agg_result> outs(n);
  std::vector<vq_stats> st(n);
  for (size_t i = 0; i < n; i++) {
    memset(&preds[i], 0, sizeof preds[i]);
    if (ks[i]->has_pred) preds[i] = ks[i]->pred;
    outs[i] = *ks[i]->out;
  }
  // (a canceled query's place in the pass is kept: the others' results do
  // not depend on it; it finishes canceled)
  bool all = true;
  for (size_t i = 0; i < n; i++) all = all && ks[i]->cancel.load();
  const int rc = all ? fail_canceled()
                     : vq_agg_dict_multi(k0.c, k0.q.table, (int32_t)n, preds.data(), k0.q.naggs,
                                         k0.aggs.empty() ? nullptr : k0.aggs.data(), outs.data(), st.data());
  const std::string err = rc ? g_err : std::string();
  for (size_t i = 0; i < n; i++) {
    Ticket* k = ks[i];
    const bool cx = k->cancel.load();
    k->out->ngroups = cx ? 0 : outs[i].ngroups;
    if (k->stats) *k->stats = st[i];
    std::lock_guard<std::mutex> lk(k->mu);
    k->status = cx ? VQ_E_CANCELED : rc;
    if (cx) k->err = "exec: gpu: query canceled";
    else if (rc) k->err = err;
    k->done = true;
    k->cv.notify_all();
  }
}

// Continuous batching: a worker takes the oldest queued query that the
// memory budget admits, and with a global aggregate also every queued one
// that can share its pass (up to kMultiMax), wherever they wait in the queue.
void worker_loop(vq_ctx* c) {
  hipSetDevice(c->device);
  const bool share = !share_disabled();
  std::vector<Ticket*> ks;
  for (;;) {
    ks.clear();
    uint64_t est = 0;
    {
      std::unique_lock<std::mutex> lk(c->qmu);
      c->qcv.wait(lk, [&] {
        if (c->stopping) return true;
        if (c->queue.empty()) return false;
        const Ticket* h = c->queue.front();
        return c->running == 0 || c->serve_reserved + h->est <= c->serve_budget;
      });
      if (c->stopping && c->queue.empty()) return;
      if (c->queue.empty()) continue;
      Ticket* k = c->queue.front();
      c->queue.pop_front();
      ks.push_back(k);
      est = k->est;
      if (share && k->q.kind == VQ_Q_AGG_DICT && k->q.nkeys == 0)
        for (auto it = c->queue.begin(); it != c->queue.end() && (int32_t)ks.size() < kMultiMax;) {
          if (shareable(*k, **it)) {
            ks.push_back(*it);
            it = c->queue.erase(it);
          } else {
            ++it;
          }
        }
      c->serve_reserved += est;
      c->running++;
      c->active.insert(c->active.end(), ks.begin(), ks.end());
    }
    if (ks.size() > 1) run_shared(ks);
    else run_ticket(ks[0]);
    {
      std::lock_guard<std::mutex> lk(c->qmu);
      c->serve_reserved -= est;
      c->running--;
      for (Ticket* k : ks) c->active.erase(std::find(c->active.begin(), c->active.end(), k));
    }
    c->qcv.notify_all();
  }
}

void stop_workers(vq_ctx* c) {
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    c->stopping = true;
  }
  c->qcv.notify_all();
  for (auto& w : c->workers) w.join();
  c->workers.clear();
  c->stopping = false;
}

}  // namespace

extern "C" int vq_ctx_set_slots(vq_ctx* c, int32_t nslots, uint64_t serve_budget) {
  if (!c || nslots < 1 || nslots > 16) return fail(VQ_E_ARG, "exec: gpu: %d query slots", nslots);
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    if (c->running || !c->queue.empty()) return fail(VQ_E_ARG, "exec: gpu: queries in flight");
  }
  stop_workers(c);
  hipSetDevice(c->device);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->idle.size() != c->slots.size()) return fail(VQ_E_ARG, "exec: gpu: queries in flight");
  while ((int32_t)c->slots.size() < nslots) {
    Slot* S = nullptr;
    if (int rc = slot_create(&S)) return rc;
    c->slots.push_back(S);
    c->idle.push_back(S);
  }
  while ((int32_t)c->slots.size() > nslots) {
    Slot* S = c->slots.back();
    c->slots.pop_back();
    c->idle.erase(std::find(c->idle.begin(), c->idle.end(), S));
    slot_destroy(c, S);
  }
  c->serve_budget = serve_budget ? serve_budget : c->budget / 2;
  return VQ_OK;
}

extern "C" int vq_submit(vq_ctx* c, const vq_query* q, vq_agg_result* out, vq_stats* stats,
                         vq_ticket** ticket) {
  if (!c || !q || !out || !ticket) return fail(VQ_E_ARG, "exec: gpu: NULL argument");
  *ticket = nullptr;
  if (q->kind != VQ_Q_AGG_DICT && q->kind != VQ_Q_JOIN_AGG && q->kind != VQ_Q_AGG_HASH)
    return fail(VQ_E_ARG, "exec: gpu: query kind %d", q->kind);
  if (!q->table || (q->kind == VQ_Q_JOIN_AGG && !q->build))
    return fail(VQ_E_ARG, "exec: gpu: query without its table");
  if (q->naggs < 0 || q->naggs > VQ_MAX_AGGS || (q->naggs > 0 && !q->aggs))
    return fail(VQ_E_ARG, "exec: gpu: %d aggregates", q->naggs);
  if (q->kind == VQ_Q_AGG_DICT && (q->nkeys < 0 || q->nkeys > 4 || (q->nkeys > 0 && !q->key_cols)))
    return fail(VQ_E_ARG, "exec: gpu: %d group keys", q->nkeys);
  if (q->kind == VQ_Q_AGG_HASH && (q->nkeys != 1 || !q->key_cols))
    return fail(VQ_E_ARG, "exec: gpu: a hash GROUP BY takes one key, not %d", q->nkeys);
  Ticket* k = new (std::nothrow) Ticket();
  vq_ticket* h = new (std::nothrow) vq_ticket();
  if (!k || !h) {
    delete k;
    delete h;
    return fail(VQ_E_NOMEM, "exec: gpu: ticket allocation");
  }
  k->c = c;
  k->q = *q;
  if (q->pred) {
    k->pred = *q->pred;
    k->has_pred = true;
  }
  if (q->build_pred) {
    k->build_pred = *q->build_pred;
    k->has_build_pred = true;
  }
  if (q->kind != VQ_Q_JOIN_AGG && q->nkeys > 0) k->key_cols.assign(q->key_cols, q->key_cols + q->nkeys);
  if (q->naggs > 0) k->aggs.assign(q->aggs, q->aggs + q->naggs);
  k->out = out;
  k->stats = stats;
  k->est = estimate_bytes(*k);
  h->t = k;
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    if (c->serve_budget == 0) c->serve_budget = c->budget / 2;
    if (c->workers.empty()) {
      const size_t n = c->slots.size();
      for (size_t i = 0; i < n; i++) c->workers.emplace_back(worker_loop, c);
    }
    c->queue.push_back(k);
  }
  c->qcv.notify_all();
  *ticket = h;
  return VQ_OK;
}

extern "C" int vq_ticket_done(const vq_ticket* h) {
  if (!h || !h->t) return 1;
  std::lock_guard<std::mutex> lk(h->t->mu);
  return h->t->done ? 1 : 0;
}

extern "C" int vq_wait(vq_ticket* h) {
  if (!h || !h->t) return fail(VQ_E_ARG, "exec: gpu: NULL ticket");
  Ticket* k = h->t;
  int rc;
  {
    std::unique_lock<std::mutex> lk(k->mu);
    k->cv.wait(lk, [&] { return k->done; });
    rc = k->status;
    if (rc) g_err = k->err;
  }
  delete k;
  delete h;
  return rc;
}

namespace {

// A queued ticket dropped: finished as canceled.  (Under c->qmu.)
void cancel_queued(Ticket* k) {
  std::lock_guard<std::mutex> lk(k->mu);
  k->out->ngroups = 0;
  k->status = VQ_E_CANCELED;
  k->err = "exec: gpu: query canceled";
  k->done = true;
  k->cv.notify_all();
}

}  // namespace

extern "C" int vq_ticket_cancel(vq_ticket* h) {
  if (!h || !h->t) return fail(VQ_E_ARG, "exec: gpu: NULL ticket");
  Ticket* k = h->t;
  vq_ctx* c = k->c;
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    k->cancel.store(true);
    auto it = std::find(c->queue.begin(), c->queue.end(), k);
    if (it != c->queue.end()) {
      c->queue.erase(it);
      cancel_queued(k);
    }
  }
  c->qcv.notify_all();
  return VQ_OK;
}

extern "C" int vq_ctx_cancel(vq_ctx* c) {
  if (!c) return fail(VQ_E_ARG, "exec: gpu: NULL context");
  {
    std::lock_guard<std::mutex> lk(c->qmu);
    for (Ticket* k : c->queue) {
      k->cancel.store(true);
      cancel_queued(k);
    }
    c->queue.clear();
    for (Ticket* k : c->active) k->cancel.store(true);
  }
  c->qcv.notify_all();
  return VQ_OK;
}

// ===========================================================================
// batches of independent queries
// ===========================================================================
namespace {

// FNV-1a over bytes, with two seeds: a 128-bit fingerprint of everything a
// batch's launches take as arguments.
struct Fp {
  uint64_t h1 = 0xcbf29ce484222325ULL, h2 = 0x84222325cbf29ce4ULL;
  void add(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) {
      h1 = (h1 ^ b[i]) * 0x100000001b3ULL;
      h2 = (h2 ^ b[i]) * 0x100000001b3ULL + 0x9e3779b97f4a7c15ULL;
    }
  }
};

int slot_branches(Slot* S) {
  if (S->fork) return VQ_OK;
  hipError_t e = hipEventCreateWithFlags(&S->fork, hipEventDisableTiming);
  for (int k = 0; k < Slot::kBranches && e == hipSuccess; k++) {
    e = hipStreamCreateWithFlags(&S->side[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->join[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) return fail(VQ_E_DEVICE, "exec: gpu: batch streams: %s", hipGetErrorString(e));
  return VQ_OK;
}

// The class an instance is tuned for: its kind, key type, and the binary
// order of magnitude of each side's rows and of its estimated keys.
std::string tune_class(const JoinJob& j) {
  auto lg = [](int64_t x) { return ilog2_ceil((uint64_t)std::max<int64_t>(x, 1)); };
  char buf[128];
  snprintf(buf, sizeof buf, "%s/k%d/p%d/b%d/e%d/a%d", j.groupby ? "hash" : "join",
           j.probe->cols[j.probe_key].type, lg(j.probe->nrows), j.groupby ? 0 : lg(j.build->nrows),
           lg(j.est), j.naggs);
  return buf;
}

bool tune_lookup(vq_ctx* c, const std::string& cls, JoinTune* t) {
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->tune.find(cls);
  if (it == c->tune.end()) return false;
  return sscanf(it->second.c_str(), "%d %d %d %d", &t->bits_delta, &t->pbuf, &t->tile_shift,
                &t->reps) == 4;
}

void tune_store(vq_ctx* c, const std::string& cls, const JoinTune& t) {
  char buf[64];
  snprintf(buf, sizeof buf, "%d %d %d %d", t.bits_delta, t.pbuf, t.tile_shift, t.reps);
  std::lock_guard<std::mutex> lk(c->mu);
  c->tune[cls] = buf;
  if (const char* f = getenv("VQ_TUNE_FILE")) {
    if (FILE* fp = fopen(f, "a")) {
      fprintf(fp, "%s %s\n", cls.c_str(), buf);
      fclose(fp);
    }
  }
}

void tune_load(vq_ctx* c) {
  const char* f = getenv("VQ_TUNE_FILE");
  FILE* fp = f ? fopen(f, "r") : nullptr;
  if (!fp) return;
  char cls[128];
  int a, b, d, r;
  std::lock_guard<std::mutex> lk(c->mu);
  while (fscanf(fp, "%127s %d %d %d %d", cls, &a, &b, &d, &r) == 5) {
    char buf[64];
    snprintf(buf, sizeof buf, "%d %d %d %d", a, b, d, r);
    c->tune[cls] = buf;
  }
  fclose(fp);
}

// The autotuner: run the instance under each candidate (radix bits one
// below / at / above the planner's choice, the largest or an 8-entry scatter
// buffer, default or 4x tiles, the planner's or a single accumulator
// replica), keep the fastest device time.  Every
// candidate computes the same result (the accumulation is exact and
// order-free), so tuning never changes an answer.
JoinTune autotune(const JoinJob& proto, vq_agg_result* out) {
  JoinTune best;
  double best_ms = 1e30;
  for (int bd = -1; bd <= 1; bd++)
    for (int pb : {-1, 8})
      for (int ts : {0, 2})
        for (int rp : {0, 1}) {
          JoinTune t;
          t.bits_delta = bd;
          t.pbuf = pb;
          t.tile_shift = ts;
          t.reps = rp;
          double ms = 1e30;
          for (int rep = 0; rep < 2; rep++) {
            // (a candidate whose first plan overflows is not taken: its
            // device time would not count the attempts before the last)
            vq_stats st;
            std::unique_ptr<JoinJob> j(new (std::nothrow) JoinJob());
            if (!j) break;
            join_job_init(*j, proto.c, proto.build, proto.build_key, proto.bp, proto.probe,
                          proto.probe_key, proto.pp, proto.naggs, proto.aggs, proto.radix_bits,
                          proto.topk, proto.groupby);
            j->tune = t;
            j->no_direct = true;
            const int rc = join_run(*j, out, &st);
            if (rc || j->attempts > 1) {
              ms = 1e30;
              break;
            }
            ms = std::min(ms, st.device_ms);
          }
          if (ms < best_ms) {
            best_ms = ms;
            best = t;
          }
        }
  return best;
}

}  // namespace

extern "C" int vq_batch_run(vq_ctx* c, int32_t n, const vq_query* qs, vq_agg_result* outs,
                            vq_stats* stats, int32_t flags) {
  if (!c || n < 0 || (n > 0 && (!qs || !outs))) return fail(VQ_E_ARG, "exec: gpu: bad batch arguments");
  for (int32_t i = 0; i < n; i++) outs[i].ngroups = 0;  // no result unless its instance succeeds
  hipSetDevice(c->device);
  std::vector<std::unique_ptr<JoinJob>> jobs(n);
  std::vector<char> direct(n, 0);  // dictionary GROUP BYs run one by one
  for (int32_t i = 0; i < n; i++) {
    const vq_query& q = qs[i];
    if (q.kind == VQ_Q_AGG_DICT) {
      direct[i] = 1;
      continue;
    }
    if (q.kind != VQ_Q_JOIN_AGG && q.kind != VQ_Q_AGG_HASH)
      return fail(VQ_E_ARG, "exec: gpu: batch query %d: kind %d", i, q.kind);
    if (q.naggs < 0 || q.naggs > VQ_MAX_AGGS || (q.naggs > 0 && !q.aggs))
      return fail(VQ_E_ARG, "exec: gpu: batch query %d: %d aggregates", i, q.naggs);
    const bool gb = q.kind == VQ_Q_AGG_HASH;
    if (gb && (q.nkeys != 1 || !q.key_cols)) return fail(VQ_E_ARG, "exec: gpu: batch query %d: one key", i);
    jobs[i].reset(new (std::nothrow) JoinJob());
    if (!jobs[i]) return fail(VQ_E_NOMEM, "exec: gpu: batch job allocation");
    join_job_init(*jobs[i], c, gb ? nullptr : q.build, q.build_key, q.build_pred, q.table,
                  gb ? q.key_cols[0] : q.probe_key, q.pred, q.naggs, q.aggs, gb ? -1 : q.radix_bits,
                  gb ? 0 : q.topk, gb);
  }
  // each instance's plan (its class's tuning first; with VQ_BATCH_TUNE a
  // class not tuned yet is tuned here, by running it)
  auto plan_one = [&](int32_t i) -> int {
    JoinJob& j = *jobs[i];
    j.no_direct = true;  // (the instances share the slot: no per-slot region)
    if (int rc = join_setup(j)) return rc;
    const std::string cls = tune_class(j);
    JoinTune t;
    if (!tune_lookup(c, cls, &t) && (flags & VQ_BATCH_TUNE)) {
      t = autotune(j, &outs[i]);
      tune_store(c, cls, t);
    }
    if (t.bits_delta || t.pbuf >= 0 || t.tile_shift || t.reps) {
      j.tune = t;
      if (int rc = join_setup(j)) return rc;  // the bits again, with the tuning
    }
    return join_plan(j);
  };
  if (flags & VQ_BATCH_TUNE) {
    for (int32_t i = 0; i < n; i++)
      if (jobs[i])
        if (int rc = plan_one(i)) return rc;
  } else {
    // planned by up to eight host threads (zone pruning, spans and kernel
    // lookups per instance); the first failure in instance order is reported
    std::vector<int> prc(n, VQ_OK);
    std::vector<std::string> perr(n);
    std::atomic<int32_t> next{0};
    auto worker = [&]() {
      hipSetDevice(c->device);
      for (int32_t i; (i = next.fetch_add(1)) < n;) {
        if (!jobs[i]) continue;
        prc[i] = plan_one(i);
        if (prc[i] != VQ_OK) perr[i] = g_err;
      }
    };
    const int nt = (int)std::min<int32_t>(8, std::max<int32_t>(1, n / 16));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) {
      try {
        th.emplace_back(worker);
      } catch (const std::system_error&) {
        break;  // (the threads there are, this one included, plan them all)
      }
    }
    worker();
    for (auto& x : th) x.join();
    for (int32_t i = 0; i < n; i++)
      if (prc[i] != VQ_OK) {
        g_err = perr[i];
        return prc[i];
      }
  }
  std::vector<int> rcs(n, VQ_OK);
  std::vector<std::string> errs(n);
  {
    SlotLease S(c);
    size_t dev = 0, host = 0;
    for (auto& j : jobs)
      if (j) {
        dev += j->bytes;
        host += (j->host_bytes + 255) & ~(size_t)255;
      }
    if (dev > 0) {
      void* ar;
      void* hp;
      if (int rc = arena_get(c, S.get(), dev, &ar)) return rc;
      if (int rc = pinned_get(S.get(), host, &hp)) return rc;
      size_t od = 0, oh = 0;
      for (auto& j : jobs)
        if (j) {
          join_bind(*j, (char*)ar + od, (char*)hp + oh);
          od += j->bytes;
          oh += (j->host_bytes + 255) & ~(size_t)255;
        }
      if (int rc = slot_branches(S.get())) return rc;
      for (auto& j : jobs)
        if (j && j->plan.lds_bytes > 64 * 1024) join_allow_lds(j->plan.lds_max);
      const int K = Slot::kBranches;
      if (flags & VQ_BATCH_NO_GRAPH) {
        int b = 0;
        for (auto& j : jobs)
          if (j) HIP_TRY(join_enqueue(*j, S->side[b++ % K], nullptr));
        for (int k = 0; k < K; k++) HIP_TRY(hipStreamSynchronize(S->side[k]));
      } else {
        // the batch's graph: instance i on branch i mod K, forked from and
        // joined back into the slot's stream; reused while the batch's
        // launch arguments are the same
        Fp fp;
        for (auto& j : jobs)
          if (j) {
            fp.add(&j->plan, sizeof j->plan);
            fp.add(&j->jd, sizeof j->jd);
            fp.add(&j->B, sizeof j->B);
            fp.add(&j->bl.args, sizeof j->bl.args);
            fp.add(&j->pl.args, sizeof j->pl.args);
            fp.add(&j->bl.grid, sizeof j->bl.grid);
            fp.add(&j->pl.grid, sizeof j->pl.grid);
            fp.add(&j->hflags, sizeof j->hflags);
          }
        hipGraphExec_t exec = nullptr;
        for (auto& g : S->graphs)
          if (g.h1 == fp.h1 && g.h2 == fp.h2) exec = g.exec;
        if (!exec) {
          hipGraph_t graph = nullptr;
          HIP_TRY(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
          hipError_t e = hipEventRecord(S->fork, S->stream);
          for (int k = 0; k < K && e == hipSuccess; k++) e = hipStreamWaitEvent(S->side[k], S->fork, 0);
          int b = 0;
          for (auto& j : jobs)
            if (j && e == hipSuccess) e = join_enqueue(*j, S->side[b++ % K], nullptr);
          for (int k = 0; k < K && e == hipSuccess; k++) {
            e = hipEventRecord(S->join[k], S->side[k]);
            if (e == hipSuccess) e = hipStreamWaitEvent(S->stream, S->join[k], 0);
          }
          const hipError_t e2 = hipStreamEndCapture(S->stream, &graph);
          if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return fail(VQ_E_DEVICE, "exec: gpu: batch capture: %s",
                        hipGetErrorString(e != hipSuccess ? e : e2));
          }
          e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          hipGraphDestroy(graph);
          if (e != hipSuccess) return fail(VQ_E_DEVICE, "exec: gpu: batch graph: %s", hipGetErrorString(e));
          if (S->graphs.size() >= 4) {
            hipGraphExecDestroy(S->graphs.front().exec);
            S->graphs.erase(S->graphs.begin());
          }
          S->graphs.push_back({fp.h1, fp.h2, exec});
        }
        HIP_TRY(hipEventRecord(S->ev[0], S->stream));
        HIP_TRY(hipGraphLaunch(exec, S->stream));
        HIP_TRY(hipEventRecord(S->ev[1], S->stream));
        HIP_TRY(hipStreamSynchronize(S->stream));
      }
      const float batch_ms = (flags & VQ_BATCH_NO_GRAPH) ? 0.0f : elapsed(S->ev[0], S->ev[1]);
      // the instances' results, assembled by a few host threads (copying and
      // ordering the groups of a large batch is host-memory bound)
      std::atomic<int32_t> next{0};
      auto assemble = [&](int k) {
        hipSetDevice(c->device);
        for (int32_t i; (i = next.fetch_add(1)) < n;) {
          if (!jobs[i]) continue;
          jobs[i]->pin = &S->rec_pin[k];
          rcs[i] = join_result(*jobs[i], &outs[i], stats ? &stats[i] : nullptr, nullptr);
          jobs[i]->pin = nullptr;
          if (rcs[i] != VQ_OK && rcs[i] != kRetry) errs[i] = g_err;
          if (stats && rcs[i] == VQ_OK) stats[i].device_ms = batch_ms;  // the whole batch's graph
        }
      };
      const int nt = std::max(1, std::min<int>({Slot::kAssemblers, n, (int)std::thread::hardware_concurrency()}));
      std::vector<std::thread> helpers;
      for (int k = 1; k < nt; k++) {
        try {
          helpers.emplace_back(assemble, k);
        } catch (const std::system_error&) {
          break;  // (the threads there are take every instance)
        }
      }
      assemble(0);
      for (auto& h : helpers) h.join();
    }
  }
  // instances that overflowed a partition run again on their own (with the
  // estimates this run has learned), and dictionary GROUP BYs run directly
  int first_rc = VQ_OK;
  std::string first_err;
  for (int32_t i = 0; i < n; i++) {
    const vq_query& q = qs[i];
    int rc = rcs[i];
    if (rc != VQ_OK && rc != kRetry) g_err = errs[i];
    if (direct[i]) {
      rc = vq_agg_dict(c, q.table, q.pred, q.nkeys, q.key_cols, q.naggs, q.aggs, &outs[i],
                       stats ? &stats[i] : nullptr);
    } else if (rc == kRetry) {
      JoinJob& j = *jobs[i];
      rc = join_impl(c, j.build, j.build_key, j.bp, j.probe, j.probe_key, j.pp, j.naggs, j.aggs,
                     j.groupby ? -1 : j.radix_bits, j.topk, &outs[i], stats ? &stats[i] : nullptr,
                     j.groupby, &j.tune);
    }
    if (rc && first_rc == VQ_OK) {
      first_rc = rc;
      first_err = g_err;
    }
  }
  if (first_rc != VQ_OK) g_err = first_err;
  return first_rc;
}
