// exec.cpp — the C++ operator layer over libvexqgpu.so (see exec.hpp).
#include "exec.hpp"

#include <algorithm>
#include <cstring>

namespace vexq {
namespace exec {

namespace {

constexpr int kBlockRows = VQ_BLOCK_ROWS;  // storage.BlockRows (storage/format.go:8)

int vq_type_of(const HostColumn& c) { return c.vq_type; }

vq_expr expr_of(const HostTable& t, const AggExpr& a, const HostTable* build = nullptr) {
  if (a.expr.nfactors > 0 || a.col_idx < 0) return a.expr;
  vq_expr e;
  memset(&e, 0, sizeof e);
  const HostTable& src = a.build_side && build ? *build : t;
  e.nfactors = 1;
  e.is_float = src.columns[a.col_idx].vq_type == VQ_F64;
  e.f[0].kind = VQ_F_COL;
  e.f[0].col = a.build_side ? (a.col_idx | VQ_BUILD_COL) : a.col_idx;
  return e;
}

vq_agg to_vq(const HostTable& t, const AggExpr& a, const HostTable* build = nullptr) {
  vq_agg g;
  memset(&g, 0, sizeof g);
  switch (a.kind) {
    case AggKind::Count: g.kind = VQ_AGG_COUNT; break;
    case AggKind::Sum: g.kind = VQ_AGG_SUM; break;
    case AggKind::Avg: g.kind = VQ_AGG_AVG; break;
    case AggKind::Min: g.kind = VQ_AGG_MIN; break;
    case AggKind::Max: g.kind = VQ_AGG_MAX; break;
  }
  g.accum = a.accum_type == DataType::Float64 ? VQ_F64 : a.accum_type == DataType::String ? VQ_STR : VQ_I64;
  // COUNT(col) of a column with NULLs counts the rows holding a value
  // (exec/aggregate.go:466-473); over any other column it is COUNT(*)
  const HostTable& src = a.build_side && build ? *build : t;
  const bool count_nullable = a.kind == AggKind::Count && a.expr.nfactors == 0 && a.col_idx >= 0 &&
                              a.col_idx < (int)src.columns.size() && src.columns[a.col_idx].valid;
  if (a.kind != AggKind::Count || count_nullable) g.arg = expr_of(t, a, build);
  return g;
}

// Output column type of an aggregate (buildOutputBatch, exec/aggregate.go:780-876):
// MIN/MAX keep the source type, AVG is float, COUNT int, SUM its accumulator.
DataType out_type(const HostTable& t, const AggExpr& a, const HostTable* build = nullptr) {
  switch (a.kind) {
    case AggKind::Count: return DataType::Int64;
    case AggKind::Avg: return DataType::Float64;
    case AggKind::Sum: return a.accum_type;
    default:
      if (a.expr.nfactors == 0 && a.col_idx >= 0)
        return (a.build_side && build ? *build : t).columns[a.col_idx].type;
      return a.accum_type;
  }
}

// Agg result rows [lo, hi) -> one batch.
Batch agg_batch(const Schema& schema, const vq_agg_result& r, int nkeys, int naggs,
                const std::vector<std::shared_ptr<const std::vector<std::string>>>& key_dicts,
                const std::vector<DataType>& key_types, const std::vector<DataType>& agg_types,
                int64_t lo, int64_t hi,
                const std::vector<std::shared_ptr<const std::vector<std::string>>>* agg_dicts = nullptr) {
  Batch b;
  b.schema = schema;
  b.length = (int)(hi - lo);
  for (int k = 0; k < nkeys; k++) {
    Vector v;
    v.type = key_types[k];
    for (int64_t i = lo; i < hi; i++) {
      if (v.type == DataType::String) {
        // a NULL key's group (key -1): a NULL row of the key column
        const int64_t code = r.keys[i * nkeys + k];
        if (code < 0 && v.valid.empty()) v.valid.assign((size_t)(i - lo), 1);
        if (!v.valid.empty()) v.valid.push_back(code >= 0);
        v.codes.push_back(code < 0 ? 0u : (uint32_t)code);
      } else if (v.type == DataType::Float64) {  // (a float column riding with a key: its bits)
        double d;
        memcpy(&d, &r.keys[i * nkeys + k], 8);
        v.f64.push_back(d);
      } else {
        v.i64.push_back(r.keys[i * nkeys + k]);
      }
    }
    v.dict = key_dicts[k];
    b.vectors.push_back(std::move(v));
  }
  for (int j = 0; j < naggs; j++) {
    Vector v;
    v.type = agg_types[j];
    for (int64_t i = lo; i < hi; i++) {
      const int64_t o = i * naggs + j;
      if (v.type == DataType::Float64) v.f64.push_back(r.fvals[o]);
      else if (v.type == DataType::String) v.codes.push_back(r.valid[o] ? (uint32_t)r.ivals[o] : 0u);
      else v.i64.push_back(r.ivals[o]);
      v.valid.push_back(r.valid[o]);
    }
    // MIN/MAX over strings: codes of the source column's dictionary
    if (v.type == DataType::String && agg_dicts) v.dict = (*agg_dicts)[j];
    b.vectors.push_back(std::move(v));
  }
  return b;
}

struct ResultBufs {
  std::vector<int64_t> keys, counts, first, ivals;
  std::vector<double> fvals;
  std::vector<uint8_t> valid;
  vq_agg_result r{};
  void alloc(int64_t cap, int nkeys, int naggs) {
    cap = std::max<int64_t>(cap, 1);
    keys.assign(cap * std::max(nkeys, 1), 0);
    counts.assign(cap, 0);
    first.assign(cap, 0);
    ivals.assign(cap * std::max(naggs, 1), 0);
    fvals.assign(cap * std::max(naggs, 1), 0);
    valid.assign(cap * std::max(naggs, 1), 0);
    r.capacity = cap;
    r.ngroups = 0;
    r.keys = keys.data();
    r.counts = counts.data();
    r.first_row = first.data();
    r.ivals = ivals.data();
    r.fvals = fvals.data();
    r.valid = valid.data();
  }
};

// ---- filter scan ---------------------------------------------------------
class GpuFilterScan : public Operator {
 public:
  GpuFilterScan(GpuTable* t, std::vector<int> cols, const vq_pred* pred)
      : t_(t), cols_(std::move(cols)), has_pred_(pred != nullptr) {
    if (pred) pred_ = *pred;
    const Schema full = t->host().schema();
    for (int c : cols_) schema_.fields.push_back(full.fields[c]);
  }

  Status Next(const Batch** out) override {
    *out = nullptr;
    if (!ran_) {
      vq_selvec* sv = nullptr;
      int rc = vq_filter(t_->device()->ctx(), t_->handle(), has_pred_ ? &pred_ : nullptr, &sv, nullptr);
      if (rc) return Status::FromLib(rc, "filter");
      rows_.resize((size_t)vq_selvec_count(sv));
      rc = vq_selvec_copy(sv, rows_.data(), (int64_t)rows_.size());
      vq_selvec_destroy(sv);
      if (rc) return Status::FromLib(rc, "filter");
      ran_ = true;
      // physical 1024-row blocks inside each row group
      int64_t st = 0;
      for (int64_t n : t_->host().row_groups) {
        for (int64_t b = 0; b < n; b += kBlockRows) blocks_.push_back({st + b, std::min<int64_t>(kBlockRows, n - b)});
        st += n;
      }
    }
    // Filter.Next skips blocks whose selection is empty (exec/filter.go:49-78).
    while (bi_ < blocks_.size()) {
      const int64_t b0 = blocks_[bi_].first, bn = blocks_[bi_].second;
      bi_++;
      size_t e = ri_;
      while (e < rows_.size() && rows_[e] < b0 + bn) e++;
      if (e == ri_) continue;
      cur_ = Batch();
      cur_.schema = schema_;
      cur_.length = (int)(e - ri_);
      for (size_t i = ri_; i < e; i++) cur_.sel.push_back((uint16_t)(rows_[i] - b0));
      ri_ = e;
      for (int c : cols_) {
        const HostColumn& hc = t_->host().columns[c];
        Vector v;
        v.type = hc.type;
        v.dict = hc.dict;
        for (int64_t r = b0; r < b0 + bn; r++) {
          switch (hc.vq_type) {
            case VQ_U8: v.codes.push_back(((const uint8_t*)hc.data)[r]); break;
            case VQ_U16: v.codes.push_back(((const uint16_t*)hc.data)[r]); break;
            case VQ_DATE: v.i64.push_back(((const int32_t*)hc.data)[r]); break;
            case VQ_I64: v.i64.push_back(((const int64_t*)hc.data)[r]); break;
            default: v.f64.push_back(((const double*)hc.data)[r]); break;
          }
        }
        if (!hc.dict && (hc.vq_type == VQ_U8 || hc.vq_type == VQ_U16)) {
          for (uint32_t x : v.codes) v.i64.push_back(x);
          v.codes.clear();
        }
        cur_.vectors.push_back(std::move(v));
      }
      *out = &cur_;
      return Status::Ok();
    }
    return Status::Ok();
  }
  const Schema& schema() const override { return schema_; }
  Status Close() override { return Status::Ok(); }

 private:
  GpuTable* t_;
  std::vector<int> cols_;
  bool has_pred_;
  vq_pred pred_{};
  Schema schema_;
  bool ran_ = false;
  std::vector<uint32_t> rows_;
  std::vector<std::pair<int64_t, int64_t>> blocks_;
  size_t bi_ = 0, ri_ = 0;
  Batch cur_;
};

// ---- dictionary GROUP BY -----------------------------------------------------
class GpuParallelHashAggregate : public Operator {
 public:
  GpuParallelHashAggregate(GpuTable* t, const vq_pred* pred, std::vector<int> group_by,
                           std::vector<AggExpr> aggs)
      : t_(t), has_pred_(pred != nullptr), group_by_(std::move(group_by)), aggs_(std::move(aggs)) {
    if (pred) pred_ = *pred;
    const HostTable& h = t->host();
    const Schema full = h.schema();
    for (int c : group_by_) {
      schema_.fields.push_back(full.fields[c]);
      key_dicts_.push_back(h.columns[c].dict);
      key_types_.push_back(h.columns[c].type);
    }
    for (const AggExpr& a : aggs_) {
      const DataType ty = out_type(h, a);
      schema_.fields.push_back({a.out_name, ty, a.kind != AggKind::Count});
      agg_types_.push_back(ty);
      agg_dicts_.push_back(ty == DataType::String && a.col_idx >= 0 ? h.columns[a.col_idx].dict : nullptr);
    }
  }

  // Work is deferred to the first Next, as in ParallelHashAggregate.Next
  // (exec/parallel.go:109-120).
  Status Next(const Batch** out) override {
    *out = nullptr;
    if (!done_) {
      const HostTable& h = t_->host();
      // at most one group per code tuple (a key with NULLs: one group more,
      // the NULL one), and per row (an integer key column: any values)
      int64_t cap = 1;
      for (int c : group_by_)
        cap = std::min<int64_t>(cap * std::max<int64_t>(1, h.columns[c].dict ? (int64_t)h.columns[c].dict->size() +
                                                                                  (h.columns[c].valid ? 1 : 0)
                                                                            : h.nrows),
                                std::max<int64_t>(1, h.nrows));
      std::vector<vq_agg> va;
      for (const AggExpr& a : aggs_) {
        va.push_back(to_vq(h, a));
        if (va.back().accum == VQ_STR && a.col_idx >= 0) {
          Status st = t_->EnsureDictOrder(a.col_idx);
          if (!st.ok()) return st;
        }
      }
      ResultBufs rb;
      rb.alloc(cap, (int)group_by_.size(), (int)aggs_.size());
      int rc = vq_agg_dict(t_->device()->ctx(), t_->handle(), has_pred_ ? &pred_ : nullptr,
                           (int32_t)group_by_.size(), group_by_.data(), (int32_t)va.size(),
                           va.data(), &rb.r, nullptr);
      if (rc) return Status::FromLib(rc, "hash aggregate");
      for (int64_t lo = 0; lo < rb.r.ngroups; lo += kBlockRows)
        batches_.push_back(agg_batch(schema_, rb.r, (int)group_by_.size(), (int)aggs_.size(),
                                     key_dicts_, key_types_, agg_types_, lo,
                                     std::min<int64_t>(rb.r.ngroups, lo + kBlockRows), &agg_dicts_));
      done_ = true;
    }
    if (pos_ < batches_.size()) *out = &batches_[pos_++];
    return Status::Ok();
  }
  const Schema& schema() const override { return schema_; }
  Status Close() override { return Status::Ok(); }

 private:
  GpuTable* t_;
  bool has_pred_;
  vq_pred pred_{};
  std::vector<int> group_by_;
  std::vector<AggExpr> aggs_;
  Schema schema_;
  std::vector<std::shared_ptr<const std::vector<std::string>>> key_dicts_, agg_dicts_;
  std::vector<DataType> key_types_, agg_types_;
  bool done_ = false;
  std::vector<Batch> batches_;
  size_t pos_ = 0;
};

}  // namespace

// ---- small pieces ------------------------------------------------------------
int Schema::IndexOf(const std::string& name) const {
  for (size_t i = 0; i < fields.size(); i++)
    if (fields[i].name == name) return (int)i;
  return -1;
}

Status Status::FromLib(int code, const char* what) {
  Status s;
  s.code = code;
  s.msg = std::string("exec: gpu ") + what + ": " + vq_last_error();
  return s;
}

DataType AccumTypeFor(AggKind kind, DataType src) {
  switch (kind) {
    case AggKind::Avg: return DataType::Float64;
    case AggKind::Sum: return src == DataType::Float64 ? DataType::Float64 : DataType::Int64;
    case AggKind::Min:
    case AggKind::Max:
      if (src == DataType::Float64) return DataType::Float64;
      if (src == DataType::String) return DataType::String;
      return DataType::Int64;
    default: return DataType::Int64;
  }
}

Schema HostTable::schema() const {
  Schema s;
  for (const HostColumn& c : columns) s.fields.push_back({c.name, c.type, false});
  return s;
}

Status Device::Open(int device, uint64_t budget, std::unique_ptr<Device>* out) {
  std::unique_ptr<Device> d(new Device());
  int rc = vq_ctx_create(device, budget, &d->ctx_);
  if (rc) return Status::FromLib(rc, "open device");
  *out = std::move(d);
  return Status::Ok();
}

Device::~Device() { vq_ctx_destroy(ctx_); }

Status GpuTable::Upload(Device* dev, const HostTable& host, std::unique_ptr<GpuTable>* out) {
  std::vector<vq_col_desc> cols;
  for (const HostColumn& c : host.columns) {
    vq_col_desc d{};
    d.type = vq_type_of(c);
    d.dict_size = c.dict ? (int32_t)c.dict->size() : 0;
    d.data = c.data;
    d.valid = c.valid;
    cols.push_back(d);
  }
  std::unique_ptr<GpuTable> t(new GpuTable());
  t->dev_ = dev;
  t->host_ = host;
  int rc = vq_table_upload(dev->ctx(), host.nrows, (int32_t)cols.size(), cols.data(),
                           (int32_t)host.row_groups.size(), host.row_groups.data(), nullptr, &t->t_);
  if (rc) return Status::FromLib(rc, "table upload");
  *out = std::move(t);
  return Status::Ok();
}

GpuTable::~GpuTable() { vq_table_destroy(t_); }

Status GpuTable::EnsureDictOrder(int col) {
  if (ordered_.size() < host_.columns.size()) ordered_.resize(host_.columns.size(), false);
  if (col < 0 || col >= (int)host_.columns.size() || !host_.columns[col].dict) {
    Status s;
    s.code = VQ_E_ARG;
    s.msg = "exec: gpu: MIN/MAX over strings of a column without a dictionary";
    return s;
  }
  if (ordered_[col]) return Status::Ok();
  const std::vector<std::string>& d = *host_.columns[col].dict;
  std::vector<uint32_t> idx(d.size()), rank(d.size());
  for (size_t k = 0; k < d.size(); k++) idx[k] = (uint32_t)k;
  // std::string compares bytes as unsigned, as Go compares strings
  std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return d[a] < d[b]; });
  for (size_t k = 0; k < idx.size(); k++) rank[idx[k]] = (uint32_t)k;
  if (int rc = vq_table_dict_order(t_, col, rank.data())) return Status::FromLib(rc, "string order");
  ordered_[col] = true;
  return Status::Ok();
}

std::unique_ptr<Operator> NewGpuFilterScan(GpuTable* table, std::vector<int> cols, const vq_pred* pred) {
  return std::unique_ptr<Operator>(new GpuFilterScan(table, std::move(cols), pred));
}

std::unique_ptr<Operator> NewGpuParallelHashAggregate(GpuTable* table, const vq_pred* pred,
                                                      std::vector<int> group_by,
                                                      std::vector<AggExpr> aggs) {
  return std::unique_ptr<Operator>(
      new GpuParallelHashAggregate(table, pred, std::move(group_by), std::move(aggs)));
}

// ---- join ----------------------------------------------------------------------
GpuParallelHashJoinAggregate::GpuParallelHashJoinAggregate(GpuTable* build, int build_key,
                                                           const vq_pred* build_pred,
                                                           GpuTable* probe, int probe_key,
                                                           const vq_pred* probe_pred,
                                                           std::vector<AggExpr> aggs,
                                                           JoinOptions opts)
    : build_(build), probe_(probe), build_key_(build_key), probe_key_(probe_key),
      has_bpred_(build_pred != nullptr), has_ppred_(probe_pred != nullptr),
      aggs_(std::move(aggs)), opts_(opts) {
  if (build_pred) bpred_ = *build_pred;
  if (probe_pred) ppred_ = *probe_pred;
  if (opts_.inner) {  // the key, then the build columns that ride with it
    const Schema bs = build->host().schema();
    schema_.fields.push_back(bs.fields[build_key]);
    for (int c : opts_.build_pay) schema_.fields.push_back(bs.fields[c]);
  } else if (opts_.group_by.empty()) {
    schema_.fields.push_back(build->host().schema().fields[build_key]);
  } else {
    const Schema ps = probe->host().schema();
    for (int c : opts_.group_by) schema_.fields.push_back(ps.fields[c]);
  }
  for (const AggExpr& a : aggs_)
    schema_.fields.push_back({a.out_name, out_type(probe->host(), a, &build->host()), a.kind != AggKind::Count});
}

Status GpuParallelHashJoinAggregate::Setup() {
  if (opts_.inner) return SetupNested();
  if (!opts_.group_by.empty()) return SetupGrouped();
  std::vector<vq_agg> va;
  for (const AggExpr& a : aggs_) va.push_back(to_vq(probe_->host(), a));
  const int64_t cap = opts_.topk > 0 ? opts_.topk
                                     : std::max<int64_t>(1, std::min(build_->host().nrows, probe_->host().nrows));
  ResultBufs rb;
  rb.alloc(cap, 1, (int)aggs_.size());
  vq_stats st;
  int rc = vq_join_agg(build_->device()->ctx(), build_->handle(), build_key_,
                       has_bpred_ ? &bpred_ : nullptr, probe_->handle(), probe_key_,
                       has_ppred_ ? &ppred_ : nullptr, (int32_t)va.size(), va.data(),
                       opts_.radix_bits, opts_.topk, &rb.r, &st);
  if (rc) return Status::FromLib(rc, "parallel join");
  stats_.rows = st.build_rows;
  stats_.keys = st.build_keys;
  stats_.partitions = st.partitions;
  stats_.elapsed_ms = st.main_kernel_ms;
  std::vector<DataType> agg_types;
  for (size_t j = 0; j < aggs_.size(); j++) agg_types.push_back(schema_.fields[j + 1].type);
  const std::vector<std::shared_ptr<const std::vector<std::string>>> kd = {nullptr};
  const std::vector<DataType> kt = {DataType::Int64};
  for (int64_t lo = 0; lo < rb.r.ngroups; lo += kBlockRows)
    out_.push_back(agg_batch(schema_, rb.r, 1, (int)aggs_.size(), kd, kt, agg_types, lo,
                             std::min<int64_t>(rb.r.ngroups, lo + kBlockRows)));
  return Status::Ok();
}

// A nested build side (vq_join_agg_nested): the key and its riding build
// columns, then the aggregates.
Status GpuParallelHashJoinAggregate::SetupNested() {
  std::vector<vq_agg> va;
  for (const AggExpr& a : aggs_) va.push_back(to_vq(probe_->host(), a));
  const int np = (int)opts_.build_pay.size(), na = (int)aggs_.size();
  const int64_t cap = opts_.topk > 0 ? opts_.topk
                                     : std::max<int64_t>(1, std::min(build_->host().nrows, probe_->host().nrows));
  ResultBufs rb;
  rb.alloc(cap, 1, na);
  std::vector<int64_t> pay((size_t)cap * std::max(np, 1));
  vq_pred ip{};
  if (opts_.inner_pred) ip = *opts_.inner_pred;
  vq_stats st;
  int rc = vq_join_agg_nested(build_->device()->ctx(), opts_.inner->handle(), opts_.inner_key,
                              opts_.inner_pred ? &ip : nullptr, build_->handle(), opts_.build_fk, build_key_,
                              has_bpred_ ? &bpred_ : nullptr, probe_->handle(), probe_key_,
                              has_ppred_ ? &ppred_ : nullptr, np, opts_.build_pay.data(), na, va.data(),
                              opts_.radix_bits, opts_.topk, &rb.r, pay.data(), &st);
  if (rc) return Status::FromLib(rc, "parallel join aggregate");
  stats_.rows = st.build_rows;
  stats_.keys = st.build_keys;
  stats_.partitions = st.partitions;
  stats_.elapsed_ms = st.device_ms;
  // the batch: key and riding columns as group-by columns of the result
  const HostTable& hb = build_->host();
  std::vector<std::shared_ptr<const std::vector<std::string>>> kd = {nullptr};
  std::vector<DataType> kt = {DataType::Int64}, agg_types;
  for (int c : opts_.build_pay) {
    kd.push_back(hb.columns[c].dict);
    kt.push_back(hb.columns[c].type);
  }
  for (int j = 0; j < na; j++) agg_types.push_back(schema_.fields[1 + np + j].type);
  const int64_t n = rb.r.ngroups;
  ResultBufs kb;  // keys widened to 1 + np columns per group
  kb.alloc(std::max<int64_t>(n, 1), 1 + np, na);
  kb.r.ngroups = n;
  for (int64_t g = 0; g < n; g++) {
    kb.keys[g * (1 + np)] = rb.r.keys[g];
    for (int c = 0; c < np; c++) kb.keys[g * (1 + np) + 1 + c] = pay[(size_t)g * np + c];
    kb.counts[g] = rb.r.counts[g];
    kb.first[g] = rb.r.first_row[g];
    for (int j = 0; j < na; j++) {
      kb.ivals[g * na + j] = rb.r.ivals[g * na + j];
      kb.fvals[g * na + j] = rb.r.fvals[g * na + j];
      kb.valid[g * na + j] = rb.r.valid[g * na + j];
    }
  }
  for (int64_t lo = 0; lo < n; lo += kBlockRows)
    out_.push_back(agg_batch(schema_, kb.r, 1 + np, na, kd, kt, agg_types, lo, std::min<int64_t>(n, lo + kBlockRows)));
  return Status::Ok();
}

// GROUP BY probe columns over the joined rows (vq_join_group), its result
// sized by the two-call capacity query.
Status GpuParallelHashJoinAggregate::SetupGrouped() {
  const HostTable& hp = probe_->host();
  std::vector<vq_agg> va;
  for (const AggExpr& a : aggs_) va.push_back(to_vq(hp, a, &build_->host()));
  const int nk = (int)opts_.group_by.size(), na = (int)aggs_.size();
  ResultBufs rb;
  vq_stats st;
  int64_t cap = 1024;
  int rc = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    rb.alloc(cap, nk, na);
    rc = vq_join_group(build_->device()->ctx(), build_->handle(), build_key_, has_bpred_ ? &bpred_ : nullptr,
                       probe_->handle(), probe_key_, has_ppred_ ? &ppred_ : nullptr, nk, opts_.group_by.data(),
                       na, va.data(), &rb.r, &st);
    if (rc != VQ_E_ARG || rb.r.ngroups <= cap) break;
    cap = rb.r.ngroups;  // the groups the result needs
  }
  if (rc) return Status::FromLib(rc, "parallel join aggregate");
  stats_.rows = st.build_rows;
  stats_.keys = -1;
  stats_.partitions = 0;
  stats_.elapsed_ms = st.device_ms;
  std::vector<std::shared_ptr<const std::vector<std::string>>> kd;
  std::vector<DataType> kt, agg_types;
  for (int c : opts_.group_by) {
    kd.push_back(hp.columns[c].dict);
    kt.push_back(hp.columns[c].type);
  }
  for (int j = 0; j < na; j++) agg_types.push_back(schema_.fields[nk + j].type);
  for (int64_t lo = 0; lo < rb.r.ngroups; lo += kBlockRows)
    out_.push_back(agg_batch(schema_, rb.r, nk, na, kd, kt, agg_types, lo,
                             std::min<int64_t>(rb.r.ngroups, lo + kBlockRows)));
  return Status::Ok();
}

Status GpuParallelHashJoinAggregate::Next(const Batch** out) {
  *out = nullptr;
  if (!done_setup_) {
    Status s = Setup();
    if (!s.ok()) return s;
    done_setup_ = true;
  }
  if (pos_ < out_.size()) *out = &out_[pos_++];
  return Status::Ok();
}

std::unique_ptr<GpuParallelHashJoinAggregate> NewGpuParallelHashJoinAggregate(
    GpuTable* build, int build_key, const vq_pred* build_pred, GpuTable* probe, int probe_key,
    const vq_pred* probe_pred, std::vector<AggExpr> aggs, JoinOptions opts) {
  return std::unique_ptr<GpuParallelHashJoinAggregate>(new GpuParallelHashJoinAggregate(
      build, build_key, build_pred, probe, probe_key, probe_pred, std::move(aggs), opts));
}

Status Drain(Operator* op, std::vector<Batch>* out) {
  for (;;) {
    const Batch* b = nullptr;
    Status s = op->Next(&b);
    if (!s.ok()) return s;
    if (!b) return Status::Ok();
    out->push_back(*b);
  }
}

}  // namespace exec
}  // namespace vexq
