// exec.hpp — C++ mirror of vexq's operator interface over the gfx950 C ABI.
//
// vexq's executor is a pull pipeline of exec.Operator values
// (exec/operator.go:25-32) exchanging exec.Batch (exec/batch.go:10-21).  The
// GPU operators here keep that contract — Next returns the next batch or
// nullptr at EOF (Go's (nil, nil)), a batch is valid until the next Next,
// Close is always safe — so a caller written against vexq's operators reads
// the same.  The constructors take the logical description the planner has
// already resolved (pushed-down predicate, group-by columns, AggExprs with
// their pre-projection), because the GPU runs the whole
// Scan -> Filter -> Project -> Aggregate pipeline as one device program
// rather than pulling batches through the host.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/vexq_gpu.h"

namespace vexq {
namespace exec {

// exec.DataType (exec/batch.go, storage/format.go:20-26).
enum class DataType { Int64, Float64, Bool, String, Date };

struct Field {
  std::string name;
  DataType type;
  bool nullable = false;
};

struct Schema {
  std::vector<Field> fields;
  int IndexOf(const std::string& name) const;
};

// One column of a batch.  Int64 and Date values are kept widened in i64,
// strings as dictionary codes into dict; valid is empty (all valid) or one
// byte per row, 1 = valid (the LSB-first validity bitmaps of batch.go:124-184
// unpacked).
struct Vector {
  DataType type = DataType::Int64;
  std::vector<int64_t> i64;
  std::vector<double> f64;
  std::vector<uint32_t> codes;
  std::shared_ptr<const std::vector<std::string>> dict;
  std::vector<uint8_t> valid;
  bool IsNull(int i) const { return !valid.empty() && !valid[i]; }
  std::string Get(int i) const { return dict ? (*dict)[codes[i]] : std::string(); }
};

// SelVec empty means every row of the batch (SelectionVector nil).
struct Batch {
  Schema schema;
  std::vector<Vector> vectors;
  int length = 0;
  std::vector<uint16_t> sel;
};

// Error value: code is a VQ_E_* status; msg follows the "exec: ..." prefix
// convention of the Go errors (exec/filter.go:42-65).
struct Status {
  int code = VQ_OK;
  std::string msg;
  bool ok() const { return code == VQ_OK; }
  static Status Ok() { return Status(); }
  static Status FromLib(int code, const char* what);
};

class Operator {
 public:
  virtual ~Operator() = default;
  // Next sets *out to the next batch, or to nullptr at end of stream.
  virtual Status Next(const Batch** out) = 0;
  virtual const Schema& schema() const = 0;
  virtual Status Close() = 0;
};

// exec.AggExpr (exec/aggregate.go:25-37).  The argument is either a column
// of the table (col_idx) or the pre-projected arithmetic the planner would
// materialise as an _agg_N column (planner/physical.go:137-168), given as a
// vq_expr with nfactors > 0.
enum class AggKind { Count, Sum, Avg, Min, Max };

struct AggExpr {
  AggKind kind = AggKind::Count;
  int col_idx = -1;
  std::string out_name;
  DataType accum_type = DataType::Int64;
  vq_expr expr{};
  // a join grouped by its joined rows (JoinOptions::group_by): col_idx names
  // a build-side column (in expr, a build column is col | VQ_BUILD_COL)
  bool build_side = false;
};

// AccumTypeFor (exec/aggregate.go:51-72).
DataType AccumTypeFor(AggKind kind, DataType src);

// ---- device context and resident tables ------------------------------------

// A host table as storage.Reader would decode it: one entry per column.
// Dictionary columns carry global codes (u8 or u16) plus the dictionary.
struct HostColumn {
  std::string name;
  DataType type = DataType::Int64;
  int vq_type = VQ_I64;  // physical type in HBM
  const void* data = nullptr;
  std::shared_ptr<const std::vector<std::string>> dict;
  // validity (the NullBitmap of exec/batch.go:113-121): bit r LSB-first, set
  // when row r holds a value; nullptr when every row does
  const uint8_t* valid = nullptr;
};

struct HostTable {
  std::vector<HostColumn> columns;
  std::vector<int64_t> row_groups;  // rows per row group
  int64_t nrows = 0;
  Schema schema() const;
};

class Device {
 public:
  static Status Open(int device, uint64_t budget_bytes, std::unique_ptr<Device>* out);
  ~Device();
  vq_ctx* ctx() const { return ctx_; }

 private:
  vq_ctx* ctx_ = nullptr;
};

class GpuTable {
 public:
  // The one bulk upload (replaces TableScan's per-block decode).
  static Status Upload(Device* dev, const HostTable& host, std::unique_ptr<GpuTable>* out);
  ~GpuTable();
  const HostTable& host() const { return host_; }
  vq_table* handle() const { return t_; }
  Device* device() const { return dev_; }
  // The byte order of dictionary column col's strings handed to the device
  // once (MIN/MAX over strings fold ranks in that order).
  Status EnsureDictOrder(int col);

 private:
  Device* dev_ = nullptr;
  HostTable host_;
  vq_table* t_ = nullptr;
  std::vector<bool> ordered_;
};

// ---- operators --------------------------------------------------------------

// TableScan(zone-pruned) -> Filter: batches of the table's physical 1024-row
// blocks that hold passing rows, each with its selection vector
// (exec/scan.go:199-236, exec/filter.go:49-78).  cols picks the output
// columns.
std::unique_ptr<Operator> NewGpuFilterScan(GpuTable* table, std::vector<int> cols,
                                           const vq_pred* pred);

// ParallelHashAggregate (exec/parallel.go:78-86): Scan -> Filter -> Project ->
// HashAggregate over dictionary group keys, emitted in <= 1024-row batches in
// first-occurrence group order.  Output schema: group-by columns then one
// column per aggregate.
std::unique_ptr<Operator> NewGpuParallelHashAggregate(GpuTable* table, const vq_pred* pred,
                                                      std::vector<int> group_by,
                                                      std::vector<AggExpr> aggs);

// JoinBuildStats (exec/parallel_join.go:697-712).
struct JoinBuildStats {
  int64_t rows = 0;
  int64_t keys = -1;
  int partitions = 0;
  bool parallel = true;
  double elapsed_ms = 0;
};

struct JoinOptions {
  int radix_bits = -1;  // < 0: choose from the build row estimate
  int topk = 0;         // > 0: ORDER BY aggregate 0 DESC LIMIT topk
  // non-empty: GROUP BY these probe columns over the joined rows (TPC-H Q12's
  // shape), aggregates over either side (AggExpr::build_side); topk unused
  std::vector<int> group_by;
  // a build side that is itself a join (TPC-H Q3's chain, the nested build of
  // planner/parallel_join.go:341-400): inner[inner_key] = build[build_fk],
  // the inner rows filtered by inner_pred; build_pay: build columns that are
  // functions of the key, output after it (Q3's o_orderdate, o_shippriority)
  GpuTable* inner = nullptr;
  int inner_key = 0;
  int build_fk = 0;
  const vq_pred* inner_pred = nullptr;
  std::vector<int> build_pay;
};

// ParallelHashJoinAggregate (exec/parallel_join.go:718-758): inner join of
// build[build_key] with probe[probe_key], each side filtered, GROUP BY the
// join key, aggregates over probe columns.  Output: key column then one
// column per aggregate.  With opts.group_by: GROUP BY those probe columns
// over the joined rows, aggregates over columns of either side; output the
// group-by columns then one column per aggregate, groups in the order of
// their first joined probe row.
class GpuParallelHashJoinAggregate : public Operator {
 public:
  GpuParallelHashJoinAggregate(GpuTable* build, int build_key, const vq_pred* build_pred,
                               GpuTable* probe, int probe_key, const vq_pred* probe_pred,
                               std::vector<AggExpr> aggs, JoinOptions opts);
  Status Next(const Batch** out) override;
  const Schema& schema() const override { return schema_; }
  Status Close() override { return Status::Ok(); }
  JoinBuildStats BuildStats() const { return stats_; }

 private:
  Status Setup();
  Status SetupGrouped();
  Status SetupNested();
  GpuTable* build_;
  GpuTable* probe_;
  int build_key_, probe_key_;
  bool has_bpred_, has_ppred_;
  vq_pred bpred_{}, ppred_{};
  std::vector<AggExpr> aggs_;
  JoinOptions opts_;
  Schema schema_;
  bool done_setup_ = false;
  std::vector<Batch> out_;
  size_t pos_ = 0;
  JoinBuildStats stats_;
};

std::unique_ptr<GpuParallelHashJoinAggregate> NewGpuParallelHashJoinAggregate(
    GpuTable* build, int build_key, const vq_pred* build_pred, GpuTable* probe, int probe_key,
    const vq_pred* probe_pred, std::vector<AggExpr> aggs, JoinOptions opts = JoinOptions());

// Drains an operator into a vector of batches' rows; helper for tests.
Status Drain(Operator* op, std::vector<Batch>* out);

}  // namespace exec
}  // namespace vexq
