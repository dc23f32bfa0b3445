// test_exec.cpp — the reference's operator-level tests, restated against the
// C++ GPU operators (vexq_amd/host/exec.hpp).  Each test names the Go test it
// follows.  Run by tests/test_gpu_operators.py on a GPU host; prints one
// "PASS name" / "FAIL name: why" line per test and exits with the failure
// count.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../vexq_amd/host/exec.hpp"

using namespace vexq::exec;

namespace {

int g_fail = 0;
std::string g_why;

#define CHECK(cond)                                              \
  do {                                                           \
    if (!(cond)) {                                               \
      g_why = std::string(#cond) + " at line " + std::to_string(__LINE__); \
      return false;                                              \
    }                                                            \
  } while (0)

#define CHECK_OK(st)                                             \
  do {                                                           \
    Status s_ = (st);                                            \
    if (!s_.ok()) {                                              \
      g_why = s_.msg;                                            \
      return false;                                              \
    }                                                            \
  } while (0)

void run(const char* name, const std::function<bool()>& fn) {
  g_why.clear();
  printf("[case] %s\n", name);
  fflush(stdout);
  if (fn()) {
    printf("PASS %s\n", name);
  } else {
    printf("FAIL %s: %s\n", name, g_why.c_str());
    g_fail++;
  }
  fflush(stdout);
}

std::unique_ptr<Device> g_dev;

std::shared_ptr<const std::vector<std::string>> dict(std::vector<std::string> v) {
  return std::make_shared<const std::vector<std::string>>(std::move(v));
}

HostColumn col(const char* name, DataType t, int vqt, const void* p,
               std::shared_ptr<const std::vector<std::string>> d = nullptr) {
  HostColumn c;
  c.name = name;
  c.type = t;
  c.vq_type = vqt;
  c.data = p;
  c.dict = std::move(d);
  return c;
}

AggExpr agg(AggKind k, int col_idx, DataType src, const char* name) {
  AggExpr a;
  a.kind = k;
  a.col_idx = col_idx;
  a.out_name = name;
  a.accum_type = AccumTypeFor(k, src);
  return a;
}

vq_pred pred1(int col, int op, int64_t iv, double fv = 0) {
  vq_pred p;
  memset(&p, 0, sizeof p);
  p.nterms = 1;
  p.terms[0].col = col;
  p.terms[0].op = op;
  p.terms[0].ival = iv;
  p.terms[0].fval = fv;
  return p;
}

// TestIntKeyMultiRowGroupDictInstability (exec/agg_intkey_test.go:221-310).
bool IntKeyMultiRowGroupDictInstability() {
  // global first-occurrence codes: R=0, A=1, N=2
  static const uint8_t flag[] = {0, 1, 2, 0, 1, 2, 0, 1};
  static const int64_t value[] = {10, 20, 30, 40, 100, 200, 300, 400};
  HostTable h;
  h.columns = {col("flag", DataType::String, VQ_U8, flag, dict({"R", "A", "N"})),
               col("value", DataType::Int64, VQ_I64, value)};
  h.row_groups = {4, 4};
  h.nrows = 8;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {0},
                                        {agg(AggKind::Sum, 1, DataType::Int64, "sum_val"),
                                         agg(AggKind::Count, -1, DataType::Int64, "cnt")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  std::map<std::string, std::pair<int64_t, int64_t>> got;
  for (const Batch& b : bs)
    for (int i = 0; i < b.length; i++) got[b.vectors[0].Get(i)] = {b.vectors[1].i64[i], b.vectors[2].i64[i]};
  CHECK(got.size() == 3);
  CHECK(got["A"].first == 520 && got["A"].second == 3);
  CHECK(got["N"].first == 230 && got["N"].second == 2);
  CHECK(got["R"].first == 350 && got["R"].second == 3);
  return true;
}

// NULLs through the operator: a NULL group key is a group of its own
// (exec/aggregate.go:424-427), emitted as a NULL row of the key column;
// SUM skips a NULL argument and COUNT(col) counts the values
// (aggregate.go:458-546), COUNT(*) every row.
bool NullKeysAndArguments() {
  static const uint8_t flag[] = {0, 1, 0, 0, 1, 0};     // R A R - A -
  static const uint8_t flag_valid[8] = {0x17};          // rows 3 and 5 NULL
  static const int64_t value[] = {10, 20, 30, 40, 50, 60};
  static const uint8_t value_valid[8] = {0x3B};         // row 2 NULL
  HostTable h;
  h.columns = {col("flag", DataType::String, VQ_U8, flag, dict({"R", "A"})),
               col("value", DataType::Int64, VQ_I64, value)};
  h.columns[0].valid = flag_valid;
  h.columns[1].valid = value_valid;
  h.row_groups = {6};
  h.nrows = 6;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {0},
                                        {agg(AggKind::Sum, 1, DataType::Int64, "s"),
                                         agg(AggKind::Count, 1, DataType::Int64, "cv"),
                                         agg(AggKind::Count, -1, DataType::Int64, "c")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 3);
  const Batch& b = bs[0];
  CHECK(!b.vectors[0].IsNull(0) && b.vectors[0].Get(0) == "R");
  CHECK(!b.vectors[0].IsNull(1) && b.vectors[0].Get(1) == "A");
  CHECK(b.vectors[0].IsNull(2));
  CHECK(b.vectors[1].i64[0] == 10 && b.vectors[2].i64[0] == 1 && b.vectors[3].i64[0] == 2);
  CHECK(b.vectors[1].i64[1] == 70 && b.vectors[2].i64[1] == 2 && b.vectors[3].i64[1] == 2);
  CHECK(b.vectors[1].i64[2] == 100 && b.vectors[2].i64[2] == 2 && b.vectors[3].i64[2] == 2);
  return true;
}

// MIN/MAX over a STRING column (minMaxString, exec/aggregate.go:670-730):
// bytewise string order, not the codes' first-occurrence order; a NULL
// string skipped.
bool StringMinMax() {
  static const uint8_t name[] = {0, 1, 2, 3, 0, 2, 1, 3};  // mango a Zebra ab mango Zebra a ab
  static const uint8_t name_valid[8] = {0xF7};              // row 3 NULL
  static const uint8_t grp[] = {0, 0, 0, 0, 1, 1, 2, 2};
  HostTable h;
  h.columns = {col("name", DataType::String, VQ_U8, name, dict({"mango", "a", "Zebra", "ab"})),
               col("grp", DataType::String, VQ_U8, grp, dict({"x", "y", "z"}))};
  h.columns[0].valid = name_valid;
  h.row_groups = {8};
  h.nrows = 8;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {1},
                                        {agg(AggKind::Min, 0, DataType::String, "mn"),
                                         agg(AggKind::Max, 0, DataType::String, "mx")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 3);
  const Batch& b = bs[0];
  // x: mango a Zebra (ab NULL) -> Zebra / mango; y: mango Zebra; z: a ab
  CHECK(b.vectors[0].Get(0) == "x" && b.vectors[1].Get(0) == "Zebra" && b.vectors[2].Get(0) == "mango");
  CHECK(b.vectors[0].Get(1) == "y" && b.vectors[1].Get(1) == "Zebra" && b.vectors[2].Get(1) == "mango");
  CHECK(b.vectors[0].Get(2) == "z" && b.vectors[1].Get(2) == "a" && b.vectors[2].Get(2) == "ab");
  CHECK(!b.vectors[1].IsNull(0) && !b.vectors[2].IsNull(2));
  return true;
}

// A composite key of a STRING and an INT64 column (the string-key path,
// exec/aggregate.go:409-557): groups in first-occurrence order.
bool CompositeKey() {
  static const uint8_t flag[] = {0, 1, 0, 1, 1};
  static const int64_t year[] = {2020, 2020, 2021, 2020, 2021};
  static const int64_t value[] = {1, 2, 4, 8, 16};
  HostTable h;
  h.columns = {col("flag", DataType::String, VQ_U8, flag, dict({"R", "A"})),
               col("year", DataType::Int64, VQ_I64, year), col("value", DataType::Int64, VQ_I64, value)};
  h.row_groups = {5};
  h.nrows = 5;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {0, 1}, {agg(AggKind::Sum, 2, DataType::Int64, "s")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 4);
  const Batch& b = bs[0];
  CHECK(b.vectors[0].Get(0) == "R" && b.vectors[1].i64[0] == 2020 && b.vectors[2].i64[0] == 1);
  CHECK(b.vectors[0].Get(1) == "A" && b.vectors[1].i64[1] == 2020 && b.vectors[2].i64[1] == 10);
  CHECK(b.vectors[0].Get(2) == "R" && b.vectors[1].i64[2] == 2021 && b.vectors[2].i64[2] == 4);
  CHECK(b.vectors[0].Get(3) == "A" && b.vectors[1].i64[3] == 2021 && b.vectors[2].i64[3] == 16);
  return true;
}

// TestAccumulateSelVecFourPaths (exec/aggregate_rows_test.go:249-322); the
// selection vectors {0,2,4,6} / {7} become the predicate sel == 1.
bool AccumulateSelVecFourPaths() {
  static const uint8_t flag[] = {0, 1, 0, 1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0};
  static const uint8_t region[] = {0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 2, 2, 2, 2, 2, 0};
  static const double price[] = {1, 100, 2, 200, 4, 400, 8, 800, -1, -1, -1, -1, -1, -1, -1, 16};
  static const uint8_t sel[] = {1, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1};
  HostTable h;
  h.columns = {col("flag", DataType::String, VQ_U8, flag, dict({"A", "B"})),
               col("region", DataType::String, VQ_U8, region, dict({"1", "2", "9"})),
               col("price", DataType::Float64, VQ_F64, price),
               col("sel", DataType::Int64, VQ_U8, sel)};
  h.row_groups = {8, 8};
  h.nrows = 16;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  vq_pred p = pred1(3, VQ_EQ, 1);
  {  // no_group_by_direct: {31, 5, 1, 16}
    auto op = NewGpuParallelHashAggregate(
        t.get(), &p, {},
        {agg(AggKind::Sum, 2, DataType::Float64, "sum_price"), agg(AggKind::Count, -1, DataType::Int64, "cnt"),
         agg(AggKind::Min, 2, DataType::Float64, "min_price"), agg(AggKind::Max, 2, DataType::Float64, "max_price")});
    std::vector<Batch> bs;
    CHECK_OK(Drain(op.get(), &bs));
    CHECK(bs.size() == 1 && bs[0].length == 1);
    CHECK(bs[0].vectors[0].f64[0] == 31 && bs[0].vectors[1].i64[0] == 5);
    CHECK(bs[0].vectors[2].f64[0] == 1 && bs[0].vectors[3].f64[0] == 16);
  }
  {  // int_key_dict_string_group: A|{31,5}
    auto op = NewGpuParallelHashAggregate(t.get(), &p, {0},
                                          {agg(AggKind::Sum, 2, DataType::Float64, "s"),
                                           agg(AggKind::Count, -1, DataType::Int64, "c")});
    std::vector<Batch> bs;
    CHECK_OK(Drain(op.get(), &bs));
    CHECK(bs.size() == 1 && bs[0].length == 1 && bs[0].vectors[0].Get(0) == "A");
    CHECK(bs[0].vectors[1].f64[0] == 31 && bs[0].vectors[2].i64[0] == 5);
  }
  {  // composite key (region, flag): 1|A {21,3}, 2|A {10,2}
    auto op = NewGpuParallelHashAggregate(t.get(), &p, {1, 0},
                                          {agg(AggKind::Sum, 2, DataType::Float64, "s"),
                                           agg(AggKind::Count, -1, DataType::Int64, "c")});
    std::vector<Batch> bs;
    CHECK_OK(Drain(op.get(), &bs));
    std::map<std::string, std::pair<double, int64_t>> got;
    for (const Batch& b : bs)
      for (int i = 0; i < b.length; i++)
        got[b.vectors[0].Get(i) + "|" + b.vectors[1].Get(i)] = {b.vectors[2].f64[i], b.vectors[3].i64[i]};
    CHECK(got.size() == 2);
    CHECK(got["1|A"].first == 21 && got["1|A"].second == 3);
    CHECK(got["2|A"].first == 10 && got["2|A"].second == 2);
  }
  return true;
}

// TestMergePartialAggNonNullCounts (exec/parallel_test.go:8-43): AVG over two
// row groups (sum 4 over 2 rows, sum 6 over 2 rows) is 2.5.
bool MergePartialAggNonNullCounts() {
  static const double v[] = {1.0, 3.0, 2.0, 4.0};
  HostTable h;
  h.columns = {col("v", DataType::Float64, VQ_F64, v)};
  h.row_groups = {2, 2};
  h.nrows = 4;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {}, {agg(AggKind::Avg, 0, DataType::Float64, "avg")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].vectors[0].f64[0] == 2.5);
  return true;
}

// Global aggregate over no rows: one row, COUNT 0, SUM/AVG/MIN NULL
// (buildEmptyGlobalResult, exec/aggregate.go:961-982; exec/exec_test.go:553-760).
bool EmptyGlobalAggregate() {
  static const int64_t x[] = {1, 2, 3};
  HostTable h;
  h.columns = {col("x", DataType::Int64, VQ_I64, x)};
  h.row_groups = {3};
  h.nrows = 3;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  vq_pred p = pred1(0, VQ_GT, 10);
  auto op = NewGpuParallelHashAggregate(t.get(), &p, {},
                                        {agg(AggKind::Count, -1, DataType::Int64, "c"),
                                         agg(AggKind::Sum, 0, DataType::Int64, "s"),
                                         agg(AggKind::Avg, 0, DataType::Int64, "a"),
                                         agg(AggKind::Min, 0, DataType::Int64, "m")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 1);
  CHECK(bs[0].vectors[0].i64[0] == 0 && !bs[0].vectors[0].IsNull(0));
  CHECK(bs[0].vectors[1].IsNull(0) && bs[0].vectors[2].IsNull(0) && bs[0].vectors[3].IsNull(0));
  // grouped aggregate over no rows emits nothing
  static const uint8_t g[] = {0, 1, 0};
  HostTable h2;
  h2.columns = {col("x", DataType::Int64, VQ_I64, x), col("g", DataType::String, VQ_U8, g, dict({"a", "b"}))};
  h2.row_groups = {3};
  h2.nrows = 3;
  std::unique_ptr<GpuTable> t2;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h2, &t2));
  auto op2 = NewGpuParallelHashAggregate(t2.get(), &p, {1}, {agg(AggKind::Count, -1, DataType::Int64, "c")});
  std::vector<Batch> bs2;
  CHECK_OK(Drain(op2.get(), &bs2));
  CHECK(bs2.empty());
  return true;
}

// TestZoneMapPruning (exec/exec_test.go:267-300): 4 row groups of ids, keep
// max(id) >= 131072 -> exactly 2 x 65536 rows read.
bool ZoneMapPruning() {
  const int64_t n = 4 * 65536;
  std::vector<int64_t> id(n);
  for (int64_t i = 0; i < n; i++) id[i] = i;
  HostTable h;
  h.columns = {col("id", DataType::Int64, VQ_I64, id.data())};
  h.row_groups = {65536, 65536, 65536, 65536};
  h.nrows = n;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  vq_pred p = pred1(0, VQ_GE, 131072);
  uint8_t keep[4];
  CHECK(vq_zone_prune(t->handle(), &p, keep) == VQ_OK);
  CHECK(!keep[0] && !keep[1] && keep[2] && keep[3]);
  auto op = NewGpuFilterScan(t.get(), {0}, &p);
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  int64_t total = 0;
  for (const Batch& b : bs) total += b.length;
  CHECK(total == 2 * 65536);
  CHECK(bs.size() == 128);  // 1024-row blocks of the two surviving row groups
  return true;
}

// Stacked filters equal one conjunction; the selection vector holds ascending
// in-block row offsets (exec/stacked_filter_test.go, exec/expr.go:1545-1562).
bool FilterSelectionVector() {
  const int64_t n = 5000;
  std::vector<int64_t> q(n);
  std::vector<double> d(n);
  for (int64_t i = 0; i < n; i++) {
    q[i] = (i * 7919) % 50 + 1;
    d[i] = (double)((i * 31) % 11) / 100.0;
  }
  HostTable h;
  h.columns = {col("q", DataType::Int64, VQ_I64, q.data()), col("d", DataType::Float64, VQ_F64, d.data())};
  h.row_groups = {n};
  h.nrows = n;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  vq_pred p;
  memset(&p, 0, sizeof p);
  p.nterms = 3;
  p.terms[0] = {0, VQ_LT, 24, 0};
  p.terms[1] = {1, VQ_GE, 0, 0.05};
  p.terms[2] = {1, VQ_LE, 0, 0.07};
  auto op = NewGpuFilterScan(t.get(), {0, 1}, &p);
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  std::vector<int64_t> got;
  int64_t block = 0;
  for (const Batch& b : bs) {
    for (size_t k = 1; k < b.sel.size(); k++) CHECK(b.sel[k] > b.sel[k - 1]);
    for (uint16_t s : b.sel) {
      CHECK(b.vectors[0].i64[s] < 24);
      CHECK(b.vectors[1].f64[s] >= 0.05 && b.vectors[1].f64[s] <= 0.07);
    }
    (void)block;
    got.push_back(b.length);
  }
  int64_t want = 0;
  for (int64_t i = 0; i < n; i++) want += q[i] < 24 && d[i] >= 0.05 && d[i] <= 0.07;
  int64_t total = 0;
  for (int64_t g : got) total += g;
  CHECK(total == want);
  return true;
}

// TestSerialHashJoinSemantics (exec/join_table_test.go:305-329) as a join
// aggregate: duplicate build keys fan out, key 0 is ordinary, unmatched probe
// keys vanish; groups come out in first-occurrence order 5, 0, 7.
bool SerialHashJoinSemantics() {
  static const int64_t bkey[] = {0, 5, 5, 0, 7, 5};
  static const int64_t pkey[] = {5, 0, 99, 7};
  static const double px[] = {1.5, 2.5, 3.5, 4.5};
  HostTable hb, hp;
  hb.columns = {col("b_key", DataType::Int64, VQ_I64, bkey)};
  hb.row_groups = {6};
  hb.nrows = 6;
  hp.columns = {col("p_key", DataType::Int64, VQ_I64, pkey), col("x", DataType::Float64, VQ_F64, px)};
  hp.row_groups = {4};
  hp.nrows = 4;
  std::unique_ptr<GpuTable> b, p;
  CHECK_OK(GpuTable::Upload(g_dev.get(), hb, &b));
  CHECK_OK(GpuTable::Upload(g_dev.get(), hp, &p));
  auto op = NewGpuParallelHashJoinAggregate(b.get(), 0, nullptr, p.get(), 0, nullptr,
                                            {agg(AggKind::Count, -1, DataType::Int64, "c"),
                                             agg(AggKind::Sum, 1, DataType::Float64, "s")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 3);
  const Vector& k = bs[0].vectors[0];
  const Vector& c = bs[0].vectors[1];
  const Vector& s = bs[0].vectors[2];
  CHECK(k.i64[0] == 5 && c.i64[0] == 3 && s.f64[0] == 4.5);
  CHECK(k.i64[1] == 0 && c.i64[1] == 2 && s.f64[1] == 5.0);
  CHECK(k.i64[2] == 7 && c.i64[2] == 1 && s.f64[2] == 4.5);
  JoinBuildStats st = op->BuildStats();
  CHECK(st.rows == 6 && st.partitions >= 1);
  return true;
}

// TestNewHashJoinShared_ProbeSemantics (exec/parallel_join_test.go:121-157)
// and TestJoinTableDuplicateMultiplicities (exec/join_table_test.go:114-154).
bool SharedProbeAndMultiplicities() {
  std::vector<int64_t> bkey;
  const int mult[3] = {1, 2, 17};
  const int64_t keys[3] = {100, 200, 300};
  for (int i = 0; i < 17; i++)
    for (int k = 0; k < 3; k++)
      if (i < mult[k]) bkey.push_back(keys[k]);
  static const int64_t pkey[] = {100, 200, 300, 999};
  HostTable hb, hp;
  hb.columns = {col("k", DataType::Int64, VQ_I64, bkey.data())};
  hb.row_groups = {(int64_t)bkey.size()};
  hb.nrows = (int64_t)bkey.size();
  hp.columns = {col("k", DataType::Int64, VQ_I64, pkey)};
  hp.row_groups = {4};
  hp.nrows = 4;
  std::unique_ptr<GpuTable> b, p;
  CHECK_OK(GpuTable::Upload(g_dev.get(), hb, &b));
  CHECK_OK(GpuTable::Upload(g_dev.get(), hp, &p));
  for (int bits : {0, 1, 4, 6, 8}) {
    JoinOptions o;
    o.radix_bits = bits;
    auto op = NewGpuParallelHashJoinAggregate(b.get(), 0, nullptr, p.get(), 0, nullptr,
                                              {agg(AggKind::Count, -1, DataType::Int64, "c")}, o);
    std::vector<Batch> bs;
    CHECK_OK(Drain(op.get(), &bs));
    CHECK(bs.size() == 1 && bs[0].length == 3);
    for (int i = 0; i < 3; i++) {
      CHECK(bs[0].vectors[0].i64[i] == keys[i]);
      CHECK(bs[0].vectors[1].i64[i] == mult[i]);
    }
    CHECK(op->BuildStats().partitions == (1 << bits));
  }
  return true;
}

// TestJoinTableSentinelAndExtremeKeys (exec/join_table_test.go:159-183).
bool SentinelAndExtremeKeys() {
  static const int64_t bkey[] = {0, -1, 1, INT64_MIN, INT64_MAX};
  static const int64_t pkey[] = {0, -1, 1, INT64_MIN, INT64_MAX, 2, -2, 999, INT64_MAX - 1};
  HostTable hb, hp;
  hb.columns = {col("k", DataType::Int64, VQ_I64, bkey)};
  hb.row_groups = {5};
  hb.nrows = 5;
  hp.columns = {col("k", DataType::Int64, VQ_I64, pkey)};
  hp.row_groups = {9};
  hp.nrows = 9;
  std::unique_ptr<GpuTable> b, p;
  CHECK_OK(GpuTable::Upload(g_dev.get(), hb, &b));
  CHECK_OK(GpuTable::Upload(g_dev.get(), hp, &p));
  auto op = NewGpuParallelHashJoinAggregate(b.get(), 0, nullptr, p.get(), 0, nullptr,
                                            {agg(AggKind::Count, -1, DataType::Int64, "c")});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 5);
  for (int i = 0; i < 5; i++) {
    CHECK(bs[0].vectors[0].i64[i] == bkey[i]);  // first-occurrence = probe order
    CHECK(bs[0].vectors[1].i64[i] == 1);
  }
  return true;
}

// TestParallelAggregateReuseMatchesRebuildAcrossWorkerCounts closed form
// (exec/morsel_reuse_test.go:410-448): SUM(2*id+1) over id < 1500 = 2,250,000.
bool ClosedFormSum() {
  std::vector<int64_t> id(8 * 256);
  for (size_t i = 0; i < id.size(); i++) id[i] = (int64_t)i;
  HostTable h;
  h.columns = {col("id", DataType::Int64, VQ_I64, id.data())};
  h.row_groups = std::vector<int64_t>(8, 256);
  h.nrows = (int64_t)id.size();
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  vq_pred p = pred1(0, VQ_LT, 1500);
  AggExpr a;
  a.kind = AggKind::Sum;
  a.out_name = "s";
  a.accum_type = DataType::Int64;
  a.expr.nfactors = 2;
  a.expr.is_float = 0;
  a.expr.f[0] = {VQ_F_COL_OP_LIT, 0, VQ_MUL, VQ_ADD, 0, 2};
  a.expr.f[1] = {VQ_F_LIT, 0, VQ_ADD, VQ_ADD, 0, 1};
  auto op = NewGpuParallelHashAggregate(t.get(), &p, {}, {a});
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].vectors[0].i64[0] == 2250000);
  return true;
}

// Errors carry the exec: prefix and a library status; nothing runs.
bool ErrorsAreReported() {
  static const int64_t x[] = {1};
  static const double f[] = {1.5};
  HostTable h;
  h.columns = {col("x", DataType::Int64, VQ_I64, x), col("f", DataType::Float64, VQ_F64, f)};
  h.row_groups = {1};
  h.nrows = 1;
  std::unique_ptr<GpuTable> t;
  CHECK_OK(GpuTable::Upload(g_dev.get(), h, &t));
  // group by a float64 column: unsupported
  auto op = NewGpuParallelHashAggregate(t.get(), nullptr, {1}, {agg(AggKind::Count, -1, DataType::Int64, "c")});
  const Batch* b = nullptr;
  Status s = op->Next(&b);
  CHECK(!s.ok() && s.code == VQ_E_UNSUPPORTED);
  CHECK(s.msg.rfind("exec: gpu", 0) == 0);
  // bad row-group accounting is an argument error
  HostTable bad = h;
  bad.row_groups = {2};
  std::unique_ptr<GpuTable> t2;
  Status s2 = GpuTable::Upload(g_dev.get(), bad, &t2);
  CHECK(!s2.ok() && s2.code == VQ_E_ARG);
  return true;
}

// TPC-H Q12's shape through the operator: GROUP BY a probe-side dictionary
// column over the joined rows, a CASE WHEN indicator sum over a build-side
// column, a duplicated build key fanning out (expected values by hand).
bool JoinGroupedByProbeColumn() {
  static const int64_t okey[] = {10, 20, 30, 40, 20};
  static const uint8_t oprio[] = {0, 1, 2, 0, 2};
  static const int64_t oval[] = {100, 200, 300, 400, 250};
  static const int64_t lkey[] = {10, 20, 20, 30, 50, 40, 10};
  static const uint8_t lmode[] = {0, 1, 0, 1, 0, 0, 1};
  static const int64_t lqty[] = {1, 2, 3, 4, 5, 6, 7};
  HostTable hb, hp;
  hb.columns = {col("o_orderkey", DataType::Int64, VQ_I64, okey),
                col("o_orderpriority", DataType::String, VQ_U8, oprio, dict({"1-URGENT", "2-HIGH", "3-MEDIUM"})),
                col("o_val", DataType::Int64, VQ_I64, oval)};
  hb.row_groups = {5};
  hb.nrows = 5;
  hp.columns = {col("l_orderkey", DataType::Int64, VQ_I64, lkey),
                col("l_shipmode", DataType::String, VQ_U8, lmode, dict({"MAIL", "SHIP"})),
                col("l_quantity", DataType::Int64, VQ_I64, lqty)};
  hp.row_groups = {4, 3};
  hp.nrows = 7;
  std::unique_ptr<GpuTable> b, p;
  CHECK_OK(GpuTable::Upload(g_dev.get(), hb, &b));
  CHECK_OK(GpuTable::Upload(g_dev.get(), hp, &p));
  // SUM(CASE WHEN o_orderpriority IN ('1-URGENT', '2-HIGH') THEN 1 ELSE 0 END)
  AggExpr high = agg(AggKind::Sum, -1, DataType::Int64, "high_line_count");
  high.expr.nfactors = 1;
  high.expr.f[0].kind = VQ_F_LIT;
  high.expr.f[0].ival = 1;
  high.expr.nwhen = 1;
  high.expr.when[0].col = 1 | VQ_BUILD_COL;
  high.expr.when[0].op = VQ_IN;
  high.expr.when[0].ival = 0x3;
  AggExpr mx = agg(AggKind::Max, 2, DataType::Int64, "max_val");
  mx.build_side = true;
  JoinOptions opts;
  opts.group_by = {1};
  auto op = NewGpuParallelHashJoinAggregate(b.get(), 0, nullptr, p.get(), 0, nullptr,
                                            {high, agg(AggKind::Count, -1, DataType::Int64, "c"),
                                             agg(AggKind::Sum, 2, DataType::Int64, "qty"), mx},
                                            opts);
  CHECK(op->schema().fields.size() == 5 && op->schema().fields[0].name == "l_shipmode");
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 2);
  const Batch& r = bs[0];
  CHECK(r.vectors[0].Get(0) == "MAIL" && r.vectors[0].Get(1) == "SHIP");
  CHECK(r.vectors[1].i64[0] == 3 && r.vectors[2].i64[0] == 4 && r.vectors[3].i64[0] == 13 &&
        r.vectors[4].i64[0] == 400);
  CHECK(r.vectors[1].i64[1] == 2 && r.vectors[2].i64[1] == 4 && r.vectors[3].i64[1] == 15 &&
        r.vectors[4].i64[1] == 300);
  return true;
}

// TPC-H Q3's chain through the operator: customers of one segment, their
// orders, those orders' lineitems; GROUP BY the order key with its date and
// ship priority riding along, top 2 by SUM(price) (expected values by hand).
bool NestedBuildChain() {
  static const int64_t ckey[] = {1, 2, 3};
  static const uint8_t cseg[] = {0, 1, 0};
  static const int64_t okey[] = {10, 20, 30, 40};
  static const int64_t ocust[] = {1, 2, 3, 1};
  static const int32_t odate[] = {100, 101, 102, 103};
  static const int64_t oship[] = {0, 0, 1, 0};
  static const int64_t lkey[] = {10, 20, 30, 40, 10, 30, 50};
  static const double lprice[] = {1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0};
  HostTable hc, ho, hl;
  hc.columns = {col("c_custkey", DataType::Int64, VQ_I64, ckey),
                col("c_mktsegment", DataType::String, VQ_U8, cseg, dict({"BUILDING", "MACHINERY"}))};
  hc.row_groups = {3};
  hc.nrows = 3;
  ho.columns = {col("o_orderkey", DataType::Int64, VQ_I64, okey), col("o_custkey", DataType::Int64, VQ_I64, ocust),
                col("o_orderdate", DataType::Date, VQ_DATE, odate),
                col("o_shippriority", DataType::Int64, VQ_I64, oship)};
  ho.row_groups = {4};
  ho.nrows = 4;
  hl.columns = {col("l_orderkey", DataType::Int64, VQ_I64, lkey),
                col("l_extendedprice", DataType::Float64, VQ_F64, lprice)};
  hl.row_groups = {4, 3};
  hl.nrows = 7;
  std::unique_ptr<GpuTable> c, o, l;
  CHECK_OK(GpuTable::Upload(g_dev.get(), hc, &c));
  CHECK_OK(GpuTable::Upload(g_dev.get(), ho, &o));
  CHECK_OK(GpuTable::Upload(g_dev.get(), hl, &l));
  const vq_pred seg = pred1(1, VQ_EQ, 0);
  JoinOptions opts;
  opts.inner = c.get();
  opts.inner_key = 0;
  opts.build_fk = 1;
  opts.inner_pred = &seg;
  opts.build_pay = {2, 3};
  opts.topk = 2;
  auto op = NewGpuParallelHashJoinAggregate(o.get(), 0, nullptr, l.get(), 0, nullptr,
                                            {agg(AggKind::Sum, 1, DataType::Float64, "revenue")}, opts);
  CHECK(op->schema().fields.size() == 4 && op->schema().fields[1].name == "o_orderdate");
  std::vector<Batch> bs;
  CHECK_OK(Drain(op.get(), &bs));
  CHECK(bs.size() == 1 && bs[0].length == 2);
  const Batch& r = bs[0];
  CHECK(r.vectors[0].i64[0] == 30 && r.vectors[1].i64[0] == 102 && r.vectors[2].i64[0] == 1 &&
        r.vectors[3].f64[0] == 9.0);
  CHECK(r.vectors[0].i64[1] == 10 && r.vectors[1].i64[1] == 100 && r.vectors[2].i64[1] == 0 &&
        r.vectors[3].f64[1] == 6.0);
  return true;
}

}  // namespace

int main() {
  Status s = Device::Open(0, 0, &g_dev);
  if (!s.ok()) {
    printf("FAIL open: %s\n", s.msg.c_str());
    return 1;
  }
  run("TestIntKeyMultiRowGroupDictInstability", IntKeyMultiRowGroupDictInstability);
  run("TestNullKeysAndArguments", NullKeysAndArguments);
  run("TestStringMinMax", StringMinMax);
  run("TestCompositeKey", CompositeKey);
  run("TestAccumulateSelVecFourPaths", AccumulateSelVecFourPaths);
  run("TestMergePartialAggNonNullCounts", MergePartialAggNonNullCounts);
  run("TestEmptyGlobalAggregate", EmptyGlobalAggregate);
  run("TestZoneMapPruning", ZoneMapPruning);
  run("TestFilterSelectionVector", FilterSelectionVector);
  run("TestSerialHashJoinSemantics", SerialHashJoinSemantics);
  run("TestSharedProbeAndMultiplicities", SharedProbeAndMultiplicities);
  run("TestJoinTableSentinelAndExtremeKeys", SentinelAndExtremeKeys);
  run("TestJoinGroupedByProbeColumn", JoinGroupedByProbeColumn);
  run("TestNestedBuildChain", NestedBuildChain);
  run("TestParallelAggregateClosedForm", ClosedFormSum);
  run("TestErrorsAreReported", ErrorsAreReported);
  g_dev.reset();
  printf("%d failed\n", g_fail);
  return g_fail;
}
