// vq_jit.hip — shape-specialized kernels compiled at run time with hipRTC.
//
// The query's shape (column types, predicate comparisons, key strides, the
// aggregate folds) is written as compile-time constants in front of
// jit_template.hip and compiled once for gfx950; the module is cached under
// that text, so every later query of the same shape — whatever its literal
// values — launches the cached kernel.  A compile failure is reported to the
// caller, which then runs the generic kernels of k_scan.hip.
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>

#include "jit_template_src.h"  // generated by vexq_amd/build.py: kFixChunksSrc, kJitTemplateSrc
#include "vq_internal.h"

namespace vq {

namespace {

struct Module {
  hipModule_t mod = nullptr;
  hipFunction_t dense = nullptr;
  std::atomic<int> dense_per_cu{0};  // occupancy of `dense` (0: not asked yet)
  hipFunction_t filter_bits = nullptr;
  hipFunction_t part_count = nullptr;
  hipFunction_t part_scatter = nullptr;
  hipFunction_t hash_local = nullptr;
  hipFunction_t dj_probe = nullptr;
  hipFunction_t dj_match = nullptr;
  hipFunction_t multi = nullptr;
  std::atomic<int> multi_per_cu{0};
};

std::mutex g_mu;

bool is_float_op(int op) { return op == ACC_SUM_F || op == ACC_MIN_F || op == ACC_MAX_F; }

void arr(std::string& s, const char* name, const int* v, int n, int size = 8) {
  s += "static constexpr int ";
  s += name;
  s += "[" + std::to_string(size) + "] = {";
  for (int i = 0; i < n; i++) s += (i ? "," : "") + std::to_string(v[i]);
  s += "};\n";
}

void arr2(std::string& s, const char* name, const int (*v)[4]) {
  s += "static constexpr int ";
  s += name;
  s += "[8][4] = {";
  for (int i = 0; i < 8; i++) {
    s += i ? ",{" : "{";
    for (int k = 0; k < 4; k++) s += (k ? "," : "") + std::to_string(v[i][k]);
    s += "}";
  }
  s += "};\n";
}

// Flattened terms (slot, op, ival, fval): the predicate's [0, nt), then each
// accumulator's CASE WHEN terms [aw0[j], aw0[j] + awn[j]).
struct Flat {
  int nt = 0, nall = 0;
  int ts[kJitTerms] = {0}, to[kJitTerms] = {0}, ts2[kJitTerms] = {0};
  int64_t ti[kJitTerms] = {0};
  double tf[kJitTerms] = {0};
  int aw0[8] = {0}, awn[8] = {0};
};

Flat flatten_terms(const QDesc& d) {
  Flat f;
  for (int c = 0; c < d.nslots; c++)
    for (int j = 0; j < d.slot[c].nterms && f.nt < 8; j++) {
      f.ts[f.nt] = c;
      f.to[f.nt] = d.slot[c].terms[j].op;
      f.ts2[f.nt] = d.slot[c].terms[j].slot2;
      f.ti[f.nt] = d.slot[c].terms[j].ival;
      f.tf[f.nt] = d.slot[c].terms[j].fval;
      f.nt++;
    }
  f.nall = f.nt;
  for (int j = 0; j < d.nacc && j < 8; j++) {
    const QExpr& e = d.acc[j].e;
    f.aw0[j] = f.nall;
    f.awn[j] = e.nwhen;
    for (int k = 0; k < e.nwhen; k++) {
      f.ts[f.nall] = e.when[k].slot;
      f.to[f.nall] = e.when[k].op;
      f.ts2[f.nall] = e.when[k].slot2;
      f.ti[f.nall] = e.when[k].ival;
      f.tf[f.nall] = e.when[k].fval;
      f.nall++;
    }
  }
  return f;
}

int rows_in_flight(int nslots) { return nslots <= 3 ? 8 : 4; }

std::string prologue(const QDesc& d, const Flat& f, int kernels = 3, int key_slot = 0,
                     int nv = 0, int probe = 0, int pbuf = 0, int bloom = 0,
                     const int* lop = nullptr, const int* lsrc = nullptr, int nag = 0,
                     int livebits = 0, const JSide* keys = nullptr) {
  int st[8] = {0}, ks[8] = {0}, aop[8] = {0}, aisf[8] = {0}, anf[8] = {0};
  int sn[8] = {0}, nc[8] = {0}, acn[8] = {0};
  int nulls = 0;
  int fk[8][4] = {{0}}, fs[8][4] = {{0}}, fo[8][4] = {{0}}, fc[8][4] = {{0}};
  for (int c = 0; c < d.nslots; c++) {
    st[c] = d.slot[c].type;
    ks[c] = d.slot[c].key_stride;
    sn[c] = d.slot[c].nul ? 1 : 0;
    nc[c] = d.slot[c].null_code;
    nulls |= sn[c];
  }
  for (int j = 0; j < d.nacc; j++) {
    acn[j] = d.acc[j].count_valid;
    aop[j] = d.acc[j].op;
    aisf[j] = is_float_op(d.acc[j].op);
    anf[j] = d.acc[j].e.nfactors;
    for (int k = 0; k < d.acc[j].e.nfactors; k++) {
      fk[j][k] = d.acc[j].e.f[k].kind;
      fs[j][k] = d.acc[j].e.f[k].slot;
      fo[j][k] = d.acc[j].e.f[k].op;
      fc[j][k] = d.acc[j].e.f[k].combine;
    }
  }
  std::string s = "#define VQ_JIT_PROLOGUE 1\n";
  s += "#define VQ_NS " + std::to_string(d.nslots) + "\n";
  s += "#define VQ_U " + std::to_string(rows_in_flight(d.nslots)) + "\n";
  s += "#define VQ_G " + std::to_string(d.ngroups) + "\n";
  s += "#define VQ_NA " + std::to_string(d.nacc) + "\n";
  s += "#define VQ_NT " + std::to_string(f.nt) + "\n";
  s += "#define VQ_KERNELS " + std::to_string(kernels) + "\n";
  s += "#define VQ_KEYSLOT " + std::to_string(key_slot) + "\n";
  s += "#define VQ_NV " + std::to_string(nv) + "\n";
  s += "#define VQ_PROBE " + std::to_string(probe) + "\n";
  s += "#define VQ_PB " + std::to_string(pbuf) + "\n";
  s += "#define VQ_BLOOM " + std::to_string(bloom) + "\n";
  s += "#define VQ_LIVEBITS " + std::to_string(livebits) + "\n";
  s += "#define VQ_DJ_SLOTS " + std::to_string(kDjSlots) + "\n";
  s += "#define VQ_NULLS " + std::to_string(nulls) + "\n";
  // the key: one column (key_slot), or several dictionary codes packed
  int nkey = 1, ksl[8] = {key_slot}, ksh[8] = {0};
  if (keys && keys->nkeys > 1) {
    nkey = keys->nkeys;
    for (int k = 0; k < nkey; k++) {
      ksl[k] = keys->key_slots[k];
      ksh[k] = keys->key_shift[k];
    }
  }
  s += "#define VQ_NKEY " + std::to_string(nkey) + "\n";
  // a join + GROUP BY of the joined rows: the probe's join key slot and the
  // build-side slots
  int sb[8] = {0};
  const bool lookup = keys && keys->lookup;
  for (int c = 0; lookup && c < 8; c++) sb[c] = (keys->build_slots >> c) & 1;
  s += "#define VQ_LOOKUP " + std::to_string(lookup ? 1 : 0) + "\n";
  s += "#define VQ_JKEY " + std::to_string(lookup ? keys->jkey_slot : 0) + "\n";
  arr(s, "kSB", sb, 8);
  arr(s, "kSN", sn, 8);
  arr(s, "kNC", nc, 8);
  arr(s, "kACN", acn, 8);
  arr(s, "kKSL", ksl, 8);
  arr(s, "kKSH", ksh, 8);
  arr(s, "kST", st, 8);
  s += "#define VQ_TERMS " + std::to_string(kJitTerms) + "\n";
  arr(s, "kTS", f.ts, f.nall, kJitTerms);
  arr(s, "kTO", f.to, f.nall, kJitTerms);
  arr(s, "kTS2", f.ts2, f.nall, kJitTerms);
  arr(s, "kAW0", f.aw0, 8);
  arr(s, "kAWN", f.awn, 8);
  arr(s, "kKS", ks, 8);
  arr(s, "kAOP", aop, 8);
  arr(s, "kAISF", aisf, 8);
  arr(s, "kANF", anf, 8);
  arr2(s, "kFK", fk);
  arr2(s, "kFS", fs);
  arr2(s, "kFO", fo);
  arr2(s, "kFC", fc);
  if (kernels & 24) {
    s += "#define VQ_NAG " + std::to_string(nag) + "\n";
    arr(s, "kLOP", lop, nag);
    arr(s, "kLSRC", lsrc, nag);
  }
  return s;
}

int fail_jit(std::string* err, const std::string& msg) {
  if (err) *err = msg;
  return -1;
}

// Compiled modules by prologue text.  A module is compiled outside the lock:
// a thread that needs a module another thread is compiling waits for that
// one only (its shared future), and queries of shapes already compiled never
// wait behind a compile.
struct Entry {
  std::shared_future<Module*> mod;
  std::string err;  // a failed compile's message (set before the future)
};
std::unordered_map<std::string, std::shared_ptr<Entry>> g_cache;

// FNV-1a, 64 bits: the on-disk cache's file name for one source text.
uint64_t fnv64(const std::string& x) {
  uint64_t h = 0xcbf29ce484222325ULL;
  for (unsigned char ch : x) {
    h ^= ch;
    h *= 0x100000001b3ULL;
  }
  return h;
}

// VQ_JIT_CACHE_DIR=<dir>: code objects also kept on disk, by a hash of the
// whole source text and options, so a new process (a restarted server) loads
// a shape's kernel instead of compiling it again.
std::string disk_path(const std::string& src, const char* const* opts, int nopts) {
  const char* dir = getenv("VQ_JIT_CACHE_DIR");
  if (!dir || !*dir) return "";
  std::string key = src;
  for (int i = 0; i < nopts; i++) key += std::string("\n") + opts[i];
  char name[40];
  snprintf(name, sizeof name, "/vq_%016llx.co", (unsigned long long)fnv64(key));
  return std::string(dir) + name;
}

bool read_file(const std::string& path, std::string& out) {
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) return false;
  fseek(fp, 0, SEEK_END);
  const long n = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  bool ok = n > 0;
  if (ok) {
    out.resize((size_t)n);
    ok = fread(&out[0], 1, (size_t)n, fp) == (size_t)n;
  }
  fclose(fp);
  return ok;
}

void write_file(const std::string& path, const std::string& data) {
  const std::string tmp = path + ".tmp" + std::to_string((unsigned long long)(uintptr_t)&data);
  FILE* fp = fopen(tmp.c_str(), "wb");
  if (!fp) return;
  const bool ok = fwrite(data.data(), 1, data.size(), fp) == data.size();
  fclose(fp);
  if (!ok || rename(tmp.c_str(), path.c_str()) != 0) remove(tmp.c_str());
}

// Compiles the module of one prologue (or reads its code from the disk
// cache) and resolves its kernels.
Module* compile_module(const std::string& pro, std::string* err) {
  const std::string src = pro + kFixChunksSrc + kJitTemplateSrc;
  // VQ_JIT_DUMP=<dir>: each module's prologue (its shape) written there, so
  // the same kernel can be compiled and inspected offline
  if (const char* dir = getenv("VQ_JIT_DUMP")) {
    const std::string path = std::string(dir) + "/module" + std::to_string(fnv64(pro)) + ".pro";
    if (FILE* fp = fopen(path.c_str(), "w")) {
      fwrite(pro.data(), 1, pro.size(), fp);
      fclose(fp);
    }
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                        "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  const int nopts = (int)(sizeof(opts) / sizeof(opts[0]));
  const std::string path = disk_path(src, opts, nopts);
  std::string code;
  if (path.empty() || !read_file(path, code)) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "vq_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
      if (err) *err = "hiprtcCreateProgram failed";
      return nullptr;
    }
    hiprtcResult rc = hiprtcCompileProgram(prog, nopts, opts);
    if (rc != HIPRTC_SUCCESS) {
      size_t ls = 0;
      hiprtcGetProgramLogSize(prog, &ls);
      std::string log(ls, '\0');
      if (ls) hiprtcGetProgramLog(prog, &log[0]);
      hiprtcDestroyProgram(&prog);
      if (err) *err = "hipRTC compile failed: " + log.substr(0, 400);
      return nullptr;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.assign(cs, '\0');
    hiprtcGetCode(prog, &code[0]);
    hiprtcDestroyProgram(&prog);
    if (!path.empty()) write_file(path, code);
  }
  Module* m = new Module();
  bool ok = hipModuleLoadData(&m->mod, code.data()) == hipSuccess;
  if (ok && pro.find("#define VQ_KERNELS 3\n") != std::string::npos)
    ok = hipModuleGetFunction(&m->dense, m->mod, "vq_dense") == hipSuccess &&
         hipModuleGetFunction(&m->filter_bits, m->mod, "vq_filter_bits") == hipSuccess;
  if (ok && pro.find("#define VQ_KERNELS 4\n") != std::string::npos)
    ok = hipModuleGetFunction(&m->part_count, m->mod, "vq_part_count") == hipSuccess &&
         hipModuleGetFunction(&m->part_scatter, m->mod, "vq_part_scatter") == hipSuccess;
  if (ok && pro.find("#define VQ_KERNELS 8\n") != std::string::npos)
    ok = hipModuleGetFunction(&m->hash_local, m->mod, "vq_hash_local") == hipSuccess;
  if (ok && pro.find("#define VQ_KERNELS 32\n") != std::string::npos)
    ok = hipModuleGetFunction(&m->multi, m->mod, "vq_dense_multi") == hipSuccess;
  if (ok && pro.find("#define VQ_KERNELS 16\n") != std::string::npos)
    ok = hipModuleGetFunction(&m->dj_probe, m->mod, "vq_dj_probe") == hipSuccess &&
         hipModuleGetFunction(&m->dj_match, m->mod, "vq_dj_match") == hipSuccess;
  if (!ok) {
    delete m;
    if (err) *err = "hipModuleLoadData failed";
    return nullptr;
  }
  return m;
}

// Compiles (or finds) the module of this shape.
Module* get_module(const std::string& pro, std::string* err) {
  std::shared_ptr<Entry> e;
  std::promise<Module*> prom;
  bool mine = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(pro);
    if (it != g_cache.end()) {
      e = it->second;
    } else {
      e = std::make_shared<Entry>();
      e->mod = prom.get_future().share();
      g_cache.emplace(pro, e);
      mine = true;
    }
  }
  if (mine) {
    std::string cerr;
    Module* m = compile_module(pro, &cerr);
    e->err = cerr;
    prom.set_value(m);
  }
  Module* m = e->mod.get();
  if (!m && err) *err = e->err.empty() ? "run-time compilation failed" : e->err;
  return m;
}

void fill_common(const QDesc& d, const Flat& f, JitArgs& a) {
  memset(&a, 0, sizeof a);
  for (int c = 0; c < d.nslots; c++) {
    a.col[c] = d.slot[c].p;
    a.nul[c] = d.slot[c].nul;
  }
  for (int t = 0; t < f.nall; t++) {
    a.ti[t] = f.ti[t];
    a.tf[t] = f.tf[t];
  }
  for (int j = 0; j < d.nacc; j++) {
    for (int k = 0; k < d.acc[j].e.nfactors; k++) {
      a.fi[j][k] = d.acc[j].e.f[k].ival;
      a.ff[j][k] = d.acc[j].e.f[k].fval;
    }
    a.ei[j] = d.acc[j].e.else_i;
    a.ef[j] = d.acc[j].e.else_f;
  }
  for (int s = 0; s < d.nspans; s++) {
    a.sp_start[s] = d.span[s].start;
    a.sp_rows[s] = d.span[s].rows;
    a.sp_tile0[s] = d.span[s].tile0;
  }
  a.nspans = d.nspans;
  a.tile_rows = d.tile_rows;
  a.ntiles = d.ntiles;
}

hipError_t launch(hipFunction_t fn, unsigned grid, size_t lds, hipStream_t s, JitArgs& a,
                  unsigned block = 64) {
  size_t sz = sizeof a;
  void* cfg[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, &a, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                 HIP_LAUNCH_PARAM_END};
  return hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)lds, s, nullptr, cfg);
}

// A join side as a QDesc: its slots and predicate, and (probe side) the
// aggregate arguments as value expressions.
QDesc side_desc(const JSide& sd, const JDesc* jd, int* nv) {
  QDesc d;
  memset(&d, 0, sizeof d);
  d.nslots = sd.nslots;
  for (int c = 0; c < sd.nslots; c++) d.slot[c] = sd.slot[c];
  d.ngroups = 1;
  *nv = 0;
  if (jd)
    for (int j = 0; j < jd->naggs; j++) {
      if (jd->agg[j].vidx != *nv) continue;  // COUNT(*), or an argument already listed
      d.acc[*nv].op = jd->agg[j].e.is_float ? ACC_SUM_F : ACC_SUM_I;
      d.acc[*nv].e = jd->agg[j].e;
      (*nv)++;
    }
  d.nacc = *nv;
  d.nspans = sd.nspans;
  for (int i = 0; i < sd.nspans; i++) d.span[i] = sd.span[i];
  d.tile_rows = sd.tile_rows;
  d.ntiles = sd.ntiles;
  return d;
}

bool jit_disabled() {
  const char* e = getenv("VQ_NO_JIT");
  return e && *e && *e != '0';
}

}  // namespace
int64_t part_chunk_rows(int nslots) { return (int64_t)1024 * rows_in_flight(nslots); }
int64_t dense_chunk_rows(int nslots) { return (int64_t)64 * rows_in_flight(nslots); }

int device_cus() {
  // (initialised once, by whichever thread asks first)
  static const int cus = [] {
    int dev = 0, n = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return n > 0 ? n : 256;
  }();
  return cus;
}

int jit_part_prepare(const JSide& sd, const JDesc* jd, int32_t P, int32_t pbuf, PartLaunch* out,
                     std::string* err, bool bloom, bool livebits) {
  if (jit_disabled()) return fail_jit(err, "JIT disabled");
  int nv = 0;
  const QDesc d = side_desc(sd, jd, &nv);
  const Flat f = flatten_terms(d);
  Module* m = get_module(prologue(d, f, 4, sd.key_slot, nv, jd != nullptr, pbuf, bloom ? 1 : 0,
                                  nullptr, nullptr, 0, livebits ? 1 : 0, &sd),
                         err);
  if (!m) return -1;
  out->count = m->part_count;
  out->scatter = m->part_scatter;
  const int32_t entry = jd ? 8 + 4 + 8 * nv : 8;
  out->count_lds = P <= 16384 ? (uint32_t)P * 4 : 0;
  out->scatter_lds = part_scatter_lds(P, pbuf, entry);
  const int cus = device_cus();
  int per_cu = 0;
  hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->part_scatter, 1024,
                                                     out->scatter_lds);
  if (per_cu <= 0) per_cu = 1;
  const int64_t g = (int64_t)cus * per_cu;
  out->grid = (int32_t)std::max<int64_t>(1, d.ntiles < g ? d.ntiles : g);
  fill_common(d, f, out->args);
  out->args.P = P;
  return 0;
}

hipError_t jit_part_launch(const PartLaunch& pl, const PartBufs& pb, int32_t P, hipStream_t s,
                           bool count) {
  JitArgs a = pl.args;
  a.counts = pb.counts;
  a.pstart = pb.pstart;
  a.perr = pb.perr;
  a.keys = pb.keys;
  a.rows = pb.rows;
  a.vals = pb.vals;
  a.vstride = pb.vstride;
  a.P = P;
  a.bloom = pb.bloom;
  a.bloom_lpp = pb.bloom_lpp;
  a.bloom_pmask = pb.bloom_pmask;
  a.livebits = pb.livebits;
  const uint32_t lds = count ? pl.count_lds : pl.scatter_lds;
  a.lds_words = lds / 4;
  return launch(count ? pl.count : pl.scatter, (unsigned)pl.grid, lds, s, a, 1024);
}

int jit_local_prepare(const JSide& sd, const JDesc& jd, int32_t lcap, LocalLaunch* out,
                      std::string* err) {
  if (jit_disabled()) return fail_jit(err, "JIT disabled");
  int nv = 0;
  const QDesc d = side_desc(sd, &jd, &nv);
  const Flat f = flatten_terms(d);
  // each aggregate with an argument: its fold and its value index
  int lop[8] = {0}, lsrc[8] = {0}, nag = 0;
  for (int j = 0; j < jd.naggs && nag < 8; j++) {
    const JAgg& g = jd.agg[j];
    if (g.vidx < 0) continue;
    if (g.kind == VQ_AGG_SUM || g.kind == VQ_AGG_AVG) lop[nag] = g.e.is_float ? 0 : 1;
    else if (g.kind == VQ_AGG_MIN) lop[nag] = g.accum == VQ_F64 ? 2 : 4;
    else lop[nag] = g.accum == VQ_F64 ? 3 : 5;
    lsrc[nag] = g.vidx;
    nag++;
  }
  Module* m = get_module(prologue(d, f, 8, sd.key_slot, nv, 1, 0, 0, lop, lsrc, nag, 0, &sd), err);
  if (!m) return -1;
  out->fn = m->hash_local;
  out->lds = kLocalLds;
  const int cus = device_cus();
  int per_cu = 0;
  hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->hash_local, 512, out->lds);
  if (per_cu <= 0) per_cu = 1;
  const int64_t g = (int64_t)cus * per_cu;
  out->grid = (int32_t)std::max<int64_t>(1, d.ntiles < g ? d.ntiles : g);
  fill_common(d, f, out->args);
  JitArgs& a = out->args;
  a.lcap = lcap;
  a.lnw = jd.nwords;
  int v = 0;
  for (int j = 0; j < jd.naggs; j++) {
    const JAgg& g = jd.agg[j];
    if (g.vidx < 0) continue;
    int op;
    if (g.kind == VQ_AGG_SUM || g.kind == VQ_AGG_AVG) op = g.e.is_float ? 0 : 1;
    else if (g.kind == VQ_AGG_MIN) op = g.accum == VQ_F64 ? 2 : 4;
    else op = g.accum == VQ_F64 ? 3 : 5;
    a.vop[v] = op;
    a.vword[v] = g.word0;
    a.vlsb[v] = g.fs.lsb;
    a.vw[v] = g.fs.w;
    a.vsrc[v] = g.vidx;
    v++;
  }
  a.nvop = v;
  return 0;
}

// The aggregates' folds and words as the run-time kernels take them (every
// aggregate with an argument: its fold, first word, fixed-point scale and
// value index); returns their count.
int jit_value_ops(const JDesc& jd, int* lop, int* lsrc, JitArgs* a) {
  int v = 0;
  for (int j = 0; j < jd.naggs && v < 8; j++) {
    const JAgg& g = jd.agg[j];
    if (g.vidx < 0) continue;
    int op;
    if (g.kind == VQ_AGG_SUM || g.kind == VQ_AGG_AVG) op = g.e.is_float ? 0 : 1;
    else if (g.kind == VQ_AGG_MIN) op = g.accum == VQ_F64 ? 2 : 4;
    else op = g.accum == VQ_F64 ? 3 : 5;
    lop[v] = op;
    lsrc[v] = g.vidx;
    if (a) {
      a->vop[v] = op;
      a->vword[v] = g.word0;
      a->vlsb[v] = g.fs.lsb;
      a->vw[v] = g.fs.w;
      a->vsrc[v] = g.vidx;
    }
    v++;
  }
  if (a) a->nvop = v;
  return v;
}

int jit_direct_prepare(const JSide& sd, const JDesc& jd, DirectLaunch* out, std::string* err) {
  if (jit_disabled()) return fail_jit(err, "JIT disabled");
  int nv = 0;
  const QDesc d = side_desc(sd, &jd, &nv);
  const Flat f = flatten_terms(d);
  int lop[8] = {0}, lsrc[8] = {0};
  const int nag = jit_value_ops(jd, lop, lsrc, nullptr);
  Module* m = get_module(prologue(d, f, 16, sd.key_slot, nv, 1, 0, 0, lop, lsrc, nag, 0, &sd), err);
  if (!m) return -1;
  out->fn = m->dj_probe;
  out->match = m->dj_match;
  const int cus = device_cus();
  int per_cu = 0;
  hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->dj_probe, 1024, 0);
  if (per_cu <= 0) per_cu = 1;
  const int64_t g = (int64_t)cus * per_cu;
  out->grid = (int32_t)std::max<int64_t>(1, d.ntiles < g ? d.ntiles : g);
  fill_common(d, f, out->args);
  jit_value_ops(jd, lop, lsrc, &out->args);
  return 0;
}

hipError_t jit_direct_launch(const DirectLaunch& dl, const DirectBufs& D, hipStream_t s) {
  JitArgs a = dl.args;
  a.dj_pres = D.pres;
  a.dj_rec = D.rec;
  a.dj_mrow = D.mrow;
  a.dj_mcount = D.mcount;
  a.dj_ent = D.ent;
  a.dj_tlist = D.tlist;
  a.dj_tcount = D.tcount;
  a.dj_base = D.base;
  a.dj_range = D.range;
  a.dj_ew = D.ew;
  hipError_t e = launch(dl.fn, (unsigned)dl.grid, 0, s, a, 1024);
  if (e != hipSuccess) return e;
  return launch(dl.match, (unsigned)dl.grid, 0, s, a, 256);
}

hipError_t jit_local_launch(const LocalLaunch& ll, const JoinBufs& B, hipStream_t s, int32_t k0) {
  JitArgs a = ll.args;
  a.lk0 = k0;
  a.lkeys = B.l.keys;
  a.lflags = B.flags;
  a.lgkey = B.l.gkey;
  a.part = B.l.part;
  a.lk_keys = B.lk.keys;
  a.lk_head = B.lk.head;
  a.lk_next = B.lk.next;
  a.lk_min = B.lk.min_head;
  a.lk_mask = B.lk.mask;
  a.lk_packed = B.lk.packed;
  a.lk_base = B.lk.base;
  a.lk_range = B.lk.range;
  a.lds_words = ll.lds / 4;
  return launch(ll.fn, (unsigned)ll.grid, ll.lds, s, a, 512);
}

int jit_dense_prepare(const QDesc& d, DenseJit* out, std::string* err) {
  if (jit_disabled()) return fail_jit(err, "JIT disabled");
  const Flat f = flatten_terms(d);
  Module* m = get_module(prologue(d, f), err);
  if (!m) return -1;
  const size_t lds = dense_lds_bytes(d);
  // the waves per CU the kernel's occupancy allows (once per module: its
  // shape fixes its LDS)
  int occ = m->dense_per_cu.load();
  if (occ <= 0) {
    hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, m->dense, 64, lds);
    if (occ <= 0) occ = 1;
    m->dense_per_cu.store(occ);
  }
  // a large query: one wave per SIMD; the prefetching waves already keep
  // each SIMD's loads in flight, more add per-wave set-up and partials
  // (measured on the 6M- and 60M-row GROUP BYs: 4 waves a CU 36.3 / 376 us,
  // 3 or 5 38.7 / 380+, 8 38.7 / 384)
  int per_cu = std::min(occ, 4);
  if (const char* e = getenv("VQ_WAVES_PER_CU")) {
    const int v = atoi(e);
    if (v > 0) per_cu = v;
  }
  out->fn = m->dense;
  out->lds = lds;
  out->slots = device_cus() * per_cu;
  out->slots_max = device_cus() * std::max(occ, per_cu);
  return 0;
}

hipError_t jit_dense_launch(const DenseJit& h, const QDesc& d, int32_t grid, uint64_t* part,
                            hipStream_t s) {
  const Flat f = flatten_terms(d);
  JitArgs a;
  fill_common(d, f, a);
  a.part = part;
  return launch(h.fn, (unsigned)grid, h.lds, s, a);
}

bool jit_available_filter(const QDesc& d, std::string* err) {
  if (jit_disabled()) return false;
  const Flat f = flatten_terms(d);
  return get_module(prologue(d, f), err) != nullptr;
}

int jit_term_literals(const QDesc& d, int64_t* lit) {
  const Flat f = flatten_terms(d);
  for (int t = 0; t < f.nt; t++) {
    const int op = f.to[t];
    const bool fl = op < 16 && op != VQ_IN && d.slot[f.ts[t]].type == VQ_F64;
    lit[t] = op >= 16 ? 0 : fl ? f64_to_bits(f.tf[t]) : f.ti[t];
  }
  return f.nt;
}

int jit_multi_prepare(const QDesc& d, int32_t nq, DenseJit* out, std::string* err) {
  if (jit_disabled()) return fail_jit(err, "JIT disabled");
  const Flat f = flatten_terms(d);
  const std::string pro = prologue(d, f, 32) + "#define VQ_NQ " + std::to_string(nq) + "\n";
  Module* m = get_module(pro, err);
  if (!m) return -1;
  int per_cu = m->multi_per_cu.load();
  if (per_cu <= 0) {
    hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, m->multi, 64, 0);
    if (per_cu <= 0) per_cu = 1;
    m->multi_per_cu.store(per_cu);
  }
  out->fn = m->multi;
  out->lds = 0;
  out->slots = device_cus() * per_cu;
  return 0;
}

hipError_t jit_multi_launch(const DenseJit& h, const QDesc& d, int32_t grid, uint64_t* part,
                            const int64_t* mlit, hipStream_t s) {
  const Flat f = flatten_terms(d);
  JitArgs a;
  fill_common(d, f, a);
  a.part = part;
  a.mlit = mlit;
  return launch(h.fn, (unsigned)grid, 0, s, a);
}

hipError_t jit_launch_filter_bits(const QDesc& d, uint64_t* bits, uint32_t* tile_counts,
                                  hipStream_t s) {
  const Flat f = flatten_terms(d);
  Module* m = get_module(prologue(d, f), nullptr);
  if (!m) return hipErrorInvalidDeviceFunction;
  JitArgs a;
  fill_common(d, f, a);
  a.bits = bits;
  a.tile_counts = tile_counts;
  return launch(m->filter_bits, (unsigned)d.ntiles, 0, s, a);
}

}  // namespace vq
