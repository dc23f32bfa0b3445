// vq_ingest.hip — a table straight from a .vxq file image: the one bulk
// upload of the raw row-group sections, then block checks and decode on the
// device.
//
// Replaces, for the columns a query needs, storage.Reader's footer parse
// (storage/reader.go:295-408), ColumnReader.NextBlock's per-block CRC check
// (storage/rowgroup.go:101-151), TableScan's decode into vectors
// (exec/scan.go:254-409) and the int-key path's per-row-group dictionary
// remap to global codes in first-seen order (getOrBuildRemap,
// exec/agg_intkey.go:118-134), narrowed to u8/u16 codes.
//
// Host: the footer and the dictionaries (a few bytes per row group and
// column) are parsed and checked here; the per-row-group remap tables are
// built from them.  Device: k_vxq_check computes every block's CRC-32 (one
// wave per block: each lane the CRC of its own stretch of the block with
// slice-by-8 tables in LDS, the stretches joined by GF(2) shifts) and counts
// its valid rows; k_vxq_decode (one workgroup per block) writes the block's
// values into the resident column, codes through their row group's remap
// table.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "vq_internal.h"

namespace vq {
namespace {

// storage/format.go:8-26
constexpr int kBlockRows = 1024;
constexpr int kNullBitmapBytes = 128;
constexpr int kTInt64 = 0x01, kTFloat64 = 0x02, kTBool = 0x03, kTString = 0x04, kTDate = 0x05;
constexpr uint64_t kTrailer = 16;  // [4 B CRC][8 B footer length][4 B magic]

int vfail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int vfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_error(code, buf);
}

// ---- CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320; the table
// hash/crc32 builds for crc32.IEEE, internal/encoding/crc.go:9): slice-by-8
// tables, T[0] the byte table, T[k][i] = (T[k-1][i] >> 8) ^ T[0][T[k-1][i] & 0xFF]
struct CrcTables {
  uint32_t t[8][256];
  CrcTables() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
      t[0][i] = c;
    }
    for (int k = 1; k < 8; k++)
      for (int i = 0; i < 256; i++) t[k][i] = (t[k - 1][i] >> 8) ^ t[0][t[k - 1][i] & 0xFF];
  }
};
const CrcTables& crc_tables() {
  static const CrcTables tabs;
  return tabs;
}

uint32_t crc32_host(const uint8_t* p, size_t n) {
  const auto& T = crc_tables().t;
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = T[0][(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

// CRC combination (the polynomial arithmetic behind joining CRCs of pieces;
// bit 31 is x^0, as the reflected register holds it):
//   multmodp(a, b) = a(x) b(x) mod P(x);  x8n(n) = x^(8n) mod P(x).
// The register after bytes M from a start value r is f(0, M) ^ r x^(8|M|),
// so a block's CRC is the XOR over its pieces M_i of f(0, M_i) shifted by
// the bytes after M_i, plus 0xFFFFFFFF shifted by the whole length, then
// 0xFFFFFFFF.
__host__ __device__ inline uint32_t multmodp(uint32_t a, uint32_t b) {
  uint32_t p = 0;
  for (int i = 0; i < 32; i++) {
    if (a & (0x80000000u >> i)) p ^= b;
    b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
  }
  return p;
}

uint32_t x8n(uint64_t n) {
  static const auto pow2 = [] {
    std::vector<uint32_t> t(32);
    t[0] = 0x40000000u;  // x^1
    for (int k = 1; k < 32; k++) t[k] = multmodp(t[k - 1], t[k - 1]);  // x^(2^k)
    return t;
  }();
  uint32_t p = 0x80000000u;  // x^0
  for (int k = 3; n; n >>= 1, k++)
    if (n & 1) p = multmodp(pow2[k & 31], p);
  return p;
}

// Bytes of one lane's stretch of an n-byte block: a multiple of 8, the
// block spread over the 64 lanes of a wave.
__host__ __device__ inline uint32_t stretch_bytes(uint32_t n) { return 8 * ((n + 511) / 512); }

uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | p[1] << 8); }
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | (uint64_t)rd32(p + 4) << 32; }

struct ColMeta {
  int64_t sec_off, sec_len, null_count, sum;
  uint64_t min, max;
  bool has_minmax;
  uint32_t dict_off, dict_len;
};
struct RgMeta {
  int64_t file_off;
  uint32_t rows;
  std::vector<ColMeta> cols;
};
struct Field {
  std::string name;
  int type, enc, flags;
};

// readFooter + parseFooter (storage/reader.go:295-408).
int parse_footer(const uint8_t* img, uint64_t n, std::vector<Field>& fields, std::vector<RgMeta>& rgs) {
  if (n < kTrailer + 4 || memcmp(img, "VXQ1", 4) != 0 || memcmp(img + n - 4, "VXQ1", 4) != 0)
    return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: bad magic");
  const uint32_t crc = rd32(img + n - kTrailer);
  const int64_t blen = (int64_t)rd64(img + n - kTrailer + 4);
  if (blen < 0 || (uint64_t)blen > n - kTrailer - 4) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: corrupt footer");
  const uint8_t* b = img + n - kTrailer - blen;
  const uint8_t* e = b + blen;
  if (crc32_host(b, (size_t)blen) != crc) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer checksum");
  auto need = [&](int64_t k) { return e - b >= k; };
  if (!need(4)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse schema");
  const uint32_t ncols = rd32(b);
  b += 4;
  fields.clear();
  for (uint32_t i = 0; i < ncols; i++) {
    if (!need(2)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse schema");
    const uint16_t nl = rd16(b);
    b += 2;
    if (!need((int64_t)nl + 3)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse schema");
    Field f;
    f.name.assign((const char*)b, nl);
    b += nl;
    f.type = b[0];
    f.enc = b[1];
    f.flags = b[2];
    b += 3;
    fields.push_back(f);
  }
  if (!need(4)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse row groups");
  const uint32_t nrg = rd32(b);
  b += 4;
  constexpr int64_t kColMeta = 8 + 8 + 8 + 8 + 8 + 8 + 1 + 4 + 4;  // 57 bytes
  rgs.clear();
  for (uint32_t g = 0; g < nrg; g++) {
    if (!need(12)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse row group");
    RgMeta r;
    r.file_off = (int64_t)rd64(b);
    r.rows = rd32(b + 8);
    b += 12;
    for (uint32_t c = 0; c < ncols; c++) {
      if (!need(kColMeta)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: footer: parse column meta");
      ColMeta m;
      m.sec_off = (int64_t)rd64(b);
      m.sec_len = (int64_t)rd64(b + 8);
      m.null_count = (int64_t)rd64(b + 16);
      m.sum = (int64_t)rd64(b + 24);
      m.min = rd64(b + 32);
      m.max = rd64(b + 40);
      m.has_minmax = b[48] != 0;
      m.dict_off = rd32(b + 49);
      m.dict_len = rd32(b + 53);
      b += kColMeta;
      r.cols.push_back(m);
    }
    rgs.push_back(std::move(r));
  }
  return VQ_OK;
}

// UnmarshalDictionary (storage/codec.go:216-247).
int parse_dict(const uint8_t* p, uint64_t n, std::vector<std::string>& out) {
  if (n < 8 || crc32_host(p, n - 4) != rd32(p + n - 4)) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: dictionary checksum");
  const uint64_t body = n - 4;
  const uint32_t cnt = rd32(p);
  if (body < 8 + 4ull * cnt) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: dictionary size");
  const uint32_t tl = rd32(p + 4 + 4ull * cnt);
  const uint8_t* data = p + 8 + 4ull * cnt;
  if (body < 8 + 4ull * cnt + tl) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: dictionary data");
  out.resize(cnt);
  for (uint32_t i = 0; i < cnt; i++) {
    const uint32_t a = rd32(p + 4 + 4ull * i), z = i + 1 < cnt ? rd32(p + 8 + 4ull * i) : tl;
    if (a > z || z > tl) return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: dictionary offsets");
    out[i].assign((const char*)data + a, z - a);
  }
  return VQ_OK;
}

// ---- device side --------------------------------------------------------------
struct VxqBlock {
  uint64_t src;       // byte offset of the block in the image
  int64_t dst_row;    // its first row in the table
  uint32_t rows;
  uint32_t col;       // table column
  uint32_t remap;     // STRING: first entry of its row group's remap table
  uint32_t ncodes;    // STRING: codes in its row group's dictionary
  uint32_t kidx;      // its row of CRC shift constants (one per block length)
  uint32_t pad_;
};

struct VxqCol {
  void* dst;
  int32_t type;  // VQ_* of the resident column
  int32_t vs;    // bytes per value in the file
  int32_t is_string;
  int32_t pad_;
};

struct VxqArgs {
  const uint8_t* img;
  const VxqBlock* blocks;
  const VxqCol* cols;
  const uint32_t* remap;
  const uint32_t* crc;  // the slice-by-8 tables, [8][256]
  // per block length n: [64] lane l's stretch shifted past the bytes after
  // it, x^(8 (n - end_l)); [64] the start value's shift, x^(8 n)
  const uint32_t* kconst;
  uint32_t* flags;      // [0] bad checksums, [1] first bad block, [2] NULL rows, [3] bad codes
  int64_t nblocks;
};

__device__ __forceinline__ uint32_t ld8(const uint8_t* p) { return p[0]; }
// 4 bytes at any byte address (sections after a dictionary are unaligned)
__device__ __forceinline__ uint32_t ldu32(const uint8_t* p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// Block checks, one wave per block: the CRC-32 of [bitmap][values] against
// the 4 bytes after them (VerifyTrailing, internal/encoding/crc.go:31-40),
// and the valid rows of the bitmap (bit i of byte i/8, LSB first;
// storage/codec.go:9-25).  Lane l takes bytes [l s, l s + s) of the block
// (s = stretch_bytes), 8 at a time through the slice-by-8 tables; the lanes'
// registers are shifted past the bytes after their stretches and XORed.
constexpr int kCheckThreads = 256;

__device__ __forceinline__ uint32_t ld32_any(const uint8_t* p, bool a4) {
  return a4 ? *(const uint32_t*)p : ldu32(p);
}

__global__ __launch_bounds__(kCheckThreads) void k_vxq_check(const VxqArgs a) {
  __shared__ uint32_t T[8][256];
  for (int i = threadIdx.x; i < 8 * 256; i += kCheckThreads) (&T[0][0])[i] = a.crc[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t waves = (int64_t)gridDim.x * (kCheckThreads / 64);
  for (int64_t b = (int64_t)blockIdx.x * (kCheckThreads / 64) + wv; b < a.nblocks; b += waves) {
    const VxqBlock k = a.blocks[b];
    const VxqCol c = a.cols[k.col];
    const uint8_t* p = a.img + k.src;
    const uint32_t n = (uint32_t)kNullBitmapBytes + k.rows * (uint32_t)c.vs;
    const uint32_t sb = stretch_bytes(n);
    const uint32_t s0 = lane * sb < n ? lane * sb : n;
    const uint32_t s1 = s0 + sb < n ? s0 + sb : n;
    const bool a4 = (((uintptr_t)p) & 3) == 0;  // (s0 is a multiple of 8)
    uint32_t crc = 0, i = s0;
    for (; i + 8 <= s1; i += 8) {
      const uint32_t lo = ld32_any(p + i, a4) ^ crc, hi = ld32_any(p + i + 4, a4);
      crc = T[7][lo & 0xFF] ^ T[6][(lo >> 8) & 0xFF] ^ T[5][(lo >> 16) & 0xFF] ^ T[4][lo >> 24] ^
            T[3][hi & 0xFF] ^ T[2][(hi >> 8) & 0xFF] ^ T[1][(hi >> 16) & 0xFF] ^ T[0][hi >> 24];
    }
    for (; i < s1; i++) crc = T[0][(crc ^ ld8(p + i)) & 0xFF] ^ (crc >> 8);
    const uint32_t* K = a.kconst + (size_t)k.kidx * 65;
    uint32_t v = crc ? multmodp(K[lane], crc) : 0u;
    // valid rows: lanes 0..15 take 8 bitmap bytes each
    uint32_t valid = 0;
    if (lane < 16)
      for (int j = 0; j < 8; j++) {
        const uint32_t q = lane * 8 + j, r0 = q * 8;
        if (r0 >= k.rows) break;
        uint32_t byte = p[q];
        if (k.rows - r0 < 8) byte &= (1u << (k.rows - r0)) - 1u;
        valid += (uint32_t)__popc(byte);
      }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      v ^= (uint32_t)__shfl_xor((int)v, off);
      valid += (uint32_t)__shfl_xor((int)valid, off);
    }
    if (lane == 0) {
      const uint32_t total = v ^ multmodp(K[64], 0xFFFFFFFFu) ^ 0xFFFFFFFFu;
      if (total != ldu32(p + n)) {
        atomicAdd(&a.flags[0], 1u);
        atomicMin(&a.flags[1], (uint32_t)b);
      }
      if (valid != k.rows) atomicAdd(&a.flags[2], k.rows - valid);
    }
  }
}

// Decode: one 256-thread workgroup per block, rows strided over the threads.
__global__ __launch_bounds__(256) void k_vxq_decode(const VxqArgs a) {
  const VxqBlock k = a.blocks[blockIdx.x];
  const VxqCol c = a.cols[k.col];
  const uint8_t* v = a.img + k.src + kNullBitmapBytes;
  const bool a4 = (((uintptr_t)v) & 3) == 0;
  for (uint32_t r = threadIdx.x; r < k.rows; r += 256) {
    const int64_t row = k.dst_row + r;
    if (c.is_string) {
      const uint32_t code = ld32_any(v + 4ull * r, a4);
      uint32_t g = 0;
      if (code < k.ncodes) g = a.remap[k.remap + code];
      else atomicAdd(&a.flags[3], 1u);
      if (c.type == VQ_U8) ((uint8_t*)c.dst)[row] = (uint8_t)g;
      else ((uint16_t*)c.dst)[row] = (uint16_t)g;
    } else if (c.vs == 8) {
      const uint64_t x = (uint64_t)ld32_any(v + 8ull * r, a4) | (uint64_t)ld32_any(v + 8ull * r + 4, a4) << 32;
      ((uint64_t*)c.dst)[row] = x;
    } else {
      ((uint32_t*)c.dst)[row] = ld32_any(v + 4ull * r, a4);
    }
  }
}

}  // namespace
}  // namespace vq

using namespace vq;

extern "C" int vq_table_upload_vxq(vq_ctx* c, const void* file, uint64_t file_len, int32_t ncols,
                                   const int32_t* cols, vq_table** out, vq_stats* stats) {
  if (!c || !file || !out || ncols <= 0 || !cols) return vfail(VQ_E_ARG, "exec: gpu: bad vxq upload arguments");
  *out = nullptr;
  const uint8_t* img = (const uint8_t*)file;
  std::vector<Field> fields;
  std::vector<RgMeta> rgs;
  if (int rc = parse_footer(img, file_len, fields, rgs)) return rc;
  const uint64_t body_end = file_len - kTrailer - rd64(img + file_len - kTrailer + 4);
  int64_t nrows = 0;
  for (const RgMeta& r : rgs) nrows += r.rows;
  if (nrows > 0xFFFFFFFELL) return vfail(VQ_E_UNSUPPORTED, "exec: gpu: vxq: more than 2^32 rows");
  TableParts tp;
  tp.nrows = nrows;
  std::vector<VxqCol> vcols(ncols);
  std::vector<VxqBlock> blocks;
  std::vector<uint32_t> remap;
  std::vector<uint32_t> kconst;                    // CRC shift constants, 65 per block length
  std::unordered_map<uint32_t, uint32_t> kidx_of;  // block length -> row of kconst
  tp.dicts.resize(ncols);
  for (int32_t i = 0; i < ncols; i++) {
    const int32_t fc = cols[i];
    if (fc < 0 || fc >= (int32_t)fields.size()) return vfail(VQ_E_ARG, "exec: gpu: vxq: column %d out of range", fc);
    const Field& f = fields[fc];
    int vs = 0, type = 0;
    switch (f.type) {
      case kTInt64: vs = 8; type = VQ_I64; break;
      case kTFloat64: vs = 8; type = VQ_F64; break;
      case kTDate: vs = 4; type = VQ_DATE; break;
      case kTString: vs = 4; type = VQ_U16; break;  // narrowed below
      case kTBool:
        return vfail(VQ_E_UNSUPPORTED, "exec: gpu: vxq: column %s is BOOL", f.name.c_str());
      default:
        return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: column %s has unknown type %d", f.name.c_str(), f.type);
    }
    vcols[i].vs = vs;
    vcols[i].is_string = f.type == kTString;
    // the global dictionary in first-seen order over the row groups, and
    // each row group's local -> global table (getOrBuildRemap)
    std::unordered_map<std::string, uint32_t> gindex;
    std::vector<std::string>& gdict = tp.dicts[i];
    int64_t row0 = 0;
    for (size_t g = 0; g < rgs.size(); g++) {
      const RgMeta& r = rgs[g];
      const ColMeta& m = r.cols[fc];
      // NULL rows: the GPU operators take only columns without them (the
      // planner's eligibility check, ZoneMap.NullCount == 0)
      if (m.null_count != 0)
        return vfail(VQ_E_UNSUPPORTED, "exec: gpu: vxq: column %s has NULLs in row group %zu", f.name.c_str(), g);
      if (m.sec_off < 4 || m.sec_len < 0 || (uint64_t)m.sec_off + (uint64_t)m.sec_len > body_end ||
          (uint64_t)m.dict_off + m.dict_len > (uint64_t)m.sec_len)
        return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: section of rg=%zu col=%s outside the file", g, f.name.c_str());
      uint32_t remap0 = 0, ncodes = 0;
      if (vcols[i].is_string) {
        std::vector<std::string> local;
        if (m.dict_len == 0 && r.rows > 0)
          return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: rg=%zu col=%s has no dictionary", g, f.name.c_str());
        if (m.dict_len)
          if (int rc = parse_dict(img + m.sec_off + m.dict_off, m.dict_len, local)) return rc;
        remap0 = (uint32_t)remap.size();
        ncodes = (uint32_t)local.size();
        for (const std::string& s : local) {
          auto it = gindex.find(s);
          uint32_t code;
          if (it == gindex.end()) {
            code = (uint32_t)gdict.size();
            gindex.emplace(s, code);
            gdict.push_back(s);
          } else {
            code = it->second;
          }
          remap.push_back(code);
        }
      }
      // the section's blocks: [bitmap][rows x vs][CRC], back to back
      uint64_t pos = (uint64_t)m.sec_off;
      const uint64_t sec_end = (uint64_t)m.sec_off + (uint64_t)m.sec_len - m.dict_len;
      for (uint32_t b0 = 0; b0 < r.rows; b0 += kBlockRows) {
        const uint32_t br = std::min<uint32_t>(kBlockRows, r.rows - b0);
        const uint64_t blen = (uint64_t)kNullBitmapBytes + (uint64_t)br * vs + 4;
        if (pos + blen > sec_end)
          return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: blocks of rg=%zu col=%s overrun their section", g, f.name.c_str());
        const uint32_t n = (uint32_t)(blen - 4);
        auto kit = kidx_of.find(n);
        if (kit == kidx_of.end()) {
          kit = kidx_of.emplace(n, (uint32_t)kconst.size() / 65).first;
          const uint32_t sb = stretch_bytes(n);
          for (uint32_t l = 0; l < 64; l++) kconst.push_back(x8n(n - std::min(n, (l + 1) * sb)));
          kconst.push_back(x8n(n));
        }
        blocks.push_back(VxqBlock{pos, row0 + b0, br, (uint32_t)i, remap0, ncodes, kit->second, 0});
        pos += blen;
      }
      row0 += r.rows;
    }
    if (vcols[i].is_string) {
      if (gdict.size() > 65536)
        return vfail(VQ_E_UNSUPPORTED, "exec: gpu: vxq: column %s has %zu distinct strings", f.name.c_str(), gdict.size());
      type = gdict.size() <= 256 ? VQ_U8 : VQ_U16;
    }
    vcols[i].type = type;
    tp.types.push_back(type);
    tp.dict_sizes.push_back(vcols[i].is_string ? (int32_t)gdict.size() : 0);
  }
  // zone maps from the footer (raw bits as storage/writer.go computes them);
  // dictionary columns span their global codes, as vq_table_upload sets them
  tp.rg_rows.resize(rgs.size());
  tp.zones.assign(rgs.size() * ncols, vq_zone{});
  for (size_t g = 0; g < rgs.size(); g++) {
    tp.rg_rows[g] = rgs[g].rows;
    for (int32_t i = 0; i < ncols; i++) {
      const ColMeta& m = rgs[g].cols[cols[i]];
      vq_zone& z = tp.zones[g * ncols + i];
      z.null_count = m.null_count;
      if (vcols[i].is_string) {
        z.has_minmax = rgs[g].rows > 0;
        z.min_bits = 0;
        z.max_bits = tp.dicts[i].empty() ? 0 : tp.dicts[i].size() - 1;
      } else {
        z.has_minmax = m.has_minmax;
        z.min_bits = m.min;
        z.max_bits = m.max;
      }
    }
  }

  // device: the image, the columns, the block list; check, then decode
  const int dev = ctx_device(c);
  hipSetDevice(dev);
  auto width = [](int t) { return t == VQ_U8 ? 1 : t == VQ_U16 ? 2 : t == VQ_DATE ? 4 : 8; };
  uint64_t col_bytes = 0;
  for (int32_t i = 0; i < ncols; i++) col_bytes += (((uint64_t)nrows * width(vcols[i].type) + 255) & ~255ULL) + 256;
  const auto& tabs = crc_tables();
  const uint64_t aux = blocks.size() * sizeof(VxqBlock) + ncols * sizeof(VxqCol) + remap.size() * 4 +
                       kconst.size() * 4 + sizeof(tabs.t) + 64 + 6 * 256;
  const uint64_t stage = ((body_end + 255) & ~255ULL) + aux;
  if (!ctx_reserve(c, col_bytes + stage))
    return vfail(VQ_E_BUDGET, "exec: gpu: vxq: %llu bytes of columns and staging exceed the device budget",
                 (unsigned long long)(col_bytes + stage));
  hipStream_t s = nullptr;
  uint8_t* dimg = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&dimg, stage);
  tp.dev.assign(ncols, nullptr);
  for (int32_t i = 0; i < ncols && e == hipSuccess; i++) {
    e = hipMalloc(&tp.dev[i], (((uint64_t)nrows * width(vcols[i].type) + 255) & ~255ULL) + 256);
    vcols[i].dst = tp.dev[i];
  }
  // staging layout: image | blocks | columns | remap | CRC tables | flags
  uint64_t off = (body_end + 255) & ~255ULL;
  auto carve = [&](uint64_t bytes) {
    uint8_t* p = dimg + off;
    off += (bytes + 255) & ~255ULL;
    return p;
  };
  VxqArgs a;
  memset(&a, 0, sizeof a);
  uint32_t hflags[4] = {0, 0xFFFFFFFFu, 0, 0};
  float ms = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) {
    a.img = dimg;
    a.blocks = (const VxqBlock*)carve(blocks.size() * sizeof(VxqBlock));
    a.cols = (const VxqCol*)carve(ncols * sizeof(VxqCol));
    a.remap = (const uint32_t*)carve(std::max<size_t>(remap.size(), 1) * 4);
    a.crc = (const uint32_t*)carve(sizeof(tabs.t));
    a.kconst = (const uint32_t*)carve(std::max<size_t>(kconst.size(), 1) * 4);
    a.flags = (uint32_t*)carve(64);
    a.nblocks = (int64_t)blocks.size();
    // the one bulk upload: the file's row-group sections as they lie on disk
    e = hipMemcpyAsync(dimg, img, body_end, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !blocks.empty())
      e = hipMemcpyAsync((void*)a.blocks, blocks.data(), blocks.size() * sizeof(VxqBlock), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync((void*)a.cols, vcols.data(), ncols * sizeof(VxqCol), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !remap.empty())
      e = hipMemcpyAsync((void*)a.remap, remap.data(), remap.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync((void*)a.crc, &tabs.t[0][0], sizeof(tabs.t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !kconst.empty())
      e = hipMemcpyAsync((void*)a.kconst, kconst.data(), kconst.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(a.flags, hflags, sizeof hflags, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess && !blocks.empty()) {
      const size_t wgs = std::min<size_t>((blocks.size() + 3) / 4, (size_t)device_cus() * 8);
      hipLaunchKernelGGL(k_vxq_check, dim3((unsigned)wgs), dim3(kCheckThreads), 0, s, a);
      hipLaunchKernelGGL(k_vxq_decode, dim3((unsigned)blocks.size()), dim3(256), 0, s, a);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hflags, a.flags, sizeof hflags, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (dimg) hipFree(dimg);
  if (s) hipStreamDestroy(s);
  ctx_release(c, stage);
  auto drop = [&]() {
    for (void* p : tp.dev)
      if (p) hipFree(p);
    ctx_release(c, col_bytes);
  };
  if (e != hipSuccess) {
    drop();
    return vfail(VQ_E_DEVICE, "exec: gpu: vxq upload: %s", hipGetErrorString(e));
  }
  if (hflags[0]) {
    const VxqBlock& k = blocks[hflags[1] < blocks.size() ? hflags[1] : 0];
    int64_t rg = 0, acc = 0;
    while (rg + 1 < (int64_t)rgs.size() && acc + rgs[rg].rows <= k.dst_row) acc += rgs[rg++].rows;
    const std::string name = fields[cols[k.col]].name;
    drop();
    return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: crc block rg=%lld col=%s off=%llu (%u bad blocks)", (long long)rg,
                 name.c_str(), (unsigned long long)k.src, hflags[0]);
  }
  if (hflags[2] || hflags[3]) {
    drop();
    return vfail(VQ_E_CORRUPT, "exec: gpu: vxq: %u NULL rows against zero NULL counts, %u codes outside their dictionary",
                 hflags[2], hflags[3]);
  }
  tp.bytes = col_bytes;
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->device_ms = ms;
    stats->main_kernel_ms = ms;
    stats->rows_scanned = nrows;
    stats->rg_total = stats->rg_scanned = (int64_t)rgs.size();
    // the blocks read once and the columns written once
    stats->bytes_alg = 0;
    for (const VxqBlock& k : blocks)
      stats->bytes_alg += kNullBitmapBytes + (int64_t)k.rows * (vcols[k.col].vs + width(vcols[k.col].type)) + 4;
  }
  std::vector<void*> owned = tp.dev;
  if (int rc = table_adopt(c, std::move(tp), out)) {
    for (void* p : owned)
      if (p) hipFree(p);
    ctx_release(c, col_bytes);
    return rc;
  }
  return VQ_OK;
}
