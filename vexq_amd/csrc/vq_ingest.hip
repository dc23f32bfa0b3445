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
// built from them.  Device: k_vxq_ingest reads every block once (one wave
// per block, coalesced 16-byte loads), computes its CRC-32 with slice-by-8
// tables in LDS (the lanes' lines joined by GF(2) shifts), counts its valid
// rows and writes its values into the resident column, codes through their
// row group's remap table.
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
inline uint32_t multmodp(uint32_t a, uint32_t b) {
  uint32_t p = 0;
  for (int i = 0; i < 32; i++) {
    if (a & (0x80000000u >> i)) p ^= b;
    b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
  }
  return p;
}

// a(x) b(x) mod P on the device without branches: bit i of a (from the top)
// selects b x^i
__device__ __forceinline__ uint32_t gf_mul(uint32_t a, uint32_t b) {
  uint32_t p = 0;
#pragma unroll
  for (int i = 0; i < 32; i++) {
    p ^= b & (uint32_t)((int32_t)(a << i) >> 31);
    b = (b >> 1) ^ (0xEDB88320u & (0u - (b & 1u)));
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

// x^(-8n) mod P: powers of x^-1 = (P(x) + 1) / x, which is 0xDB710641 with
// bit 31 as x^0 (x times it is P(x) + 1 = 1 mod P)
uint32_t xinv8n(uint64_t n) {
  uint32_t p = 0x80000000u, b = 0xDB710641u;
  for (uint64_t e = 8 * n; e; e >>= 1) {
    if (e & 1) p = multmodp(b, p);
    b = multmodp(b, b);
  }
  return p;
}

// The device's tables: the slice-by-8 tables, then x^8192 mod P times each
// byte value at each of the register's four byte positions; and, on the host,
// lane l's shift to the end of a window, x^(8 (1008 - 16 l)).
struct DeviceCrcTables {
  uint32_t t[12][256];
  uint32_t lane[64];
  DeviceCrcTables() {
    memcpy(t, crc_tables().t, sizeof crc_tables().t);
    const uint32_t x1024 = x8n(1024);
    for (int k = 0; k < 4; k++)
      for (uint32_t b = 0; b < 256; b++) t[8 + k][b] = multmodp(x1024, b << (8 * k));
    for (int l = 0; l < 64; l++) lane[l] = x8n(1008 - 16 * l);
  }
};
const DeviceCrcTables& device_crc_tables() {
  static const DeviceCrcTables tabs;
  return tabs;
}

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
// What a block's values become in the resident column
enum : uint32_t { kCopy32 = 0, kCopy64 = 1, kCodeU8 = 2, kCodeU16 = 3 };

struct VxqBlock {
  uint64_t src;     // byte offset of the block in the image
  uint64_t dst;     // device address of its first value in the column
  uint32_t rows;
  uint32_t kind;    // kCopy32 .. kCodeU16
  uint32_t remap;   // codes: first entry of its row group's remap table
  uint32_t ncodes;  // codes: entries of its row group's dictionary
  uint32_t krow;    // its row of lane constants (one row per t, the zero bytes after it in its window)
  uint32_t initc;   // x^(8n) 0xFFFFFFFF ^ 0xFFFFFFFF: the start and end values' share
};

struct VxqArgs {
  const uint8_t* img;
  const VxqBlock* blocks;
  const uint32_t* remap;
  const uint32_t* crc;     // [12][256]: the slice-by-8 tables, then the 1024-byte shift's byte tables
  const uint32_t* kconst;  // [rows][64]: lane l's register moved to its window's end, then back
                           // past the window's t zero bytes: x^(8 (1008 - 16 l - t))
  uint32_t* flags;         // [0] bad checksums, [1] first bad block, [2] NULL rows, [3] bad codes
  int64_t nblocks;
};

// One wave per block.  A block's window is the block rounded out to 16-byte
// lines: [d masked bytes][bitmap][rows x vs][CRC], d = its offset mod 16, in
// L = 5 or 9 lines of 1024 bytes (64 lanes x 16 bytes).  Lane l loads lines
// l, l + 64, l + 128, ... (coalesced 16-byte loads, every byte of the image
// read once) and the wave keeps the window in LDS.  The waves are a
// software pipeline: a block's lines are loaded while the wave checks and
// decodes the block before it.
//
// CRC-32 (VerifyTrailing, internal/encoding/crc.go:31-40) by linearity: lane
// l's register R starts at 0 and, for each of its lines, moves 1024 bytes on
// (its line and the other lanes' 63 lines; x^8192 mod P applied through four
// byte tables) and takes the line's own register from zero (two slice-by-8
// steps that do not depend on R, so a lane's lines overlap).  R is then
// moved to the end of the window (x^(8 (1008 - 16 l))): the XOR of the lanes'
// registers is the register of the whole window, with the bytes outside
// [bitmap, CRC) masked to zero, which is x^(8t) times the block's register
// from zero, t = 1024 L - d - n for a block of n bytes under the CRC.  So each
// lane multiplies R by x^(8 (1008 - 16 l - t)) (a row of constants per t,
// computed on the host) and the XOR of the lanes is the block's register
// from zero; the block passes when it equals stored ^ x^(8n) 0xFFFFFFFF ^
// 0xFFFFFFFF.
// The bitmap's valid rows are counted (bit i of byte i/8, LSB first;
// storage/codec.go:9-25) and the values decoded from LDS into the resident
// column, codes through their row group's remap table.
constexpr int kIngestThreads = 256;
constexpr int kMaxLines = 9;    // 16-byte lines per lane
constexpr int kShortLines = 5;  // enough for a block of 4-byte values
constexpr int kRowIters = kBlockRows / 64;
static_assert(kMaxLines * 1024 >= 15 + kNullBitmapBytes + kBlockRows * 8 + 4, "a block's window fits");
static_assert(kShortLines * 1024 >= 15 + kNullBitmapBytes + kBlockRows * 4, "4-byte blocks fit the short window");

// lines of 1024 bytes in the window of a block of n bytes under the CRC at
// offset d mod 16
__host__ __device__ inline uint32_t window_lines(uint32_t d, uint32_t n) {
  return d + n <= 1024u * kShortLines ? kShortLines : kMaxLines;
}

// 4 bytes at any offset of a wave's LDS window (the dword after the read
// must lie inside the window's array)
__device__ __forceinline__ uint32_t win32(const uint32_t* w, uint32_t off) {
  const uint32_t i = off >> 2;
  return __builtin_amdgcn_alignbyte(w[i + 1], w[i], off & 3);
}

// v's bytes at window offsets [o, o + 4) kept where they lie in [lo, hi)
__device__ __forceinline__ uint32_t keep_bytes(uint32_t v, int o, int lo, int hi) {
  const int a = min(max(lo - o, 0), 4), b = min(max(hi - o, 0), 4);
  return v & (uint32_t)(((1ull << (8 * b)) - 1) & ~((1ull << (8 * a)) - 1));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void load_lines(const VxqArgs& a, const VxqBlock& k, uint32_t vs, int lane, uint4* v) {
  const uint32_t d = (uint32_t)(k.src & 15);
  const uint8_t* base = a.img + (k.src - d);
  const uint32_t wlen = d + kNullBitmapBytes + k.rows * vs + 4;
#pragma unroll
  for (int j = 0; j < kMaxLines; j++) {
    const uint32_t q = 16u * (lane + 64 * j);
    v[j] = q < wlen ? *(const uint4*)(base + q) : make_uint4(0, 0, 0, 0);
  }
}

__device__ __forceinline__ uint32_t slice8(const uint32_t (*T)[256], uint32_t lo, uint32_t hi) {
  return T[7][lo & 0xFF] ^ T[6][(lo >> 8) & 0xFF] ^ T[5][(lo >> 16) & 0xFF] ^ T[4][lo >> 24] ^
         T[3][hi & 0xFF] ^ T[2][(hi >> 8) & 0xFF] ^ T[1][(hi >> 16) & 0xFF] ^ T[0][hi >> 24];
}

// the lane's register over L lines of the window, moved to the window's end
template <int L>
__device__ __forceinline__ uint32_t lane_crc(const uint32_t (*T)[256], const uint4* win, int lane, uint32_t d,
                                             uint32_t n, uint32_t wlen, uint32_t kc) {
  uint32_t raw[L];
#pragma unroll
  for (int j = 0; j < L; j++) {
    const int q = 16 * (lane + 64 * j);
    uint4 x = 1024u * j < wlen ? win[lane + 64 * j] : make_uint4(0, 0, 0, 0);
    if ((j == 0 && d) || 1024u * (j + 1) > d + n) {  // the line holds an edge of the block
      x.x = keep_bytes(x.x, q, d, d + n);
      x.y = keep_bytes(x.y, q + 4, d, d + n);
      x.z = keep_bytes(x.z, q + 8, d, d + n);
      x.w = keep_bytes(x.w, q + 12, d, d + n);
    }
    raw[j] = slice8(T, x.z ^ slice8(T, x.x, x.y), x.w);
  }
  uint32_t R = 0;
#pragma unroll
  for (int j = 0; j < L; j++) R = T[8][R & 0xFF] ^ T[9][(R >> 8) & 0xFF] ^ T[10][(R >> 16) & 0xFF] ^ T[11][R >> 24] ^ raw[j];
  return gf_mul(kc, R);
}

// blocks: the descriptors, read-only and wave-uniform (scalar loads)
__global__ __launch_bounds__(kIngestThreads, 3) void k_vxq_ingest(const VxqArgs a, const VxqBlock* __restrict__ blocks) {
  __shared__ uint32_t T[12][256];
  __shared__ uint4 W[kIngestThreads / 64][kMaxLines * 64 + 1];
  for (int i = threadIdx.x; i < 12 * 256; i += kIngestThreads) (&T[0][0])[i] = a.crc[i];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  __syncthreads();
  uint4* win = W[wv];
  const uint32_t* w32 = (const uint32_t*)win;
  const int64_t waves = (int64_t)gridDim.x * (kIngestThreads / 64);
  int64_t b = (int64_t)blockIdx.x * (kIngestThreads / 64) + wv;
  if (b >= a.nblocks) return;
  VxqBlock k = blocks[b];
  uint4 v[kMaxLines];
  load_lines(a, k, k.kind == kCopy64 ? 8 : 4, lane, v);
  uint32_t kc = a.kconst[(size_t)k.krow * 64 + lane];
  for (;;) {
    const int64_t nb = b + waves;
    const bool more = nb < a.nblocks;
    VxqBlock kn = k;
    if (more) kn = blocks[nb];
    const uint32_t vs = k.kind == kCopy64 ? 8 : 4;
    const uint32_t d = (uint32_t)(k.src & 15), n = (uint32_t)kNullBitmapBytes + k.rows * vs, wlen = d + n + 4;
    wave_sync();  // the previous block's reads of the window are done
#pragma unroll
    for (int j = 0; j < kMaxLines; j++)
      if (1024u * j < wlen) win[lane + 64 * j] = v[j];
    wave_sync();  // the window is in LDS
    // decode: every lane reads its 16 rows (in the window's array whatever
    // the block's length), stores those below the block's rows
    const uint32_t v0 = d + (uint32_t)kNullBitmapBytes;
    if (k.kind == kCopy64) {
      uint64_t x[kRowIters];
#pragma unroll
      for (int i = 0; i < kRowIters; i++) {
        const uint32_t o = v0 + 8u * (lane + 64 * i);
        x[i] = (uint64_t)win32(w32, o) | (uint64_t)win32(w32, o + 4) << 32;
      }
#pragma unroll
      for (int i = 0; i < kRowIters; i++)
        if (lane + 64u * i < k.rows) ((uint64_t*)k.dst)[lane + 64 * i] = x[i];
    } else if (k.kind == kCopy32) {
      uint32_t x[kRowIters];
#pragma unroll
      for (int i = 0; i < kRowIters; i++) x[i] = win32(w32, v0 + 4u * (lane + 64 * i));
#pragma unroll
      for (int i = 0; i < kRowIters; i++)
        if (lane + 64u * i < k.rows) ((uint32_t*)k.dst)[lane + 64 * i] = x[i];
    } else {
      uint32_t g[kRowIters], bad = 0;
#pragma unroll
      for (int i = 0; i < kRowIters; i++) {
        const uint32_t code = win32(w32, v0 + 4u * (lane + 64 * i));
        const bool live = lane + 64u * i < k.rows;
        bad += live && code >= k.ncodes;
        g[i] = a.remap[k.remap + (code < k.ncodes ? code : 0u)];  // the table ends in a spare entry
      }
#pragma unroll
      for (int i = 0; i < kRowIters; i++)
        if (lane + 64u * i < k.rows) {
          if (k.kind == kCodeU8) ((uint8_t*)k.dst)[lane + 64 * i] = (uint8_t)g[i];
          else ((uint16_t*)k.dst)[lane + 64 * i] = (uint16_t)g[i];
        }
      if (bad) atomicAdd(&a.flags[3], bad);
    }
    // the next block's lines, in flight while this one is checked
    const uint32_t kcur = kc;
    if (more) {
      load_lines(a, kn, kn.kind == kCopy64 ? 8 : 4, lane, v);
      kc = a.kconst[(size_t)kn.krow * 64 + lane];
    }
    uint32_t valid = 0;
    if (lane < kNullBitmapBytes / 4 && 32u * lane < k.rows) {
      uint32_t bits = win32(w32, d + 4 * lane);
      if (k.rows - 32u * lane < 32) bits &= (1u << (k.rows - 32u * lane)) - 1u;
      valid = (uint32_t)__popc(bits);
    }
    uint32_t x = window_lines(d, n) == kShortLines ? lane_crc<kShortLines>(T, win, lane, d, n, wlen, kcur)
                                                    : lane_crc<kMaxLines>(T, win, lane, d, n, wlen, kcur);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      x ^= (uint32_t)__shfl_xor((int)x, off);
      valid += (uint32_t)__shfl_xor((int)valid, off);
    }
    if (lane == 0) {
      if (x != (win32(w32, d + n) ^ k.initc)) {
        atomicAdd(&a.flags[0], 1u);
        atomicMin(&a.flags[1], (uint32_t)b);
      }
      if (valid != k.rows) atomicAdd(&a.flags[2], k.rows - valid);
    }
    if (!more) break;
    k = kn;
    b = nb;
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
  struct OutCol {
    int type = 0, vs = 0;
    bool is_string = false;
  };
  struct BlockAt {
    int32_t col;
    int64_t row;
  };
  std::vector<OutCol> vcols(ncols);
  std::vector<VxqBlock> blocks;
  std::vector<BlockAt> where;  // each block's column and first row
  std::vector<uint32_t> remap;
  const auto& tabs = device_crc_tables();
  std::vector<uint32_t> kconst;                      // [row][64] lane constants, one row per t
  std::vector<int32_t> krow_of(1024 * kMaxLines, -1);  // t -> row
  std::unordered_map<uint32_t, uint32_t> initc_of;  // block length n -> x^(8n) 0xFFFFFFFF ^ 0xFFFFFFFF
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
        // the block's window on the device: d bytes before it (its offset
        // mod 16), then lines of 1024 bytes per lane pass to its end
        const uint32_t n = (uint32_t)(blen - 4), d = (uint32_t)(pos & 15);
        const uint32_t t = 1024 * window_lines(d, n) - d - n;
        if (krow_of[t] < 0) {
          krow_of[t] = (int32_t)(kconst.size() / 64);
          const uint32_t back = xinv8n(t);
          for (int l = 0; l < 64; l++) kconst.push_back(multmodp(tabs.lane[l], back));
        }
        auto kit = initc_of.find(n);
        if (kit == initc_of.end()) kit = initc_of.emplace(n, multmodp(x8n(n), 0xFFFFFFFFu) ^ 0xFFFFFFFFu).first;
        // dst and kind are set once the columns exist
        blocks.push_back(VxqBlock{pos, 0, br, 0, remap0, ncodes, (uint32_t)krow_of[t], kit->second});
        where.push_back(BlockAt{i, row0 + b0});
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
  remap.push_back(0);  // a spare entry: the kernel reads entry 0 of an empty dictionary
  const uint64_t aux = blocks.size() * sizeof(VxqBlock) + remap.size() * 4 + sizeof(tabs.t) +
                       std::max<size_t>(kconst.size(), 1) * 4 + 64 + 6 * 256;
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
  }
  if (e == hipSuccess)
    for (size_t j = 0; j < blocks.size(); j++) {
      const OutCol& oc = vcols[where[j].col];
      blocks[j].dst = (uint64_t)(uintptr_t)tp.dev[where[j].col] + (uint64_t)where[j].row * width(oc.type);
      blocks[j].kind = !oc.is_string ? (oc.vs == 8 ? kCopy64 : kCopy32) : oc.type == VQ_U8 ? kCodeU8 : kCodeU16;
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
  const int reps = ctx_kernel_repeats(c);
  float ms = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) {
    a.img = dimg;
    a.blocks = (const VxqBlock*)carve(blocks.size() * sizeof(VxqBlock));
    a.remap = (const uint32_t*)carve(std::max<size_t>(remap.size(), 1) * 4);
    a.crc = (const uint32_t*)carve(sizeof(tabs.t));
    a.kconst = (const uint32_t*)carve(std::max<size_t>(kconst.size(), 1) * 4);
    a.flags = (uint32_t*)carve(64);
    a.nblocks = (int64_t)blocks.size();
    // the one bulk upload: the file's row-group sections as they lie on disk
    e = hipMemcpyAsync(dimg, img, body_end, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !blocks.empty())
      e = hipMemcpyAsync((void*)a.blocks, blocks.data(), blocks.size() * sizeof(VxqBlock), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync((void*)a.remap, remap.data(), remap.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync((void*)a.crc, &tabs.t[0][0], sizeof(tabs.t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !kconst.empty())
      e = hipMemcpyAsync((void*)a.kconst, kconst.data(), kconst.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(a.flags, hflags, sizeof hflags, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess && !blocks.empty()) {
      // persistent: three workgroups of four waves fit a CU's LDS
      const size_t wgs = std::min<size_t>((blocks.size() + 3) / 4, (size_t)device_cus() * 3);
      for (int rep = 0; rep < reps; rep++)  // the same columns each time; the flags count each pass
        hipLaunchKernelGGL(k_vxq_ingest, dim3((unsigned)wgs), dim3(kIngestThreads), 0, s, a, a.blocks);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipMemcpyAsync(hflags, a.flags, sizeof hflags, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    hflags[0] /= reps;
    hflags[2] /= reps;
    hflags[3] /= reps;
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
    const size_t j = hflags[1] < blocks.size() ? hflags[1] : 0;
    const VxqBlock& k = blocks[j];
    int64_t rg = 0, acc = 0;
    while (rg + 1 < (int64_t)rgs.size() && acc + rgs[rg].rows <= where[j].row) acc += rgs[rg++].rows;
    const std::string name = fields[cols[where[j].col]].name;
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
    for (size_t j = 0; j < blocks.size(); j++) {
      const OutCol& oc = vcols[where[j].col];
      stats->bytes_alg += kNullBitmapBytes + (int64_t)blocks[j].rows * (oc.vs + width(oc.type)) + 4;
    }
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
