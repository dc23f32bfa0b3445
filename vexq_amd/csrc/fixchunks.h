// fixchunks.h — the addend side of exact fixed-point summation: a float64
// split into three signed w-bit chunks on a per-aggregate scale.  No includes
// and only built-in types, so the same text is compiled into the static
// kernels, the host code and the run-time compiled kernels (hipRTC has no
// standard headers).  The rounding side is in fixpoint.h.
#ifndef VQ_FIXCHUNKS_H
#define VQ_FIXCHUNKS_H

#if defined(__HIPCC__) || defined(__HIP__) || defined(__HIPCC_RTC__)
#define VQ_FX_HD __host__ __device__ __forceinline__
#else
#define VQ_FX_HD inline
#endif

namespace vqfx {

struct FixScale {
  int lsb;  // weight of limb 0's least significant bit: 2^lsb
  int w;    // bits per chunk (at most 40)
};

// The three w-bit chunks of float64 t on scale fs, each carrying t's sign:
// t = (c0 + c1 2^w + c2 2^2w) 2^lsb for the bits of t at or above 2^lsb.
// Returns false for a zero addend (nothing to add).  I is a signed 64-bit
// integer type.
template <class I>
VQ_FX_HD bool fix_chunks(double t, FixScale fs, I& c0, I& c1, I& c2) {
  typedef unsigned long long U;
  union {
    double d;
    U u;
  } cv;
  cv.d = t;
  const U b = cv.u;
  const int ex = (int)((b >> 52) & 0x7FF);
  U m = b & ((1ULL << 52) - 1);
  if (ex == 0 && m == 0) return false;
  int e;
  if (ex == 0) {
    e = -1074;
  } else {
    m |= 1ULL << 52;
    e = ex - 1075;
  }
  const int sh = e - fs.lsb;  // |t| = m * 2^sh units of 2^lsb
  U lo, hi;
  if (sh >= 0) {
    if (sh >= 128) return false;  // outside the planned range (the planner's bound excludes it)
    lo = sh >= 64 ? 0 : m << sh;
    hi = sh == 0 ? 0 : (sh >= 64 ? m << (sh - 64) : m >> (64 - sh));
  } else {
    if (-sh >= 64) return false;
    lo = m >> (-sh);
    hi = 0;
  }
  const int w = fs.w;
  const U wm = (1ULL << w) - 1;
  const U u0 = lo & wm;
  const U u1 = ((lo >> w) | (hi << (64 - w))) & wm;
  const U u2 = (2 * w >= 64 ? hi >> (2 * w - 64) : (lo >> (2 * w)) | (hi << (64 - 2 * w))) & wm;
  const bool neg = (b >> 63) != 0;
  c0 = neg ? -(I)u0 : (I)u0;
  c1 = neg ? -(I)u1 : (I)u1;
  c2 = neg ? -(I)u2 : (I)u2;
  return true;
}

}  // namespace vqfx

#endif
