// fixpoint.h — exact float64 summation as fixed-point integers.
//
// A float64 SUM over a join's groups is accumulated as three integer limbs:
//   value = sum_i limb[i] * 2^(lsb + i*w),   0 < w <= 40,
// where every addend is split into three w-bit chunks on the same scale.
// Adding chunks with integer atomics is associative, so the sum does not
// depend on the order in which rows arrive; the limbs are only normalised
// (carried) once, when the group's result is rounded to float64.
//
// The scale is chosen per aggregate from a bound B on |addend| (zone maps):
// with 2^(E-1) <= B < 2^E, lsb = E - 3w, so every addend's bits at or above
// 2^lsb are kept (bits below it cannot occur for addends whose magnitude
// spans fewer than 3w - 53 binades below 2^E; smaller bits are dropped, and
// the planner keeps w large enough that this does not happen for the data's
// range).  Each limb takes at most n addends of w bits: with n < 2^L rows,
// |limb| < 2^(w+L) <= 2^62.
//
// Header-only and free of device-only calls so that the same code is built
// into the gfx950 kernels and into the host-side unit test.  No local array
// is indexed with a run-time value: on the device such arrays live in
// scratch memory.
#pragma once
#include <stdint.h>

#include <cmath>

#include "fixchunks.h"

namespace vq {

using vqfx::FixScale;
using vqfx::fix_chunks;

// 256-bit two's complement integer, least significant word first.
struct U256 {
  uint64_t w0, w1, w2, w3;
};

VQ_FX_HD uint64_t u256_word(const U256& x, int i) {
  return i == 0 ? x.w0 : i == 1 ? x.w1 : i == 2 ? x.w2 : i == 3 ? x.w3 : 0;
}

// x << s, 0 <= s < 256 (bits shifted past 255 are dropped)
VQ_FX_HD U256 u256_shl(const U256& x, int s) {
  const int ws = s >> 6, bs = s & 63;
  auto part = [&](int k) -> uint64_t {
    const uint64_t hi = u256_word(x, k - ws);
    const uint64_t lo = u256_word(x, k - ws - 1);
    return bs ? (hi << bs) | (lo >> (64 - bs)) : hi;
  };
  return U256{part(0), part(1), part(2), part(3)};
}

// x >> s (logical), 0 <= s < 256
VQ_FX_HD U256 u256_shr(const U256& x, int s) {
  const int ws = s >> 6, bs = s & 63;
  auto part = [&](int k) -> uint64_t {
    const uint64_t lo = u256_word(x, k + ws);
    const uint64_t hi = u256_word(x, k + ws + 1);
    return bs ? (lo >> bs) | (hi << (64 - bs)) : lo;
  };
  return U256{part(0), part(1), part(2), part(3)};
}

VQ_FX_HD U256 u256_add(const U256& a, const U256& b) {
  U256 r;
  uint64_t c;
  r.w0 = a.w0 + b.w0;
  c = r.w0 < a.w0;
  const uint64_t s1 = a.w1 + b.w1;
  r.w1 = s1 + c;
  c = (s1 < a.w1) | (r.w1 < s1);
  const uint64_t s2 = a.w2 + b.w2;
  r.w2 = s2 + c;
  c = (s2 < a.w2) | (r.w2 < s2);
  r.w3 = a.w3 + b.w3 + c;
  return r;
}

VQ_FX_HD U256 u256_neg(const U256& a) {
  const U256 inv{~a.w0, ~a.w1, ~a.w2, ~a.w3};
  return u256_add(inv, U256{1, 0, 0, 0});
}

VQ_FX_HD U256 u256_sext(int64_t v) {
  const uint64_t e = v < 0 ? ~0ULL : 0ULL;
  return U256{(uint64_t)v, e, e, e};
}

VQ_FX_HD bool u256_zero(const U256& a) { return (a.w0 | a.w1 | a.w2 | a.w3) == 0; }

// 64 x 64 -> 128 bits from 32-bit halves.
VQ_FX_HD void mul64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  const uint64_t a0 = a & 0xFFFFFFFFULL, a1 = a >> 32;
  const uint64_t b0 = b & 0xFFFFFFFFULL, b1 = b >> 32;
  const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  const uint64_t mid = (p00 >> 32) + (p01 & 0xFFFFFFFFULL) + (p10 & 0xFFFFFFFFULL);
  lo = (p00 & 0xFFFFFFFFULL) | (mid << 32);
  hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// a * m for a non-negative a, truncated to 256 bits.
VQ_FX_HD U256 u256_mul64(const U256& a, uint64_t m) {
  uint64_t l0, h0, l1, h1, l2, h2, l3, h3;
  mul64(a.w0, m, l0, h0);
  mul64(a.w1, m, l1, h1);
  mul64(a.w2, m, l2, h2);
  mul64(a.w3, m, l3, h3);
  (void)h3;
  return u256_add(U256{l0, l1, l2, l3}, U256{0, h0, h1, h2});
}

// Index of the highest set bit, -1 for zero.
VQ_FX_HD int u256_top(const U256& a) {
  if (a.w3) return 192 + 63 - __builtin_clzll(a.w3);
  if (a.w2) return 128 + 63 - __builtin_clzll(a.w2);
  if (a.w1) return 64 + 63 - __builtin_clzll(a.w1);
  if (a.w0) return 63 - __builtin_clzll(a.w0);
  return -1;
}

// (limb0 + limb1 2^w + limb2 2^2w) * 2^lsb * mult, rounded once to the
// nearest float64 (ties to even).  mult is a non-negative integer < 2^32
// (the build side's multiplicity of the group's key).
VQ_FX_HD double fix_to_f64(int64_t l0, int64_t l1, int64_t l2, FixScale fs, uint64_t mult) {
  U256 a = u256_sext(l0);
  a = u256_add(a, u256_shl(u256_sext(l1), fs.w));
  a = u256_add(a, u256_shl(u256_sext(l2), 2 * fs.w));
  const bool neg = (int64_t)a.w3 < 0;
  if (neg) a = u256_neg(a);
  if (mult != 1) a = u256_mul64(a, mult);
  const int top = u256_top(a);
  if (top < 0) return 0.0;
  double r;
  if (top < 53) {
    r = std::ldexp((double)a.w0, fs.lsb);
  } else {
    const int sh = top - 52;  // drop sh bits: mantissa = a >> sh (53 bits)
    uint64_t mant = u256_shr(a, sh).w0;
    const uint64_t rb = u256_shr(a, sh - 1).w0 & 1;  // first dropped bit
    const bool sticky = sh >= 2 && !u256_zero(u256_shl(a, 256 - (sh - 1)));  // the rest
    if (rb && (sticky || (mant & 1))) mant++;
    r = std::ldexp((double)mant, sh + fs.lsb);
  }
  return neg ? -r : r;
}

}  // namespace vq
