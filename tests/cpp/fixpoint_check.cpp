// Host build of vexq_amd/csrc/fixpoint.h for the CPU test suite: the same
// chunking and rounding code that the join kernels run, driven from ctypes.
#include <stdint.h>

#include "../../vexq_amd/csrc/fixpoint.h"

extern "C" {

// The three chunks of t on scale (lsb, w); returns 0 for a zero addend.
int vqfx_chunks(double t, int32_t lsb, int32_t w, int64_t* c) {
  vq::FixScale fs{lsb, w};
  return vq::fix_chunks(t, fs, c[0], c[1], c[2]) ? 1 : 0;
}

double vqfx_to_f64(int64_t l0, int64_t l1, int64_t l2, int32_t lsb, int32_t w, uint64_t mult) {
  return vq::fix_to_f64(l0, l1, l2, vq::FixScale{lsb, w}, mult);
}

// Sum of x[0..n) the way a join accumulator does it: chunk each addend,
// add the chunks into three wrapping int64 limbs (in the given order), round
// once at the end.
double vqfx_sum(const double* x, int64_t n, int32_t lsb, int32_t w, uint64_t mult) {
  vq::FixScale fs{lsb, w};
  uint64_t l0 = 0, l1 = 0, l2 = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t c0, c1, c2;
    if (!vq::fix_chunks(x[i], fs, c0, c1, c2)) continue;
    l0 += (uint64_t)c0;
    l1 += (uint64_t)c1;
    l2 += (uint64_t)c2;
  }
  return vq::fix_to_f64((int64_t)l0, (int64_t)l1, (int64_t)l2, fs, mult);
}

}  // extern "C"
