// Cloud-in-cell geometry shared by the deposit and gather kernels.
//
// A particle at position x sits at grid coordinate g = x/dx; its base cell is
// ix = floor(g) and its fractional offset f = g - ix (reference deposit.py:62-70, 101-109).
// Corner c of the cell, c in [0, 8) read as bits (cx, cy, cz) = (c>>2, (c>>1)&1, c&1),
// is node (ix+cx, iy+cy, iz+cz) with weight wx*wy*wz, w = f for the +1 corner and 1-f
// for the base corner. Out-of-box corners are dropped.
//
// All arithmetic is single rounding per operation (the library is built with
// -ffp-contract=off): the division x/dx is IEEE correctly rounded, so the weights are
// bitwise equal to numpy float32 arithmetic with the same operation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm {

struct Cic {
  float fx, fy, fz;
  int ix, iy, iz;
};

// g = x / dx, the IEEE quotient, with the reciprocal of dx refined once per thread: the
// steps of the compiler's float32 division (reciprocal, one Newton step, quotient and two
// remainder corrections by fma) without its operand range scaling, which never applies
// when dx lies in [2^-30, 2^30] and 2^-60 <= |x| < 2^60 (no operand, remainder or quotient
// near the under- or overflow range), so the result is the IEEE quotient there. Larger |x|
// give |g| >= 2^29, infinity or NaN, outside the grid like the IEEE quotient; for smaller
// |x| (and dx outside that range) callers take the division itself.
__device__ __forceinline__ float quot(float x, float dx, float r1) {
  const float q0 = x * r1;
  const float e1 = fmaf(-dx, q0, x);
  const float q1 = fmaf(e1, r1, q0);
  const float e2 = fmaf(-dx, q1, x);
  return fmaf(e2, r1, q1);
}

// The refined reciprocal quot() takes, and whether dx lies in its range.
__device__ __forceinline__ float quot_recip(float dx) {
  const float r0 = __builtin_amdgcn_rcpf(dx);
  return fmaf(fmaf(-dx, r0, 1.0f), r0, r0);
}
__device__ __forceinline__ bool quot_dx_ok(float dx) { return dx >= 0x1p-30f && dx <= 0x1p30f; }

// Base cell and offsets of one coordinate. Returns false when no corner of the cell can
// lie in [0, n) along this axis, i.e. unless -1 <= g < n (NaN also returns false); this
// keeps the float->int conversion in range for any input.
__device__ __forceinline__ bool cic_axis_g(float g, int n, int& i, float& f) {
  if (!(g >= -1.0f && g < static_cast<float>(n))) return false;
  const float fl = floorf(g);
  i = static_cast<int>(fl);
  f = g - static_cast<float>(i);
  return true;
}
__device__ __forceinline__ bool cic_axis(float x, float dx, int n, int& i, float& f) {
  return cic_axis_g(x / dx, n, i, f);
}

// Extended-cell key: cells with base index -1..n-1 per axis shifted by +1, C order.
// Keys lie in [0, (n+1)^3); the value (n+1)^3 marks a particle with no in-box corner.
// Edge (in nodes) of the node bricks the deposit marks for the gradient and assembly passes.
constexpr int kNodeBrick = 8;

__device__ __forceinline__ uint32_t cell_key(int ix, int iy, int iz, int n) {
  const uint32_t m = static_cast<uint32_t>(n) + 1u;
  return (static_cast<uint32_t>(ix + 1) * m + static_cast<uint32_t>(iy + 1)) * m +
         static_cast<uint32_t>(iz + 1);
}

}  // namespace pm
