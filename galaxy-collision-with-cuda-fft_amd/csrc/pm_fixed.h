// 64-bit fixed-point corner terms of the order-independent deposits (PM_DEPOSIT_FIXED and
// PM_DEPOSIT_BRICKED).
//
// A corner term t (a positive finite float32) becomes q = round-half-even(t * 2^s), the
// value llrint(ldexp((double)t, s)) gives in the default rounding mode, and node sums of
// q are exact. Scale: with m'_max in [2^a, 2^(a+1)) (largest finite positive
// m' = mass/dx^3) and N < 2^b particles, s = 61 - a - b keeps every q below 2^(62-b) and
// every node sum below 2^62 (a particle adds to a node at most once).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

namespace pm {

// s for the largest m' given as its float bits (0 when there is none) and b.
__host__ __device__ inline int fixed_scale(unsigned int maxbits, int b) {
  if (maxbits == 0u) return 0;
  float v;
#if defined(__HIP_DEVICE_COMPILE__)
  v = __uint_as_float(maxbits);
#else
  std::memcpy(&v, &maxbits, sizeof v);
#endif
  int e;
  (void)frexp(static_cast<double>(v), &e);  // v in [2^(e-1), 2^e)
  return 61 - (e - 1) - b;
}

// round-half-even(t * 2^s) for a positive finite float t whose scaled value is below 2^62,
// in integer arithmetic: t = mant * 2^E exactly (mant < 2^24), so t * 2^s = mant * 2^(E+s);
// for a negative exponent the dropped bits decide the rounding against one half.
__host__ __device__ inline uint64_t fixed_q(float t, int s) {
  uint32_t u;
#if defined(__HIP_DEVICE_COMPILE__)
  u = __float_as_uint(t);
#else
  std::memcpy(&u, &t, sizeof u);
#endif
  const uint32_t e = u >> 23;
  const uint32_t mant = e ? ((u & 0x7fffffu) | 0x800000u) : (u & 0x7fffffu);
  const int sh = (e ? static_cast<int>(e) - 150 : -149) + s;
  if (sh >= 0) return static_cast<uint64_t>(mant) << sh;
  const int r = -sh;
  if (r > 24) return 0ull;  // mant < 2^24 <= 2^(r-1): below one half
  const uint32_t q = mant >> r;
  const uint32_t rem = mant & ((1u << r) - 1u);
  const uint32_t half = 1u << (r - 1);
  return static_cast<uint64_t>(q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u));
}

// Number of bits of v >= 0 (b with v < 2^b).
inline int bit_length(int64_t v) {
  int b = 0;
  while (v > 0) {
    ++b;
    v >>= 1;
  }
  return b;
}

}  // namespace pm
