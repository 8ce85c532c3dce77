// Shared helpers for the pmchain HIP sources: error state, launch checks, grid math.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "pmchain.h"

namespace pm {

// Thread-local last-error message (read back through pm_last_error()).
void set_error(const std::string& msg);
void clear_error();

inline pm_status invalid(const std::string& msg) {
  set_error(msg);
  return PM_ERR_INVALID;
}

inline pm_status hip_fail(const char* what, hipError_t e) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP;
}

constexpr int kBlock = 256;  // threads per workgroup for the streaming kernels

inline unsigned blocks_for(int64_t work, int per_block = kBlock) {
  return static_cast<unsigned>((work + per_block - 1) / per_block);
}

// Grid-stride launch size for streaming kernels: enough workgroups to fill 256 CUs
// several times over without one workgroup per element for very large arrays.
inline unsigned stream_blocks(int64_t work, int per_block = kBlock) {
  int64_t b = (work + per_block - 1) / per_block;
  const int64_t cap = 256 * 64;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

// Logical workgroup index for 1-D launches, arranged so that the workgroups one XCD runs
// cover one contiguous range of the index space. The hardware hands workgroup w to XCD
// w % 8; with the identity mapping a stencil's neighbouring rows would sit in other XCDs'
// L2 caches and be fetched again from HBM. The mapping is a bijection on [0, gridDim.x),
// so it changes only which XCD does what, never the result.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned nb = gridDim.x, w = blockIdx.x;
  const unsigned per = nb >> 3, rem = nb & 7u, x = w & 7u;
  return static_cast<int64_t>(x * per + (x < rem ? x : rem) + (w >> 3));
}

// Contiguous element range [begin, end) of this workgroup when `count` elements are split
// into gridDim.x chunks (multiples of 256), in XCD-aware order.
__device__ __forceinline__ void block_range(int64_t count, int64_t& begin, int64_t& end) {
  const int64_t nb = gridDim.x;
  const int64_t chunk = ((count + nb - 1) / nb + 255) / 256 * 256;
  begin = xcd_block() * chunk;
  end = begin + chunk < count ? begin + chunk : count;
}

// (i, j, k) of flat node index c on an n^3 grid (k fastest) with 32-bit arithmetic; every
// grid entry point limits n to 1024, so n^3 < 2^32.
struct Idx3 {
  int i, j, k;
};
__device__ __forceinline__ Idx3 idx3(int64_t c, int n) {
  const uint32_t u = static_cast<uint32_t>(c), un = static_cast<uint32_t>(n);
  const uint32_t q = u / un;
  Idx3 r;
  r.k = static_cast<int>(u - q * un);
  r.j = static_cast<int>(q % un);
  r.i = static_cast<int>(q / un);
  return r;
}

// Correctly rounded float32 square root. The hardware estimate (v_sqrt_f32) is within one
// ulp; each candidate is then checked against the exact midpoints to its neighbours: a
// midpoint of two floats has at most 25 significant bits and its square at most 50, so both
// are exact in float64 and the comparison with x decides the rounding exactly (no ties are
// possible). Two correction rounds cover an estimate up to two ulps off.
__device__ __forceinline__ float sqrt_rn(float x) {
  float s = __builtin_sqrtf(x);
  if (!(x > 0.0f) || !(x < __builtin_huge_valf())) return s;
  const double xd = static_cast<double>(x);
#pragma unroll
  for (int round = 0; round < 2; ++round) {
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const double mlo = (static_cast<double>(lo) + static_cast<double>(s)) * 0.5;
    if (xd < mlo * mlo) {
      s = lo;
      continue;
    }
    const float hi = __int_as_float(__float_as_int(s) + 1);
    const double mhi = (static_cast<double>(s) + static_cast<double>(hi)) * 0.5;
    if (xd > mhi * mhi) s = hi;
  }
  return s;
}

// Exact x / d for 32-bit unsigned x and a divisor fixed for a launch (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI 1994, Fig. 4.1):
// with l = ceil(log2 d) and M = floor(2^32 (2^l - d) / d) + 1 (< 2^32),
// x / d = (t + ((x - t) >> min(l, 1))) >> max(l - 1, 0), t = mulhi(M, x). A mulhi and a few
// adds and shifts instead of the ~15-instruction division sequence; t <= x, so the sum
// cannot overflow.
struct FastDiv {
  uint32_t d, mul;
  int s1, s2;
};

inline FastDiv make_fast_div(uint32_t d) {
  int l = 0;
  while ((uint64_t{1} << l) < d) ++l;
  FastDiv f;
  f.d = d;
  f.mul = static_cast<uint32_t>(((uint64_t{1} << 32) * ((uint64_t{1} << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

__device__ __forceinline__ uint32_t fast_div(uint32_t x, const FastDiv& f) {
  const uint32_t t = __umulhi(f.mul, x);
  return (t + ((x - t) >> f.s1)) >> f.s2;
}

// idx3 with the divisions by n taken by a FastDiv of n.
__device__ __forceinline__ Idx3 idx3(int64_t c, const FastDiv& fn) {
  const uint32_t u = static_cast<uint32_t>(c), un = fn.d;
  const uint32_t q = fast_div(u, fn);
  const uint32_t i = fast_div(q, fn);
  Idx3 r;
  r.k = static_cast<int>(u - q * un);
  r.j = static_cast<int>(q - i * un);
  r.i = static_cast<int>(i);
  return r;
}

}  // namespace pm

#define PM_TRY_HIP(expr)                                   \
  do {                                                     \
    hipError_t pm_e_ = (expr);                             \
    if (pm_e_ != hipSuccess) return pm::hip_fail(#expr, pm_e_); \
  } while (0)

// Check the launch that was just enqueued.
#define PM_CHECK_LAUNCH(name)                              \
  do {                                                     \
    hipError_t pm_e_ = hipGetLastError();                  \
    if (pm_e_ != hipSuccess) return pm::hip_fail(name, pm_e_); \
  } while (0)

#define PM_REQUIRE(cond, msg)                              \
  do {                                                     \
    if (!(cond)) return pm::invalid(msg);                  \
  } while (0)
