// BF16 convolution primitives shared by the C ABI (pm_conv.hip) and the warm-start network
// (pm_wsnet.hip). Shapes and layouts are described at the top of pm_conv.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pm_common.h"

namespace pm {
namespace conv {

// float32 -> bfloat16 bits, round to nearest even (NaN stays a quiet NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

int rows_pad(int cout);        // output channels rounded up to 16 (packed weight rows)
int kpad(int cin);             // 27*cin rounded up to 32 (packed weight columns)
int store_channels(int c);     // channels rounded up to 8 (stored bf16 tensors)
int64_t packed_elems(int cout, int cin, int transpose);
int64_t wgrad_partial_floats(int cin, int cout, int r);
void wgrad_plan(int cin, int cout, int r, int* splits, int* chunks_per_split);

pm_status pack(const float* w, int cout, int cin, int transpose, uint16_t* wp, hipStream_t s);
// scratch (may be null): float32 space for split-K partial sums; with less than
// fwd_scratch_floats(cin, cout, r) the reduction is not split.
int64_t fwd_scratch_floats(int cin, int cout, int r);
pm_status fwd(const uint16_t* xa, int ca, const uint16_t* xb, int cb, int r, int up,
              const uint16_t* wp, const float* bias, int cout, int act, float slope, uint16_t* y,
              float* y32, const char* timer, hipStream_t s, float* scratch,
              int64_t scratch_floats);
pm_status wgrad(const uint16_t* xa, int ca, const uint16_t* xb, int cb, int r, int up,
                const uint16_t* dz, int cout, float* dw, float* db, int accumulate,
                float* partial, int64_t partial_floats, const char* timer, hipStream_t s);
pm_status pool2(const uint16_t* x, int r, int c, uint16_t* y, hipStream_t s);
pm_status act_bwd(const float* g1, int g1_stride, const float* g2, const uint16_t* y, float slope,
                  int r, int gc, int dz_c, uint16_t* dz, hipStream_t s);
pm_status upsum(const float* dx, int r, int cstride, int ca, float* ga, hipStream_t s);
pm_status to_bf16(const float* x, int64_t count, uint16_t* y, hipStream_t s);

}  // namespace conv
}  // namespace pm
