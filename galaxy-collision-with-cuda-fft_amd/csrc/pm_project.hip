// Device 2D mass projection (reference viz/project.py:26-83): nearest-bin (NGP) mass per
// bin over two in-plane axes, the third projected out; out-of-extent particles dropped.
//
// The reference scatters with float atomics, so its sums depend on the arrival order. Here
// each particle's mass is converted to a fixed-point integer (scale 2^s chosen from the
// largest mass so that any sum of up to 2^28 particles fits in 63 bits) and accumulated with
// 64-bit integer atomics, which commute: the image is identical run to run, and the only
// error is the quantization of each mass (relative <= 2^-34 of the largest mass).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pm_common.h"
#include "pm_timing.h"

namespace {

// bit pattern of the largest mass (masses are non-negative, so their float bits order like
// the values); one atomic per workgroup
__global__ void __launch_bounds__(256) max_mass_kernel(const float* __restrict__ mass, int64_t n,
                                                       unsigned int* __restrict__ out) {
  __shared__ unsigned int s[256];
  unsigned int m = 0u;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float v = mass[i];
    const unsigned int b = v > 0.0f ? __float_as_uint(v) : 0u;
    m = b > m ? b : m;
  }
  s[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off && s[threadIdx.x + off] > s[threadIdx.x])
      s[threadIdx.x] = s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, s[0]);
}

__global__ void __launch_bounds__(256) zero_u64_kernel(unsigned long long* __restrict__ p,
                                                       int64_t count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

// scale exponent: largest mass * 2^s < 2^34 (s from the max mass's binary exponent)
__device__ __forceinline__ int scale_exp(unsigned int maxbits) {
  if (maxbits == 0u) return 0;
  const int e = static_cast<int>((maxbits >> 23) & 0xffu) - 127;  // max in [2^e, 2^(e+1))
  return 33 - e;
}

__global__ void __launch_bounds__(256) project_kernel(const float* __restrict__ pa,
                                                      const float* __restrict__ pb,
                                                      const float* __restrict__ mass, int64_t n,
                                                      float lo_a, float lo_b, float binw_a,
                                                      float binw_b, int bins,
                                                      const unsigned int* __restrict__ maxbits,
                                                      unsigned long long* __restrict__ acc) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float fa = floorf((pa[i] - lo_a) / binw_a);
  const float fb = floorf((pb[i] - lo_b) / binw_b);
  // the range test on the float value also rejects NaN and values beyond int range
  if (!(fa >= 0.0f && fa < static_cast<float>(bins) && fb >= 0.0f && fb < static_cast<float>(bins)))
    return;
  const int ia = static_cast<int>(fa), ib = static_cast<int>(fb);
  const double q = ldexp(static_cast<double>(mass[i]), scale_exp(*maxbits));
  if (!(q > 0.0 && q < 1.8e16)) return;  // zero, negative or non-finite masses add nothing
  atomicAdd(&acc[static_cast<int64_t>(ia) * bins + ib],
            static_cast<unsigned long long>(llrint(q)));
}

__global__ void __launch_bounds__(256) unscale_kernel(const unsigned long long* __restrict__ acc,
                                                      int64_t count,
                                                      const unsigned int* __restrict__ maxbits,
                                                      float* __restrict__ img) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count)
    img[i] = static_cast<float>(ldexp(static_cast<double>(acc[i]), -scale_exp(*maxbits)));
}

}  // namespace

extern "C" {

pm_status pm_project_mass(const float* pos_a, const float* pos_b, const float* mass, int64_t n,
                          float lo_a, float lo_b, float binw_a, float binw_b, int32_t bins,
                          float* img, void* scratch, hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(img && scratch, "pm_project_mass: null output or scratch");
  PM_REQUIRE(n >= 0 && n <= (int64_t(1) << 28), "pm_project_mass: particle count out of range");
  PM_REQUIRE(n == 0 || (pos_a && pos_b && mass), "pm_project_mass: null particle arrays");
  PM_REQUIRE(bins >= 1 && bins <= 8192, "pm_project_mass: bins must be in [1, 8192]");
  PM_REQUIRE(binw_a > 0.0f && binw_b > 0.0f, "pm_project_mass: bin widths must be positive");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(scratch) & 7u) == 0,
             "pm_project_mass: scratch must be 8-byte aligned");
  const int64_t cells = static_cast<int64_t>(bins) * bins;
  auto* acc = static_cast<unsigned long long*>(scratch);
  auto* maxbits = reinterpret_cast<unsigned int*>(acc + cells);
  PM_TIMED("project", 16.0 * static_cast<double>(n) + 12.0 * static_cast<double>(cells), s);
  zero_u64_kernel<<<pm::blocks_for(cells + 1), 256, 0, s>>>(acc, cells + 1);
  PM_CHECK_LAUNCH("zero_u64_kernel");
  if (n > 0) {
    max_mass_kernel<<<pm::stream_blocks(n), 256, 0, s>>>(mass, n, maxbits);
    PM_CHECK_LAUNCH("max_mass_kernel");
    project_kernel<<<pm::blocks_for(n), 256, 0, s>>>(pos_a, pos_b, mass, n, lo_a, lo_b, binw_a,
                                                     binw_b, bins, maxbits, acc);
    PM_CHECK_LAUNCH("project_kernel");
  }
  unscale_kernel<<<pm::blocks_for(cells), 256, 0, s>>>(acc, cells, maxbits, img);
  PM_CHECK_LAUNCH("unscale_kernel");
  return PM_OK;
}

pm_status pm_project_scratch_bytes(int32_t bins, int64_t* out) {
  PM_REQUIRE(out && bins >= 1 && bins <= 8192, "pm_project_scratch_bytes: bad argument");
  *out = (static_cast<int64_t>(bins) * bins + 1) * 8;
  return PM_OK;
}

}  // extern "C"
