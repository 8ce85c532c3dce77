// Warm-start kernels of the multigrid solve: linear extrapolation of phi in time, the
// learned correction's input (pooled, normalized fine fields) and output (trilinear
// interpolation onto the fine grid), the residual check that judges it, and the training
// target written after a solve.
//
// Coarse voxel V = (Z, Y, X) of the network grid (R = n / p per axis) covers the fine nodes
// p*V .. p*V + p-1 along each axis; its value is their mean. The fine node i sits at coarse
// coordinate u = (i - (p-1)/2) / p, so interpolation back is the cell-centred trilinear
// rule with edge clamping.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pm_common.h"
#include "pm_conv.h"
#include "pm_reduce.h"
#include "pm_timing.h"
#include "pm_warm.h"

namespace {

using pm::conv::f2bf;

// phi_new = 2*phi - hist on the interior (faces keep phi's values); hist = phi (old) on
// every node; cand = phi_new if requested.
__global__ void __launch_bounds__(256) extrap_kernel(float* __restrict__ phi,
                                                     float* __restrict__ hist,
                                                     float* __restrict__ cand, int n) {
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  const int64_t total = static_cast<int64_t>(n) * n * n;
  if (c >= total) return;
  const pm::Idx3 q = pm::idx3(c, n);
  const float p = phi[c];
  float v = p;
  if (q.i >= 1 && q.i <= n - 2 && q.j >= 1 && q.j <= n - 2 && q.k >= 1 && q.k <= n - 2)
    v = 2.0f * p - hist[c];
  phi[c] = v;
  hist[c] = p;
  if (cand) cand[c] = v;
}

// Pooled channels [r, phi - hist, f] at coarse voxel V -> pooled[V][3]; float64 partial sums
// of squares per channel. Grid-stride over coarse voxels (a fixed order for a given launch
// size); a wave reads p*64 contiguous floats per fine row.
__global__ void __launch_bounds__(256) pool_input_kernel(const float* __restrict__ r32,
                                                         const float* __restrict__ phi,
                                                         const float* __restrict__ hist,
                                                         const float* __restrict__ f, int n,
                                                         int p, int R,
                                                         float* __restrict__ pooled,
                                                         double* __restrict__ partials) {
  __shared__ double s_red[4 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  for (int64_t V = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; V < nv;
       V += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int X = static_cast<int>(V % R), Y = static_cast<int>((V / R) % R),
              Z = static_cast<int>(V / (static_cast<int64_t>(R) * R));
    float sr = 0.0f, sd = 0.0f, sf = 0.0f;
    for (int a = 0; a < p; ++a)
      for (int b = 0; b < p; ++b) {
        const int64_t base = (static_cast<int64_t>(p * Z + a) * n + (p * Y + b)) * n + p * X;
        for (int c = 0; c < p; ++c) {
          sr += r32[base + c];
          sd += phi[base + c] - hist[base + c];
          sf += f[base + c];
        }
      }
    const float inv = 1.0f / static_cast<float>(p * p * p);
    const float o[3] = {sr * inv, sd * inv, sf * inv};
    for (int q = 0; q < 3; ++q) {
      pooled[V * 3 + q] = o[q];
      v[q] = v[q] + static_cast<double>(o[q]) * static_cast<double>(o[q]);
    }
  }
  pm::block_sum<3>(v, s_red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) partials[blockIdx.x * pm::kRedMaxQ + q] = v[q];
}

// scales[q] = RMS of pooled channel q (q < 3); scales[3] = RMS(r) * (p*h)^2, the output
// scale; scales[4] = 1 / scales[3] (0 when the residual vanishes).
__global__ void scales_kernel(const double* __restrict__ sums, int64_t nv, double ph2,
                              float* __restrict__ scales) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = 0; q < 3; ++q) scales[q] = static_cast<float>(std::sqrt(sums[q] / nv));
  const double so = std::sqrt(sums[0] / nv) * ph2;
  scales[3] = static_cast<float>(so);
  scales[4] = so > 0.0 ? static_cast<float>(1.0 / so) : 0.0f;
}

// input[V][0..7] = bf16(pooled[V][q] / scales[q]) for q < 3, 0 for the pad channels.
__global__ void __launch_bounds__(256) normalize_kernel(const float* __restrict__ pooled,
                                                        int64_t nv,
                                                        const float* __restrict__ scales,
                                                        uint16_t* __restrict__ input) {
  const int64_t V = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (V >= nv) return;
  uint16_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    const float s = scales[q];
    h[q] = f2bf(s > 0.0f ? pooled[V * 3 + q] / s : 0.0f);
  }
  uint4 w;
  w.x = h[0] | (static_cast<uint32_t>(h[1]) << 16);
  w.y = h[2] | (static_cast<uint32_t>(h[3]) << 16);
  w.z = h[4] | (static_cast<uint32_t>(h[5]) << 16);
  w.w = h[6] | (static_cast<uint32_t>(h[7]) << 16);
  *reinterpret_cast<uint4*>(input + V * 8) = w;
}

__device__ __forceinline__ void interp_axis(int i, int p, int R, int& i0, int& i1, float& w) {
  const float u = (static_cast<float>(i) - 0.5f * static_cast<float>(p - 1)) / static_cast<float>(p);
  const float f = floorf(u);
  w = u - f;
  const int a = static_cast<int>(f);
  i0 = a < 0 ? 0 : (a > R - 1 ? R - 1 : a);
  i1 = a + 1 < 0 ? 0 : (a + 1 > R - 1 ? R - 1 : a + 1);
}

// e[c] = scales[3] * trilinear(out)(node c) on the interior, 0 on the faces.
__global__ void __launch_bounds__(256) apply_kernel(const float* __restrict__ out, int R, int p,
                                                    int n, const float* __restrict__ scales,
                                                    float* __restrict__ e) {
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  const int64_t total = static_cast<int64_t>(n) * n * n;
  if (c >= total) return;
  const pm::Idx3 q = pm::idx3(c, n);
  float val = 0.0f;
  if (q.i >= 1 && q.i <= n - 2 && q.j >= 1 && q.j <= n - 2 && q.k >= 1 && q.k <= n - 2) {
    int z0, z1, y0, y1, x0, x1;
    float wz, wy, wx;
    interp_axis(q.i, p, R, z0, z1, wz);
    interp_axis(q.j, p, R, y0, y1, wy);
    interp_axis(q.k, p, R, x0, x1, wx);
    auto at = [&](int z, int y, int x) {
      return out[(static_cast<int64_t>(z) * R + y) * R + x];
    };
    const float c00 = at(z0, y0, x0) * (1.0f - wx) + at(z0, y0, x1) * wx;
    const float c01 = at(z0, y1, x0) * (1.0f - wx) + at(z0, y1, x1) * wx;
    const float c10 = at(z1, y0, x0) * (1.0f - wx) + at(z1, y0, x1) * wx;
    const float c11 = at(z1, y1, x0) * (1.0f - wx) + at(z1, y1, x1) * wx;
    const float c0 = c00 * (1.0f - wy) + c01 * wy;
    const float c1 = c10 * (1.0f - wy) + c11 * wy;
    val = scales[3] * (c0 * (1.0f - wz) + c1 * wz);
  }
  e[c] = val;
}

// The judge of a learned correction e: per network voxel V, the mean over its fine nodes
// of the residual before (r) and after (r - lap(e)) the correction, evaluated in float64 on
// the interior; float64 partial sums of the squares of both means. The pooled residual
// measures the smooth error the correction targets (the full-resolution residual norm is
// dominated by high-frequency components that the smoother removes anyway).
__global__ void __launch_bounds__(256) guard_kernel(const float* __restrict__ r32,
                                                    const float* __restrict__ e, int n, int p,
                                                    int R, double inv_h2,
                                                    double* __restrict__ partials) {
  __shared__ double s_red[4 * 2];
  double v[2] = {0.0, 0.0};
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  for (int64_t V = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; V < nv;
       V += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int X = static_cast<int>(V % R), Y = static_cast<int>((V / R) % R),
              Z = static_cast<int>(V / (static_cast<int64_t>(R) * R));
    double sb = 0.0, sa = 0.0;
    for (int a = 0; a < p; ++a) {
      const int i = p * Z + a;
      for (int b = 0; b < p; ++b) {
        const int j = p * Y + b;
        for (int c = 0; c < p; ++c) {
          const int k = p * X + c;
          if (i < 1 || i > n - 2 || j < 1 || j > n - 2 || k < 1 || k > n - 2) continue;
          const int64_t q = (static_cast<int64_t>(i) * n + j) * n + k;
          const double sum = static_cast<double>(e[q + n2]) + static_cast<double>(e[q - n2]) +
                             static_cast<double>(e[q + n]) + static_cast<double>(e[q - n]) +
                             static_cast<double>(e[q + 1]) + static_cast<double>(e[q - 1]);
          const double r = static_cast<double>(r32[q]);
          sb += r;
          sa += r - (sum - 6.0 * static_cast<double>(e[q])) * inv_h2;
        }
      }
    }
    const double inv = 1.0 / (static_cast<double>(p) * p * p);
    sb *= inv;
    sa *= inv;
    v[0] = v[0] + sb * sb;
    v[1] = v[1] + sa * sa;
  }
  pm::block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * pm::kRedMaxQ] = v[0];
    partials[blockIdx.x * pm::kRedMaxQ + 1] = v[1];
  }
}

// tgt[V] = mean over V's fine nodes of (phi - cand), times scales[4].
__global__ void __launch_bounds__(256) target_kernel(const float* __restrict__ phi,
                                                     const float* __restrict__ cand, int n, int p,
                                                     int R, const float* __restrict__ scales,
                                                     float* __restrict__ tgt) {
  const int64_t V = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  if (V >= nv) return;
  const int X = static_cast<int>(V % R), Y = static_cast<int>((V / R) % R),
            Z = static_cast<int>(V / (static_cast<int64_t>(R) * R));
  float s = 0.0f;
  for (int a = 0; a < p; ++a)
    for (int b = 0; b < p; ++b) {
      const int64_t base = (static_cast<int64_t>(p * Z + a) * n + (p * Y + b)) * n + p * X;
      for (int c = 0; c < p; ++c) s += phi[base + c] - cand[base + c];
    }
  tgt[V] = s / static_cast<float>(p * p * p) * scales[4];
}

}  // namespace

namespace pm {
namespace warm {

pm_status extrapolate(float* phi, float* hist, float* cand, int n, hipStream_t s) {
  const int64_t total = static_cast<int64_t>(n) * n * n;
  PM_TIMED("warm.extrapolate", (cand ? 16.0 : 12.0) * static_cast<double>(total), s);
  extrap_kernel<<<pm::blocks_for(total), pm::kBlock, 0, s>>>(phi, hist, cand, n);
  PM_CHECK_LAUNCH("extrap_kernel");
  return PM_OK;
}

pm_status make_input(const float* r32, const float* phi, const float* hist, const float* f, int n,
                     int p, double h, float* pooled, uint16_t* input, double* partials,
                     double* red, float* scales, hipStream_t s) {
  PM_REQUIRE(p >= 1 && n % p == 0, "warm start: grid size must be a multiple of the pool factor");
  const int R = n / p;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  const unsigned blocks = pm::red_blocks(nv);
  {
    PM_TIMED("warm.input", 16.0 * static_cast<double>(n) * n * n, s);
    pool_input_kernel<<<blocks, pm::kBlock, 0, s>>>(r32, phi, hist, f, n, p, R, pooled, partials);
    PM_CHECK_LAUNCH("pool_input_kernel");
    pm::fold_partials_kernel<3><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(blocks), red);
    PM_CHECK_LAUNCH("fold_partials_kernel<3>");
    scales_kernel<<<1, 64, 0, s>>>(red, nv, (p * h) * (p * h), scales);
    PM_CHECK_LAUNCH("scales_kernel");
    normalize_kernel<<<pm::blocks_for(nv), pm::kBlock, 0, s>>>(pooled, nv, scales, input);
    PM_CHECK_LAUNCH("normalize_kernel");
  }
  return PM_OK;
}

pm_status apply(const float* out, int n, int p, const float* scales, float* e, hipStream_t s) {
  const int64_t total = static_cast<int64_t>(n) * n * n;
  PM_TIMED("warm.apply", 4.0 * static_cast<double>(total), s);
  apply_kernel<<<pm::blocks_for(total), pm::kBlock, 0, s>>>(out, n / p, p, n, scales, e);
  PM_CHECK_LAUNCH("apply_kernel");
  return PM_OK;
}

pm_status guard(const float* r32, const float* e, int n, int p, double h, double* partials,
                double* red, hipStream_t s) {
  const int R = n / p;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  const unsigned b = pm::red_blocks(nv);
  {
    PM_TIMED("warm.guard", 8.0 * static_cast<double>(n) * n * n, s);
    guard_kernel<<<b, pm::kBlock, 0, s>>>(r32, e, n, p, R, 1.0 / (h * h), partials);
    PM_CHECK_LAUNCH("guard_kernel");
  }
  pm::fold_partials_kernel<2><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(b), red);
  PM_CHECK_LAUNCH("fold_partials_kernel<2>");
  return PM_OK;
}

pm_status target(const float* phi, const float* cand, int n, int p, const float* scales, float* tgt,
                 hipStream_t s) {
  const int R = n / p;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  PM_TIMED("warm.target", 8.0 * static_cast<double>(n) * n * n, s);
  target_kernel<<<pm::blocks_for(nv), pm::kBlock, 0, s>>>(phi, cand, n, p, R, scales, tgt);
  PM_CHECK_LAUNCH("target_kernel");
  return PM_OK;
}

}  // namespace warm
}  // namespace pm
