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

#include <algorithm>
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

// Slab decomposition shared by the kernels below: slab (Z, Y, seg) is the p x p x 64 block
// of fine nodes i = p*Z + a, j = p*Y + b, k = 64*seg + lane (a, b < p), i.e. the fine nodes
// of 64/p consecutive network voxels along x. Lane = k offset (coalesced rows), wave w takes
// the (a, b) pairs w, w+4, ...; p divides 64.
constexpr int kSeg = 64;

struct Slab {
  int Z, Y, seg;
};

__device__ __forceinline__ Slab slab_of(int64_t slab, int R, int nseg) {
  Slab q;
  q.seg = static_cast<int>(slab % nseg);
  q.Y = static_cast<int>((slab / nseg) % R);
  q.Z = static_cast<int>(slab / (static_cast<int64_t>(nseg) * R));
  return q;
}

// Sums Q per-thread values over the block's 4 waves (LDS) and then over aligned groups of p
// lanes (shuffles): on return lane X*p of wave 0 holds the sum over the p^3 fine nodes of
// network voxel X of the slab. Contains __syncthreads(): uniform control flow only.
template <int Q>
__device__ __forceinline__ void slab_reduce(float (&t)[Q], int p, float (*s)[kSeg][Q]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < Q; ++q) s[wave][lane][q] = t[q];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) t[q] = ((s[0][lane][q] + s[1][lane][q]) + s[2][lane][q]) + s[3][lane][q];
    for (int off = 1; off < p; off <<= 1) {
#pragma unroll
      for (int q = 0; q < Q; ++q) t[q] += __shfl_down(t[q], off, 64);
    }
  }
  __syncthreads();
}

// Pooled channels [r, phi - hist, f] of every network voxel -> pooled[V][3]; float64
// partial sums of their squares (grid-stride over slabs, fixed order per launch size).
__global__ void __launch_bounds__(256) pool_input_kernel(const float* __restrict__ r32,
                                                         const float* __restrict__ phi,
                                                         const float* __restrict__ hist,
                                                         const float* __restrict__ f, int n,
                                                         int p, int R,
                                                         float* __restrict__ pooled,
                                                         double* __restrict__ partials) {
  __shared__ float s[4][kSeg][3];
  __shared__ double s_red[4 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nseg = (n + kSeg - 1) / kSeg;
  const int64_t slabs = static_cast<int64_t>(R) * R * nseg;
  const float inv = 1.0f / static_cast<float>(p * p * p);
  for (int64_t sl = blockIdx.x; sl < slabs; sl += gridDim.x) {
    const Slab q = slab_of(sl, R, nseg);
    const int k = q.seg * kSeg + lane;
    float t[3] = {0.0f, 0.0f, 0.0f};
    if (k < n) {
      for (int ab = wave; ab < p * p; ab += 4) {
        const int i = p * q.Z + ab / p, j = p * q.Y + ab % p;
        const int64_t c = (static_cast<int64_t>(i) * n + j) * n + k;
        t[0] += r32[c];
        t[1] += phi[c] - hist[c];
        t[2] += f[c];
      }
    }
    slab_reduce<3>(t, p, s);
    if (wave == 0 && k < n && lane % p == 0) {
      const int64_t V = (static_cast<int64_t>(q.Z) * R + q.Y) * R + k / p;
      for (int c = 0; c < 3; ++c) {
        const float o = t[c] * inv;
        pooled[V * 3 + c] = o;
        v[c] = v[c] + static_cast<double>(o) * static_cast<double>(o);
      }
    }
  }
  pm::block_sum<3>(v, s_red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) partials[blockIdx.x * pm::kRedMaxQ + c] = v[c];
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
  // u = (i - (p-1)/2) / p: the node's coordinate in network voxels (cell centres)
  const float u = static_cast<float>(2 * i - (p - 1)) / static_cast<float>(2 * p);
  const float f = floorf(u);
  w = u - f;
  const int a = static_cast<int>(f);
  i0 = a < 0 ? 0 : (a > R - 1 ? R - 1 : a);
  i1 = a + 1 < 0 ? 0 : (a + 1 > R - 1 ? R - 1 : a + 1);
}

// e = scales[3] * trilinear(out) on the interior, 0 on the faces; one slab per block.
__global__ void __launch_bounds__(256) apply_kernel(const float* __restrict__ out, int R, int p,
                                                    int n, const float* __restrict__ scales,
                                                    float* __restrict__ e) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nseg = (n + kSeg - 1) / kSeg;
  const Slab q = slab_of(blockIdx.x, R, nseg);
  const int k = q.seg * kSeg + lane;
  if (k >= n) return;  // no barriers below
  int x0, x1;
  float wx;
  interp_axis(k, p, R, x0, x1, wx);
  const float so = scales[3];
  for (int ab = wave; ab < p * p; ab += 4) {
    const int i = p * q.Z + ab / p, j = p * q.Y + ab % p;
    const int64_t c = (static_cast<int64_t>(i) * n + j) * n + k;
    float val = 0.0f;
    if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
      int z0, z1, y0, y1;
      float wz, wy;
      interp_axis(i, p, R, z0, z1, wz);
      interp_axis(j, p, R, y0, y1, wy);
      const float* r00 = out + (static_cast<int64_t>(z0) * R + y0) * R;
      const float* r01 = out + (static_cast<int64_t>(z0) * R + y1) * R;
      const float* r10 = out + (static_cast<int64_t>(z1) * R + y0) * R;
      const float* r11 = out + (static_cast<int64_t>(z1) * R + y1) * R;
      const float c00 = r00[x0] * (1.0f - wx) + r00[x1] * wx;
      const float c01 = r01[x0] * (1.0f - wx) + r01[x1] * wx;
      const float c10 = r10[x0] * (1.0f - wx) + r10[x1] * wx;
      const float c11 = r11[x0] * (1.0f - wx) + r11[x1] * wx;
      const float c0 = c00 * (1.0f - wy) + c01 * wy;
      const float c1 = c10 * (1.0f - wy) + c11 * wy;
      val = so * (c0 * (1.0f - wz) + c1 * wz);
    }
    e[c] = val;
  }
}

// The judge of a learned correction e. Per fine node: r' = r - lap(e) in float64 on the
// interior (0 on the faces), written to rnew (the residual of the corrected guess); per
// network voxel: the mean residual before (r) and after (r'). Float64 partial sums of
// [mean before^2, mean after^2, r'^2]. The pooled residual measures the smooth error the
// correction targets (the full-resolution residual norm is dominated by high-frequency
// components that the smoother removes anyway).
__global__ void __launch_bounds__(256) guard_kernel(const float* __restrict__ r32,
                                                    const float* __restrict__ e, int n, int p,
                                                    int R, double inv_h2,
                                                    float* __restrict__ rnew,
                                                    double* __restrict__ partials) {
  __shared__ float s[4][kSeg][2];
  __shared__ double s_red[4 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nseg = (n + kSeg - 1) / kSeg;
  const int64_t slabs = static_cast<int64_t>(R) * R * nseg;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const float inv = 1.0f / static_cast<float>(p * p * p);
  for (int64_t sl = blockIdx.x; sl < slabs; sl += gridDim.x) {
    const Slab q = slab_of(sl, R, nseg);
    const int k = q.seg * kSeg + lane;
    float t[2] = {0.0f, 0.0f};
    if (k < n) {
      for (int ab = wave; ab < p * p; ab += 4) {
        const int i = p * q.Z + ab / p, j = p * q.Y + ab % p;
        const int64_t c = (static_cast<int64_t>(i) * n + j) * n + k;
        double ra = 0.0, rb = 0.0;
        if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
          const double sum = static_cast<double>(e[c + n2]) + static_cast<double>(e[c - n2]) +
                             static_cast<double>(e[c + n]) + static_cast<double>(e[c - n]) +
                             static_cast<double>(e[c + 1]) + static_cast<double>(e[c - 1]);
          rb = static_cast<double>(r32[c]);
          ra = rb - (sum - 6.0 * static_cast<double>(e[c])) * inv_h2;
        }
        rnew[c] = static_cast<float>(ra);
        v[2] = v[2] + ra * ra;
        t[0] += static_cast<float>(rb);
        t[1] += static_cast<float>(ra);
      }
    }
    slab_reduce<2>(t, p, s);
    if (wave == 0 && k < n && lane % p == 0) {
      const double mb = static_cast<double>(t[0] * inv), ma = static_cast<double>(t[1] * inv);
      v[0] = v[0] + mb * mb;
      v[1] = v[1] + ma * ma;
    }
  }
  pm::block_sum<3>(v, s_red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) partials[blockIdx.x * pm::kRedMaxQ + c] = v[c];
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
  PM_TIMED("warm.extrapolate", (cand ? 20.0 : 16.0) * static_cast<double>(total), s);
  extrap_kernel<<<pm::blocks_for(total), pm::kBlock, 0, s>>>(phi, hist, cand, n);
  PM_CHECK_LAUNCH("extrap_kernel");
  return PM_OK;
}

pm_status make_input(const float* r32, const float* phi, const float* hist, const float* f, int n,
                     int p, double h, float* pooled, uint16_t* input, double* partials,
                     double* red, float* scales, hipStream_t s) {
  PM_REQUIRE(p >= 1 && p <= kSeg && kSeg % p == 0 && n % p == 0,
             "warm start: the pool factor must divide 64 and the grid size");
  const int R = n / p;
  const int64_t nv = static_cast<int64_t>(R) * R * R;
  const int nseg = (n + kSeg - 1) / kSeg;
  const int64_t slabs = static_cast<int64_t>(R) * R * nseg;
  const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(slabs, pm::kRedBlocks));
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
  const int R = n / p;
  const int64_t slabs = static_cast<int64_t>(R) * R * ((n + kSeg - 1) / kSeg);
  PM_TIMED("warm.apply", 4.0 * static_cast<double>(n) * n * n, s);
  apply_kernel<<<static_cast<unsigned>(slabs), pm::kBlock, 0, s>>>(out, R, p, n, scales, e);
  PM_CHECK_LAUNCH("apply_kernel");
  return PM_OK;
}

pm_status guard(const float* r32, const float* e, int n, int p, double h, float* rnew,
                double* partials, double* red, hipStream_t s) {
  const int R = n / p;
  const int64_t slabs = static_cast<int64_t>(R) * R * ((n + kSeg - 1) / kSeg);
  const unsigned b = static_cast<unsigned>(std::min<int64_t>(slabs, pm::kRedBlocks));
  {
    PM_TIMED("warm.guard", 12.0 * static_cast<double>(n) * n * n, s);
    guard_kernel<<<b, pm::kBlock, 0, s>>>(r32, e, n, p, R, 1.0 / (h * h), rnew, partials);
    PM_CHECK_LAUNCH("guard_kernel");
  }
  pm::fold_partials_kernel<3><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(b), red);
  PM_CHECK_LAUNCH("fold_partials_kernel<3>");
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
