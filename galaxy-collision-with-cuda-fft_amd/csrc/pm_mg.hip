// Open-boundary geometric multigrid for lap(phi) = 4*pi*G*rho on the node grid
// (reference solver/multigrid.py:46-494).
//
// PLAIN mode follows the reference solve step for step: multipole Dirichlet faces from
// float64 moments, red-black Gauss-Seidel V(nu1,nu2) cycles with coarse J <-> fine 2J,
// full-weighting restriction, trilinear prolongation, coarse_sweeps on the coarsest level,
// and an adaptive stop on ||r|| < tol*||rhs||. Each kernel keeps the reference's
// per-node float32 operation order, so single passes are bitwise equal to a numpy
// restatement.
//
// MIXED mode is defect correction (mixed-precision iterative refinement): the residual
// r = f - A phi is evaluated in float64 arithmetic from the float32 iterate; each outer
// iteration solves A e = r with one float32 V-cycle (float64 coarsest solve) on an aligned
// hierarchy and adds e into phi. The float32 iterate bounds the attainable error near the
// float32 resolution of phi (~1e-7 relative), far below the 1e-5 target. f is the same
// float32 4*pi*G*rho as PLAIN mode and the faces are the same float32 values, so both
// modes solve the identical discrete problem.
#include <cmath>
#include <utility>
#include <vector>

#include <cstdio>

#include "pm_reduce.h"
#include "pm_timing.h"
#include "pm_warm.h"

namespace {

constexpr int kCoarseLocal = 13;  // coarsest levels up to this size run in one workgroup (LDS)
constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------------------
// float32 level kernels
// ---------------------------------------------------------------------------------------

// One colour of red-black Gauss-Seidel (multigrid.py:46-59). Thread (i, j, kk) owns the
// node of colour `color` among k = 2kk, 2kk+1.
__global__ void __launch_bounds__(256) rb_pass_kernel(float* __restrict__ phi,
                                                      const float* __restrict__ rhs, int n,
                                                      float h2, int color) {
  const int half = (n + 1) >> 1;
  const int64_t total = static_cast<int64_t>(n) * n * half;
  const int64_t t = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const uint32_t u = static_cast<uint32_t>(t), q = u / static_cast<uint32_t>(half);
  const int kk = static_cast<int>(u - q * static_cast<uint32_t>(half));
  const int j = static_cast<int>(q % static_cast<uint32_t>(n));
  const int i = static_cast<int>(q / static_cast<uint32_t>(n));
  const int k = 2 * kk + ((i + j + color) & 1);
  if (i < 1 || i > n - 2 || j < 1 || j > n - 2 || k < 1 || k > n - 2) return;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t c = i * n2 + static_cast<int64_t>(j) * n + k;
  const float s = phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] + phi[c - 1];
  phi[c] = (s - h2 * rhs[c]) / 6.0f;
}

// One full red-black sweep (colour 0 then colour 1) in a single pass over the level, out of
// place: phi_in is only read and every node of phi_out is written, so workgroups never see
// each other's updates. A workgroup owns a (kSwTJ x kSwTK) tile of a (j,k) plane and marches
// along i through a segment of kSwSeg planes, keeping four planes of phi (tile + 2-node ring) and two planes
// of rhs (tile + 1-node ring) in LDS. Plane p+2 is loaded, the red nodes of plane p+1 are
// updated on the tile and its 1-node ring (from old black values), then the black nodes of
// plane p are updated on the tile (from new red values) and plane p is written back. Red
// nodes on a ring or at a segment's first plane are recomputed by the neighbouring
// workgroup from the same inputs, so each node gets exactly the values of two separate
// colour passes (rb_pass_kernel), bit for bit.
constexpr int kSwTJ = 32, kSwTK = 64, kSwSeg = 16;
constexpr int kSwRJ = kSwTJ + 4, kSwRK = kSwTK + 4;  // ring-2 region
constexpr int kSwHJ = kSwTJ + 2, kSwHK = kSwTK + 2;  // ring-1 region

__global__ void __launch_bounds__(256) rb_sweep_kernel(const float* __restrict__ phi,
                                                       float* __restrict__ phi_out,
                                                       const float* __restrict__ rhs, int n,
                                                       float h2) {
  __shared__ float s_phi[4][kSwRJ * kSwRK];
  __shared__ float s_rhs[2][kSwHJ * kSwHK];
  const int k0 = blockIdx.x * kSwTK;
  const int j0 = blockIdx.y * kSwTJ;
  const int i0 = 1 + static_cast<int>(blockIdx.z) * kSwSeg;
  if (i0 > n - 2) return;  // uniform across the workgroup
  const int i1 = i0 + kSwSeg < n - 1 ? i0 + kSwSeg : n - 1;  // exclusive
  const int64_t n2 = static_cast<int64_t>(n) * n;

  auto load_phi = [&](int plane) {
    float* dst = s_phi[plane & 3];
    const float* src = phi + plane * n2;
    for (int t = threadIdx.x; t < kSwRJ * kSwRK; t += blockDim.x) {
      const int j = j0 - 2 + t / kSwRK, k = k0 - 2 + t % kSwRK;
      dst[t] = (j >= 0 && j < n && k >= 0 && k < n) ? src[static_cast<int64_t>(j) * n + k] : 0.0f;
    }
  };
  auto load_rhs = [&](int plane) {
    float* dst = s_rhs[plane & 1];
    const float* src = rhs + plane * n2;
    for (int t = threadIdx.x; t < kSwHJ * kSwHK; t += blockDim.x) {
      const int j = j0 - 1 + t / kSwHK, k = k0 - 1 + t % kSwHK;
      dst[t] = (j >= 0 && j < n && k >= 0 && k < n) ? src[static_cast<int64_t>(j) * n + k] : 0.0f;
    }
  };
  // colour 0 nodes of `plane` on the tile and its 1-node ring
  auto red = [&](int plane) {
    if (plane < 1 || plane > n - 2) return;
    float* c = s_phi[plane & 3];
    const float* dn = s_phi[(plane - 1) & 3];
    const float* up = s_phi[(plane + 1) & 3];
    const float* r = s_rhs[plane & 1];
    for (int t = threadIdx.x; t < kSwHJ * kSwHK; t += blockDim.x) {
      const int hj = t / kSwHK, hk = t % kSwHK;
      const int j = j0 - 1 + hj, k = k0 - 1 + hk;
      if (j < 1 || j > n - 2 || k < 1 || k > n - 2 || ((plane + j + k) & 1) != 0) continue;
      const int q = (hj + 1) * kSwRK + (hk + 1);
      const float sum = up[q] + dn[q] + c[q + kSwRK] + c[q - kSwRK] + c[q + 1] + c[q - 1];
      c[q] = (sum - h2 * r[t]) / 6.0f;
    }
  };
  // colour 1 nodes of `plane` on the tile
  auto black = [&](int plane) {
    float* c = s_phi[plane & 3];
    const float* dn = s_phi[(plane - 1) & 3];
    const float* up = s_phi[(plane + 1) & 3];
    const float* r = s_rhs[plane & 1];
    for (int t = threadIdx.x; t < kSwTJ * kSwTK; t += blockDim.x) {
      const int tj = t / kSwTK, tk = t % kSwTK;
      const int j = j0 + tj, k = k0 + tk;
      if (j < 1 || j > n - 2 || k < 1 || k > n - 2 || ((plane + j + k) & 1) != 1) continue;
      const int q = (tj + 2) * kSwRK + (tk + 2);
      const float sum = up[q] + dn[q] + c[q + kSwRK] + c[q - kSwRK] + c[q + 1] + c[q - 1];
      c[q] = (sum - h2 * r[(tj + 1) * kSwHK + (tk + 1)]) / 6.0f;
    }
  };
  // every node of the tile (boundary nodes pass through unchanged)
  auto store = [&](int plane) {
    const float* c = s_phi[plane & 3];
    float* dst = phi_out + plane * n2;
    for (int t = threadIdx.x; t < kSwTJ * kSwTK; t += blockDim.x) {
      const int tj = t / kSwTK, tk = t % kSwTK;
      const int j = j0 + tj, k = k0 + tk;
      if (j >= n || k >= n) continue;
      dst[static_cast<int64_t>(j) * n + k] = c[(tj + 2) * kSwRK + (tk + 2)];
    }
  };

  // prologue: planes i0-2 .. i0+1 and the red nodes of planes i0-1 and i0
  for (int p = i0 - 2; p <= i0 + 1; ++p)
    if (p >= 0 && p <= n - 1) load_phi(p);
  for (int p = i0 - 1; p <= i0; ++p)
    if (p >= 0 && p <= n - 1) load_rhs(p);
  __syncthreads();
  red(i0 - 1);
  __syncthreads();
  red(i0);
  __syncthreads();
  if (i0 == 1) store(0);  // the first segment also copies the i = 0 face
  for (int p = i0; p < i1; ++p) {
    if (p + 2 <= n - 1) load_phi(p + 2);
    if (p + 1 <= n - 1) load_rhs(p + 1);
    __syncthreads();
    red(p + 1);
    __syncthreads();
    black(p);
    __syncthreads();
    store(p);
  }
  if (i1 == n - 1) store(n - 1);  // the last segment also copies the i = n-1 face
}

// res = rhs - lap(phi) on the interior, 0 on the boundary (multigrid.py:62-79).
__device__ __forceinline__ float residual_at(const float* __restrict__ phi,
                                             const float* __restrict__ rhs, int64_t c,
                                             int64_t n, int64_t n2, float inv_h2) {
  const float lap = (phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] +
                     phi[c - 1] - 6.0f * phi[c]) *
                    inv_h2;
  return rhs[c] - lap;
}

__global__ void __launch_bounds__(256) residual_kernel(const float* __restrict__ phi,
                                                       const float* __restrict__ rhs,
                                                       float* __restrict__ res, int n,
                                                       float inv_h2) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= n2 * n) return;
  const pm::Idx3 q3 = pm::idx3(c, n);
  const int i = q3.i, j = q3.j, k = q3.k;
  float r = 0.0f;
  if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2)
    r = residual_at(phi, rhs, c, n, n2, inv_h2);
  res[c] = r;
}

// Sum of squares (float64) of the float32 residual, without storing it.
__global__ void __launch_bounds__(256) residual_sumsq_kernel(const float* __restrict__ phi,
                                                             const float* __restrict__ rhs,
                                                             int n, float inv_h2,
                                                             double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t total = n2 * n;
  int64_t r_begin, r_end;
  pm::block_range(total, r_begin, r_end);
  for (int64_t c = r_begin + threadIdx.x; c < r_end; c += blockDim.x) {
    const pm::Idx3 q3 = pm::idx3(c, n);
    const int i = q3.i, j = q3.j, k = q3.k;
    if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
      const double r = static_cast<double>(residual_at(phi, rhs, c, n, n2, inv_h2));
      v[0] = v[0] + r * r;
    }
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

// Full weighting (1,2,1)^3/64 at fine 2J (multigrid.py:82-97); 0 on the coarse boundary.
// Coarse interior J in [1, nc-2] reads fine 2J-1..2J+1 <= 2nc-3 <= nf-1 for both hierarchies.
__global__ void __launch_bounds__(256) restrict_kernel(const float* __restrict__ res_f,
                                                       float* __restrict__ rhs_c,
                                                       float* __restrict__ phi_c, int nf,
                                                       int nc) {
  const int64_t nc2 = static_cast<int64_t>(nc) * nc;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nc2 * nc) return;
  const pm::Idx3 q3 = pm::idx3(c, nc);
  const int ic = q3.i, jc = q3.j, kc = q3.k;
  float out = 0.0f;
  if (ic >= 1 && ic <= nc - 2 && jc >= 1 && jc <= nc - 2 && kc >= 1 && kc <= nc - 2) {
    const int64_t nf2 = static_cast<int64_t>(nf) * nf;
    const int64_t f0 = (2 * ic) * nf2 + static_cast<int64_t>(2 * jc) * nf + 2 * kc;
    float acc = 0.0f;
#pragma unroll
    for (int a = -1; a <= 1; ++a) {
#pragma unroll
      for (int b = -1; b <= 1; ++b) {
#pragma unroll
        for (int d = -1; d <= 1; ++d) {
          const float w = (a == 0 ? 2.0f : 1.0f) * (b == 0 ? 2.0f : 1.0f) * (d == 0 ? 2.0f : 1.0f);
          acc += w * res_f[f0 + a * nf2 + b * nf + d];
        }
      }
    }
    out = acc / 64.0f;
  }
  rhs_c[c] = out;
  if (phi_c) phi_c[c] = 0.0f;  // the coarse correction starts from zero (multigrid.py:437)
}

// Trilinear prolongation of the coarse correction added into the fine interior
// (multigrid.py:100-130): an even fine index takes coarse i/2 with weight 1, an odd one
// takes i/2 and i/2+1 with weight 1/2 each; terms are summed in (di, dj, dk) order.
__global__ void __launch_bounds__(256) prolong_add_kernel(const float* __restrict__ phi_c,
                                                          float* __restrict__ phi_f, int nf,
                                                          int nc) {
  const int64_t nf2 = static_cast<int64_t>(nf) * nf;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nf2 * nf) return;
  const pm::Idx3 q3 = pm::idx3(c, nf);
  const int i = q3.i, j = q3.j, k = q3.k;
  if (i < 1 || i > nf - 2 || j < 1 || j > nf - 2 || k < 1 || k > nf - 2) return;
  const int ci = i >> 1, cj = j >> 1, ck = k >> 1;
  const int oi = i & 1, oj = j & 1, ok = k & 1;
  const int64_t nc2 = static_cast<int64_t>(nc) * nc;
  float val = 0.0f;
#pragma unroll
  for (int di = 0; di < 2; ++di) {
    const float wi = oi ? 0.5f : (di == 0 ? 1.0f : 0.0f);
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const float wj = oj ? 0.5f : (dj == 0 ? 1.0f : 0.0f);
#pragma unroll
      for (int dk = 0; dk < 2; ++dk) {
        const float wk = ok ? 0.5f : (dk == 0 ? 1.0f : 0.0f);
        const float w = wi * wj * wk;
        if (w > 0.0f)
          val += w * phi_c[(ci + di) * nc2 + static_cast<int64_t>(cj + dj) * nc + (ck + dk)];
      }
    }
  }
  phi_f[c] = phi_f[c] + val;
}

__global__ void __launch_bounds__(256) fill_f32_kernel(float* __restrict__ p, int64_t count,
                                                       float v) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// rhs = factor * rho (multigrid.py:249-252)
__global__ void __launch_bounds__(256) scale_kernel(const float* __restrict__ src,
                                                    float* __restrict__ dst, int64_t count,
                                                    float factor) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = factor * src[i];
}

// Multipole face potential (multigrid.py:261-295); p = [G, M, cx, cy, cz, sxx, syy, szz,
// sxy, sxz, syz]. Thread (face, a, b) with face = 2*axis + side.
__device__ __forceinline__ float face_value(int i, int j, int k, float dx, const float* p) {
  const float g = p[0], mass = p[1], cmx = p[2], cmy = p[3], cmz = p[4];
  const float sxx = p[5], syy = p[6], szz = p[7], sxy = p[8], sxz = p[9], syz = p[10];
  const float tr_s = sxx + syy + szz;
  const float rx = static_cast<float>(i) * dx - cmx;
  const float ry = static_cast<float>(j) * dx - cmy;
  const float rz = static_cast<float>(k) * dx - cmz;
  const float d2 = rx * rx + ry * ry + rz * rz + 1e-12f;
  const float d = pm::sqrt_rn(d2);
  const float nx = rx / d, ny = ry / d, nz = rz / d;
  const float n_s_n =
      nx * nx * sxx + ny * ny * syy + nz * nz * szz + 2.0f * (nx * ny * sxy + nx * nz * sxz + ny * nz * syz);
  const float quad = 3.0f * n_s_n - tr_s;
  return -g * (mass / d + quad / (2.0f * d2 * d));
}

__global__ void __launch_bounds__(256) faces_kernel(float* __restrict__ phi, int n, float dx,
                                                    const float* __restrict__ params) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= 6 * n2) return;
  float p[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) p[q] = params[q];
  const int face = static_cast<int>(t / n2);
  const int a = static_cast<int>((t / n) % n);
  const int b = static_cast<int>(t % n);
  const int axis = face >> 1;
  const int v = (face & 1) ? n - 1 : 0;
  int i, j, k;
  if (axis == 0) { i = v; j = a; k = b; }
  else if (axis == 1) { i = a; j = v; k = b; }
  else { i = a; j = b; k = v; }
  phi[i * n2 + static_cast<int64_t>(j) * n + k] = face_value(i, j, k, dx, p);
}

// Coarsest-level relaxation in one workgroup: `sweeps` red-black sweeps (PLAIN mode,
// float32, multigrid.py:431) or, with Real = double, the float64 coarsest solve of MIXED
// mode. The level is staged through LDS and written back once.
template <typename Real>
__global__ void __launch_bounds__(256) coarse_solve_kernel(float* __restrict__ phi,
                                                           const float* __restrict__ rhs,
                                                           int n, Real h2, int sweeps) {
  __shared__ Real s_phi[kCoarseLocal * kCoarseLocal * kCoarseLocal];
  __shared__ Real s_rhs[kCoarseLocal * kCoarseLocal * kCoarseLocal];
  const int total = n * n * n;
  const int n2 = n * n;
  for (int c = threadIdx.x; c < total; c += blockDim.x) {
    s_phi[c] = static_cast<Real>(phi[c]);
    s_rhs[c] = static_cast<Real>(rhs[c]);
  }
  __syncthreads();
  const int half = (n + 1) >> 1;
  const int pairs = n * n * half;
  for (int s = 0; s < sweeps; ++s) {
    for (int color = 0; color < 2; ++color) {
      for (int t = threadIdx.x; t < pairs; t += blockDim.x) {
        const int kk = t % half;
        const int j = (t / half) % n;
        const int i = t / (half * n);
        const int k = 2 * kk + ((i + j + color) & 1);
        if (i < 1 || i > n - 2 || j < 1 || j > n - 2 || k < 1 || k > n - 2) continue;
        const int c = i * n2 + j * n + k;
        const Real sum = s_phi[c + n2] + s_phi[c - n2] + s_phi[c + n] + s_phi[c - n] +
                         s_phi[c + 1] + s_phi[c - 1];
        s_phi[c] = (sum - h2 * s_rhs[c]) / static_cast<Real>(6);
      }
      __syncthreads();
    }
  }
  for (int c = threadIdx.x; c < total; c += blockDim.x) phi[c] = static_cast<float>(s_phi[c]);
}

// ---------------------------------------------------------------------------------------
// Bottom solver: the V-cycle of every level of size <= kBottomMax in one workgroup, with
// all of those levels resident in LDS. Each phase repeats the per-node arithmetic of the
// global kernels above (RB pass, residual + restriction, prolongation) in the same order,
// so the result is bitwise that of the multi-kernel V-cycle; the residual is recomputed
// inside the restriction instead of being stored.
// ---------------------------------------------------------------------------------------
constexpr int kBottomMax = 17;
constexpr int kBottomLevels = 8;
constexpr int kBottomThreads = 1024;
constexpr int kCoarse64 = 5;  // coarsest levels up to this size are solved in float64 (MIXED)

struct BottomDesc {
  int levels;
  int n[kBottomLevels];
  int off[kBottomLevels];  // LDS offset (floats) of phi; rhs follows at off + n^3
  float h2[kBottomLevels];
  float inv_h2[kBottomLevels];
  double h2_coarsest;
  int pre, post, coarse_sweeps, coarse64_sweeps;
};

__device__ __forceinline__ void lds_rb_pass(float* phi, const float* rhs, int n, float h2,
                                            int color) {
  const int half = (n + 1) >> 1;
  const int pairs = n * n * half;
  const int n2 = n * n;
  for (int t = threadIdx.x; t < pairs; t += blockDim.x) {
    const int kk = t % half;
    const int j = (t / half) % n;
    const int i = t / (half * n);
    const int k = 2 * kk + ((i + j + color) & 1);
    if (i < 1 || i > n - 2 || j < 1 || j > n - 2 || k < 1 || k > n - 2) continue;
    const int c = i * n2 + j * n + k;
    const float sum = phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] + phi[c - 1];
    phi[c] = (sum - h2 * rhs[c]) / 6.0f;
  }
  __syncthreads();
}

__device__ __forceinline__ void lds_smooth(float* phi, const float* rhs, int n, float h2,
                                           int sweeps) {
  for (int q = 0; q < sweeps; ++q) {
    lds_rb_pass(phi, rhs, n, h2, 0);
    lds_rb_pass(phi, rhs, n, h2, 1);
  }
}

__device__ __forceinline__ float lds_residual(const float* phi, const float* rhs, int n, int i,
                                              int j, int k, float inv_h2) {
  if (i < 1 || i > n - 2 || j < 1 || j > n - 2 || k < 1 || k > n - 2) return 0.0f;
  const int n2 = n * n;
  const int c = i * n2 + j * n + k;
  const float lap = (phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] +
                     phi[c - 1] - 6.0f * phi[c]) *
                    inv_h2;
  return rhs[c] - lap;
}

template <bool kMixed>
__global__ void __launch_bounds__(kBottomThreads) bottom_vcycle_kernel(float* __restrict__ phi_top,
                                                                       const float* __restrict__ rhs_top,
                                                                       BottomDesc d) {
  extern __shared__ float lds[];
  __shared__ double s64_phi[kCoarse64 * kCoarse64 * kCoarse64];
  __shared__ double s64_rhs[kCoarse64 * kCoarse64 * kCoarse64];
  const int L = d.levels;
  {
    const int nn = d.n[0] * d.n[0] * d.n[0];
    float* phi = lds + d.off[0];
    float* rhs = phi + nn;
    for (int c = threadIdx.x; c < nn; c += blockDim.x) {
      phi[c] = phi_top[c];
      rhs[c] = rhs_top[c];
    }
    __syncthreads();
  }
  // down
  for (int l = 0; l + 1 < L; ++l) {
    const int nf = d.n[l], nc = d.n[l + 1];
    float* phi = lds + d.off[l];
    float* rhs = phi + nf * nf * nf;
    lds_smooth(phi, rhs, nf, d.h2[l], d.pre);
    float* cphi = lds + d.off[l + 1];
    float* crhs = cphi + nc * nc * nc;
    const int ncc = nc * nc * nc;
    for (int c = threadIdx.x; c < ncc; c += blockDim.x) {
      const int kc = c % nc, jc = (c / nc) % nc, ic = c / (nc * nc);
      float out = 0.0f;
      if (ic >= 1 && ic <= nc - 2 && jc >= 1 && jc <= nc - 2 && kc >= 1 && kc <= nc - 2) {
        float acc = 0.0f;
        for (int a = -1; a <= 1; ++a)
          for (int b = -1; b <= 1; ++b)
            for (int e = -1; e <= 1; ++e) {
              const float w = (a == 0 ? 2.0f : 1.0f) * (b == 0 ? 2.0f : 1.0f) * (e == 0 ? 2.0f : 1.0f);
              acc += w * lds_residual(phi, rhs, nf, 2 * ic + a, 2 * jc + b, 2 * kc + e, d.inv_h2[l]);
            }
        out = acc / 64.0f;
      }
      crhs[c] = out;
      cphi[c] = 0.0f;
    }
    __syncthreads();
  }
  // coarsest
  {
    const int nc = d.n[L - 1];
    float* phi = lds + d.off[L - 1];
    float* rhs = phi + nc * nc * nc;
    if (kMixed && nc <= kCoarse64) {
      const int nn = nc * nc * nc, n2 = nc * nc, half = (nc + 1) >> 1;
      for (int c = threadIdx.x; c < nn; c += blockDim.x) {
        s64_phi[c] = phi[c];
        s64_rhs[c] = rhs[c];
      }
      __syncthreads();
      for (int q = 0; q < d.coarse64_sweeps; ++q) {
        for (int color = 0; color < 2; ++color) {
          for (int t = threadIdx.x; t < nc * nc * half; t += blockDim.x) {
            const int kk = t % half, j = (t / half) % nc, i = t / (half * nc);
            const int k = 2 * kk + ((i + j + color) & 1);
            if (i < 1 || i > nc - 2 || j < 1 || j > nc - 2 || k < 1 || k > nc - 2) continue;
            const int c = i * n2 + j * nc + k;
            const double sum = s64_phi[c + n2] + s64_phi[c - n2] + s64_phi[c + nc] +
                               s64_phi[c - nc] + s64_phi[c + 1] + s64_phi[c - 1];
            s64_phi[c] = (sum - d.h2_coarsest * s64_rhs[c]) / 6.0;
          }
          __syncthreads();
        }
      }
      for (int c = threadIdx.x; c < nn; c += blockDim.x) phi[c] = static_cast<float>(s64_phi[c]);
      __syncthreads();
    } else {
      lds_smooth(phi, rhs, nc, d.h2[L - 1], d.coarse_sweeps);
    }
  }
  // up
  for (int l = L - 2; l >= 0; --l) {
    const int nf = d.n[l], nc = d.n[l + 1];
    float* phi = lds + d.off[l];
    float* rhs = phi + nf * nf * nf;
    const float* cphi = lds + d.off[l + 1];
    const int nf2 = nf * nf, nc2 = nc * nc;
    for (int c = threadIdx.x; c < nf2 * nf; c += blockDim.x) {
      const int k = c % nf, j = (c / nf) % nf, i = c / nf2;
      if (i < 1 || i > nf - 2 || j < 1 || j > nf - 2 || k < 1 || k > nf - 2) continue;
      const int ci = i >> 1, cj = j >> 1, ck = k >> 1;
      const int oi = i & 1, oj = j & 1, ok = k & 1;
      float val = 0.0f;
      for (int di = 0; di < 2; ++di) {
        const float wi = oi ? 0.5f : (di == 0 ? 1.0f : 0.0f);
        for (int dj = 0; dj < 2; ++dj) {
          const float wj = oj ? 0.5f : (dj == 0 ? 1.0f : 0.0f);
          for (int dk = 0; dk < 2; ++dk) {
            const float wk = ok ? 0.5f : (dk == 0 ? 1.0f : 0.0f);
            const float w = wi * wj * wk;
            if (w > 0.0f) val += w * cphi[(ci + di) * nc2 + (cj + dj) * nc + (ck + dk)];
          }
        }
      }
      phi[c] = phi[c] + val;
    }
    __syncthreads();
    lds_smooth(phi, rhs, nf, d.h2[l], d.post);
  }
  {
    const int nn = d.n[0] * d.n[0] * d.n[0];
    const float* phi = lds + d.off[0];
    for (int c = threadIdx.x; c < nn; c += blockDim.x) phi_top[c] = phi[c];
  }
}

// ---------------------------------------------------------------------------------------
// float64 defect-correction kernels (MIXED mode)
// ---------------------------------------------------------------------------------------

// r = f - lap(phi) evaluated in float64 from the float32 iterate (interior; 0 on the
// boundary); r32 = (float)r; partial sums of r^2 (float64).
__global__ void __launch_bounds__(256) residual64_kernel(const float* __restrict__ phi,
                                                         const float* __restrict__ f, int n,
                                                         double inv_h2,
                                                         float* __restrict__ r32,
                                                         double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t total = n2 * n;
  int64_t r_begin, r_end;
  pm::block_range(total, r_begin, r_end);
  for (int64_t c = r_begin + threadIdx.x; c < r_end; c += blockDim.x) {
    const pm::Idx3 q3 = pm::idx3(c, n);
    const int i = q3.i, j = q3.j, k = q3.k;
    double r = 0.0;
    if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
      const double sum = static_cast<double>(phi[c + n2]) + static_cast<double>(phi[c - n2]) +
                         static_cast<double>(phi[c + n]) + static_cast<double>(phi[c - n]) +
                         static_cast<double>(phi[c + 1]) + static_cast<double>(phi[c - 1]);
      const double lap = (sum - 6.0 * static_cast<double>(phi[c])) * inv_h2;
      r = static_cast<double>(f[c]) - lap;
    }
    r32[c] = static_cast<float>(r);
    v[0] = v[0] + r * r;
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

// phi += e (e is 0 on the faces); partial sums of e^2 and phi^2 (float64).
__global__ void __launch_bounds__(256) update_kernel(float* __restrict__ phi,
                                                     const float* __restrict__ e,
                                                     int64_t count,
                                                     double* __restrict__ partials) {
  __shared__ double s_red[4 * 2];
  double v[2] = {0.0, 0.0};
  int64_t r_begin, r_end;
  pm::block_range(count, r_begin, r_end);
  for (int64_t c = r_begin + threadIdx.x; c < r_end; c += blockDim.x) {
    const float d = e[c];
    const float p = phi[c] + d;
    phi[c] = p;
    v[0] = v[0] + static_cast<double>(d) * static_cast<double>(d);
    v[1] = v[1] + static_cast<double>(p) * static_cast<double>(p);
  }
  pm::block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * pm::kRedMaxQ] = v[0];
    partials[blockIdx.x * pm::kRedMaxQ + 1] = v[1];
  }
}

struct Level {
  int n = 0;
  double h = 0.0;
  float h2 = 0.0f;
  float inv_h2 = 0.0f;
  float* phi = nullptr;  // level 0: the solve's phi (caller's in PLAIN, e0 in MIXED)
  float* rhs = nullptr;  // level 0: 4*pi*G*rho (PLAIN) or the float32 defect (MIXED)
  float* res = nullptr;
  float* tmp = nullptr;  // ping-pong partner of phi for the fused sweep (large levels only)
  int64_t nodes() const { return static_cast<int64_t>(n) * n * n; }
};

// Level sizes. PLAIN: the reference rule (halve while even and >= 3). MIXED: the aligned
// rule (n even -> n/2+1, n odd -> (n-1)/2+1, down to 3 nodes).
std::vector<int> level_sizes(int n, int mode) {
  std::vector<int> sizes{n};
  if (mode == PM_MG_PLAIN) {
    while (sizes.back() % 2 == 0 && sizes.back() / 2 >= 3) sizes.push_back(sizes.back() / 2);
    return sizes;
  }
  while (sizes.back() > 3) {
    const int m = sizes.back();
    const int nc = (m % 2 == 0) ? m / 2 + 1 : (m - 1) / 2 + 1;
    if (nc < 3 || nc >= m) break;
    sizes.push_back(nc);
  }
  return sizes;
}

// Algorithmic bytes per launch (compulsory traffic per node, float32 fields).
constexpr double kRbBytes = 12.0;        // read phi + rhs, write phi (per colour pass)
constexpr double kResidualBytes = 12.0;  // read phi + rhs, write res
constexpr double kRestrictBytes = 5.0;   // per fine node: read res, write 1/8 coarse rhs + phi
constexpr double kProlongBytes = 8.5;    // per fine node: read+write phi, read 1/8 coarse
constexpr double kFillBytes = 4.0;
constexpr double kRes64Bytes = 12.0;     // read phi + f, write r32
constexpr double kUpdBytes = 12.0;       // read phi + e, write phi

}  // namespace

struct pm_mg {
  pm_mg_desc d{};
  std::vector<Level> lv;
  float* f0 = nullptr;        // 4*pi*G*rho, float32 (both modes)
  float* e0 = nullptr;        // MIXED: finest correction
  float* r32 = nullptr;       // MIXED: finest float32 defect
  double* partials = nullptr;
  double* red = nullptr;      // device reduction results [kRedMaxQ]
  double* host = nullptr;     // pinned mirror
  double* moments = nullptr;  // 10 raw moments
  float* face_params = nullptr;
  float* last_phi = nullptr;  // caller phi of the last solve
  int bottom_from = -1;       // first level handled by the one-workgroup bottom solver
  bool solved = false;
  // warm-start state (allocated on the first pm_mg_solve_warm)
  float* hist = nullptr;      // phi of the previous warm solve (the time history)
  float* cand = nullptr;      // extrapolated guess of the current solve
  bool hist_valid = false;
  int warm_pool = 0;          // network pool factor the buffers below are sized for
  float* pooled = nullptr;    // [R^3][3] pooled input channels
  uint16_t* net_in = nullptr; // [R^3][8] bf16 network input
  float* net_out = nullptr;   // [R^3] network output
  float* scales = nullptr;    // [8] input/output scales
  std::vector<void*> owned;
};

namespace {

pm_status dalloc(pm_mg* mg, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes > 0 ? bytes : 1);
  if (e != hipSuccess) return pm::hip_fail("hipMalloc (multigrid)", e);
  mg->owned.push_back(*p);
  return PM_OK;
}

void destroy(pm_mg* mg) {
  if (!mg) return;
  for (void* p : mg->owned) (void)hipFree(p);
  if (mg->host) (void)hipHostFree(mg->host);
  delete mg;
}

const char* level_name(const char* base, int n) {
  // stable per-level kernel names for the timing report
  static thread_local char buf[64];
  std::snprintf(buf, sizeof(buf), "%s[%d]", base, n);
  return buf;
}

pm_status rb_pass(float* phi, const float* rhs, int n, float h2, int color, hipStream_t s) {
  const int half = (n + 1) >> 1;
  const int64_t total = static_cast<int64_t>(n) * n * half;
  PM_TIMED(level_name("mg.rb_pass", n), kRbBytes * n * n * n, s);
  rb_pass_kernel<<<pm::blocks_for(total), pm::kBlock, 0, s>>>(phi, rhs, n, h2, color);
  PM_CHECK_LAUNCH("rb_pass_kernel");
  return PM_OK;
}

// Levels at least this large would use the fused sweep in the solve. Measured on MI355X at
// 256^3 the streaming sweep (221 us) is slower than two colour passes (2 x 54 us): its
// three barriers per plane leave it latency-bound, so the solve keeps the colour passes.
constexpr int kSweepMin = 1 << 30;

pm_status rb_sweep(const float* phi_in, float* phi_out, const float* rhs, int n, float h2,
                   hipStream_t s) {
  const dim3 grid((n + kSwTK - 1) / kSwTK, (n + kSwTJ - 1) / kSwTJ, (n - 2 + kSwSeg - 1) / kSwSeg);
  PM_TIMED(level_name("mg.rb_sweep", n), 12.0 * n * n * n, s);
  rb_sweep_kernel<<<grid, 256, 0, s>>>(phi_in, phi_out, rhs, n, h2);
  PM_CHECK_LAUNCH("rb_sweep_kernel");
  return PM_OK;
}

// `sweeps` red-black sweeps in place on L.phi. Large levels ping-pong pairs of fused sweeps
// through L.tmp (ending back in L.phi); a leftover single sweep uses two colour passes.
pm_status smooth(const Level& L, int sweeps, hipStream_t s) {
  pm_status st;
  int q = 0;
  if (L.tmp) {
    for (; q + 2 <= sweeps; q += 2) {
      if ((st = rb_sweep(L.phi, L.tmp, L.rhs, L.n, L.h2, s))) return st;
      if ((st = rb_sweep(L.tmp, L.phi, L.rhs, L.n, L.h2, s))) return st;
    }
  }
  for (; q < sweeps; ++q) {
    if ((st = rb_pass(L.phi, L.rhs, L.n, L.h2, 0, s))) return st;
    if ((st = rb_pass(L.phi, L.rhs, L.n, L.h2, 1, s))) return st;
  }
  return PM_OK;
}

pm_status fill(float* p, int64_t count, float v, hipStream_t s) {
  PM_TIMED("mg.fill", kFillBytes * count, s);
  fill_f32_kernel<<<pm::blocks_for(count), pm::kBlock, 0, s>>>(p, count, v);
  PM_CHECK_LAUNCH("fill_f32_kernel");
  return PM_OK;
}

pm_status coarsest(pm_mg* mg, const Level& L, hipStream_t s) {
  const bool mixed = mg->d.mode == PM_MG_MIXED;
  if (L.n <= kCoarseLocal) {
    PM_TIMED("mg.coarse_solve", 8.0 * L.nodes(), s);
    if (mixed) {
      // float64 solve: enough sweeps to converge the small coarsest system
      const int sweeps = mg->d.coarse_sweeps > 200 ? mg->d.coarse_sweeps : 200;
      coarse_solve_kernel<double><<<1, 256, 0, s>>>(L.phi, L.rhs, L.n, L.h * L.h, sweeps);
    } else {
      coarse_solve_kernel<float><<<1, 256, 0, s>>>(L.phi, L.rhs, L.n, L.h2,
                                                    mg->d.coarse_sweeps);
    }
    PM_CHECK_LAUNCH("coarse_solve_kernel");
    return PM_OK;
  }
  return smooth(L, mg->d.coarse_sweeps, s);
}

// Levels l..end in one workgroup (all of them are <= kBottomMax).
pm_status bottom(pm_mg* mg, size_t l, hipStream_t s) {
  BottomDesc d{};
  d.levels = static_cast<int>(mg->lv.size() - l);
  int off = 0;
  double bytes = 0.0;
  for (int q = 0; q < d.levels; ++q) {
    const Level& L = mg->lv[l + q];
    d.n[q] = L.n;
    d.h2[q] = L.h2;
    d.inv_h2[q] = L.inv_h2;
    d.off[q] = off;
    off += 2 * L.n * L.n * L.n;
  }
  const Level& top = mg->lv[l];
  d.h2_coarsest = mg->lv.back().h * mg->lv.back().h;
  d.pre = mg->d.pre_sweeps;
  d.post = mg->d.post_sweeps;
  d.coarse_sweeps = mg->d.coarse_sweeps;
  d.coarse64_sweeps = 60;
  bytes = 12.0 * top.nodes();  // read phi + rhs, write phi of the top level
  const size_t lds = sizeof(float) * static_cast<size_t>(off);
  PM_TIMED(level_name("mg.bottom", top.n), bytes, s);
  if (mg->d.mode == PM_MG_MIXED)
    bottom_vcycle_kernel<true><<<1, kBottomThreads, lds, s>>>(top.phi, top.rhs, d);
  else
    bottom_vcycle_kernel<false><<<1, kBottomThreads, lds, s>>>(top.phi, top.rhs, d);
  PM_CHECK_LAUNCH("bottom_vcycle_kernel");
  return PM_OK;
}

pm_status vcycle(pm_mg* mg, size_t l, hipStream_t s) {
  const Level& L = mg->lv[l];
  if (mg->bottom_from >= 0 && static_cast<int>(l) >= mg->bottom_from) return bottom(mg, l, s);
  if (l + 1 == mg->lv.size()) return coarsest(mg, L, s);
  pm_status st;
  if ((st = smooth(L, mg->d.pre_sweeps, s))) return st;
  {
    PM_TIMED(level_name("mg.residual", L.n), kResidualBytes * L.nodes(), s);
    residual_kernel<<<pm::blocks_for(L.nodes()), pm::kBlock, 0, s>>>(L.phi, L.rhs, L.res, L.n,
                                                                      L.inv_h2);
    PM_CHECK_LAUNCH("residual_kernel");
  }
  const Level& C = mg->lv[l + 1];
  {
    PM_TIMED(level_name("mg.restrict", L.n), kRestrictBytes * L.nodes(), s);
    restrict_kernel<<<pm::blocks_for(C.nodes()), pm::kBlock, 0, s>>>(L.res, C.rhs, C.phi, L.n,
                                                                      C.n);
    PM_CHECK_LAUNCH("restrict_kernel");
  }
  if ((st = vcycle(mg, l + 1, s))) return st;
  {
    PM_TIMED(level_name("mg.prolong", L.n), kProlongBytes * L.nodes(), s);
    prolong_add_kernel<<<pm::blocks_for(L.nodes()), pm::kBlock, 0, s>>>(C.phi, L.phi, L.n, C.n);
    PM_CHECK_LAUNCH("prolong_add_kernel");
  }
  return smooth(L, mg->d.post_sweeps, s);
}

// Copies `count` float64 results from the device (synchronizes the stream).
pm_status read_scalars(pm_mg* mg, const double* dev, int count, double* out, hipStream_t s) {
  PM_TRY_HIP(hipMemcpyAsync(mg->host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  PM_TRY_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < count; ++q) out[q] = mg->host[q];
  return PM_OK;
}

// ||rhs - A phi|| over the finest level (float32 residual, float64 sum), synchronizing.
pm_status plain_residual_norm(pm_mg* mg, const float* phi, double* out, hipStream_t s) {
  const Level& L = mg->lv[0];
  const unsigned b = pm::red_blocks(L.nodes());
  {
    PM_TIMED("mg.residual_norm", 8.0 * L.nodes(), s);
    residual_sumsq_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->f0, L.n, L.inv_h2, mg->partials);
    PM_CHECK_LAUNCH("residual_sumsq_kernel");
  }
  pm::fold_partials_kernel<1><<<1, pm::kBlock, 0, s>>>(mg->partials, static_cast<int>(b),
                                                        mg->red);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  double ss;
  pm_status st = read_scalars(mg, mg->red, 1, &ss, s);
  if (st) return st;
  *out = std::sqrt(ss);
  return PM_OK;
}

// r = f - A phi (float64 arithmetic) -> r32; its sum of squares lands in red[2] (no sync).
pm_status mixed_residual(pm_mg* mg, const float* phi, hipStream_t s) {
  const Level& L = mg->lv[0];
  const unsigned b = pm::red_blocks(L.nodes());
  {
    PM_TIMED("mg.residual64", kRes64Bytes * L.nodes(), s);
    residual64_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->f0, L.n, 1.0 / (L.h * L.h),
                                                mg->r32, mg->partials);
    PM_CHECK_LAUNCH("residual64_kernel");
  }
  pm::fold_partials_kernel<1><<<1, pm::kBlock, 0, s>>>(mg->partials, static_cast<int>(b),
                                                        mg->red + 2);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  return PM_OK;
}

// One defect-correction iteration: e = V(r32); phi += e. Returns ||e||, ||phi||
// (synchronizing).
pm_status mixed_correct(pm_mg* mg, float* phi, double* e_norm, double* phi_norm,
                        hipStream_t s) {
  Level& L0 = mg->lv[0];
  pm_status st;
  if ((st = fill(mg->e0, L0.nodes(), 0.0f, s))) return st;
  if ((st = vcycle(mg, 0, s))) return st;
  const unsigned b = pm::red_blocks(L0.nodes());
  {
    PM_TIMED("mg.update", kUpdBytes * L0.nodes(), s);
    update_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->e0, L0.nodes(), mg->partials);
    PM_CHECK_LAUNCH("update_kernel");
  }
  pm::fold_partials_kernel<2><<<1, pm::kBlock, 0, s>>>(mg->partials, static_cast<int>(b),
                                                        mg->red);
  PM_CHECK_LAUNCH("fold_partials_kernel<2>");
  double v[2];
  if ((st = read_scalars(mg, mg->red, 2, v, s))) return st;
  *e_norm = std::sqrt(v[0]);
  *phi_norm = std::sqrt(v[1]);
  return PM_OK;
}

// Moments, f = 4*pi*G*rho, the initial guess (zero unless warm) with multipole faces, and
// ||f|| (into red[3]). Enqueued only.
pm_status solve_setup(pm_mg* mg, const float* rho, float* phi, bool warm, hipStream_t s) {
  Level& L0 = mg->lv[0];
  const int n = L0.n;
  const int64_t nodes = L0.nodes();
  pm_status st;
  // boundary moments (multigrid.py:458) -> float32 face parameters, on the device
  {
    PM_TIMED("mg.moments", 4.0 * nodes, s);
    if ((st = pm::launch_moments(rho, n, mg->d.dx, mg->d.grav, mg->partials, mg->moments,
                                 mg->face_params, s)))
      return st;
  }
  // f = 4*pi*G*rho (multigrid.py:463)
  {
    const float factor = static_cast<float>(4.0 * kPi * mg->d.grav);
    PM_TIMED("mg.scale", 8.0 * nodes, s);
    scale_kernel<<<pm::blocks_for(nodes), pm::kBlock, 0, s>>>(rho, mg->f0, nodes, factor);
    PM_CHECK_LAUNCH("scale_kernel");
  }
  if (!warm && (st = fill(phi, nodes, 0.0f, s))) return st;
  {
    PM_TIMED("mg.faces", 4.0 * 6 * n * n, s);
    faces_kernel<<<pm::blocks_for(6 * static_cast<int64_t>(n) * n), pm::kBlock, 0, s>>>(
        phi, n, mg->d.dx, mg->face_params);
    PM_CHECK_LAUNCH("faces_kernel");
  }
  // ||rhs|| (multigrid.py:477)
  {
    PM_TIMED("mg.rhs_norm", 4.0 * nodes, s);
    if ((st = pm::launch_sumsq_f32(mg->f0, nodes, mg->partials, mg->red + 3, s))) return st;
  }
  return PM_OK;
}

// The V-cycle loop of a solve whose setup has run (PLAIN: reference stop rule; MIXED:
// defect correction with the error-estimate stop). Fills stats. residual_ready: MIXED only,
// r32 already holds the defect of phi and red[2] its squared norm.
pm_status solve_iterate(pm_mg* mg, float* phi, pm_mg_stats* stats, hipStream_t s,
                        bool residual_ready = false) {
  Level& L0 = mg->lv[0];
  const bool mixed = mg->d.mode == PM_MG_MIXED;
  pm_status st;
  int cycles = mg->d.n_cycles;
  double resid = -1.0, err_est = -1.0, rhs_norm = 0.0;

  if (!mixed) {
    double rhs_ss;
    if ((st = read_scalars(mg, mg->red + 3, 1, &rhs_ss, s))) return st;
    rhs_norm = std::sqrt(rhs_ss);
    L0.phi = phi;
    L0.rhs = mg->f0;
    mg->last_phi = phi;
    for (int c = 0; c < mg->d.n_cycles; ++c) {
      if ((st = vcycle(mg, 0, s))) return st;
      if (c + 1 >= mg->d.min_cycles && rhs_norm > 0.0) {
        if ((st = plain_residual_norm(mg, phi, &resid, s))) return st;
        if (resid < mg->d.tol * rhs_norm) {
          cycles = c + 1;
          break;
        }
      }
    }
  } else {
    L0.phi = mg->e0;
    L0.rhs = mg->r32;
    mg->last_phi = phi;
    if (!residual_ready && (st = mixed_residual(mg, phi, s))) return st;
    double prev_e = -1.0;
    for (int c = 0; c < mg->d.n_cycles; ++c) {
      double e_norm, phi_norm;
      if ((st = mixed_correct(mg, phi, &e_norm, &phi_norm, s))) return st;
      // contraction estimate from successive corrections; 0.5 before two are known
      double rate = prev_e > 0.0 ? e_norm / prev_e : 0.5;
      if (rate > 0.95) rate = 0.95;
      prev_e = e_norm;
      err_est = phi_norm > 0.0 ? e_norm * rate / (1.0 - rate) / phi_norm : 0.0;
      if (c + 1 >= mg->d.min_cycles && err_est < mg->d.tol) {
        cycles = c + 1;
        break;
      }
      if (c + 1 < mg->d.n_cycles && (st = mixed_residual(mg, phi, s))) return st;
    }
    double v[2];
    if ((st = read_scalars(mg, mg->red + 2, 2, v, s))) return st;
    resid = std::sqrt(v[0]);
    rhs_norm = std::sqrt(v[1]);
  }
  mg->solved = true;
  if (stats) {
    stats->cycles = cycles;
    stats->levels = static_cast<int32_t>(mg->lv.size());
    stats->resid_norm = resid;
    stats->rhs_norm = rhs_norm;
    stats->err_est = mixed ? err_est : -1.0;
  }
  return PM_OK;
}

}  // namespace

extern "C" {

pm_status pm_mg_create(const pm_mg_desc* desc, pm_mg** out) {
  PM_REQUIRE(desc && out, "pm_mg_create: null pointer");
  *out = nullptr;
  const pm_mg_desc& d = *desc;
  PM_REQUIRE(d.grid_size >= 3 && d.grid_size <= 1024, "pm_mg_create: grid_size must be in [3, 1024]");
  PM_REQUIRE(d.dx > 0.0f, "pm_mg_create: dx must be positive");
  PM_REQUIRE(d.n_cycles >= 0 && d.pre_sweeps >= 0 && d.post_sweeps >= 0 && d.coarse_sweeps >= 0,
             "pm_mg_create: cycle and sweep counts must be non-negative");
  PM_REQUIRE(d.mode == PM_MG_PLAIN || d.mode == PM_MG_MIXED, "pm_mg_create: unknown mode");
  auto* mg = new pm_mg();
  mg->d = d;
  if (mg->d.min_cycles < 1) mg->d.min_cycles = 1;
  const std::vector<int> sizes = level_sizes(d.grid_size, d.mode);
  pm_status st = PM_OK;
  for (size_t l = 0; l < sizes.size(); ++l) {
    Level L;
    L.n = sizes[l];
    L.h = static_cast<double>(d.dx) * static_cast<double>(int64_t(1) << l);
    L.h2 = static_cast<float>(L.h * L.h);
    L.inv_h2 = static_cast<float>(1.0 / (L.h * L.h));
    const size_t bytes = sizeof(float) * static_cast<size_t>(L.nodes());
    if (l > 0) {
      if ((st = dalloc(mg, reinterpret_cast<void**>(&L.phi), bytes)) ||
          (st = dalloc(mg, reinterpret_cast<void**>(&L.rhs), bytes))) {
        destroy(mg);
        return st;
      }
    }
    if (l + 1 < sizes.size()) {
      if ((st = dalloc(mg, reinterpret_cast<void**>(&L.res), bytes))) {
        destroy(mg);
        return st;
      }
    }
    if (L.n >= kSweepMin) {
      if ((st = dalloc(mg, reinterpret_cast<void**>(&L.tmp), bytes))) {
        destroy(mg);
        return st;
      }
    }
    mg->lv.push_back(L);
  }
  // the bottom solver takes over at the first level <= kBottomMax when every level below it
  // is small too and the hierarchy fits its descriptor
  for (size_t l = 0; l < mg->lv.size(); ++l) {
    if (mg->lv[l].n <= kBottomMax && mg->lv.size() - l <= static_cast<size_t>(kBottomLevels)) {
      mg->bottom_from = static_cast<int>(l);
      break;
    }
  }
  const int64_t n0 = mg->lv[0].nodes();
  if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->f0), sizeof(float) * n0)) ||
      (st = dalloc(mg, reinterpret_cast<void**>(&mg->partials),
                   sizeof(double) * pm::kRedBlocks * pm::kRedMaxQ)) ||
      (st = dalloc(mg, reinterpret_cast<void**>(&mg->red), sizeof(double) * pm::kRedMaxQ)) ||
      (st = dalloc(mg, reinterpret_cast<void**>(&mg->moments), sizeof(double) * 10)) ||
      (st = dalloc(mg, reinterpret_cast<void**>(&mg->face_params), sizeof(float) * 11))) {
    destroy(mg);
    return st;
  }
  if (d.mode == PM_MG_MIXED) {
    if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->e0), sizeof(float) * n0)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->r32), sizeof(float) * n0))) {
      destroy(mg);
      return st;
    }
  }
  hipError_t he = hipHostMalloc(&mg->host, sizeof(double) * pm::kRedMaxQ, 0);
  if (he != hipSuccess) {
    destroy(mg);
    return pm::hip_fail("hipHostMalloc", he);
  }
  *out = mg;
  return PM_OK;
}

pm_status pm_mg_destroy(pm_mg* mg) {
  destroy(mg);
  return PM_OK;
}

pm_status pm_mg_num_levels(pm_mg* mg, int32_t* out) {
  PM_REQUIRE(mg && out, "pm_mg_num_levels: null pointer");
  *out = static_cast<int32_t>(mg->lv.size());
  return PM_OK;
}

pm_status pm_mg_solve(pm_mg* mg, const float* rho, float* phi, int32_t warm,
                      pm_mg_stats* stats, hipStream_t s) {
  PM_REQUIRE(mg && rho && phi, "pm_mg_solve: null pointer");
  PM_REQUIRE(rho != phi, "pm_mg_solve: rho and phi must not alias");
  mg->hist_valid = false;  // a solve outside the warm-start sequence ends its history
  pm_status st;
  if ((st = solve_setup(mg, rho, phi, warm != 0, s))) return st;
  return solve_iterate(mg, phi, stats, s);
}

pm_status pm_mg_warm_reset(pm_mg* mg) {
  PM_REQUIRE(mg, "pm_mg_warm_reset: null handle");
  mg->hist_valid = false;
  return PM_OK;
}

pm_status pm_mg_solve_warm(pm_mg* mg, const float* rho, float* phi, pm_mg_warm* w,
                           pm_mg_stats* stats, hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(mg && rho && phi && w, "pm_mg_solve_warm: null pointer");
  PM_REQUIRE(rho != phi, "pm_mg_solve_warm: rho and phi must not alias");
  PM_REQUIRE(w->mode == PM_WARM_PREVIOUS || w->mode == PM_WARM_EXTRAPOLATE ||
                 w->mode == PM_WARM_LEARNED,
             "pm_mg_solve_warm: unknown mode");
  const Level& L0 = mg->lv[0];
  const int n = L0.n;
  const int64_t nodes = L0.nodes();
  const bool learned = w->mode == PM_WARM_LEARNED;
  const bool pairs = w->pair_input || w->pair_target;
  const bool want_net_grid = learned || pairs;
  w->extrapolated = 0;
  w->used_net = 0;
  w->resid_before = -1.0;
  w->resid_after = -1.0;
  pm_status st;
  if (want_net_grid) {
    PM_REQUIRE(w->pool >= 1 && n % w->pool == 0 && (n / w->pool) % 8 == 0,
               "pm_mg_solve_warm: grid_size / pool must be a multiple of 8");
  }
  if (learned) {
    PM_REQUIRE(w->net, "pm_mg_solve_warm: PM_WARM_LEARNED needs a network");
    pm_wsnet_layer info;
    if ((st = pm_wsnet_layer_info(w->net, 0, &info))) return st;
    PM_REQUIRE(info.res == n / w->pool,
               "pm_mg_solve_warm: network resolution must equal grid_size / pool");
    PM_REQUIRE(w->fallback_ratio > 0.0, "pm_mg_solve_warm: fallback_ratio must be positive");
  }
  // buffers: time history always; the network grid only when used
  if (!mg->hist) {
    if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->hist), sizeof(float) * nodes)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->cand), sizeof(float) * nodes)))
      return st;
  }
  if (want_net_grid && !mg->r32) {
    if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->r32), sizeof(float) * nodes)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->e0), sizeof(float) * nodes)))
      return st;
  }
  if (want_net_grid && mg->warm_pool != w->pool) {
    const int R = n / w->pool;
    const int64_t nv = static_cast<int64_t>(R) * R * R;
    if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->pooled), sizeof(float) * 3 * nv)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->net_in), sizeof(uint16_t) * 8 * nv)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->net_out), sizeof(float) * nv)) ||
        (st = dalloc(mg, reinterpret_cast<void**>(&mg->scales), sizeof(float) * 8)))
      return st;
    mg->warm_pool = w->pool;
  }

  if ((st = solve_setup(mg, rho, phi, true, s))) return st;
  const bool extrap = w->mode != PM_WARM_PREVIOUS && mg->hist_valid;
  if (!extrap) {
    // first solve of a sequence (or PREVIOUS mode): phi itself is the guess; remember it
    PM_TRY_HIP(hipMemcpyAsync(mg->hist, phi, sizeof(float) * nodes, hipMemcpyDeviceToDevice, s));
    mg->hist_valid = true;
    return solve_iterate(mg, phi, stats, s);
  }
  w->extrapolated = 1;
  if ((st = pm::warm::extrapolate(phi, mg->hist, (learned || w->pair_target) ? mg->cand : nullptr,
                                  n, s)))
    return st;
  if (want_net_grid) {
    if ((st = mixed_residual(mg, phi, s))) return st;  // r32, ||r||^2 -> red[2]
    if ((st = pm::warm::make_input(mg->r32, phi, mg->hist, mg->f0, n, w->pool, L0.h, mg->pooled,
                                   mg->net_in, mg->partials, mg->red + 4, mg->scales, s)))
      return st;
    if (w->pair_input) {
      const int R = n / w->pool;
      PM_TRY_HIP(hipMemcpyAsync(w->pair_input, mg->net_in,
                                sizeof(uint16_t) * 8 * static_cast<size_t>(R) * R * R,
                                hipMemcpyDeviceToDevice, s));
    }
  }
  if (learned) {
    if ((st = pm_wsnet_forward(w->net, mg->net_in, mg->net_out, s))) return st;
    if ((st = pm::warm::apply(mg->net_out, n, w->pool, mg->scales, mg->e0, s))) return st;
    // the judge also writes the corrected guess's residual into the level-0 scratch
    Level& lv0 = mg->lv[0];
    if ((st = pm::warm::guard(mg->r32, mg->e0, n, w->pool, L0.h, lv0.res, mg->partials,
                              mg->red + 8, s)))
      return st;
    double v[11];  // red[8], red[9]: pooled residual before / after; red[10]: ||r after||^2
    if ((st = read_scalars(mg, mg->red, 11, v, s))) return st;
    w->resid_before = std::sqrt(v[8]);
    w->resid_after = std::sqrt(v[9]);
    if (std::isfinite(w->resid_after) && w->resid_after < w->fallback_ratio * w->resid_before) {
      const unsigned b = pm::red_blocks(nodes);
      {
        PM_TIMED("mg.update", kUpdBytes * nodes, s);
        update_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->e0, nodes, mg->partials);
        PM_CHECK_LAUNCH("update_kernel");
      }
      std::swap(mg->r32, lv0.res);  // the defect of the corrected phi
      PM_TRY_HIP(hipMemcpyAsync(mg->red + 2, mg->red + 10, sizeof(double),
                                hipMemcpyDeviceToDevice, s));
      w->used_net = 1;
    }
  }
  // the MIXED loop can start from the defect computed above (of phi, corrected or not)
  const bool ready = want_net_grid && mg->d.mode == PM_MG_MIXED;
  if ((st = solve_iterate(mg, phi, stats, s, ready))) return st;
  if (w->pair_target)
    return pm::warm::target(phi, mg->cand, n, w->pool, mg->scales, w->pair_target, s);
  return PM_OK;
}

pm_status pm_mg_vcycle(pm_mg* mg, hipStream_t s) {
  PM_REQUIRE(mg, "pm_mg_vcycle: null handle");
  PM_REQUIRE(mg->solved, "pm_mg_vcycle: no solve has run on this handle");
  if (mg->d.mode == PM_MG_MIXED) {
    pm_status st = mixed_residual(mg, mg->last_phi, s);
    if (st) return st;
    double e, p;
    return mixed_correct(mg, mg->last_phi, &e, &p, s);
  }
  return vcycle(mg, 0, s);
}

pm_status pm_mg_residual_norm(pm_mg* mg, double* out, hipStream_t s) {
  PM_REQUIRE(mg && out, "pm_mg_residual_norm: null pointer");
  PM_REQUIRE(mg->solved, "pm_mg_residual_norm: no solve has run on this handle");
  if (mg->d.mode == PM_MG_MIXED) {
    pm_status st = mixed_residual(mg, mg->last_phi, s);
    if (st) return st;
    double ss;
    if ((st = read_scalars(mg, mg->red + 2, 1, &ss, s))) return st;
    *out = std::sqrt(ss);
    return PM_OK;
  }
  return plain_residual_norm(mg, mg->last_phi, out, s);
}

pm_status pm_mg_moments(pm_mg* mg, const float* rho, double out[10], hipStream_t s) {
  PM_REQUIRE(mg && rho && out, "pm_mg_moments: null pointer");
  pm_status st = pm::launch_moments(rho, mg->lv[0].n, mg->d.dx, mg->d.grav, mg->partials,
                                    mg->moments, nullptr, s);
  if (st) return st;
  return read_scalars(mg, mg->moments, 10, out, s);
}

// ---- single-kernel hooks ----
pm_status pm_rb_pass(float* phi, const float* rhs, int32_t n, float h2, int32_t color,
                     hipStream_t s) {
  PM_REQUIRE(phi && rhs, "pm_rb_pass: null pointer");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_rb_pass: n must be in [1, 1024]");
  PM_REQUIRE(color == 0 || color == 1, "pm_rb_pass: color must be 0 or 1");
  return rb_pass(phi, rhs, n, h2, color, s);
}

pm_status pm_rb_sweep(const float* phi_in, float* phi_out, const float* rhs, int32_t n,
                      float h2, hipStream_t s) {
  PM_REQUIRE(phi_in && phi_out && rhs, "pm_rb_sweep: null pointer");
  PM_REQUIRE(phi_in != phi_out, "pm_rb_sweep: phi_in and phi_out must differ");
  PM_REQUIRE(n >= 3 && n <= 1024, "pm_rb_sweep: n must be in [3, 1024]");
  return rb_sweep(phi_in, phi_out, rhs, n, h2, s);
}

pm_status pm_residual(const float* phi, const float* rhs, float* res, int32_t n, float inv_h2,
                      hipStream_t s) {
  PM_REQUIRE(phi && rhs && res, "pm_residual: null pointer");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_residual: n must be in [1, 1024]");
  const int64_t nodes = static_cast<int64_t>(n) * n * n;
  residual_kernel<<<pm::blocks_for(nodes), pm::kBlock, 0, s>>>(phi, rhs, res, n, inv_h2);
  PM_CHECK_LAUNCH("residual_kernel");
  return PM_OK;
}

pm_status pm_restrict(const float* res_f, float* rhs_c, int32_t nf, int32_t nc, hipStream_t s) {
  PM_REQUIRE(res_f && rhs_c, "pm_restrict: null pointer");
  PM_REQUIRE(nc >= 1 && nc <= 1025 && nf >= 2 * nc - 2 && nf <= 2 * nc,
             "pm_restrict: need 2*nc-2 <= nf <= 2*nc (coarse J at fine 2J)");
  const int64_t nodes = static_cast<int64_t>(nc) * nc * nc;
  restrict_kernel<<<pm::blocks_for(nodes), pm::kBlock, 0, s>>>(res_f, rhs_c, nullptr, nf, nc);
  PM_CHECK_LAUNCH("restrict_kernel");
  return PM_OK;
}

pm_status pm_prolong_add(const float* phi_c, float* phi_f, int32_t nf, int32_t nc, hipStream_t s) {
  PM_REQUIRE(phi_c && phi_f, "pm_prolong_add: null pointer");
  PM_REQUIRE(nf >= 3 && nf <= 1024 && nc >= 2 && 2 * (nc - 1) >= nf - 2 && nf >= 2 * nc - 2,
             "pm_prolong_add: need 2*nc-2 <= nf <= 2*nc (coarse J at fine 2J)");
  const int64_t nodes = static_cast<int64_t>(nf) * nf * nf;
  prolong_add_kernel<<<pm::blocks_for(nodes), pm::kBlock, 0, s>>>(phi_c, phi_f, nf, nc);
  PM_CHECK_LAUNCH("prolong_add_kernel");
  return PM_OK;
}

pm_status pm_set_faces(float* phi, int32_t n, float dx, const float params[11], hipStream_t s) {
  PM_REQUIRE(phi && params, "pm_set_faces: null pointer");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_set_faces: n must be in [1, 1024]");
  pm::RedScratch* rs = nullptr;
  pm_status st = pm::red_scratch(&rs);
  if (st) return st;
  // stage the 11 parameters through the pinned mirror into device scratch
  float* dev = reinterpret_cast<float*>(rs->result);
  float* host = reinterpret_cast<float*>(rs->host);
  PM_TRY_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < 11; ++q) host[q] = params[q];
  PM_TRY_HIP(hipMemcpyAsync(dev, host, sizeof(float) * 11, hipMemcpyHostToDevice, s));
  faces_kernel<<<pm::blocks_for(6 * static_cast<int64_t>(n) * n), pm::kBlock, 0, s>>>(phi, n, dx,
                                                                                     dev);
  PM_CHECK_LAUNCH("faces_kernel");
  PM_TRY_HIP(hipStreamSynchronize(s));
  return PM_OK;
}

}  // extern "C"
