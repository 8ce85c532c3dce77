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
#include <algorithm>
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

// New colour-0 value of interior node c from the current values (rb_pass_kernel's
// expression; all six neighbours have colour 1).
__device__ __forceinline__ float red_value(const float* __restrict__ phi,
                                           const float* __restrict__ rhs, int64_t c, int64_t n2,
                                           int n, float h2) {
  const float s = phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] + phi[c - 1];
  return (s - h2 * rhs[c]) / 6.0f;
}

// One red-black sweep (colour 0 then colour 1) of a small level in one launch, out of place,
// one thread per node: a colour-0 node takes its new value from the old colour-1 values, a
// colour-1 node recomputes its six colour-0 neighbours' new values (faces keep theirs) and
// then takes its own, each with rb_pass_kernel's expression, so phi_out is bitwise what the
// two colour passes leave. Each colour-1 node repeats six neighbour updates: worth it only
// where a launch costs more than that work (small levels).
__global__ void __launch_bounds__(256) rb_sweep_small_kernel(const float* __restrict__ phi,
                                                             float* __restrict__ phi_out,
                                                             const float* __restrict__ rhs,
                                                             int n, float h2) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= n2 * n) return;
  const pm::Idx3 q3 = pm::idx3(c, n);
  const int i = q3.i, j = q3.j, k = q3.k;
  auto inner = [n](int a, int b, int d) {
    return a >= 1 && a <= n - 2 && b >= 1 && b <= n - 2 && d >= 1 && d <= n - 2;
  };
  if (!inner(i, j, k)) {
    phi_out[c] = phi[c];
    return;
  }
  if (((i + j + k) & 1) == 0) {
    phi_out[c] = red_value(phi, rhs, c, n2, n, h2);
    return;
  }
  auto nb = [&](int64_t cc, int a, int b, int d) {
    return inner(a, b, d) ? red_value(phi, rhs, cc, n2, n, h2) : phi[cc];
  };
  const float s = nb(c + n2, i + 1, j, k) + nb(c - n2, i - 1, j, k) + nb(c + n, i, j + 1, k) +
                  nb(c - n, i, j - 1, k) + nb(c + 1, i, j, k + 1) + nb(c - 1, i, j, k - 1);
  phi_out[c] = (s - h2 * rhs[c]) / 6.0f;
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
__device__ __forceinline__ float prolong_value(const float* __restrict__ phi_c, int nc, int i,
                                               int j, int k) {
  const int ci = i >> 1, cj = j >> 1, ck = k >> 1;
  const int oi = i & 1, oj = j & 1, ok = k & 1;
  const int64_t nc2 = static_cast<int64_t>(nc) * nc;
  // the 8 loads issue together; a zero-weight term (possibly one past the coarse grid's
  // edge, so its index is clamped) is loaded but not added
  float cv[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int a = ci + (c >> 2), b = cj + ((c >> 1) & 1), d = ck + (c & 1);
    cv[c] = phi_c[(a < nc ? a : nc - 1) * nc2 + static_cast<int64_t>(b < nc ? b : nc - 1) * nc +
                  (d < nc ? d : nc - 1)];
  }
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
        const float t = val + w * cv[di * 4 + dj * 2 + dk];
        val = w > 0.0f ? t : val;
      }
    }
  }
  return val;
}

__global__ void __launch_bounds__(256) prolong_add_kernel(const float* __restrict__ phi_c,
                                                          float* __restrict__ phi_f, int nf,
                                                          int nc) {
  const int64_t nf2 = static_cast<int64_t>(nf) * nf;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= nf2 * nf) return;
  const pm::Idx3 q3 = pm::idx3(c, nf);
  const int i = q3.i, j = q3.j, k = q3.k;
  if (i < 1 || i > nf - 2 || j < 1 || j > nf - 2 || k < 1 || k > nf - 2) return;
  phi_f[c] = phi_f[c] + prolong_value(phi_c, nc, i, j, k);
}

__global__ void __launch_bounds__(256) fill_f32_kernel(float* __restrict__ p, int64_t count,
                                                       float v) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
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
// so the result is bitwise that of the multi-kernel V-cycle; the residual lives in an LDS
// scratch of the finest bottom level's size between its evaluation and the restriction.
// ---------------------------------------------------------------------------------------
constexpr int kBottomMax = 17;
constexpr int kBottomLevels = 8;
constexpr int kBottomThreads = 1024;
constexpr int kBottomLoads = (kBottomMax * kBottomMax * kBottomMax + kBottomThreads - 1) / kBottomThreads;
constexpr int kCoarse64 = 5;  // coarsest levels up to this size are solved in float64 (MIXED)

struct BottomDesc {
  int levels;
  int n[kBottomLevels];
  int off[kBottomLevels];  // LDS offset (floats) of phi; rhs follows at off + n^3
  int res_off;             // LDS offset (floats) of the residual scratch ((n[0]-1)^3 floats)
  float h2[kBottomLevels];
  float inv_h2[kBottomLevels];
  float inv_n[kBottomLevels];     // 1/n, 1/((n+1)/2), 1/(n-1) as float: node indices are
  float inv_half[kBottomLevels];  // split by qdiv, a few VALU ops instead of an integer
  float inv_m[kBottomLevels];     // division per index
  float inv_cells[kBottomLevels];  // 1 / (((n-2)>>1) + 1): coarse cells per axis under level n
  float inv_in[kBottomLevels];     // 1 / (n-2): interior nodes per axis
  float inv_hi[kBottomLevels];     // 1 / ((n-1)>>1): most nodes of one colour per interior row
  double h2_coarsest;
  int pre, post, coarse_sweeps, coarse64_sweeps;
};

// x / d for 0 <= x < 2^15 and 1 <= d <= 64, given inv = 1.0f / d rounded to float: exact,
// since (x + 0.5) / d lies at least 0.5 / d from an integer and the two roundings move it
// by less than 2^-22 * x / d (checked exhaustively over that range)
__device__ __forceinline__ int qdiv(int x, float inv) {
  return static_cast<int>((static_cast<float>(x) + 0.5f) * inv);
}
static_assert(kBottomMax * kBottomMax * kBottomMax < (1 << 15), "qdiv range");

// One colour of interior nodes: lanes cover the interior rows (i, j) and, per row, the
// nodes k in 1..n-2 whose i + j + k has the colour's parity (k = 2 kk + 2 - p, p the
// parity of i + j + color). Nodes of one colour do not read each other, so the order in
// which lanes update them does not change the result.
__device__ __forceinline__ void lds_rb_pass(float* phi, const float* rhs, int n, float h2,
                                            int color, float inv_in, float inv_hi) {
  const int m = n - 2;
  const int hi = (n - 1) >> 1;  // most nodes of one colour in an interior row
  const int pairs = m * m * hi;
  const int n2 = n * n;
  for (int t = threadIdx.x; t < pairs; t += blockDim.x) {
    const int q = qdiv(t, inv_hi);  // t / hi
    const int kk = t - q * hi;
    const int ii = qdiv(q, inv_in);
    const int i = ii + 1;
    const int j = q - ii * m + 1;
    const int k = 2 * kk + 2 - ((i + j + color) & 1);
    if (k > n - 2) continue;
    const int c = i * n2 + j * n + k;
    const float sum = phi[c + n2] + phi[c - n2] + phi[c + n] + phi[c - n] + phi[c + 1] + phi[c - 1];
    phi[c] = (sum - h2 * rhs[c]) / 6.0f;
  }
  __syncthreads();
}

__device__ __forceinline__ void lds_smooth(float* phi, const float* rhs, int n, float h2,
                                           int sweeps, float inv_in, float inv_hi) {
  for (int q = 0; q < sweeps; ++q) {
    lds_rb_pass(phi, rhs, n, h2, 0, inv_in, inv_hi);
    lds_rb_pass(phi, rhs, n, h2, 1, inv_in, inv_hi);
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
  float* res = lds + d.res_off;  // residual scratch of the finest bottom level
  __shared__ double s64_phi[kCoarse64 * kCoarse64 * kCoarse64];
  __shared__ double s64_rhs[kCoarse64 * kCoarse64 * kCoarse64];
  const int L = d.levels;
  {
    // every load is issued before the first LDS store: one memory round trip, not one
    // per iteration (kBottomMax^3 <= kBottomLoads * kBottomThreads)
    const int nn = d.n[0] * d.n[0] * d.n[0];
    float* phi = lds + d.off[0];
    float* rhs = phi + nn;
    float vp[kBottomLoads], vr[kBottomLoads];
#pragma unroll
    for (int q = 0; q < kBottomLoads; ++q) {
      const int c = threadIdx.x + q * kBottomThreads;
      vp[q] = c < nn ? phi_top[c] : 0.0f;
      vr[q] = c < nn ? rhs_top[c] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kBottomLoads; ++q) {
      const int c = threadIdx.x + q * kBottomThreads;
      if (c < nn) {
        phi[c] = vp[q];
        rhs[c] = vr[q];
      }
    }
    __syncthreads();
  }
  // down
  for (int l = 0; l + 1 < L; ++l) {
    const int nf = d.n[l], nc = d.n[l + 1];
    float* phi = lds + d.off[l];
    float* rhs = phi + nf * nf * nf;
    lds_smooth(phi, rhs, nf, d.h2[l], d.pre, d.inv_in[l], d.inv_hi[l]);
    // residual of fine nodes 1..nf-1 per axis once into LDS (packed (nf-1)^3; 0 at the
    // face nf-1, which the restriction of an even level reaches; node 0 it never reads),
    // then the 27-point restriction reads it: the same values and the same summation
    // order as evaluating the residual inside the restriction
    const int m = nf - 1, m2 = m * m;
    for (int c = threadIdx.x; c < m2 * m; c += blockDim.x) {
      const int q = qdiv(c, d.inv_m[l]);
      const int k = c - q * m, i = qdiv(q, d.inv_m[l]), j = q - i * m;
      res[c] = lds_residual(phi, rhs, nf, i + 1, j + 1, k + 1, d.inv_h2[l]);
    }
    __syncthreads();
    float* cphi = lds + d.off[l + 1];
    float* crhs = cphi + nc * nc * nc;
    const int ncc = nc * nc * nc;
    for (int c = threadIdx.x; c < ncc; c += blockDim.x) {
      const int q = qdiv(c, d.inv_n[l + 1]);
      const int kc = c - q * nc, ic = qdiv(q, d.inv_n[l + 1]), jc = q - ic * nc;
      float out = 0.0f;
      if (ic >= 1 && ic <= nc - 2 && jc >= 1 && jc <= nc - 2 && kc >= 1 && kc <= nc - 2) {
        const float* r = res + (2 * ic - 1) * m2 + (2 * jc - 1) * m + (2 * kc - 1);
        float v[27];
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
          for (int b = -1; b <= 1; ++b)
#pragma unroll
            for (int e = -1; e <= 1; ++e) v[(a + 1) * 9 + (b + 1) * 3 + (e + 1)] = r[a * m2 + b * m + e];
        float acc = 0.0f;
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
          for (int b = -1; b <= 1; ++b)
#pragma unroll
            for (int e = -1; e <= 1; ++e) {
              const float w = (a == 0 ? 2.0f : 1.0f) * (b == 0 ? 2.0f : 1.0f) * (e == 0 ? 2.0f : 1.0f);
              acc += w * v[(a + 1) * 9 + (b + 1) * 3 + (e + 1)];
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
      // a 3^3 level has one interior node, whose neighbours are all (zero) faces: the first
      // sweep gives its exact value and every further sweep recomputes that same value
      const int sweeps = nc == 3 && d.coarse64_sweeps > 1 ? 1 : d.coarse64_sweeps;
      for (int q = 0; q < sweeps; ++q) {
        for (int color = 0; color < 2; ++color) {
          for (int t = threadIdx.x; t < nc * nc * half; t += blockDim.x) {
            const int q = qdiv(t, d.inv_half[L - 1]);
            const int kk = t - q * half, i = qdiv(q, d.inv_n[L - 1]), j = q - i * nc;
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
      lds_smooth(phi, rhs, nc, d.h2[L - 1], nc == 3 && d.coarse_sweeps > 1 ? 1 : d.coarse_sweeps,
                 d.inv_in[L - 1], d.inv_hi[L - 1]);
    }
  }
  // up
  for (int l = L - 2; l >= 0; --l) {
    const int nf = d.n[l], nc = d.n[l + 1];
    float* phi = lds + d.off[l];
    float* rhs = phi + nf * nf * nf;
    const float* cphi = lds + d.off[l + 1];
    const int nf2 = nf * nf, nc2 = nc * nc;
    // A fine node on an odd axis takes both coarse corners along it with weight 1/2, on
    // an even axis the lower one with weight 1; each node sums its nonzero terms in
    // (di, dj, dk) order, as the per-level kernel does, so the result is bitwise the same.
    // Corner ci+1 is clamped: on a halving level it lies past the grid only where its
    // weight is 0 (the odd fine node beyond it is a face).
    if (nf2 * nf <= kBottomThreads) {
      // small level: one lane per fine node (a single pass over the workgroup); all eight
      // corners are loaded first and terms of weight 0 leave val unchanged, as if skipped
      const int c = threadIdx.x;
      const int q = qdiv(c, d.inv_n[l]);
      const int k = c - q * nf, i = qdiv(q, d.inv_n[l]), j = q - i * nf;
      if (c < nf2 * nf && i >= 1 && i <= nf - 2 && j >= 1 && j <= nf - 2 && k >= 1 &&
          k <= nf - 2) {
        const int ci = i >> 1, cj = j >> 1, ck = k >> 1;
        const int oi = i & 1, oj = j & 1, ok = k & 1;
        const int i1 = min(ci + 1, nc - 1), j1 = min(cj + 1, nc - 1), k1 = min(ck + 1, nc - 1);
        float cv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          cv[u] = cphi[((u & 4) ? i1 : ci) * nc2 + ((u & 2) ? j1 : cj) * nc + ((u & 1) ? k1 : ck)];
        float val = 0.0f;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int di = u >> 2, dj = (u >> 1) & 1, dk = u & 1;
          const float wi = oi ? 0.5f : (di == 0 ? 1.0f : 0.0f);
          const float wj = oj ? 0.5f : (dj == 0 ? 1.0f : 0.0f);
          const float wk = ok ? 0.5f : (dk == 0 ? 1.0f : 0.0f);
          const float w = wi * wj * wk;
          const float t = val + w * cv[u];
          val = w > 0.0f ? t : val;
        }
        phi[c] = phi[c] + val;
      }
    } else {
      // larger level: one lane per coarse cell (ci, cj, ck), which loads the cell's eight
      // corners once and corrects the eight fine nodes (2ci+oi, 2cj+oj, 2ck+ok)
      const int cc = ((nf - 2) >> 1) + 1;  // cells per axis that hold interior fine nodes
      const float inv_cc = d.inv_cells[l];
      for (int t = threadIdx.x; t < cc * cc * cc; t += blockDim.x) {
        const int q = qdiv(t, inv_cc);
        const int ck = t - q * cc, ci = qdiv(q, inv_cc), cj = q - ci * cc;
        const int i1 = min(ci + 1, nc - 1), j1 = min(cj + 1, nc - 1), k1 = min(ck + 1, nc - 1);
        float cv[8];
  #pragma unroll
        for (int u = 0; u < 8; ++u)
          cv[u] = cphi[((u & 4) ? i1 : ci) * nc2 + ((u & 2) ? j1 : cj) * nc + ((u & 1) ? k1 : ck)];
  #pragma unroll
        for (int o = 0; o < 8; ++o) {
          const int oi = o >> 2, oj = (o >> 1) & 1, ok = o & 1;
          const int i = 2 * ci + oi, j = 2 * cj + oj, k = 2 * ck + ok;
          if (i < 1 || i > nf - 2 || j < 1 || j > nf - 2 || k < 1 || k > nf - 2) continue;
          float val = 0.0f;
  #pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int di = u >> 2, dj = (u >> 1) & 1, dk = u & 1;
            if ((di && !oi) || (dj && !oj) || (dk && !ok)) continue;  // weight 0: skipped
            const float w = (oi ? 0.5f : 1.0f) * (oj ? 0.5f : 1.0f) * (ok ? 0.5f : 1.0f);
            val += w * cv[u];
          }
          const int c = i * nf2 + j * nf + k;
          phi[c] = phi[c] + val;
        }
      }
    }
    __syncthreads();
    lds_smooth(phi, rhs, nf, d.h2[l], d.post, d.inv_in[l], d.inv_hi[l]);
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
                                                         const float* __restrict__ rho,
                                                         float factor, int n,
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
      r = static_cast<double>(factor * rho[c]) - lap;
    }
    r32[c] = static_cast<float>(r);
    v[0] = v[0] + r * r;
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

// phi += e (e is 0 on the faces); partial sums of e^2 and phi^2 (float64).
//
// Carried guess (the extrapolated warm start without its own pass): the next solve's guess
// 2*phi_t - phi_(t-1) lives in `guess`, and the correction updates it too.
//   kCarryNone:  phi = phi + e
//   kCarryFirst: phi = guess + e (the guess this solve started from); guess = 2*phi - old phi
//   kCarryMore:  phi = phi + e; guess = guess + 2*e (a further correction of the same solve)
enum { kCarryNone = 0, kCarryFirst = 1, kCarryMore = 2 };

template <int CARRY>
__device__ __forceinline__ float carry_update(float old, float& g, float d) {
  if (CARRY == kCarryFirst) {
    const float p = g + d;
    g = 2.0f * p - old;
    return p;
  }
  if (CARRY == kCarryMore) g = g + 2.0f * d;
  return old + d;
}

template <int CARRY>
__global__ void __launch_bounds__(256) update_kernel(float* __restrict__ phi,
                                                     const float* __restrict__ e,
                                                     float* __restrict__ guess, int64_t count,
                                                     double* __restrict__ partials) {
  __shared__ double s_red[4 * 2];
  double v[2] = {0.0, 0.0};
  int64_t r_begin, r_end;
  pm::block_range(count, r_begin, r_end);
  for (int64_t c = r_begin + threadIdx.x; c < r_end; c += blockDim.x) {
    const float d = e[c];
    float g = CARRY != kCarryNone ? guess[c] : 0.0f;
    const float p = carry_update<CARRY>(phi[c], g, d);
    phi[c] = p;
    if (CARRY != kCarryNone) guess[c] = g;
    v[0] = v[0] + static_cast<double>(d) * static_cast<double>(d);
    v[1] = v[1] + static_cast<double>(p) * static_cast<double>(p);
  }
  pm::block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * pm::kRedMaxQ] = v[0];
    partials[blockIdx.x * pm::kRedMaxQ + 1] = v[1];
  }
}

// ---------------------------------------------------------------------------------------
// Finest level of the MIXED correction in colour-split layout
// ---------------------------------------------------------------------------------------
// The correction e and the defect r of the finest level are internal to the solver, so they
// are stored split by colour: node (i, j, k) of colour q = (i+j+k)&1 sits at
//   q * plane + (i*n + j) * hn + (k >> 1),   hn = (n+1)/2, plane = n*n*hn
// (odd n leaves one pad slot in half of the rows). Every neighbour of a node has the other
// colour, and (i+-1, j, k), (i, j+-1, k) sit at the same k>>1, so a colour pass reads half
// of e and half of r, contiguously, instead of all of both. Each kernel below evaluates the
// same float expressions in the same order as its natural-layout counterpart, so the
// solve is bit for bit that of the natural-layout V-cycle (fill, rb passes, residual,
// prolongation, update), with three launches folded away:
//   * the correction starts at zero, so the first colour-0 pass needs no neighbour loads
//     (their sum is +0) and no fill precedes it: the first colour-0 and colour-1 passes
//     rewrite every interior node, and the faces are zero from allocation on;
//   * the colour-0 pass after the prolongation overwrites colour-0 nodes without reading
//     them, so the prolongation is added to colour-1 nodes only;
//   * the last colour-1 pass also applies phi += e to both colours and accumulates the
//     norms, so e is not written back or read again.
struct SplitIdx {
  int i, j, kk;
};
__device__ __forceinline__ SplitIdx split_idx(int64_t t, int n, int hn) {
  const uint32_t u = static_cast<uint32_t>(t), uh = static_cast<uint32_t>(hn);
  const uint32_t q = u / uh;
  SplitIdx r;
  r.kk = static_cast<int>(u - q * uh);
  r.j = static_cast<int>(q % static_cast<uint32_t>(n));
  r.i = static_cast<int>(q / static_cast<uint32_t>(n));
  return r;
}

// Sum of the six neighbours (other colour) of the node at position t, in the order of
// rb_pass_kernel: i+1, i-1, j+1, j-1, k+1, k-1.
__device__ __forceinline__ float split_nbr_sum(const float* __restrict__ eo, int64_t t, int kk,
                                               int k, int64_t nh, int hn) {
  const int64_t row = t - kk;
  return eo[t + nh] + eo[t - nh] + eo[t + hn] + eo[t - hn] + eo[row + ((k + 1) >> 1)] +
         eo[row + ((k - 1) >> 1)];
}

// One colour pass (rb_pass_kernel's update); ZERO: every neighbour is zero.
template <bool ZERO>
__global__ void __launch_bounds__(256) rbs_kernel(float* __restrict__ e,
                                                  const float* __restrict__ r, int n, int hn,
                                                  float h2, int color) {
  const int64_t plane = static_cast<int64_t>(n) * n * hn;
  const int64_t t = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  const SplitIdx q = split_idx(t, n, hn);
  const int k = 2 * q.kk + ((q.i + q.j + color) & 1);
  if (q.i < 1 || q.i > n - 2 || q.j < 1 || q.j > n - 2 || k < 1 || k > n - 2) return;
  const int64_t own = static_cast<int64_t>(color) * plane;
  const float sum =
      ZERO ? 0.0f
           : split_nbr_sum(e + (plane - own), t, q.kk, k, static_cast<int64_t>(n) * hn, hn);
  e[own + t] = (sum - h2 * r[own + t]) / 6.0f;
}

// residual_kernel from the split e and r into the natural-layout res.
__global__ void __launch_bounds__(256) residual_s_kernel(const float* __restrict__ e,
                                                         const float* __restrict__ r,
                                                         float* __restrict__ res, int n,
                                                         int hn, float inv_h2) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= n2 * n) return;
  const pm::Idx3 q3 = pm::idx3(c, n);
  const int i = q3.i, j = q3.j, k = q3.k;
  float out = 0.0f;
  if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2 && k >= 1 && k <= n - 2) {
    const int64_t plane = n2 * hn;
    const int64_t own = static_cast<int64_t>((i + j + k) & 1) * plane;
    const int64_t t = (static_cast<int64_t>(i) * n + j) * hn + (k >> 1);
    const float sum = split_nbr_sum(e + (plane - own), t, k >> 1, k,
                                    static_cast<int64_t>(n) * hn, hn);
    const float lap = (sum - 6.0f * e[own + t]) * inv_h2;
    out = r[own + t] - lap;
  }
  res[c] = out;
}

// prolong_add_kernel on the colour-1 nodes of the split e.
__global__ void __launch_bounds__(256) prolong_s_kernel(const float* __restrict__ phi_c,
                                                        float* __restrict__ e, int n, int hn,
                                                        int nc) {
  const int64_t plane = static_cast<int64_t>(n) * n * hn;
  const int64_t t = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (t >= plane) return;
  const SplitIdx q = split_idx(t, n, hn);
  const int k = 2 * q.kk + ((q.i + q.j + 1) & 1);
  if (q.i < 1 || q.i > n - 2 || q.j < 1 || q.j > n - 2 || k < 1 || k > n - 2) return;
  e[plane + t] = e[plane + t] + prolong_value(phi_c, nc, q.i, q.j, k);
}

// The last colour-1 pass fused with update_kernel: phi += e over every node (e = 0 on the
// faces), partial sums of e^2 and phi^2.
template <int CARRY>
__global__ void __launch_bounds__(256) black_update_kernel(const float* __restrict__ e,
                                                           const float* __restrict__ r,
                                                           float* __restrict__ phi,
                                                           float* __restrict__ guess, int n,
                                                           int hn, float h2,
                                                           double* __restrict__ partials) {
  __shared__ double s_red[4 * 2];
  double v[2] = {0.0, 0.0};
  const int64_t plane = static_cast<int64_t>(n) * n * hn;
  const int64_t nh = static_cast<int64_t>(n) * hn;
  int64_t r_begin, r_end;
  pm::block_range(plane, r_begin, r_end);
  for (int64_t t = r_begin + threadIdx.x; t < r_end; t += blockDim.x) {
    const SplitIdx q = split_idx(t, n, hn);
    const int64_t row = (static_cast<int64_t>(q.i) * n + q.j) * n;
    const int k0 = 2 * q.kk + ((q.i + q.j) & 1);
    if (k0 < n) {
      const float d = e[t];
      float g = CARRY != kCarryNone ? guess[row + k0] : 0.0f;
      const float p = carry_update<CARRY>(phi[row + k0], g, d);
      phi[row + k0] = p;
      if (CARRY != kCarryNone) guess[row + k0] = g;
      v[0] = v[0] + static_cast<double>(d) * static_cast<double>(d);
      v[1] = v[1] + static_cast<double>(p) * static_cast<double>(p);
    }
    const int k1 = 2 * q.kk + ((q.i + q.j + 1) & 1);
    if (k1 < n) {
      float d = 0.0f;
      if (q.i >= 1 && q.i <= n - 2 && q.j >= 1 && q.j <= n - 2 && k1 >= 1 && k1 <= n - 2)
        d = (split_nbr_sum(e, t, q.kk, k1, nh, hn) - h2 * r[plane + t]) / 6.0f;
      float g = CARRY != kCarryNone ? guess[row + k1] : 0.0f;
      const float p = carry_update<CARRY>(phi[row + k1], g, d);
      phi[row + k1] = p;
      if (CARRY != kCarryNone) guess[row + k1] = g;
      v[0] = v[0] + static_cast<double>(d) * static_cast<double>(d);
      v[1] = v[1] + static_cast<double>(p) * static_cast<double>(p);
    }
  }
  pm::block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * pm::kRedMaxQ] = v[0];
    partials[blockIdx.x * pm::kRedMaxQ + 1] = v[1];
  }
}

// residual64_kernel writing r32 in split layout (same sums, same partials).
__global__ void __launch_bounds__(256) residual64_s_kernel(const float* __restrict__ phi,
                                                           const float* __restrict__ rho,
                                                           float factor, int n,
                                                           int hn, double inv_h2,
                                                           float* __restrict__ r32s,
                                                           double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t total = n2 * n;
  const int64_t plane = n2 * hn;
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
      r = static_cast<double>(factor * rho[c]) - lap;
    }
    r32s[static_cast<int64_t>((i + j + k) & 1) * plane +
         (static_cast<int64_t>(i) * n + j) * hn + (k >> 1)] = static_cast<float>(r);
    v[0] = v[0] + r * r;
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

// --- 4-wide variants (n % 8 == 0): a thread owns positions kk0 .. kk0+3 of one row ------
// The index math runs once per 4 nodes and the loads are 16-byte vectors: the (i+-1) and
// (j+-1) rows, the node's own row of the other colour (its k+-1 neighbours, plus one
// scalar past either end) and the rhs. The per-node expressions are those of the scalar
// kernels, in the same order.
struct Quad {
  int i, j, kk0;
  uint32_t t0;  // position of kk0 within a colour plane
};
__device__ __forceinline__ Quad quad_idx(uint32_t u, int n, int hq) {
  // u = (i*n + j) * hq + kq, hq = hn / 4
  Quad q;
  q.t0 = u * 4u;
  if ((n & (n - 1)) == 0 && (hq & (hq - 1)) == 0) {
    // power-of-two n and hq (uniform over the launch): shifts and masks instead of the
    // integer divisions
    const uint32_t row = u >> __builtin_ctz(static_cast<uint32_t>(hq));
    q.kk0 = static_cast<int>(u & static_cast<uint32_t>(hq - 1)) * 4;
    q.j = static_cast<int>(row & static_cast<uint32_t>(n - 1));
    q.i = static_cast<int>(row >> __builtin_ctz(static_cast<uint32_t>(n)));
    return q;
  }
  const uint32_t row = u / static_cast<uint32_t>(hq);
  q.kk0 = static_cast<int>(u - row * static_cast<uint32_t>(hq)) * 4;
  q.j = static_cast<int>(row % static_cast<uint32_t>(n));
  q.i = static_cast<int>(row / static_cast<uint32_t>(n));
  return q;
}
__device__ __forceinline__ float4 ld4(const float* __restrict__ p, uint32_t off) {
  return *reinterpret_cast<const float4*>(p + off);
}
__device__ __forceinline__ float at4(const float4& v, int e) {
  return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}

// Neighbour sums of the 4 nodes of parity p = (i+j+color)&1 in a row of the colour-plane
// `eo` (the other colour), in rb_pass_kernel's order. m = the other colour's same-row quad,
// edge = eo[t0-1] (p == 0) or eo[t0+4] (p == 1).
// the loads of quad_nbr_sums, and the sums from them (split so that a kernel can issue
// several quads' loads before any arithmetic)
struct NbrQuad {
  float4 a, b, c, d, m;
  float edge;
};

__device__ __forceinline__ NbrQuad quad_nbr_load(const float* __restrict__ eo, uint32_t t0,
                                                 uint32_t nh, uint32_t hn, int p) {
  NbrQuad v;
  v.a = ld4(eo, t0 + nh);
  v.b = ld4(eo, t0 - nh);
  v.c = ld4(eo, t0 + hn);
  v.d = ld4(eo, t0 - hn);
  v.m = ld4(eo, t0);
  v.edge = p ? eo[t0 + 4] : eo[t0 - 1];
  return v;
}

__device__ __forceinline__ void quad_nbr_sum(const NbrQuad& v, int p, float s[4]) {
  // k = 2(kk0+e)+p: the k+1 neighbour sits at kk0+e+p, the k-1 neighbour at kk0+e+p-1
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float kp = p ? (e < 3 ? at4(v.m, e + 1) : v.edge) : at4(v.m, e);
    const float km = p ? at4(v.m, e) : (e > 0 ? at4(v.m, e - 1) : v.edge);
    s[e] = at4(v.a, e) + at4(v.b, e) + at4(v.c, e) + at4(v.d, e) + kp + km;
  }
}

__device__ __forceinline__ void quad_nbr_sums(const float* __restrict__ eo, uint32_t t0,
                                              uint32_t nh, uint32_t hn, int p, float s[4]) {
  quad_nbr_sum(quad_nbr_load(eo, t0, nh, hn, p), p, s);
}

template <bool ZERO>
__global__ void __launch_bounds__(256) rbs4_kernel(float* __restrict__ e,
                                                   const float* __restrict__ r, int n, int hn,
                                                   float h2, int color) {
  const int hq = hn >> 2;
  const uint32_t quads = static_cast<uint32_t>(n) * n * hq;
  const int64_t u = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (u >= quads) return;
  const Quad q = quad_idx(static_cast<uint32_t>(u), n, hq);
  if (q.i < 1 || q.i > n - 2 || q.j < 1 || q.j > n - 2) return;
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  const uint32_t own = static_cast<uint32_t>(color) * plane;
  const int p = (q.i + q.j + color) & 1;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  NbrQuad nb;
  if (!ZERO)
    nb = quad_nbr_load(e + (plane - own), q.t0, static_cast<uint32_t>(n) * hn,
                       static_cast<uint32_t>(hn), p);
  const float4 rr = ld4(r, own + q.t0);
  __builtin_amdgcn_sched_barrier(0);  // all loads leave before any arithmetic
  if (!ZERO) quad_nbr_sum(nb, p, s);
  float out[4];
#pragma unroll
  for (int k4 = 0; k4 < 4; ++k4) {
    const int k = 2 * (q.kk0 + k4) + p;
    // faces (k = 0, k = n-1) and pad slots keep their zero
    out[k4] = (k >= 1 && k <= n - 2) ? (s[k4] - h2 * at4(rr, k4)) / 6.0f : 0.0f;
  }
  *reinterpret_cast<float4*>(e + own + q.t0) = make_float4(out[0], out[1], out[2], out[3]);
}

// The first sweep of a cycle's pre-smoothing in one pass: rbs4_kernel<true> on colour 0
// (zero guess: e0 = (0 - h2*r0)/6, 0 on faces) then rbs4_kernel<false> on colour 1, whose
// colour-0 neighbours are those same values, computed here from r0 instead of read back
// (r is 0 on faces and pad slots, so they come out 0 as written). Bitwise the two passes.
__device__ __forceinline__ float zero_guess_value(float h2, float r) {
  return (0.0f - h2 * r) / 6.0f;
}

__global__ void __launch_bounds__(256) rbs4_first_kernel(float* __restrict__ e,
                                                         const float* __restrict__ r, int n,
                                                         int hn, float h2) {
  const int hq = hn >> 2;
  const uint32_t quads = static_cast<uint32_t>(n) * n * hq;
  const int64_t u = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (u >= quads) return;
  const Quad q = quad_idx(static_cast<uint32_t>(u), n, hq);
  if (q.i < 1 || q.i > n - 2 || q.j < 1 || q.j > n - 2) return;
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  const uint32_t nh = static_cast<uint32_t>(n) * hn;
  const int p0 = (q.i + q.j) & 1, p1 = p0 ^ 1;
  // colour-0 defect around the quad (as quad_nbr_sums reads e) and colour-1 defect
  const float4 a = ld4(r, q.t0 + nh), b = ld4(r, q.t0 - nh), c = ld4(r, q.t0 + hn),
               d = ld4(r, q.t0 - hn), m = ld4(r, q.t0);
  const float edge = p1 ? r[q.t0 + 4] : r[q.t0 - 1];
  const float4 r1 = ld4(r, plane + q.t0);
  float out0[4], out1[4];
#pragma unroll
  for (int k4 = 0; k4 < 4; ++k4) {
    const int k0 = 2 * (q.kk0 + k4) + p0;
    out0[k4] = (k0 >= 1 && k0 <= n - 2) ? zero_guess_value(h2, at4(m, k4)) : 0.0f;
    const float kp = p1 ? (k4 < 3 ? zero_guess_value(h2, at4(m, k4 + 1)) : zero_guess_value(h2, edge))
                        : zero_guess_value(h2, at4(m, k4));
    const float km = p1 ? zero_guess_value(h2, at4(m, k4))
                        : (k4 > 0 ? zero_guess_value(h2, at4(m, k4 - 1)) : zero_guess_value(h2, edge));
    const float sum = zero_guess_value(h2, at4(a, k4)) + zero_guess_value(h2, at4(b, k4)) +
                      zero_guess_value(h2, at4(c, k4)) + zero_guess_value(h2, at4(d, k4)) + kp + km;
    const int k1 = 2 * (q.kk0 + k4) + p1;
    out1[k4] = (k1 >= 1 && k1 <= n - 2) ? (sum - h2 * at4(r1, k4)) / 6.0f : 0.0f;
  }
  *reinterpret_cast<float4*>(e + q.t0) = make_float4(out0[0], out0[1], out0[2], out0[3]);
  *reinterpret_cast<float4*>(e + plane + q.t0) = make_float4(out1[0], out1[1], out1[2], out1[3]);
}

// residual_s_kernel, 4 positions of both colours (8 consecutive natural nodes) per thread.
__global__ void __launch_bounds__(256) residual_s4_kernel(const float* __restrict__ e,
                                                          const float* __restrict__ r,
                                                          float* __restrict__ res, int n,
                                                          int hn, float inv_h2) {
  const int hq = hn >> 2;
  const uint32_t quads = static_cast<uint32_t>(n) * n * hq;
  const int64_t u = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (u >= quads) return;
  const Quad q = quad_idx(static_cast<uint32_t>(u), n, hq);
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  float out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (q.i >= 1 && q.i <= n - 2 && q.j >= 1 && q.j <= n - 2) {
    const uint32_t nh = static_cast<uint32_t>(n) * hn;
    // both colours' loads issue before either colour is evaluated
    float s[2][4];
    NbrQuad nb[2];
    float4 ev[2], rv[2];
#pragma unroll
    for (int color = 0; color < 2; ++color) {
      const uint32_t own = static_cast<uint32_t>(color) * plane;
      nb[color] = quad_nbr_load(e + (plane - own), q.t0, nh, static_cast<uint32_t>(hn),
                                (q.i + q.j + color) & 1);
      ev[color] = ld4(e, own + q.t0);
      rv[color] = ld4(r, own + q.t0);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int color = 0; color < 2; ++color) quad_nbr_sum(nb[color], (q.i + q.j + color) & 1, s[color]);
    // every node is evaluated and the faces masked afterwards (a branch around the face
    // terms would let the compiler move the loads only they use behind it)
#pragma unroll
    for (int color = 0; color < 2; ++color) {
      const int p = (q.i + q.j + color) & 1;
#pragma unroll
      for (int k4 = 0; k4 < 4; ++k4) {
        const int k = 2 * (q.kk0 + k4) + p;
        const float lap = (s[color][k4] - 6.0f * at4(ev[color], k4)) * inv_h2;
        const float val = at4(rv[color], k4) - lap;
        out[2 * k4 + p] = (k >= 1 && k <= n - 2) ? val : 0.0f;
      }
    }
  }
  float* dst = res + (static_cast<int64_t>(q.i) * n + q.j) * n + 2 * q.kk0;
  if (2 * q.kk0 + 8 <= n) {
    *reinterpret_cast<float4*>(dst) = make_float4(out[0], out[1], out[2], out[3]);
    *reinterpret_cast<float4*>(dst + 4) = make_float4(out[4], out[5], out[6], out[7]);
  } else {
    for (int k = 0; k < 8 && 2 * q.kk0 + k < n; ++k) dst[k] = out[k];
  }
}

// prolong_s_kernel, 4 colour-1 positions per thread. The 4 nodes share i, j and the parity
// of k, so their coarse terms come from the same (di, dj) rows at 4 or 5 consecutive coarse
// k, loaded once; each node's sum keeps prolong_value's order.
// Coarse rows (ci0 + a, cj0 + b), a < 2, b < kProlongMaxCj, of a prolongation workgroup,
// staged whole (nc floats each) in LDS.
constexpr int kProlongMaxLds = 4200;  // 2*(2048/n + 1)*(n/2 + 1) <= 4100 floats for n <= 1024

// Quads per thread of prolong_s4_kernel: quads u0 + tid + 256*m, m < kProlongQ.
constexpr int kProlongQ = 2;
constexpr int kProlongBlock = 256;  // threads of a prolong_s4_kernel workgroup
constexpr int kProlongStageIt = (kProlongMaxLds + kProlongBlock - 1) / kProlongBlock;
static_assert(kProlongBlock == pm::kBlock, "prolong_s4_kernel is launched with pm::kBlock");

__global__ void __launch_bounds__(kProlongBlock) prolong_s4_kernel(const float* __restrict__ phi_c,
                                                         float* __restrict__ e, int n, int hn,
                                                         int nc) {
  __shared__ float s_c[kProlongMaxLds];
  const int hq = hn >> 2;
  const uint32_t quads = static_cast<uint32_t>(n) * n * hq;
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  const int64_t u0 = pm::xcd_block() * (blockDim.x * kProlongQ);
  const uint32_t nc2 = static_cast<uint32_t>(nc) * nc;
  // this thread's quads and their colour-1 values, loaded before the staging below
  Quad q[kProlongQ];
  bool live[kProlongQ];
  float4 v[kProlongQ];
#pragma unroll
  for (int m = 0; m < kProlongQ; ++m) {
    const int64_t u = u0 + threadIdx.x + static_cast<int64_t>(m) * blockDim.x;
    live[m] = u < quads;
    q[m] = quad_idx(live[m] ? static_cast<uint32_t>(u) : 0u, n, hq);
    live[m] = live[m] && q[m].i >= 1 && q[m].i <= n - 2 && q[m].j >= 1 && q[m].j <= n - 2;
    v[m] = live[m] ? ld4(e, plane + q[m].t0) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // The workgroup's quads cover fine rows r0..r1 (row = i*n + j). When they share one i,
  // the coarse values they read lie in coarse rows ci0..ci0+1 x cj0..cj1, staged in LDS
  // with coalesced loads (each quad would otherwise issue 20 strided scalar loads); the
  // values and the arithmetic are those of the direct loads.
  const uint32_t span = blockDim.x * kProlongQ;
  const uint32_t last = static_cast<uint32_t>(u0) + span - 1 < quads
                            ? static_cast<uint32_t>(u0) + span - 1
                            : quads - 1;
  const uint32_t r0 = static_cast<uint32_t>(u0) / static_cast<uint32_t>(hq);
  const uint32_t r1 = last / static_cast<uint32_t>(hq);
  const int i0 = static_cast<int>(r0 / static_cast<uint32_t>(n));
  const int ci0 = i0 >> 1;
  const int cj0 = static_cast<int>(r0 % static_cast<uint32_t>(n)) >> 1;
  const int ncj = (static_cast<int>(r1 % static_cast<uint32_t>(n)) >> 1) + 2 - cj0;
  const bool staged = static_cast<uint32_t>(u0) < quads &&
                      static_cast<int>(r1 / static_cast<uint32_t>(n)) == i0 &&
                      ci0 + 1 < nc && cj0 + ncj <= nc && 2 * ncj * nc <= kProlongMaxLds;
  if (staged) {
    // rows cj0..cj0+ncj-1 of coarse plane ci0 + a are one contiguous run of per floats:
    // the workgroup's values are two runs, loaded all at once (every load issued before
    // the first LDS store)
    const int per = ncj * nc;
    const uint32_t run0 = static_cast<uint32_t>(ci0) * nc2 + static_cast<uint32_t>(cj0) * nc;
    float tmp[kProlongStageIt];
#pragma unroll
    for (int it = 0; it < kProlongStageIt; ++it) {
      const int t = static_cast<int>(threadIdx.x) + it * kProlongBlock;
      const uint32_t src = run0 + static_cast<uint32_t>(t < per ? t : t - per + nc2);
      tmp[it] = t < 2 * per ? phi_c[src] : 0.0f;
    }
#pragma unroll
    for (int it = 0; it < kProlongStageIt; ++it) {
      const int t = static_cast<int>(threadIdx.x) + it * kProlongBlock;
      if (t < 2 * per) s_c[t] = tmp[it];
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kProlongQ; ++m) {
    if (!live[m]) continue;
    const Quad& qq = q[m];
    const int p = (qq.i + qq.j + 1) & 1;  // parity of k; k>>1 = kk0 + e
    const int ci = qq.i >> 1, cj = qq.j >> 1, oi = qq.i & 1, oj = qq.j & 1;
    float c[2][2][5];
#pragma unroll
    for (int di = 0; di < 2; ++di) {
#pragma unroll
      for (int dj = 0; dj < 2; ++dj) {
        const bool need = (di == 0 || oi) && (dj == 0 || oj);
        if (staged) {
          const int row = ((ci + di - ci0) * ncj + (cj + dj - cj0)) * nc + qq.kk0;
#pragma unroll
          for (int x = 0; x < 5; ++x)
            c[di][dj][x] = (need && (x < 4 || p)) ? s_c[row + x] : 0.0f;
        } else {
          const uint32_t row = (ci + di) * nc2 + static_cast<uint32_t>(cj + dj) * nc + qq.kk0;
#pragma unroll
          for (int x = 0; x < 5; ++x)
            c[di][dj][x] = (need && (x < 4 || p)) ? phi_c[row + x] : 0.0f;
        }
      }
    }
    float out[4] = {v[m].x, v[m].y, v[m].z, v[m].w};
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
      const int k = 2 * (qq.kk0 + k4) + p;
      if (k < 1 || k > n - 2) continue;
      float val = 0.0f;
#pragma unroll
      for (int di = 0; di < 2; ++di) {
        const float wi = oi ? 0.5f : (di == 0 ? 1.0f : 0.0f);
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
          const float wj = oj ? 0.5f : (dj == 0 ? 1.0f : 0.0f);
#pragma unroll
          for (int dk = 0; dk < 2; ++dk) {
            const float wk = p ? 0.5f : (dk == 0 ? 1.0f : 0.0f);
            const float w = wi * wj * wk;
            if (w > 0.0f) val += w * c[di][dj][k4 + dk];
          }
        }
      }
      out[k4] = out[k4] + val;
    }
    *reinterpret_cast<float4*>(e + plane + qq.t0) = make_float4(out[0], out[1], out[2], out[3]);
  }
}

// residual64_s_kernel, 8 consecutive natural nodes per thread (n % 8 == 0): the r values
// are the same; the squared norm is summed in another order. A thread takes kRes64Q such
// runs (quads u, u + 256, ...) and issues the loads of all of them before any arithmetic
// (2 measured slower than 1 at 256^3: 140 VGPRs halve the waves in flight).
constexpr int kRes64Q = 1;
__global__ void __launch_bounds__(256) residual64_s4_kernel(const float* __restrict__ phi,
                                                            const float* __restrict__ rho,
                                                            float factor, int n,
                                                            int hn, double inv_h2,
                                                            float* __restrict__ r32s,
                                                            double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  const int hq = hn >> 2;
  const int64_t quads = static_cast<int64_t>(n) * n * hq;
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  const uint32_t n2 = static_cast<uint32_t>(n) * n;
  int64_t r_begin, r_end;
  pm::block_range(quads, r_begin, r_end);
  const int64_t step = static_cast<int64_t>(blockDim.x);
  for (int64_t u0 = r_begin + threadIdx.x; u0 < r_end; u0 += kRes64Q * step) {
    Quad q[kRes64Q];
    bool live[kRes64Q], in[kRes64Q];
    uint32_t c0[kRes64Q];
    float ctr[kRes64Q][10], ip[kRes64Q][8], im[kRes64Q][8], jp[kRes64Q][8], jm[kRes64Q][8],
        fv[kRes64Q][8];
    auto load8 = [&](float* dst, const float* src, uint32_t off) {
      const float4 a = ld4(src, off), b = ld4(src, off + 4);
      dst[0] = a.x; dst[1] = a.y; dst[2] = a.z; dst[3] = a.w;
      dst[4] = b.x; dst[5] = b.y; dst[6] = b.z; dst[7] = b.w;
    };
#pragma unroll
    for (int m = 0; m < kRes64Q; ++m) {
      const int64_t u = u0 + m * step;
      live[m] = u < r_end;
      q[m] = quad_idx(live[m] ? static_cast<uint32_t>(u) : 0u, n, hq);
      c0[m] = (static_cast<uint32_t>(q[m].i) * n + q[m].j) * n + 2 * q[m].kk0;
      in[m] = live[m] && q[m].i >= 1 && q[m].i <= n - 2 && q[m].j >= 1 && q[m].j <= n - 2;
      if (in[m]) {
        const uint32_t c = c0[m];
        ctr[m][0] = phi[c - 1];
        load8(&ctr[m][1], phi, c);
        ctr[m][9] = phi[c + 8];
        load8(ip[m], phi, c + n2);
        load8(im[m], phi, c - n2);
        load8(jp[m], phi, c + n);
        load8(jm[m], phi, c - n);
        load8(fv[m], rho, c);
      }
    }
    // all loads leave before any arithmetic: the scheduler would otherwise hold some back
    // to save registers, a second round trip for them
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int m = 0; m < kRes64Q; ++m) {
      if (!live[m]) continue;
      float r8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (in[m]) {
#pragma unroll
        for (int x = 0; x < 8; ++x) fv[m][x] = factor * fv[m][x];  // f = 4*pi*G*rho, as the setup
        // every node is evaluated and the faces (k = 0, n-1) are masked afterwards: were the
        // face terms skipped by a branch, the compiler would move the loads that only they
        // use behind it, a second round trip to memory for most rows
#pragma unroll
        for (int x = 0; x < 8; ++x) {
          const int k = 2 * q[m].kk0 + x;
          const bool inner = k >= 1 && k <= n - 2;
          const double sum = static_cast<double>(ip[m][x]) + static_cast<double>(im[m][x]) +
                             static_cast<double>(jp[m][x]) + static_cast<double>(jm[m][x]) +
                             static_cast<double>(ctr[m][x + 2]) + static_cast<double>(ctr[m][x]);
          const double lap = (sum - 6.0 * static_cast<double>(ctr[m][x + 1])) * inv_h2;
          const double r = static_cast<double>(fv[m][x]) - lap;
          r8[x] = inner ? static_cast<float>(r) : 0.0f;
          const double acc = v[0] + r * r;
          v[0] = inner ? acc : v[0];
        }
      }
      // colour of natural k = 2*kk0 + x is (p0 + x) & 1; its position kk0 + (x >> 1)
      const int p0 = (q[m].i + q[m].j) & 1;
      float a0[4], a1[4];
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        if (((p0 + x) & 1) == 0) a0[x >> 1] = r8[x]; else a1[x >> 1] = r8[x];
      }
      *reinterpret_cast<float4*>(r32s + q[m].t0) = make_float4(a0[0], a0[1], a0[2], a0[3]);
      *reinterpret_cast<float4*>(r32s + plane + q[m].t0) =
          make_float4(a1[0], a1[1], a1[2], a1[3]);
    }
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];  // one column (pm::fold_columns_kernel)
}

// black_update_kernel, 4 positions of both colours (8 natural nodes) per thread (n % 8 == 0).
template <int CARRY>
__global__ void __launch_bounds__(256) black_update4_kernel(const float* __restrict__ e,
                                                            const float* __restrict__ r,
                                                            float* __restrict__ phi,
                                                            float* __restrict__ guess, int n,
                                                            int hn, float h2,
                                                            double* __restrict__ partials) {
  __shared__ double s_red[4 * 2];
  double v[2] = {0.0, 0.0};
  const int hq = hn >> 2;
  const int64_t quads = static_cast<int64_t>(n) * n * hq;
  const uint32_t plane = static_cast<uint32_t>(n) * n * hn;
  const uint32_t nh = static_cast<uint32_t>(n) * hn;
  int64_t r_begin, r_end;
  pm::block_range(quads, r_begin, r_end);
  for (int64_t u = r_begin + threadIdx.x; u < r_end; u += blockDim.x) {
    const Quad q = quad_idx(static_cast<uint32_t>(u), n, hq);
    const bool inner_row = q.i >= 1 && q.i <= n - 2 && q.j >= 1 && q.j <= n - 2;
    const int p0 = (q.i + q.j) & 1, p1 = p0 ^ 1;
    const int64_t row0 = (static_cast<int64_t>(q.i) * n + q.j) * n + 2 * q.kk0;
    float* dst = phi + row0;
    float* gdst = CARRY != kCarryNone ? guess + row0 : nullptr;
    float ph[8], gv[8];
    // rows are whole quads (the kernel runs for n % 8 == 0 only)
    auto load8 = [&](float* out, const float* src) {
      const float4 a = *reinterpret_cast<const float4*>(src);
      const float4 b = *reinterpret_cast<const float4*>(src + 4);
      out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
      out[4] = b.x; out[5] = b.y; out[6] = b.z; out[7] = b.w;
    };
    auto store8 = [&](float* out, const float* val) {
      *reinterpret_cast<float4*>(out) = make_float4(val[0], val[1], val[2], val[3]);
      *reinterpret_cast<float4*>(out + 4) = make_float4(val[4], val[5], val[6], val[7]);
    };
    // every load of the quad issues before any is used: phi (and the guess), e, and the
    // neighbour sums, which a face row takes from row (1, 1) and does not use
    load8(ph, dst);
    if (CARRY != kCarryNone) load8(gv, gdst);
    const float4 e0 = ld4(e, q.t0);
    const uint32_t tn = inner_row ? q.t0 : nh + static_cast<uint32_t>(hn) + q.kk0;
    float s[4];
    quad_nbr_sums(e, tn, nh, static_cast<uint32_t>(hn), p1, s);
    const float4 rv = ld4(r, plane + q.t0);
    __builtin_amdgcn_sched_barrier(0);  // all loads leave before any arithmetic
    // every colour-1 value is evaluated and faces masked afterwards (a branch around them
    // would let the compiler move the loads only they use behind it)
    float d1[4];
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) {
      const int k = 2 * (q.kk0 + k4) + p1;
      const float val = (s[k4] - h2 * at4(rv, k4)) / 6.0f;
      d1[k4] = (inner_row && k >= 1 && k <= n - 2) ? val : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int k4 = k >> 1;
      const float d = ((k & 1) == p0) ? at4(e0, k4) : d1[k4];
      float g = CARRY != kCarryNone ? gv[k] : 0.0f;
      const float pv = carry_update<CARRY>(ph[k], g, d);
      ph[k] = pv;
      if (CARRY != kCarryNone) gv[k] = g;
      v[0] = v[0] + static_cast<double>(d) * static_cast<double>(d);
      v[1] = v[1] + static_cast<double>(pv) * static_cast<double>(pv);
    }
    store8(dst, ph);
    if (CARRY != kCarryNone) store8(gdst, gv);
  }
  pm::block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {  // two columns of gridDim.x partials (pm::fold_columns_kernel)
    partials[blockIdx.x] = v[0];
    partials[gridDim.x + blockIdx.x] = v[1];
  }
}

// natural -> split copy (the learned warm start hands over its defect in natural layout).
__global__ void __launch_bounds__(256) to_split_kernel(const float* __restrict__ src,
                                                       float* __restrict__ dst, int n, int hn) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t c = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (c >= n2 * n) return;
  const pm::Idx3 q3 = pm::idx3(c, n);
  dst[static_cast<int64_t>((q3.i + q3.j + q3.k) & 1) * n2 * hn +
      (static_cast<int64_t>(q3.i) * n + q3.j) * hn + (q3.k >> 1)] = src[c];
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
  float* f0 = nullptr;        // 4*pi*G*rho, float32 (PLAIN, and the network's input)
  const float* cur_rho = nullptr;  // rho of the last setup (the float64 defect scales it)
  float factor = 0.0f;             // 4*pi*G as float32
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
  // carried guess (EXTRAPOLATE without pairs, MIXED): the next solve's guess
  // 2*phi_t - phi_(t-1), written by the updates of the solve before (no extrapolation pass)
  float* guess = nullptr;
  bool guess_valid = false;
  const float* guess_phi = nullptr;  // the phi buffer the carried guess belongs to
  int warm_pool = 0;          // network pool factor the buffers below are sized for
  float* pooled = nullptr;    // [R^3][3] pooled input channels
  uint16_t* net_in = nullptr; // [R^3][8] bf16 network input
  float* net_out = nullptr;   // [R^3] network output
  float* scales = nullptr;    // [8] input/output scales
  std::vector<void*> owned;
  // V-cycle launches replayed from HIP graphs (captured on a private stream, launched on the
  // caller's): one graph per kind and set of buffer addresses it bakes in.
  struct CycleGraph {
    int kind;
    const void* key[4];
    hipGraphExec_t exec;
  };
  std::vector<CycleGraph> graphs;
  hipStream_t capture = nullptr;
  bool use_graphs = true;
  // MIXED: the finest level's correction and defect in colour-split layout (allocated when
  // the hierarchy has a coarser level below the finest and pre/post sweeps are >= 1)
  float* e0s = nullptr;
  float* r32s = nullptr;
  int hn = 0;
  bool use_split = true;
  bool use_bottom = true;  // levels from bottom_from on run in the one-workgroup solver
  bool use_small_sweep = true;  // levels <= kSmallSweepMax smooth with rb_sweep_small
  // the 4-wide split reductions (float64 defect, fused update) run one quad per thread:
  // quad_blocks workgroups with their own partial sums
  unsigned quad_blocks = 0;
  unsigned res_blocks = 0;  // workgroups (and partials) of residual64_s4_kernel
  double* partials_q = nullptr;
  // the float64 defect's partials when their fold is left to the cycle's own fold launch
  double* partials_r = nullptr;
  bool split_active() const { return e0s != nullptr && use_split; }
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
  for (auto& g : mg->graphs) (void)hipGraphExecDestroy(g.exec);
  if (mg->capture) (void)hipStreamDestroy(mg->capture);
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

// Levels up to this size (below the finest) smooth with rb_sweep_small_kernel, one launch
// per sweep ping-ponging through L.tmp, instead of two colour passes: there a launch costs
// more than the repeated neighbour updates (PM_MG_OPT_SMALL_SWEEP).
constexpr int kSmallSweepMax = 65;

pm_status rb_sweep_small(const float* phi_in, float* phi_out, const float* rhs, int n,
                         float h2, hipStream_t s) {
  const int64_t nodes = static_cast<int64_t>(n) * n * n;
  PM_TIMED(level_name("mg.rb_sweep_small", n), 12.0 * nodes, s);
  rb_sweep_small_kernel<<<pm::blocks_for(nodes), pm::kBlock, 0, s>>>(phi_in, phi_out, rhs, n,
                                                                      h2);
  PM_CHECK_LAUNCH("rb_sweep_small_kernel");
  return PM_OK;
}

// `sweeps` sweeps from *cur (L.phi or L.tmp), each into the other buffer; *cur becomes the
// buffer holding the result.
pm_status smooth_pp(const Level& L, int sweeps, float** cur, hipStream_t s) {
  pm_status st;
  for (int q = 0; q < sweeps; ++q) {
    float* out = *cur == L.phi ? L.tmp : L.phi;
    if ((st = rb_sweep_small(*cur, out, L.rhs, L.n, L.h2, s))) return st;
    *cur = out;
  }
  return PM_OK;
}

// `sweeps` red-black sweeps in place on L.phi. Large levels ping-pong pairs of fused sweeps
// through L.tmp (ending back in L.phi); a leftover single sweep uses two colour passes.
pm_status smooth(const Level& L, int sweeps, hipStream_t s) {
  pm_status st;
  int q = 0;
  if (L.tmp && L.n >= kSweepMin) {
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
    d.inv_n[q] = 1.0f / static_cast<float>(L.n);
    d.inv_half[q] = 1.0f / static_cast<float>((L.n + 1) >> 1);
    d.inv_m[q] = 1.0f / static_cast<float>(L.n - 1);
    d.inv_cells[q] = 1.0f / static_cast<float>(((L.n - 2) >> 1) + 1);
    d.inv_in[q] = 1.0f / static_cast<float>(L.n - 2);
    d.inv_hi[q] = 1.0f / static_cast<float>((L.n - 1) >> 1);
    d.off[q] = off;
    off += 2 * L.n * L.n * L.n;
  }
  const Level& top = mg->lv[l];
  d.res_off = off;
  off += (top.n - 1) * (top.n - 1) * (top.n - 1);
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
  if (mg->use_bottom && mg->bottom_from >= 0 && static_cast<int>(l) >= mg->bottom_from)
    return bottom(mg, l, s);
  if (l + 1 == mg->lv.size()) return coarsest(mg, L, s);
  pm_status st;
  // small levels: one launch per sweep through L.tmp; an even number of sweeps in the cycle
  // leaves the result in L.phi
  const bool pp = mg->use_small_sweep && L.tmp && L.n <= kSmallSweepMax &&
                  (mg->d.pre_sweeps + mg->d.post_sweeps) % 2 == 0;
  float* cur = L.phi;
  if (pp) {
    if ((st = smooth_pp(L, mg->d.pre_sweeps, &cur, s))) return st;
  } else if ((st = smooth(L, mg->d.pre_sweeps, s))) {
    return st;
  }
  {
    PM_TIMED(level_name("mg.residual", L.n), kResidualBytes * L.nodes(), s);
    residual_kernel<<<pm::blocks_for(L.nodes()), pm::kBlock, 0, s>>>(cur, L.rhs, L.res, L.n,
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
    prolong_add_kernel<<<pm::blocks_for(L.nodes()), pm::kBlock, 0, s>>>(C.phi, cur, L.n, C.n);
    PM_CHECK_LAUNCH("prolong_add_kernel");
  }
  if (pp) return smooth_pp(L, mg->d.post_sweeps, &cur, s);  // ends in L.phi
  return smooth(L, mg->d.post_sweeps, s);
}

// Workgroups of the split path's fused update (the partials the fold has to sum).
unsigned split_update_blocks(const pm_mg* mg, const float* phi, double** partials) {
  const int n = mg->lv[0].n;
  const int64_t plane = static_cast<int64_t>(n) * n * mg->hn;
  if (mg->quad_blocks && (reinterpret_cast<uintptr_t>(phi) & 15u) == 0) {
    *partials = mg->partials_q;
    return mg->quad_blocks;
  }
  *partials = mg->partials;
  return pm::red_blocks(plane);
}

// The finest level of the MIXED correction in split layout: e0s = V(r32s) from a zero
// start (pre/post sweeps >= 1), then phi += e0s with the norms in partials.
pm_status vcycle_split0(pm_mg* mg, float* phi, int carry, hipStream_t s) {
  const Level& L = mg->lv[0];
  const Level& C = mg->lv[1];
  const int n = L.n, hn = mg->hn;
  const int64_t plane = static_cast<int64_t>(n) * n * hn;
  const double nodes = static_cast<double>(L.nodes());
  float* e = mg->e0s;
  const float* r = mg->r32s;
  // 4 positions per thread when rows split into aligned quads (n % 8 == 0)
  const bool quad = n % 8 == 0;
  const bool phi_aligned = (reinterpret_cast<uintptr_t>(phi) & 15u) == 0;
  const int64_t quads = plane / 4;
  auto pass = [&](int color, bool zero) -> pm_status {
    PM_TIMED(level_name(zero ? "mg.rbs0" : "mg.rbs", n), (zero ? 4.0 : 6.0) * nodes, s);
    if (quad && zero)
      rbs4_kernel<true><<<pm::blocks_for(quads), pm::kBlock, 0, s>>>(e, r, n, hn, L.h2, color);
    else if (quad)
      rbs4_kernel<false><<<pm::blocks_for(quads), pm::kBlock, 0, s>>>(e, r, n, hn, L.h2, color);
    else if (zero)
      rbs_kernel<true><<<pm::blocks_for(plane), pm::kBlock, 0, s>>>(e, r, n, hn, L.h2, color);
    else
      rbs_kernel<false><<<pm::blocks_for(plane), pm::kBlock, 0, s>>>(e, r, n, hn, L.h2, color);
    PM_CHECK_LAUNCH("rbs_kernel");
    return PM_OK;
  };
  pm_status st;
  for (int q = 0; q < mg->d.pre_sweeps; ++q) {
    if (q == 0 && quad) {  // zero-guess colour 0 and colour 1 in one pass
      PM_TIMED(level_name("mg.rbs_first", n), 8.0 * nodes, s);
      rbs4_first_kernel<<<pm::blocks_for(quads), pm::kBlock, 0, s>>>(e, r, n, hn, L.h2);
      PM_CHECK_LAUNCH("rbs4_first_kernel");
      continue;
    }
    if ((st = pass(0, q == 0)) || (st = pass(1, false))) return st;
  }
  {
    PM_TIMED(level_name("mg.residual_s", n), kResidualBytes * nodes, s);
    if (quad)
      residual_s4_kernel<<<pm::blocks_for(quads), pm::kBlock, 0, s>>>(e, r, L.res, n, hn,
                                                                      L.inv_h2);
    else
      residual_s_kernel<<<pm::blocks_for(L.nodes()), pm::kBlock, 0, s>>>(e, r, L.res, n, hn,
                                                                          L.inv_h2);
    PM_CHECK_LAUNCH("residual_s_kernel");
  }
  {
    PM_TIMED(level_name("mg.restrict", n), kRestrictBytes * nodes, s);
    restrict_kernel<<<pm::blocks_for(C.nodes()), pm::kBlock, 0, s>>>(L.res, C.rhs, C.phi, n,
                                                                      C.n);
    PM_CHECK_LAUNCH("restrict_kernel");
  }
  if ((st = vcycle(mg, 1, s))) return st;
  {
    PM_TIMED(level_name("mg.prolong_s", n), 4.0 * nodes, s);
    if (quad)
      prolong_s4_kernel<<<pm::blocks_for((quads + kProlongQ - 1) / kProlongQ), pm::kBlock, 0,
                          s>>>(C.phi, e, n, hn, C.n);
    else
      prolong_s_kernel<<<pm::blocks_for(plane), pm::kBlock, 0, s>>>(C.phi, e, n, hn, C.n);
    PM_CHECK_LAUNCH("prolong_s_kernel");
  }
  for (int q = 0; q < mg->d.post_sweeps; ++q) {
    if ((st = pass(0, false))) return st;
    if (q + 1 < mg->d.post_sweeps && (st = pass(1, false))) return st;
  }
  {
    PM_TIMED("mg.black_update", (carry != kCarryNone ? 20.0 : 12.0) * nodes, s);
    double* partials;
    const unsigned blocks = split_update_blocks(mg, phi, &partials);
    float* g = mg->guess;
    const bool vec = quad && phi_aligned && (carry == kCarryNone ||
                                             (reinterpret_cast<uintptr_t>(g) & 15u) == 0);
#define PM_BLACK_UPDATE(K, C)                                                        \
  K<C><<<blocks, pm::kBlock, 0, s>>>(e, r, phi, g, n, hn, L.h2, partials)
    if (vec && carry == kCarryFirst)
      PM_BLACK_UPDATE(black_update4_kernel, kCarryFirst);
    else if (vec && carry == kCarryMore)
      PM_BLACK_UPDATE(black_update4_kernel, kCarryMore);
    else if (vec)
      PM_BLACK_UPDATE(black_update4_kernel, kCarryNone);
    else if (carry == kCarryFirst)
      PM_BLACK_UPDATE(black_update_kernel, kCarryFirst);
    else if (carry == kCarryMore)
      PM_BLACK_UPDATE(black_update_kernel, kCarryMore);
    else
      PM_BLACK_UPDATE(black_update_kernel, kCarryNone);
#undef PM_BLACK_UPDATE
    PM_CHECK_LAUNCH("black_update_kernel");
  }
  return PM_OK;
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

// r = f - A phi (float64 arithmetic) -> r32 (natural), or r32s (split layout) when `split`;
// its sum of squares lands in red[2] (no sync).
// defer: the 8-node path only (the caller checks quad_defect_ok): the partials go to
// partials_r and the cycle's fold launch folds them into red[2] (kCycleCarry*Fold).
pm_status mixed_residual(pm_mg* mg, const float* phi, hipStream_t s, bool split = false,
                         bool defer = false) {
  const Level& L = mg->lv[0];
  // 8 nodes per thread on split grids with n % 8 == 0 and 16-byte aligned phi and rho
  const bool quad = split && mg->quad_blocks && (reinterpret_cast<uintptr_t>(phi) & 15u) == 0 &&
                    (reinterpret_cast<uintptr_t>(mg->cur_rho) & 15u) == 0;
  const unsigned b = quad ? mg->res_blocks : pm::red_blocks(L.nodes());
  double* partials = quad ? (defer ? mg->partials_r : mg->partials_q) : mg->partials;
  {
    PM_TIMED("mg.residual64", kRes64Bytes * L.nodes(), s);
    if (quad)
      residual64_s4_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->cur_rho, mg->factor, L.n, mg->hn,
                                                     1.0 / (L.h * L.h),
                                                     mg->r32s, partials);
    else if (split)
      residual64_s_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->cur_rho, mg->factor, L.n, mg->hn,
                                                    1.0 / (L.h * L.h),
                                                    mg->r32s, mg->partials);
    else
      residual64_kernel<<<b, pm::kBlock, 0, s>>>(phi, mg->cur_rho, mg->factor, L.n,
                                                  1.0 / (L.h * L.h),
                                                  mg->r32, mg->partials);
    PM_CHECK_LAUNCH("residual64_kernel");
  }
  if (quad && defer) return PM_OK;
  if (quad)
    pm::fold_columns_kernel<1><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(b),
                                                         mg->red + 2);
  else
    pm::fold_partials_kernel<1><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(b),
                                                          mg->red + 2);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  return PM_OK;
}

// The launches of one V-cycle step. kCycleCorrect (MIXED): e0 = 0, e0 = V(r32), phi += e0,
// red[0] = ||e0||^2, red[1] = ||phi||^2. kCyclePlain (PLAIN): phi = V(phi) on f0.
// kCycleCarryFirst / kCycleCarryMore: kCycleCorrect with the carried guess (update_kernel).
// kCycleCarry*Fold: also fold the deferred defect partials (partials_r) into red[2], in the
// launch that folds the update's norms.
enum {
  kCycleCorrect = 0,
  kCyclePlain = 1,
  kCycleCarryFirst = 2,
  kCycleCarryMore = 3,
  kCycleCarryFirstFold = 4,
  kCycleCarryMoreFold = 5
};
constexpr size_t kMaxGraphs = 6;

pm_status enqueue_cycle(pm_mg* mg, int kind, float* phi, hipStream_t s) {
  Level& L0 = mg->lv[0];
  pm_status st;
  if (kind == kCyclePlain) return vcycle(mg, 0, s);
  const bool fold_defect = kind == kCycleCarryFirstFold || kind == kCycleCarryMoreFold;
  const int carry = (kind == kCycleCarryFirst || kind == kCycleCarryFirstFold) ? kCarryFirst
                    : (kind == kCycleCarryMore || kind == kCycleCarryMoreFold) ? kCarryMore
                                                                                : kCarryNone;
  if (mg->split_active()) {
    if ((st = vcycle_split0(mg, phi, carry, s))) return st;
    double* partials;
    const unsigned blocks = split_update_blocks(mg, phi, &partials);
    if (fold_defect) {
      PM_REQUIRE(partials == mg->partials_q, "multigrid: deferred defect fold off the 8-node path");
      pm::fold_columns_pair_kernel<2, 1><<<1, pm::kBlock, 0, s>>>(
          partials, static_cast<int>(blocks), mg->red, mg->partials_r,
          static_cast<int>(mg->res_blocks), mg->red + 2);
    } else if (partials == mg->partials_q)  // black_update4_kernel's columns
      pm::fold_columns_kernel<2><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(blocks),
                                                           mg->red);
    else
      pm::fold_partials_kernel<2><<<1, pm::kBlock, 0, s>>>(partials, static_cast<int>(blocks),
                                                            mg->red);
    PM_CHECK_LAUNCH("fold_partials_kernel<2>");
    return PM_OK;
  }
  if ((st = fill(mg->e0, L0.nodes(), 0.0f, s))) return st;
  if ((st = vcycle(mg, 0, s))) return st;
  const unsigned b = pm::red_blocks(L0.nodes());
  {
    PM_TIMED("mg.update", (carry != kCarryNone ? kUpdBytes + 8.0 : kUpdBytes) * L0.nodes(), s);
    if (carry == kCarryFirst)
      update_kernel<kCarryFirst><<<b, pm::kBlock, 0, s>>>(phi, mg->e0, mg->guess, L0.nodes(),
                                                          mg->partials);
    else if (carry == kCarryMore)
      update_kernel<kCarryMore><<<b, pm::kBlock, 0, s>>>(phi, mg->e0, mg->guess, L0.nodes(),
                                                         mg->partials);
    else
      update_kernel<kCarryNone><<<b, pm::kBlock, 0, s>>>(phi, mg->e0, nullptr, L0.nodes(),
                                                         mg->partials);
    PM_CHECK_LAUNCH("update_kernel");
  }
  pm::fold_partials_kernel<2><<<1, pm::kBlock, 0, s>>>(mg->partials, static_cast<int>(b),
                                                        mg->red);
  PM_CHECK_LAUNCH("fold_partials_kernel<2>");
  return PM_OK;
}

// Runs enqueue_cycle on stream s, from a HIP graph when graphs are on and per-kernel timing
// is off (the timer's events need eager launches). A cycle is some 40 launches, most of
// them on coarse levels that finish faster than the host can launch them; a graph replays
// them back to back. The graph bakes in the addresses of phi and of the level-0 buffers
// (the learned warm start swaps r32 with the level-0 residual), so it is keyed on them.
pm_status run_cycle(pm_mg* mg, int kind, float* phi, hipStream_t s) {
  if (!mg->use_graphs || pm::timing_enabled()) return enqueue_cycle(mg, kind, phi, s);
  const Level& L0 = mg->lv[0];
  const void* key[4] = {phi, L0.phi, L0.rhs, L0.res};
  hipGraphExec_t exec = nullptr;
  for (const auto& g : mg->graphs)
    if (g.kind == kind && std::equal(key, key + 4, g.key)) exec = g.exec;
  if (!exec) {
    if (!mg->capture) PM_TRY_HIP(hipStreamCreateWithFlags(&mg->capture, hipStreamNonBlocking));
    PM_TRY_HIP(hipStreamBeginCapture(mg->capture, hipStreamCaptureModeThreadLocal));
    const pm_status st = enqueue_cycle(mg, kind, phi, mg->capture);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(mg->capture, &graph);
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return pm::hip_fail("hipStreamEndCapture (multigrid cycle)", e);
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return pm::hip_fail("hipGraphInstantiate (multigrid cycle)", ei);
    if (mg->graphs.size() >= kMaxGraphs) {
      (void)hipGraphExecDestroy(mg->graphs.front().exec);
      mg->graphs.erase(mg->graphs.begin());
    }
    pm_mg::CycleGraph g{kind, {key[0], key[1], key[2], key[3]}, exec};
    mg->graphs.push_back(g);
  }
  PM_TRY_HIP(hipGraphLaunch(exec, s));
  return PM_OK;
}

// One defect-correction iteration: e = V(r32); phi += e. Returns (synchronizing) red[0..3]:
// ||e||^2, ||phi||^2, the squared defect norm this correction started from, ||f||^2.
pm_status mixed_correct(pm_mg* mg, float* phi, double v[4], hipStream_t s,
                        int kind = kCycleCorrect) {
  pm_status st;
  if ((st = run_cycle(mg, kind, phi, s))) return st;
  return read_scalars(mg, mg->red, 4, v, s);
}

// Moments, f = 4*pi*G*rho, the initial guess (zero unless warm) with multipole faces, and
// ||f|| (into red[3]). Enqueued only. f is stored in f0 for the PLAIN cycle and the
// network's input (store_f); the MIXED defect scales rho itself with the same float product.
pm_status solve_setup(pm_mg* mg, const float* rho, float* phi, bool warm, hipStream_t s,
                      bool store_f) {
  Level& L0 = mg->lv[0];
  const int n = L0.n;
  const int64_t nodes = L0.nodes();
  pm_status st;
  // boundary moments (multigrid.py:458) -> float32 face parameters, on the device, with
  // f = 4*pi*G*rho (multigrid.py:463) and ||f||^2 (multigrid.py:477, into red[3]) from the
  // same pass over rho
  {
    const float factor = static_cast<float>(4.0 * kPi * mg->d.grav);
    mg->cur_rho = rho;
    mg->factor = factor;
    PM_TIMED("mg.moments", (store_f ? 8.0 : 4.0) * nodes, s);
    if ((st = pm::launch_moments(rho, n, mg->d.dx, mg->d.grav, mg->partials, mg->moments,
                                 mg->face_params, s, store_f ? mg->f0 : nullptr, factor,
                                 mg->red + 3)))
      return st;
  }
  if (!warm && (st = fill(phi, nodes, 0.0f, s))) return st;
  {
    PM_TIMED("mg.faces", 4.0 * 6 * n * n, s);
    faces_kernel<<<pm::blocks_for(6 * static_cast<int64_t>(n) * n), pm::kBlock, 0, s>>>(
        phi, n, mg->d.dx, mg->face_params);
    PM_CHECK_LAUNCH("faces_kernel");
  }
  return PM_OK;
}

// The V-cycle loop of a solve whose setup has run (PLAIN: reference stop rule; MIXED:
// defect correction with the error-estimate stop). Fills stats. residual_ready: MIXED only,
// r32 already holds the defect of phi and red[2] its squared norm. carry: MIXED only, the
// solve starts from mg->guess and its corrections carry the next guess (kCycleCarry*).
pm_status solve_iterate(pm_mg* mg, float* phi, pm_mg_stats* stats, hipStream_t s,
                        bool residual_ready = false, bool carry = false) {
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
      if ((st = run_cycle(mg, kCyclePlain, phi, s))) return st;
      if (c + 1 >= mg->d.min_cycles && rhs_norm > 0.0) {
        if ((st = plain_residual_norm(mg, phi, &resid, s))) return st;
        if (resid < mg->d.tol * rhs_norm) {
          cycles = c + 1;
          break;
        }
      }
    }
  } else {
    const bool split = mg->split_active();
    L0.phi = split ? mg->e0s : mg->e0;
    L0.rhs = split ? mg->r32s : mg->r32;
    mg->last_phi = phi;
    // carried solves on the 8-node path fold the defect norm in the cycle's fold launch
    auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
    const bool defer = carry && split && mg->quad_blocks && aligned(phi) && aligned(mg->guess) &&
                       aligned(mg->cur_rho);
    if (!residual_ready) {
      if ((st = mixed_residual(mg, carry ? mg->guess : phi, s, split, defer))) return st;
    } else if (split) {
      PM_TIMED("mg.to_split", 8.0 * L0.nodes(), s);
      to_split_kernel<<<pm::blocks_for(L0.nodes()), pm::kBlock, 0, s>>>(mg->r32, mg->r32s,
                                                                        L0.n, mg->hn);
      PM_CHECK_LAUNCH("to_split_kernel");
    }
    double prev_e = -1.0;
    double v[4] = {0.0, 0.0, -1.0, 0.0};
    bool read = false;
    for (int c = 0; c < mg->d.n_cycles; ++c) {
      const int kind = !carry ? kCycleCorrect
                       : c == 0   ? (defer ? kCycleCarryFirstFold : kCycleCarryFirst)
                                  : (defer ? kCycleCarryMoreFold : kCycleCarryMore);
      if ((st = mixed_correct(mg, phi, v, s, kind))) return st;
      read = true;
      const double e_norm = std::sqrt(v[0]);
      const double phi_norm = std::sqrt(v[1]);
      // contraction estimate from successive corrections; 0.5 before two are known
      double rate = prev_e > 0.0 ? e_norm / prev_e : 0.5;
      if (rate > 0.95) rate = 0.95;
      prev_e = e_norm;
      err_est = phi_norm > 0.0 ? e_norm * rate / (1.0 - rate) / phi_norm : 0.0;
      if (c + 1 >= mg->d.min_cycles && err_est < mg->d.tol) {
        cycles = c + 1;
        break;
      }
      if (c + 1 < mg->d.n_cycles && (st = mixed_residual(mg, phi, s, split, defer))) return st;
    }
    // no correction ran: phi is the guess (as an in-place guess would leave it)
    if (!read && carry)
      PM_TRY_HIP(hipMemcpyAsync(phi, mg->guess, sizeof(float) * L0.nodes(),
                                hipMemcpyDeviceToDevice, s));
    // the last correction's read holds the defect it started from and ||f||^2
    if (!read && (st = read_scalars(mg, mg->red, 4, v, s))) return st;
    resid = std::sqrt(v[2]);
    rhs_norm = std::sqrt(v[3]);
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
    if (L.n >= kSweepMin || (l > 0 && l + 1 < sizes.size() && L.n <= kSmallSweepMax)) {
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
    if (mg->lv.size() > 1 && mg->bottom_from != 0 && d.pre_sweeps >= 1 && d.post_sweeps >= 1) {
      const int n = mg->lv[0].n;
      mg->hn = (n + 1) / 2;
      const size_t bytes = sizeof(float) * 2 * static_cast<size_t>(n) * n * mg->hn;
      if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->e0s), bytes)) ||
          (st = dalloc(mg, reinterpret_cast<void**>(&mg->r32s), bytes))) {
        destroy(mg);
        return st;
      }
      // faces and pad slots of e0s stay zero for good (no kernel writes them)
      hipError_t ze = hipMemset(mg->e0s, 0, bytes);
      if (ze == hipSuccess) ze = hipMemset(mg->r32s, 0, bytes);
      if (ze == hipSuccess) ze = hipDeviceSynchronize();
      if (ze != hipSuccess) {
        destroy(mg);
        return pm::hip_fail("hipMemset (multigrid split buffers)", ze);
      }
      if (n % 8 == 0) {
        const int64_t quads = static_cast<int64_t>(n) * n * mg->hn / 4;
        mg->quad_blocks = static_cast<unsigned>((quads + pm::kBlock - 1) / pm::kBlock);
        mg->res_blocks = static_cast<unsigned>((quads + pm::kBlock * kRes64Q - 1) /
                                               (pm::kBlock * kRes64Q));
        if ((st = dalloc(mg, reinterpret_cast<void**>(&mg->partials_q),
                         sizeof(double) * pm::kRedMaxQ * mg->quad_blocks)) ||
            (st = dalloc(mg, reinterpret_cast<void**>(&mg->partials_r),
                         sizeof(double) * mg->quad_blocks))) {
          destroy(mg);
          return st;
        }
      }
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
  mg->guess_valid = false;
  pm_status st;
  if ((st = solve_setup(mg, rho, phi, warm != 0, s, mg->d.mode == PM_MG_PLAIN))) return st;
  return solve_iterate(mg, phi, stats, s);
}

pm_status pm_mg_warm_reset(pm_mg* mg) {
  PM_REQUIRE(mg, "pm_mg_warm_reset: null handle");
  mg->hist_valid = false;
  mg->guess_valid = false;
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

  // Extrapolation without training pairs on the MIXED solver: the guess 2*phi_t - phi_(t-1)
  // was written by the previous solve's last update (bitwise the extrapolation pass's
  // guess after a one-cycle solve), so this solve reads it where it lies.
  if (w->mode == PM_WARM_EXTRAPOLATE && !pairs && mg->d.mode == PM_MG_MIXED) {
    if (!mg->guess && (st = dalloc(mg, reinterpret_cast<void**>(&mg->guess),
                                   sizeof(float) * nodes)))
      return st;
    const bool have = mg->guess_valid && mg->guess_phi == phi;
    if (!have)  // first solve of a sequence: phi itself is the guess
      PM_TRY_HIP(hipMemcpyAsync(mg->guess, phi, sizeof(float) * nodes, hipMemcpyDeviceToDevice,
                                s));
    mg->hist_valid = false;  // the explicit time history is not kept alongside
    mg->guess_valid = false;
    if ((st = solve_setup(mg, rho, mg->guess, true, s, false))) return st;  // faces into it
    w->extrapolated = have ? 1 : 0;
    if ((st = solve_iterate(mg, phi, stats, s, false, true))) return st;
    mg->guess_valid = mg->d.n_cycles > 0;
    mg->guess_phi = phi;
    return PM_OK;
  }
  mg->guess_valid = false;
  if ((st = solve_setup(mg, rho, phi, true, s, mg->d.mode == PM_MG_PLAIN || want_net_grid)))
    return st;
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
        update_kernel<kCarryNone><<<b, pm::kBlock, 0, s>>>(phi, mg->e0, nullptr, nodes,
                                                           mg->partials);
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
    pm_status st = mixed_residual(mg, mg->last_phi, s, mg->split_active());
    if (st) return st;
    double v[4];
    return mixed_correct(mg, mg->last_phi, v, s);
  }
  return vcycle(mg, 0, s);
}

pm_status pm_mg_set_option(pm_mg* mg, int32_t option, int32_t value) {
  pm::clear_error();
  PM_REQUIRE(mg, "pm_mg_set_option: null handle");
  PM_REQUIRE(option == PM_MG_OPT_GRAPHS || option == PM_MG_OPT_SPLIT_FINE ||
                 option == PM_MG_OPT_BOTTOM || option == PM_MG_OPT_SMALL_SWEEP,
             "pm_mg_set_option: unknown option");
  // captured cycles bake in the launch sequence the options select
  for (auto& g : mg->graphs) (void)hipGraphExecDestroy(g.exec);
  mg->graphs.clear();
  switch (option) {
    case PM_MG_OPT_GRAPHS:
      mg->use_graphs = value != 0;
      break;
    case PM_MG_OPT_SPLIT_FINE:
      mg->use_split = value != 0;
      break;
    case PM_MG_OPT_SMALL_SWEEP:
      mg->use_small_sweep = value != 0;
      break;
    default:
      mg->use_bottom = value != 0;
      break;
  }
  return PM_OK;
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
