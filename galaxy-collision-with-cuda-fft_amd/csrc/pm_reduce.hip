// Float64 reductions: sums of squares, rho moments (multigrid faces), and the
// conservation diagnostics of reference diagnostics_device.py:41-87, 152-184.
#include <map>
#include <mutex>

#include "pm_reduce.h"

namespace {

__global__ void __launch_bounds__(256) sumsq_f32_kernel(const float* __restrict__ x,
                                                        int64_t count,
                                                        double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  int64_t r_begin, r_end;
  pm::block_range(count, r_begin, r_end);
  for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x) {
    const double t = static_cast<double>(x[i]);
    v[0] = v[0] + t * t;
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

__global__ void __launch_bounds__(256) sumsq_f64_kernel(const double* __restrict__ x,
                                                        int64_t count,
                                                        double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  int64_t r_begin, r_end;
  pm::block_range(count, r_begin, r_end);
  for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x) {
    const double t = x[i];
    v[0] = v[0] + t * t;
  }
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

// Ten raw moments (reference multigrid.py:133-178): m = rho*dx^3, (x,y,z) = (i,j,k)*dx.
// With F, the same pass also sums f^2 (f = factor * rho, multigrid.py:249-252) as quantity
// 10 and, with STORE, writes f. Workgroup b stores quantity q at partials[q * blocks + b]
// (quantity-major, read coalesced by fold_moments_kernel).
template <bool F, bool STORE>
__global__ void __launch_bounds__(256) moments_kernel(const float* __restrict__ rho,
                                                      pm::FastDiv fn, double dxd, double vol,
                                                      double* __restrict__ partials,
                                                      float* __restrict__ f, float factor) {
  const int n = static_cast<int>(fn.d);
  constexpr int Q = F ? 11 : 10;
  __shared__ double s_red[4 * Q];
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = 0.0;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t total = n2 * n;
  int64_t r_begin, r_end;
  pm::block_range(total, r_begin, r_end);
  for (int64_t idx = r_begin + threadIdx.x; idx < r_end; idx += blockDim.x) {
    const pm::Idx3 q3 = pm::idx3(idx, fn);
    const int i = q3.i, j = q3.j, k = q3.k;
    const float r = rho[idx];
    if (F) {
      const float fv = factor * r;
      if (STORE) f[idx] = fv;
      const double t = static_cast<double>(fv);
      v[Q - 1] = v[Q - 1] + t * t;
    }
    const double m = static_cast<double>(r) * vol;
    const double x = static_cast<double>(i) * dxd;
    const double y = static_cast<double>(j) * dxd;
    const double z = static_cast<double>(k) * dxd;
    v[0] = v[0] + m;
    v[1] = v[1] + m * x;
    v[2] = v[2] + m * y;
    v[3] = v[3] + m * z;
    v[4] = v[4] + m * x * x;
    v[5] = v[5] + m * y * y;
    v[6] = v[6] + m * z * z;
    v[7] = v[7] + m * x * y;
    v[8] = v[8] + m * x * z;
    v[9] = v[9] + m * y * z;
  }
  pm::block_sum<Q>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) partials[q * gridDim.x + blockIdx.x] = v[q];
  }
}

// Centre of mass and second moments about it (parallel axis), as reference
// multigrid.py:378-389 does on the host in float64, rounded to float32 face parameters.
__device__ __forceinline__ void moments_finalize(const double* s, double grav,
                                                 float* __restrict__ params) {
  const double total = s[0];
  double cx = 0, cy = 0, cz = 0, sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
  double mass = 0.0;
  if (total > 0.0) {
    mass = total;
    cx = s[1] / total;
    cy = s[2] / total;
    cz = s[3] / total;
    sxx = s[4] - total * cx * cx;
    syy = s[5] - total * cy * cy;
    szz = s[6] - total * cz * cz;
    sxy = s[7] - total * cx * cy;
    sxz = s[8] - total * cx * cz;
    syz = s[9] - total * cy * cz;
  }
  params[0] = static_cast<float>(grav);
  params[1] = static_cast<float>(mass);
  params[2] = static_cast<float>(cx);
  params[3] = static_cast<float>(cy);
  params[4] = static_cast<float>(cz);
  params[5] = static_cast<float>(sxx);
  params[6] = static_cast<float>(syy);
  params[7] = static_cast<float>(szz);
  params[8] = static_cast<float>(sxy);
  params[9] = static_cast<float>(sxz);
  params[10] = static_cast<float>(syz);
}

// The moments' stage-2 fold (fold_partials_kernel's order and tree over quantity-major
// partials, Q = 10 or 11 with the sum of f^2 as quantity 10) with the face parameters
// computed from its result on thread 0.
template <int Q>
__global__ void __launch_bounds__(256) fold_moments_kernel(const double* __restrict__ partials,
                                                           int blocks, double grav,
                                                           double* __restrict__ sums,
                                                           double* __restrict__ f_sumsq,
                                                           float* __restrict__ params) {
  __shared__ double s_red[4 * Q];
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = 0.0;
  pm::fold_loop<Q>(partials, blocks, 1, blocks, v);  // the plain loop's order
  pm::block_sum<Q>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 10; ++q) sums[q] = v[q];
    if (Q > 10 && f_sumsq) f_sumsq[0] = v[Q - 1];
    if (params) moments_finalize(v, grav, params);
  }
}

// [M, KE, Px, Py, Pz, Lx, Ly, Lz, Smx, Smy, Smz] (reference diagnostics_device.py:41-74).
__global__ void __launch_bounds__(256) particle_scalars_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const float* __restrict__ mass, int64_t np, double* __restrict__ partials) {
  __shared__ double s_red[4 * 11];
  double v[11];
#pragma unroll
  for (int q = 0; q < 11; ++q) v[q] = 0.0;
  int64_t r_begin, r_end;
  pm::block_range(np, r_begin, r_end);
  for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x) {
    const double m = mass[i];
    const double x = px[i], y = py[i], z = pz[i];
    const double ux = vx[i], uy = vy[i], uz = vz[i];
    v[0] = v[0] + m;
    v[1] = v[1] + 0.5 * m * (ux * ux + uy * uy + uz * uz);
    v[2] = v[2] + m * ux;
    v[3] = v[3] + m * uy;
    v[4] = v[4] + m * uz;
    v[5] = v[5] + m * (y * uz - z * uy);
    v[6] = v[6] + m * (z * ux - x * uz);
    v[7] = v[7] + m * (x * uy - y * ux);
    v[8] = v[8] + m * x;
    v[9] = v[9] + m * y;
    v[10] = v[10] + m * z;
  }
  pm::block_sum<11>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 11; ++q) partials[blockIdx.x * pm::kRedMaxQ + q] = v[q];
  }
}

// sum rho*phi (reference diagnostics_device.py:77-87)
__global__ void __launch_bounds__(256) grid_pe_kernel(const float* __restrict__ rho,
                                                      const float* __restrict__ phi,
                                                      int64_t count,
                                                      double* __restrict__ partials) {
  __shared__ double s_red[4];
  double v[1] = {0.0};
  int64_t r_begin, r_end;
  pm::block_range(count, r_begin, r_end);
  for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x)
    v[0] = v[0] + static_cast<double>(rho[i]) * static_cast<double>(phi[i]);
  pm::block_sum<1>(v, s_red);
  if (threadIdx.x == 0) partials[blockIdx.x * pm::kRedMaxQ] = v[0];
}

__global__ void __launch_bounds__(256) zero_f64_kernel(double* __restrict__ p, int64_t count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0.0;
}

// Mass histogram of radius about (cx,cy,cz) (reference diagnostics_device.py:152-184).
// Workgroup-private float64 bins in LDS, then one global add per non-empty bin.
__global__ void __launch_bounds__(256) radial_hist_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ mass, int64_t np, double cx, double cy, double cz,
    double inv_binw, int nbins, double* __restrict__ hist) {
  extern __shared__ double s_hist[];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0.0;
  __syncthreads();
  int64_t r_begin, r_end;
  pm::block_range(np, r_begin, r_end);
  for (int64_t i = r_begin + threadIdx.x; i < r_end; i += blockDim.x) {
    const double rx = static_cast<double>(px[i]) - cx;
    const double ry = static_cast<double>(py[i]) - cy;
    const double rz = static_cast<double>(pz[i]) - cz;
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    const double t = r * inv_binw;
    int b;
    if (!(t >= 0.0)) b = 0;
    else if (t >= static_cast<double>(nbins)) b = nbins - 1;
    else b = static_cast<int>(t);
    atomicAdd(&s_hist[b], static_cast<double>(mass[i]));
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (s_hist[b] != 0.0) atomicAdd(&hist[b], s_hist[b]);
}

std::mutex g_scratch_mu;
std::map<int, pm::RedScratch*> g_scratch;

pm_status finish_to_host(pm::RedScratch* rs, int q, double* out, hipStream_t stream) {
  PM_TRY_HIP(hipMemcpyAsync(rs->host, rs->result, sizeof(double) * q, hipMemcpyDeviceToHost,
                            stream));
  PM_TRY_HIP(hipStreamSynchronize(stream));
  for (int i = 0; i < q; ++i) out[i] = rs->host[i];
  return PM_OK;
}

}  // namespace

namespace pm {

pm_status red_scratch(RedScratch** out) {
  int dev = 0;
  PM_TRY_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) {
    *out = it->second;
    return PM_OK;
  }
  auto* rs = new RedScratch();
  hipError_t e = hipMalloc(&rs->partials, sizeof(double) * kRedBlocks * kRedMaxQ);
  if (e == hipSuccess) e = hipMalloc(&rs->result, sizeof(double) * kRedMaxQ);
  if (e == hipSuccess) e = hipHostMalloc(&rs->host, sizeof(double) * kRedMaxQ, 0);
  if (e != hipSuccess) {
    if (rs->partials) (void)hipFree(rs->partials);
    if (rs->result) (void)hipFree(rs->result);
    delete rs;
    return hip_fail("reduction scratch allocation", e);
  }
  g_scratch[dev] = rs;
  *out = rs;
  return PM_OK;
}

// moments_kernel with 4 consecutive nodes (one row, n % 4 == 0) per iteration and 16-byte
// loads and stores: a quarter of the dependent iterations per thread. (Issuing four
// iterations' loads before summing them measured slower: 20.1 -> 22.6 us at 256^3.)
template <bool F, bool STORE>
__global__ void __launch_bounds__(256) moments4_kernel(const float* __restrict__ rho, FastDiv fn,
                                                       double dxd, double vol,
                                                       double* __restrict__ partials,
                                                       float* __restrict__ f, float factor) {
  const int n = static_cast<int>(fn.d);
  constexpr int Q = F ? 11 : 10;
  __shared__ double s_red[4 * Q];
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = 0.0;
  const int64_t total = static_cast<int64_t>(n) * n * n;
  int64_t r_begin, r_end;
  pm::block_range(total, r_begin, r_end);  // multiples of 256 nodes
  for (int64_t c = r_begin + 4 * threadIdx.x; c < r_end; c += 4 * blockDim.x) {
    const pm::Idx3 q3 = pm::idx3(c, fn);
    const float4 r4 = *reinterpret_cast<const float4*>(rho + c);
    const float r[4] = {r4.x, r4.y, r4.z, r4.w};
    if (F) {
      const float4 f4 = make_float4(factor * r[0], factor * r[1], factor * r[2], factor * r[3]);
      if (STORE) *reinterpret_cast<float4*>(f + c) = f4;
      const float fv[4] = {f4.x, f4.y, f4.z, f4.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double t = static_cast<double>(fv[e]);
        v[Q - 1] = v[Q - 1] + t * t;
      }
    }
    // the 4 nodes share x and y: sum m, m*z, m*z^2 over them, then form the moments (a
    // quarter of the float64 products of the per-node form)
    const double x = static_cast<double>(q3.i) * dxd;
    const double y = static_cast<double>(q3.j) * dxd;
    double s0 = 0.0, sz = 0.0, szz = 0.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double m = static_cast<double>(r[e]) * vol;
      const double z = static_cast<double>(q3.k + e) * dxd;
      const double mz = m * z;
      s0 = s0 + m;
      sz = sz + mz;
      szz = szz + mz * z;
    }
    const double sx = s0 * x, sy = s0 * y;
    v[0] = v[0] + s0;
    v[1] = v[1] + sx;
    v[2] = v[2] + sy;
    v[3] = v[3] + sz;
    v[4] = v[4] + sx * x;
    v[5] = v[5] + sy * y;
    v[6] = v[6] + szz;
    v[7] = v[7] + sx * y;
    v[8] = v[8] + sz * x;
    v[9] = v[9] + sz * y;
  }
  pm::block_sum<Q>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) partials[q * gridDim.x + blockIdx.x] = v[q];
  }
}

pm_status launch_sumsq_f32(const float* x, int64_t count, double* partials, double* out_dev,
                           hipStream_t stream) {
  const unsigned b = red_blocks(count);
  sumsq_f32_kernel<<<b, kBlock, 0, stream>>>(x, count, partials);
  PM_CHECK_LAUNCH("sumsq_f32_kernel");
  fold_partials_kernel<1><<<1, kBlock, 0, stream>>>(partials, static_cast<int>(b), out_dev);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  return PM_OK;
}

pm_status launch_sumsq_f64(const double* x, int64_t count, double* partials, double* out_dev,
                           hipStream_t stream) {
  const unsigned b = red_blocks(count);
  sumsq_f64_kernel<<<b, kBlock, 0, stream>>>(x, count, partials);
  PM_CHECK_LAUNCH("sumsq_f64_kernel");
  fold_partials_kernel<1><<<1, kBlock, 0, stream>>>(partials, static_cast<int>(b), out_dev);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  return PM_OK;
}

pm_status launch_moments(const float* rho, int n, float dx, double grav, double* partials,
                         double* sums_dev, float* face_params, hipStream_t stream, float* f,
                         float factor, double* f_sumsq_dev) {
  const int64_t total = static_cast<int64_t>(n) * n * n;
  const unsigned b = red_blocks(total);
  const double dxd = static_cast<double>(dx);
  const bool vec = n % 4 == 0 && (reinterpret_cast<uintptr_t>(rho) & 15u) == 0 &&
                   (!f || (reinterpret_cast<uintptr_t>(f) & 15u) == 0);
  const bool sq = f_sumsq_dev != nullptr;
  const double vol = dxd * dxd * dxd;
#define PM_MOMENTS(K, F_, S_) \
  K<F_, S_><<<b, kBlock, 0, stream>>>(rho, make_fast_div(static_cast<uint32_t>(n)), dxd, vol, \
                                      partials, f, factor)
  if (vec && sq && f)
    PM_MOMENTS(moments4_kernel, true, true);
  else if (vec && sq)
    PM_MOMENTS(moments4_kernel, true, false);
  else if (vec)
    PM_MOMENTS(moments4_kernel, false, false);
  else if (sq && f)
    PM_MOMENTS(moments_kernel, true, true);
  else if (sq)
    PM_MOMENTS(moments_kernel, true, false);
  else
    PM_MOMENTS(moments_kernel, false, false);
#undef PM_MOMENTS
  PM_CHECK_LAUNCH("moments_kernel");
  if (sq)
    fold_moments_kernel<11><<<1, kBlock, 0, stream>>>(partials, static_cast<int>(b), grav,
                                                      sums_dev, f_sumsq_dev, face_params);
  else
    fold_moments_kernel<10><<<1, kBlock, 0, stream>>>(partials, static_cast<int>(b), grav,
                                                      sums_dev, nullptr, face_params);
  PM_CHECK_LAUNCH("fold_moments_kernel");
  return PM_OK;
}

}  // namespace pm

extern "C" {

pm_status pm_reduce_particles_f64(const float* px, const float* py, const float* pz,
                                  const float* vx, const float* vy, const float* vz,
                                  const float* mass, int64_t np, double out[11],
                                  hipStream_t stream) {
  PM_REQUIRE(out, "pm_reduce_particles_f64: null output");
  PM_REQUIRE(np >= 0, "pm_reduce_particles_f64: negative particle count");
  if (np == 0) {
    for (int q = 0; q < 11; ++q) out[q] = 0.0;
    return PM_OK;
  }
  PM_REQUIRE(px && py && pz && vx && vy && vz && mass, "pm_reduce_particles_f64: null pointer");
  pm::RedScratch* rs = nullptr;
  pm_status st = pm::red_scratch(&rs);
  if (st != PM_OK) return st;
  const unsigned b = pm::red_blocks(np);
  particle_scalars_kernel<<<b, pm::kBlock, 0, stream>>>(px, py, pz, vx, vy, vz, mass, np,
                                                         rs->partials);
  PM_CHECK_LAUNCH("particle_scalars_kernel");
  pm::fold_partials_kernel<11><<<1, pm::kBlock, 0, stream>>>(rs->partials, static_cast<int>(b),
                                                              rs->result);
  PM_CHECK_LAUNCH("fold_partials_kernel<11>");
  return finish_to_host(rs, 11, out, stream);
}

pm_status pm_reduce_grid_pe_f64(const float* rho, const float* phi, int64_t count, double* out,
                                hipStream_t stream) {
  PM_REQUIRE(out, "pm_reduce_grid_pe_f64: null output");
  PM_REQUIRE(count >= 0, "pm_reduce_grid_pe_f64: negative count");
  if (count == 0) {
    *out = 0.0;
    return PM_OK;
  }
  PM_REQUIRE(rho && phi, "pm_reduce_grid_pe_f64: null pointer");
  pm::RedScratch* rs = nullptr;
  pm_status st = pm::red_scratch(&rs);
  if (st != PM_OK) return st;
  const unsigned b = pm::red_blocks(count);
  grid_pe_kernel<<<b, pm::kBlock, 0, stream>>>(rho, phi, count, rs->partials);
  PM_CHECK_LAUNCH("grid_pe_kernel");
  pm::fold_partials_kernel<1><<<1, pm::kBlock, 0, stream>>>(rs->partials, static_cast<int>(b),
                                                             rs->result);
  PM_CHECK_LAUNCH("fold_partials_kernel<1>");
  return finish_to_host(rs, 1, out, stream);
}

pm_status pm_radial_histogram_f64(const float* px, const float* py, const float* pz,
                                  const float* mass, int64_t np, float cx, float cy, float cz,
                                  float inv_binw, int32_t nbins, double* hist,
                                  hipStream_t stream) {
  PM_REQUIRE(hist, "pm_radial_histogram_f64: null histogram");
  PM_REQUIRE(nbins >= 1 && nbins <= 4096, "pm_radial_histogram_f64: nbins must be in [1, 4096]");
  PM_REQUIRE(np >= 0, "pm_radial_histogram_f64: negative particle count");
  zero_f64_kernel<<<pm::blocks_for(nbins), pm::kBlock, 0, stream>>>(hist, nbins);
  PM_CHECK_LAUNCH("zero_f64_kernel");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && mass, "pm_radial_histogram_f64: null pointer");
  radial_hist_kernel<<<pm::red_blocks(np), pm::kBlock, sizeof(double) * nbins, stream>>>(
      px, py, pz, mass, np, cx, cy, cz, inv_binw, nbins, hist);
  PM_CHECK_LAUNCH("radial_hist_kernel");
  return PM_OK;
}

}  // extern "C"
