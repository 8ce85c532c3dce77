// Streaming particle and node kernels: -grad(phi), CIC gather, kick, drift, their fused
// forms, and the small-N direct force.
//
// Every kernel keeps the reference's per-element operation order with one rounding per
// operation, so the fused variants are bitwise equal to the unfused sequence and to a
// numpy float32 restatement.
#include "pm_cic.h"
#include "pm_common.h"
#include "pm_timing.h"

namespace {

// ---------------------------------------------------------------------------------------
// g = -grad(phi) (reference deposit.py:132-174). One thread per node, k fastest, so each
// wavefront reads and writes 64 consecutive floats of a z-row.
// ---------------------------------------------------------------------------------------
// -grad(phi) at node idx = (i, j, k): central differences inside, one-sided on the faces.
// Central differences inside, one-sided on the faces; the neighbour offsets and the scale
// are selected without branches so the six loads issue together (same a, b, s as the
// branch-per-face form).
// The six phi values the gradient at (i, j, k) reads (one-sided at the faces), and the
// gradient from them (split so that a kernel can issue several nodes' loads first).
struct GradTaps {
  float ax, bx, ay, by, az, bz;
};

__device__ __forceinline__ GradTaps grad_taps(const float* __restrict__ phi, int64_t idx, int i,
                                              int j, int k, int n, int64_t n2) {
  const int64_t px = i == n - 1 ? 0 : n2, mx = i == 0 ? 0 : n2;
  const int64_t py = j == n - 1 ? 0 : n, my = j == 0 ? 0 : n;
  const int64_t pz = k == n - 1 ? 0 : 1, mz = k == 0 ? 0 : 1;
  GradTaps t;
  t.ax = phi[idx + px];
  t.bx = phi[idx - mx];
  t.ay = phi[idx + py];
  t.by = phi[idx - my];
  t.az = phi[idx + pz];
  t.bz = phi[idx - mz];
  return t;
}

__device__ __forceinline__ void grad_from_taps(const GradTaps& t, int i, int j, int k, int n,
                                               float inv_dx, float inv_2dx, float& vx,
                                               float& vy, float& vz) {
  const float sx = (i == 0 || i == n - 1) ? inv_dx : inv_2dx;
  const float sy = (j == 0 || j == n - 1) ? inv_dx : inv_2dx;
  const float sz = (k == 0 || k == n - 1) ? inv_dx : inv_2dx;
  vx = -(t.ax - t.bx) * sx;
  vy = -(t.ay - t.by) * sy;
  vz = -(t.az - t.bz) * sz;
}

__device__ __forceinline__ void grad_at_ijk(const float* __restrict__ phi, int64_t idx, int i,
                                            int j, int k, int n, int64_t n2, float inv_dx,
                                            float inv_2dx, float& vx, float& vy, float& vz) {
  grad_from_taps(grad_taps(phi, idx, i, j, k, n, n2), i, j, k, n, inv_dx, inv_2dx, vx, vy, vz);
}

__device__ __forceinline__ void grad_at(const float* __restrict__ phi, int64_t idx, int n,
                                        int64_t n2, float inv_dx, float inv_2dx, float& vx,
                                        float& vy, float& vz) {
  const pm::Idx3 q3 = pm::idx3(idx, n);
  grad_at_ijk(phi, idx, q3.i, q3.j, q3.k, n, n2, inv_dx, inv_2dx, vx, vy, vz);
}

__global__ void __launch_bounds__(256) grad_kernel(const float* __restrict__ phi,
                                                   float* __restrict__ gx,
                                                   float* __restrict__ gy,
                                                   float* __restrict__ gz, int n,
                                                   float inv_dx, float inv_2dx) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t idx = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= n2 * n) return;
  float vx, vy, vz;
  grad_at(phi, idx, n, n2, inv_dx, inv_2dx, vx, vy, vz);
  gx[idx] = vx;
  gy[idx] = vy;
  gz[idx] = vz;
}

// The same gradient written interleaved, g4[node] = (gx, gy, gz, 0): one 16-byte load per
// corner in the gather (the fused step's private layout).
__global__ void __launch_bounds__(256) grad4_kernel(const float* __restrict__ phi,
                                                    float4* __restrict__ g4, int n,
                                                    float inv_dx, float inv_2dx) {
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const int64_t idx = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= n2 * n) return;
  float vx, vy, vz;
  grad_at(phi, idx, n, n2, inv_dx, inv_2dx, vx, vy, vz);
  g4[idx] = make_float4(vx, vy, vz, 0.0f);
}

// CIC interpolation of the three node fields at one particle (reference deposit.py:100-129).
// The 8 corner loads are issued together (an out-of-box corner reads node 0 and is not
// added), so a particle waits for one round of memory latency instead of eight; the terms
// are added in (cx, cy, cz) order exactly as before.
__device__ __forceinline__ void cic_gather(float x, float y, float z,
                                           const float* __restrict__ gx,
                                           const float* __restrict__ gy,
                                           const float* __restrict__ gz, int n, float dx,
                                           float& ax, float& ay, float& az) {
  ax = 0.0f;
  ay = 0.0f;
  az = 0.0f;
  int ix, iy, iz;
  float fx, fy, fz;
  if (!pm::cic_axis(x, dx, n, ix, fx) || !pm::cic_axis(y, dx, n, iy, fy) ||
      !pm::cic_axis(z, dx, n, iz, fz))
    return;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  bool ok[8];
  float vx[8], vy[8], vz[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int jx = ix + (c >> 2), jy = iy + ((c >> 1) & 1), jz = iz + (c & 1);
    ok[c] = jx >= 0 && jx < n && jy >= 0 && jy < n && jz >= 0 && jz < n;
    const int64_t node = ok[c] ? jx * n2 + static_cast<int64_t>(jy) * n + jz : 0;
    vx[c] = gx[node];
    vy[c] = gy[node];
    vz[c] = gz[node];
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float wx = (c >> 2) ? fx : 1.0f - fx;
    const float wy = ((c >> 1) & 1) ? fy : 1.0f - fy;
    const float wz = (c & 1) ? fz : 1.0f - fz;
    const float w = wx * wy * wz;
    const float tx = ax + w * vx[c], ty = ay + w * vy[c], tz = az + w * vz[c];
    ax = ok[c] ? tx : ax;
    ay = ok[c] ? ty : ay;
    az = ok[c] ? tz : az;
  }
}

// grad4_kernel on the nodes of marked 8^3 bricks: one workgroup per brick (2 nodes per
// thread, k fastest); unmarked bricks return at once and keep their old g4.
constexpr int kGradBrick = pm::kNodeBrick;
static_assert(kGradBrick == 8, "a workgroup of 256 threads covers 8^3 nodes, 2 per thread");

__global__ void __launch_bounds__(256) grad4_bricks_kernel(const float* __restrict__ phi,
                                                           float4* __restrict__ g4, int n,
                                                           float inv_dx, float inv_2dx,
                                                           const uint8_t* __restrict__ bricks) {
  const int nb = (n + kGradBrick - 1) / kGradBrick;
  const int64_t b = pm::xcd_block();
  if (b >= static_cast<int64_t>(nb) * nb * nb || bricks[b] == 0) return;
  const uint32_t ub = static_cast<uint32_t>(b), unb = static_cast<uint32_t>(nb);  // b < 2^21
  const int bk = static_cast<int>(ub % unb), bj = static_cast<int>((ub / unb) % unb),
            bi = static_cast<int>(ub / (unb * unb));
  const int64_t n2 = static_cast<int64_t>(n) * n;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = threadIdx.x + 256 * h;
    const int i = bi * kGradBrick + (q >> 6), j = bj * kGradBrick + ((q >> 3) & 7),
              k = bk * kGradBrick + (q & 7);
    if (i < n && j < n && k < n) {
      const int64_t idx = i * n2 + static_cast<int64_t>(j) * n + k;
      float vx, vy, vz;
      grad_at_ijk(phi, idx, i, j, k, n, n2, inv_dx, inv_2dx, vx, vy, vz);
      g4[idx] = make_float4(vx, vy, vz, 0.0f);
    }
  }
}

// grad4_bricks_kernel over a list of bricks (the deposit's marked bricks, compacted),
// workgroups looping over the list: no workgroups are spent on unmarked bricks.
__global__ void __launch_bounds__(256) grad4_listed_kernel(const float* __restrict__ phi,
                                                           float4* __restrict__ g4, int n,
                                                           float inv_dx, float inv_2dx,
                                                           const int32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ count) {
  const int nb = (n + kGradBrick - 1) / kGradBrick;
  const int64_t n2 = static_cast<int64_t>(n) * n;
  const uint32_t cnt = *count;
  const uint32_t unb = static_cast<uint32_t>(nb);
  uint32_t b = blockIdx.x < cnt ? static_cast<uint32_t>(list[blockIdx.x]) : 0u;
  for (uint32_t e = blockIdx.x; e < cnt; e += gridDim.x) {
    // the next brick's list entry and both nodes' taps leave before any arithmetic. A node
    // past the grid (n % 8 != 0) is clamped to the last node on that axis, whose gradient
    // it computes and stores: the value the node's own lane stores, so the store need not
    // be skipped (a skipped store would let the compiler move the loads behind the test)
    const uint32_t en = e + gridDim.x;
    const uint32_t b_next = en < cnt ? static_cast<uint32_t>(list[en]) : 0u;
    const int bk = static_cast<int>(b % unb), bj = static_cast<int>((b / unb) % unb),
              bi = static_cast<int>(b / (unb * unb));  // brick index < nb^3 <= 2^21
    int ii[2], jj[2], kk[2];
    GradTaps t[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = threadIdx.x + 256 * h;
      ii[h] = min(bi * kGradBrick + (q >> 6), n - 1);
      jj[h] = min(bj * kGradBrick + ((q >> 3) & 7), n - 1);
      kk[h] = min(bk * kGradBrick + (q & 7), n - 1);
      t[h] = grad_taps(phi, ii[h] * n2 + static_cast<int64_t>(jj[h]) * n + kk[h], ii[h], jj[h],
                       kk[h], n, n2);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float vx, vy, vz;
      grad_from_taps(t[h], ii[h], jj[h], kk[h], n, inv_dx, inv_2dx, vx, vy, vz);
      g4[ii[h] * n2 + static_cast<int64_t>(jj[h]) * n + kk[h]] = make_float4(vx, vy, vz, 0.0f);
    }
    b = b_next;
  }
}

// CIC interpolation from the interleaved field (same weights, order and roundings as
// cic_gather, so bitwise equal). g = x/dx by pm::quot where that is the IEEE quotient (dx
// and every |coordinate| in its range), else by the division; the corner nodes are
// addressed from the base node by 32-bit offsets (n^3 < 2^32, and the sum wraps back into
// range for a base index of -1 along an axis).
__device__ __forceinline__ void cic_gather4(float x, float y, float z,
                                            const float4* __restrict__ g4, int n, float dx,
                                            float& ax, float& ay, float& az) {
  ax = 0.0f;
  ay = 0.0f;
  az = 0.0f;
  float gx, gy, gz;
  if (pm::quot_dx_ok(dx) && fabsf(x) >= 0x1p-60f && fabsf(y) >= 0x1p-60f &&
      fabsf(z) >= 0x1p-60f) {
    const float r1 = pm::quot_recip(dx);
    gx = pm::quot(x, dx, r1);
    gy = pm::quot(y, dx, r1);
    gz = pm::quot(z, dx, r1);
  } else {
    gx = x / dx;
    gy = y / dx;
    gz = z / dx;
  }
  int ix, iy, iz;
  float fx, fy, fz;
  if (!pm::cic_axis_g(gx, n, ix, fx) || !pm::cic_axis_g(gy, n, iy, fy) ||
      !pm::cic_axis_g(gz, n, iz, fz))
    return;
  // base index ix >= -1 and ix <= n - 1 per axis: corner 0 lies in the box along it when
  // ix >= 0, corner 1 when ix + 1 < n
  const uint32_t un = static_cast<uint32_t>(n);
  const uint32_t base =
      (static_cast<uint32_t>(ix) * un + static_cast<uint32_t>(iy)) * un + static_cast<uint32_t>(iz);
  const bool okx[2] = {ix >= 0, ix + 1 < n}, oky[2] = {iy >= 0, iy + 1 < n},
             okz[2] = {iz >= 0, iz + 1 < n};
  bool ok[8];
  float4 g[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    ok[c] = okx[c >> 2] && oky[(c >> 1) & 1] && okz[c & 1];
    const uint32_t off = (c >> 2) * un * un + ((c >> 1) & 1) * un + (c & 1);
    g[c] = g4[ok[c] ? base + off : 0u];
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float wx = (c >> 2) ? fx : 1.0f - fx;
    const float wy = ((c >> 1) & 1) ? fy : 1.0f - fy;
    const float wz = (c & 1) ? fz : 1.0f - fz;
    const float w = wx * wy * wz;
    const float tx = ax + w * g[c].x, ty = ay + w * g[c].y, tz = az + w * g[c].z;
    ax = ok[c] ? tx : ax;
    ay = ok[c] ? ty : ay;
    az = ok[c] ? tz : az;
  }
}

__global__ void __launch_bounds__(256) gather_kick4_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float4* __restrict__ g4, float* __restrict__ ax, float* __restrict__ ay,
    float* __restrict__ az, float* __restrict__ vx, float* __restrict__ vy,
    float* __restrict__ vz, int64_t np, int n, float dx, float h) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float x = px[p], y = py[p], z = pz[p];
  const float u0 = vx[p], u1 = vy[p], u2 = vz[p];  // issued with the position loads
  float a0, a1, a2;
  cic_gather4(x, y, z, g4, n, dx, a0, a1, a2);
  ax[p] = a0;
  ay[p] = a1;
  az[p] = a2;
  vx[p] = u0 + a0 * h;
  vy[p] = u1 + a1 * h;
  vz[p] = u2 + a2 * h;
}

// Step t's closing gather+kick(h_close) joined with step t+1's opening kick(h_open) and
// drift(dt): the same float operations in the same order as gather_kick4 followed by
// kick_drift, without the round trip of a and v through memory.
// With KEYS it also writes each particle's extended cell key for the drifted position, as
// the deposit's key pass would compute it from the stored coordinates.
template <bool KEYS>
__global__ void __launch_bounds__(256) gather_kick_drift4_kernel(
    float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
    const float4* __restrict__ g4, float* __restrict__ vx, float* __restrict__ vy,
    float* __restrict__ vz, int64_t np, int n, float dx, float h_close, float h_open,
    float dt, uint32_t* __restrict__ keys) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  // the velocity loads issue with the position loads, not after the gather
  const float x = px[p], y = py[p], z = pz[p];
  const float u0 = vx[p], u1 = vy[p], u2 = vz[p];
  float a0, a1, a2;
  cic_gather4(x, y, z, g4, n, dx, a0, a1, a2);
  const float v0 = u0 + a0 * h_close;
  const float v1 = u1 + a1 * h_close;
  const float v2 = u2 + a2 * h_close;
  const float w0 = v0 + a0 * h_open;
  const float w1 = v1 + a1 * h_open;
  const float w2 = v2 + a2 * h_open;
  vx[p] = w0;
  vy[p] = w1;
  vz[p] = w2;
  const float x1 = x + w0 * dt, y1 = y + w1 * dt, z1 = z + w2 * dt;
  px[p] = x1;
  py[p] = y1;
  pz[p] = z1;
  if (KEYS) {
    int ix = 0, iy = 0, iz = 0;
    float fx, fy, fz;
    const bool in = pm::cic_axis(x1, dx, n, ix, fx) && pm::cic_axis(y1, dx, n, iy, fy) &&
                    pm::cic_axis(z1, dx, n, iz, fz);
    const uint32_t m = static_cast<uint32_t>(n) + 1u;
    keys[p] = in ? pm::cell_key(ix, iy, iz, n) : m * m * m;
  }
}

// gather_kick_drift4_kernel<true> that also moves particle order[p] of `src` to slot p of
// `dst`: the same float operations per particle, only the slots change.
struct ParticleSlots {
  float* x;
  float* y;
  float* z;
  float* vx;
  float* vy;
  float* vz;
  float* m;
  int32_t* gid;
  int32_t* pid;
};

template <bool PID>
__global__ void __launch_bounds__(256) gather_kick_drift4_reorder_kernel(
    const int32_t* __restrict__ order, ParticleSlots src, ParticleSlots dst,
    const float4* __restrict__ g4, int64_t np, int n, float dx, float h_close, float h_open,
    float dt, uint32_t* __restrict__ keys) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int64_t q = order[p];
  if (q < 0 || q >= np) q = p;  // not a permutation entry: keep the slot's own particle
  const float x = src.x[q], y = src.y[q], z = src.z[q];
  const float u0 = src.vx[q], u1 = src.vy[q], u2 = src.vz[q];
  const float m = src.m[q];
  const int32_t gid = src.gid[q];
  const int32_t pid = PID ? src.pid[q] : 0;
  float a0, a1, a2;
  cic_gather4(x, y, z, g4, n, dx, a0, a1, a2);
  const float v0 = u0 + a0 * h_close;
  const float v1 = u1 + a1 * h_close;
  const float v2 = u2 + a2 * h_close;
  const float w0 = v0 + a0 * h_open;
  const float w1 = v1 + a1 * h_open;
  const float w2 = v2 + a2 * h_open;
  dst.vx[p] = w0;
  dst.vy[p] = w1;
  dst.vz[p] = w2;
  const float x1 = x + w0 * dt, y1 = y + w1 * dt, z1 = z + w2 * dt;
  dst.x[p] = x1;
  dst.y[p] = y1;
  dst.z[p] = z1;
  dst.m[p] = m;
  dst.gid[p] = gid;
  if (PID) dst.pid[p] = pid;
  int ix = 0, iy = 0, iz = 0;
  float fx, fy, fz;
  const bool in = pm::cic_axis(x1, dx, n, ix, fx) && pm::cic_axis(y1, dx, n, iy, fy) &&
                  pm::cic_axis(z1, dx, n, iz, fz);
  const uint32_t mm = static_cast<uint32_t>(n) + 1u;
  keys[p] = in ? pm::cell_key(ix, iy, iz, n) : mm * mm * mm;
}

__global__ void __launch_bounds__(256) gather_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ gx, const float* __restrict__ gy, const float* __restrict__ gz,
    float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az, int64_t np, int n,
    float dx) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float a0, a1, a2;
  cic_gather(px[p], py[p], pz[p], gx, gy, gz, n, dx, a0, a1, a2);
  ax[p] = a0;
  ay[p] = a1;
  az[p] = a2;
}

__global__ void __launch_bounds__(256) gather_kick_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ gx, const float* __restrict__ gy, const float* __restrict__ gz,
    float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az,
    float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz, int64_t np, int n,
    float dx, float h) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  float a0, a1, a2;
  cic_gather(px[p], py[p], pz[p], gx, gy, gz, n, dx, a0, a1, a2);
  ax[p] = a0;
  ay[p] = a1;
  az[p] = a2;
  vx[p] = vx[p] + a0 * h;
  vy[p] = vy[p] + a1 * h;
  vz[p] = vz[p] + a2 * h;
}

// v += a*dt (reference integrator.py:27-42)
__global__ void __launch_bounds__(256) kick_kernel(float* __restrict__ vx,
                                                   float* __restrict__ vy,
                                                   float* __restrict__ vz,
                                                   const float* __restrict__ ax,
                                                   const float* __restrict__ ay,
                                                   const float* __restrict__ az, int64_t np,
                                                   float dt) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  vx[p] = vx[p] + ax[p] * dt;
  vy[p] = vy[p] + ay[p] * dt;
  vz[p] = vz[p] + az[p] * dt;
}

// x += v*dt (reference integrator.py:45-60)
__global__ void __launch_bounds__(256) drift_kernel(float* __restrict__ px,
                                                    float* __restrict__ py,
                                                    float* __restrict__ pz,
                                                    const float* __restrict__ vx,
                                                    const float* __restrict__ vy,
                                                    const float* __restrict__ vz, int64_t np,
                                                    float dt) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  px[p] = px[p] + vx[p] * dt;
  py[p] = py[p] + vy[p] * dt;
  pz[p] = pz[p] + vz[p] * dt;
}

__global__ void __launch_bounds__(256) kick_drift_kernel(
    float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz,
    float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz,
    const float* __restrict__ ax, const float* __restrict__ ay, const float* __restrict__ az,
    int64_t np, float h, float dt) {
  const int64_t p = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float v0 = vx[p] + ax[p] * h;
  const float v1 = vy[p] + ay[p] * h;
  const float v2 = vz[p] + az[p] * h;
  vx[p] = v0;
  vy[p] = v1;
  vz[p] = v2;
  px[p] = px[p] + v0 * dt;
  py[p] = py[p] + v1 * dt;
  pz[p] = pz[p] + v2 * dt;
}

// a_i = G sum_{j != i} m_j (x_j - x_i) / (|x_i - x_j|^2 + eps^2)^{3/2}
// (reference integrator.py:63-97); positions of a tile of j staged through LDS.
__global__ void __launch_bounds__(256) direct_accel_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ m, float* __restrict__ ax, float* __restrict__ ay,
    float* __restrict__ az, int64_t np, float grav, float eps2) {
  __shared__ float sx[256], sy[256], sz[256], sm[256];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = i < np;
  const float xi = active ? px[i] : 0.0f;
  const float yi = active ? py[i] : 0.0f;
  const float zi = active ? pz[i] : 0.0f;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (int64_t base = 0; base < np; base += 256) {
    const int64_t j = base + threadIdx.x;
    __syncthreads();
    if (j < np) {
      sx[threadIdx.x] = px[j];
      sy[threadIdx.x] = py[j];
      sz[threadIdx.x] = pz[j];
      sm[threadIdx.x] = m[j];
    }
    __syncthreads();
    const int cnt = static_cast<int>(np - base < 256 ? np - base : 256);
    for (int t = 0; t < cnt; ++t) {
      if (base + t == i) continue;
      const float rx = sx[t] - xi;
      const float ry = sy[t] - yi;
      const float rz = sz[t] - zi;
      const float r2 = rx * rx + ry * ry + rz * rz + eps2;
      const float inv_r3 = 1.0f / (r2 * pm::sqrt_rn(r2));
      const float f = grav * sm[t] * inv_r3;
      a0 += f * rx;
      a1 += f * ry;
      a2 += f * rz;
    }
  }
  if (active) {
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
  }
}

// dst = src, one float per thread: the known-size streaming kernel used to calibrate
// profiler byte counters.
__global__ void __launch_bounds__(256) copy_f32_kernel(float* __restrict__ dst,
                                                       const float* __restrict__ src,
                                                       int64_t count) {
  const int64_t i = pm::xcd_block() * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

}  // namespace

extern "C" {

pm_status pm_copy_f32(float* dst, const float* src, int64_t count, hipStream_t stream) {
  PM_REQUIRE(count >= 0, "pm_copy_f32: negative count");
  if (count == 0) return PM_OK;
  PM_REQUIRE(dst && src, "pm_copy_f32: null pointer");
  copy_f32_kernel<<<pm::blocks_for(count), pm::kBlock, 0, stream>>>(dst, src, count);
  PM_CHECK_LAUNCH("copy_f32_kernel");
  return PM_OK;
}

pm_status pm_grad(const float* phi, float* gx, float* gy, float* gz, int32_t n, float dx,
                  hipStream_t stream) {
  PM_REQUIRE(phi && gx && gy && gz, "pm_grad: null pointer");
  PM_REQUIRE(n >= 2 && n <= 1024, "pm_grad: grid_size must be in [2, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_grad: dx must be positive");
  const float inv_dx = 1.0f / dx;
  const float inv_2dx = 1.0f / (2.0f * dx);
  const int64_t total = static_cast<int64_t>(n) * n * n;
  {
    PM_TIMED("grad", 16.0 * total, stream);
    grad_kernel<<<pm::blocks_for(total), pm::kBlock, 0, stream>>>(phi, gx, gy, gz, n, inv_dx,
                                                                   inv_2dx);
    PM_CHECK_LAUNCH("grad_kernel");
  }
  return PM_OK;
}

pm_status pm_grad4(const float* phi, float* g4, int32_t n, float dx, hipStream_t stream) {
  PM_REQUIRE(phi && g4, "pm_grad4: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0, "pm_grad4: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 2 && n <= 1024, "pm_grad4: grid_size must be in [2, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_grad4: dx must be positive");
  const float inv_dx = 1.0f / dx;
  const float inv_2dx = 1.0f / (2.0f * dx);
  const int64_t total = static_cast<int64_t>(n) * n * n;
  {
    PM_TIMED("grad4", 20.0 * total, stream);
    grad4_kernel<<<pm::blocks_for(total), pm::kBlock, 0, stream>>>(
        phi, reinterpret_cast<float4*>(g4), n, inv_dx, inv_2dx);
    PM_CHECK_LAUNCH("grad4_kernel");
  }
  return PM_OK;
}

pm_status pm_grad4_bricks(const float* phi, float* g4, int32_t n, float dx,
                         const uint8_t* bricks, hipStream_t stream) {
  PM_REQUIRE(phi && g4 && bricks, "pm_grad4_bricks: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0,
             "pm_grad4_bricks: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 2 && n <= 1024, "pm_grad4_bricks: grid_size must be in [2, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_grad4_bricks: dx must be positive");
  const float inv_dx = 1.0f / dx;
  const float inv_2dx = 1.0f / (2.0f * dx);
  const int64_t nb = (n + kGradBrick - 1) / kGradBrick;
  {
    PM_TIMED("grad4", 0.0, stream);
    grad4_bricks_kernel<<<static_cast<unsigned>(nb * nb * nb), 256, 0, stream>>>(
        phi, reinterpret_cast<float4*>(g4), n, inv_dx, inv_2dx, bricks);
    PM_CHECK_LAUNCH("grad4_bricks_kernel");
  }
  return PM_OK;
}

pm_status pm_grad4_listed(const float* phi, float* g4, int32_t n, float dx, const int32_t* list,
                          const uint32_t* count, int64_t max_count, hipStream_t stream) {
  PM_REQUIRE(phi && g4 && list && count, "pm_grad4_listed: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0, "pm_grad4_listed: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 2 && n <= 1024, "pm_grad4_listed: grid_size must be in [2, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_grad4_listed: dx must be positive");
  const int64_t nb = (n + kGradBrick - 1) / kGradBrick;
  PM_REQUIRE(max_count >= 0 && max_count <= nb * nb * nb,
             "pm_grad4_listed: max_count must be in [0, number of bricks]");
  if (max_count == 0) return PM_OK;
  const float inv_dx = 1.0f / dx;
  const float inv_2dx = 1.0f / (2.0f * dx);
  const unsigned blocks = static_cast<unsigned>(max_count < 2048 ? max_count : 2048);
  {
    PM_TIMED("grad4", 0.0, stream);
    grad4_listed_kernel<<<blocks, 256, 0, stream>>>(phi, reinterpret_cast<float4*>(g4), n,
                                                     inv_dx, inv_2dx, list, count);
    PM_CHECK_LAUNCH("grad4_listed_kernel");
  }
  return PM_OK;
}

pm_status pm_gather_kick4(const float* px, const float* py, const float* pz, const float* g4,
                          float* ax, float* ay, float* az, float* vx, float* vy, float* vz,
                          int64_t np, int32_t n, float dx, float h, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_gather_kick4: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && g4 && ax && ay && az && vx && vy && vz,
             "pm_gather_kick4: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0,
             "pm_gather_kick4: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_gather_kick4: grid_size must be in [1, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_gather_kick4: dx must be positive");
  {
    PM_TIMED("gather_kick", 48.0 * np + 12.0 * n * n * n, stream);
    gather_kick4_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
        px, py, pz, reinterpret_cast<const float4*>(g4), ax, ay, az, vx, vy, vz, np, n, dx, h);
    PM_CHECK_LAUNCH("gather_kick4_kernel");
  }
  return PM_OK;
}

pm_status pm_gather_kick_drift4(float* px, float* py, float* pz, const float* g4, float* vx,
                                float* vy, float* vz, int64_t np, int32_t n, float dx,
                                float h_close, float h_open, float dt, uint32_t* keys_out,
                                hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_gather_kick_drift4: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && g4 && vx && vy && vz, "pm_gather_kick_drift4: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0,
             "pm_gather_kick_drift4: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_gather_kick_drift4: grid_size must be in [1, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_gather_kick_drift4: dx must be positive");
  {
    PM_TIMED("gather_kick_drift", (keys_out ? 52.0 : 48.0) * np + 12.0 * n * n * n, stream);
    if (keys_out) {
      gather_kick_drift4_kernel<true><<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
          px, py, pz, reinterpret_cast<const float4*>(g4), vx, vy, vz, np, n, dx, h_close,
          h_open, dt, keys_out);
    } else {
      gather_kick_drift4_kernel<false><<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
          px, py, pz, reinterpret_cast<const float4*>(g4), vx, vy, vz, np, n, dx, h_close,
          h_open, dt, nullptr);
    }
    PM_CHECK_LAUNCH("gather_kick_drift4_kernel");
  }
  return PM_OK;
}

pm_status pm_gather_kick_drift4_reorder(const int32_t* order, const pm_particle_arrays* src,
                                        const pm_particle_arrays* dst, const float* g4,
                                        int64_t np, int32_t n, float dx, float h_close,
                                        float h_open, float dt, uint32_t* keys_out,
                                        hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_gather_kick_drift4_reorder: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(order && src && dst && g4 && keys_out, "pm_gather_kick_drift4_reorder: null pointer");
  PM_REQUIRE((reinterpret_cast<uintptr_t>(g4) & 15u) == 0,
             "pm_gather_kick_drift4_reorder: g4 must be 16-byte aligned");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_gather_kick_drift4_reorder: grid_size must be in [1, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_gather_kick_drift4_reorder: dx must be positive");
  const void* sp[9] = {src->pos_x, src->pos_y, src->pos_z, src->vel_x, src->vel_y,
                       src->vel_z, src->mass, src->gid, src->pid};
  const void* dp[9] = {dst->pos_x, dst->pos_y, dst->pos_z, dst->vel_x, dst->vel_y,
                       dst->vel_z, dst->mass, dst->gid, dst->pid};
  const bool pid = src->pid != nullptr;
  PM_REQUIRE(pid == (dst->pid != nullptr),
             "pm_gather_kick_drift4_reorder: pid must be given in both src and dst or in neither");
  const int count = pid ? 9 : 8;
  for (int k = 0; k < count; ++k)
    PM_REQUIRE(sp[k] && dp[k], "pm_gather_kick_drift4_reorder: null particle array");
  // out of place: no destination array may overlap a source array (or another destination)
  const uint64_t bytes = static_cast<uint64_t>(np) * 4u;
  auto overlap = [bytes](const void* a, const void* b) {
    const uint64_t x = reinterpret_cast<uint64_t>(a), y = reinterpret_cast<uint64_t>(b);
    return x < y + bytes && y < x + bytes;
  };
  for (int k = 0; k < count; ++k) {
    for (int j = 0; j < count; ++j) {
      PM_REQUIRE(!overlap(dp[k], sp[j]),
                 "pm_gather_kick_drift4_reorder: a destination array overlaps a source array");
      PM_REQUIRE(j == k || !overlap(dp[k], dp[j]),
                 "pm_gather_kick_drift4_reorder: destination arrays overlap");
    }
    PM_REQUIRE(!overlap(dp[k], keys_out) && !overlap(dp[k], order),
               "pm_gather_kick_drift4_reorder: a destination array overlaps keys_out or order");
  }
  ParticleSlots s{src->pos_x, src->pos_y, src->pos_z, src->vel_x, src->vel_y, src->vel_z,
                  src->mass, src->gid, src->pid};
  ParticleSlots d{dst->pos_x, dst->pos_y, dst->pos_z, dst->vel_x, dst->vel_y, dst->vel_z,
                  dst->mass, dst->gid, dst->pid};
  {
    PM_TIMED("gather_kick_drift_reorder", (pid ? 80.0 : 72.0) * np + 12.0 * n * n * n, stream);
    if (pid) {
      gather_kick_drift4_reorder_kernel<true><<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
          order, s, d, reinterpret_cast<const float4*>(g4), np, n, dx, h_close, h_open, dt,
          keys_out);
    } else {
      gather_kick_drift4_reorder_kernel<false><<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
          order, s, d, reinterpret_cast<const float4*>(g4), np, n, dx, h_close, h_open, dt,
          keys_out);
    }
    PM_CHECK_LAUNCH("gather_kick_drift4_reorder_kernel");
  }
  return PM_OK;
}

pm_status pm_gather(const float* px, const float* py, const float* pz, const float* gx,
                    const float* gy, const float* gz, float* ax, float* ay, float* az,
                    int64_t np, int32_t n, float dx, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_gather: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && gx && gy && gz && ax && ay && az, "pm_gather: null pointer");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_gather: grid_size must be in [1, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_gather: dx must be positive");
  {
    PM_TIMED("gather", 24.0 * np + 12.0 * n * n * n, stream);
    gather_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(px, py, pz, gx, gy, gz, ax, ay,
                                                                  az, np, n, dx);
    PM_CHECK_LAUNCH("gather_kernel");
  }
  return PM_OK;
}

pm_status pm_gather_kick(const float* px, const float* py, const float* pz, const float* gx,
                         const float* gy, const float* gz, float* ax, float* ay, float* az,
                         float* vx, float* vy, float* vz, int64_t np, int32_t n, float dx,
                         float h, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_gather_kick: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && gx && gy && gz && ax && ay && az && vx && vy && vz,
             "pm_gather_kick: null pointer");
  PM_REQUIRE(n >= 1 && n <= 1024, "pm_gather_kick: grid_size must be in [1, 1024]");
  PM_REQUIRE(dx > 0.0f, "pm_gather_kick: dx must be positive");
  {
    PM_TIMED("gather_kick", 48.0 * np + 12.0 * n * n * n, stream);
    gather_kick_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(
        px, py, pz, gx, gy, gz, ax, ay, az, vx, vy, vz, np, n, dx, h);
    PM_CHECK_LAUNCH("gather_kick_kernel");
  }
  return PM_OK;
}

pm_status pm_kick(float* vx, float* vy, float* vz, const float* ax, const float* ay,
                  const float* az, int64_t np, float dt, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_kick: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(vx && vy && vz && ax && ay && az, "pm_kick: null pointer");
  {
    PM_TIMED("kick", 36.0 * np, stream);
    kick_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(vx, vy, vz, ax, ay, az, np, dt);
    PM_CHECK_LAUNCH("kick_kernel");
  }
  return PM_OK;
}

pm_status pm_drift(float* px, float* py, float* pz, const float* vx, const float* vy,
                   const float* vz, int64_t np, float dt, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_drift: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && vx && vy && vz, "pm_drift: null pointer");
  {
    PM_TIMED("drift", 36.0 * np, stream);
    drift_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(px, py, pz, vx, vy, vz, np, dt);
    PM_CHECK_LAUNCH("drift_kernel");
  }
  return PM_OK;
}

pm_status pm_kick_drift(float* px, float* py, float* pz, float* vx, float* vy, float* vz,
                        const float* ax, const float* ay, const float* az, int64_t np, float h,
                        float dt, hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_kick_drift: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(px && py && pz && vx && vy && vz && ax && ay && az, "pm_kick_drift: null pointer");
  {
    PM_TIMED("kick_drift", 60.0 * np, stream);
    kick_drift_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(px, py, pz, vx, vy, vz, ax,
                                                                      ay, az, np, h, dt);
    PM_CHECK_LAUNCH("kick_drift_kernel");
  }
  return PM_OK;
}

pm_status pm_direct_accel(const float* px, const float* py, const float* pz, const float* m,
                          float* ax, float* ay, float* az, int64_t np, float grav, float eps2,
                          hipStream_t stream) {
  PM_REQUIRE(np >= 0, "pm_direct_accel: negative particle count");
  if (np == 0) return PM_OK;
  PM_REQUIRE(np <= (int64_t(1) << 24), "pm_direct_accel: O(N^2) path limited to 2^24 particles");
  PM_REQUIRE(px && py && pz && m && ax && ay && az, "pm_direct_accel: null pointer");
  {
    PM_TIMED("direct_accel", 16.0 * np, stream);
    direct_accel_kernel<<<pm::blocks_for(np), pm::kBlock, 0, stream>>>(px, py, pz, m, ax, ay, az,
                                                                        np, grav, eps2);
    PM_CHECK_LAUNCH("direct_accel_kernel");
  }
  return PM_OK;
}

}  // extern "C"
