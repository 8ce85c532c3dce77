// BF16 3x3x3 convolutions on the gfx950 matrix cores, for the learned multigrid warm start.
//
// Tensors are channels-last: element (voxel v, channel c) of an r^3 tensor with C channels
// sits at v*C + c, v = (z*r + y)*r + x (x fastest, the grid's k axis). BF16 values travel as
// their 16-bit patterns (uint16_t). Every convolution is "same" (zero padding 1, stride 1)
// and cross-correlates like torch.nn.functional.conv3d:
//     y[v][co] = bias[co] + sum_{ci, tap} w[co][ci][tap] * x[v + off(tap)][ci],
//     tap = (dz+1)*9 + (dy+1)*3 + (dx+1), off(tap) = (dz, dy, dx).
// The input may be the concatenation of two tensors: xa (ca channels) then xb (cb channels);
// with up_a set, xa lives at resolution r/2 and is read nearest-neighbour upsampled, which
// is the decoder's "upsample, concatenate the skip, convolve" fused into the operand loads.
//
// Forward is an implicit GEMM: rows = output voxels, columns = output channels, reduction
// k = tap*Cin + ci over packed weights w_p[co][k] (bf16, rows padded to 16, k to 32). Each
// wave computes 64 voxels x Cout with v_mfma_f32_16x16x32_bf16 (float32 accumulation).
// Backward-data is the forward kernel applied to dZ with the weights flipped in space and
// transposed in channels. Backward-weight reduces over voxels: dW[co][k] = sum_v dZ[v][co] *
// im2col(x)[v][k]; both operands are staged transposed through LDS so that the 32-voxel
// reduction chunk is contiguous per lane, partial sums per voxel split are written and
// folded in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "pm_common.h"
#include "pm_conv.h"
#include "pm_timing.h"

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

using pm::conv::bf2f;
using pm::conv::f2bf;

__device__ __forceinline__ f32x4_t mfma_bf16(u32x4_t a, u32x4_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                  __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

__device__ __forceinline__ u32x4_t load16(const uint16_t* p) {
  return *reinterpret_cast<const u32x4_t*>(p);
}

struct ConvFwdArgs {
  const uint16_t* xa;
  const uint16_t* xb;
  int ca, cb, cin, r, ra, up;
  const uint16_t* w;  // packed [rows][kpad]
  int kpad, kreal, rows;
  int kc_per;         // 32-wide reduction chunks per split (blockIdx.z)
  float* partial;     // split-K: raw sums [split][r^3][rows] (epilogue in a second kernel)
  const float* bias;
  int cout, cout_st, act;
  float slope;
  uint16_t* y;   // [r^3][cout_st] or null
  float* y32;    // [r^3][cout] or null
};

// Source element (voxel (z, y, x) of the concatenated input, channels ci..ci+7) or null
// when the voxel lies outside the r^3 box (zero padding).
__device__ __forceinline__ const uint16_t* conv_src(const uint16_t* xa, const uint16_t* xb, int ca,
                                                    int cb, int r, int ra, int up, int z, int y,
                                                    int x, int ci) {
  if (static_cast<unsigned>(z) >= static_cast<unsigned>(r) ||
      static_cast<unsigned>(y) >= static_cast<unsigned>(r) ||
      static_cast<unsigned>(x) >= static_cast<unsigned>(r))
    return nullptr;
  if (ci < ca) {
    const int idx = up ? ((z >> 1) * ra + (y >> 1)) * ra + (x >> 1) : (z * r + y) * r + x;
    return xa + static_cast<int64_t>(idx) * ca + ci;
  }
  return xb + static_cast<int64_t>((z * r + y) * r + x) * cb + (ci - ca);
}

// One wave: 64 output voxels (4 MFMA row blocks) x output channels co0 .. co0 + 16*NB - 1,
// co0 = blockIdx.y * 16 * NB (column blocks at or beyond the packed weight rows are
// skipped), over the reduction chunks of split blockIdx.z.
template <int NB>
__global__ void __launch_bounds__(256) conv_fwd_kernel(ConvFwdArgs a) {
  const int co0 = static_cast<int>(blockIdx.y) * 16 * NB;
  const int kbeg = static_cast<int>(blockIdx.z) * a.kc_per * 32;
  const int kend = min(a.kpad, kbeg + a.kc_per * 32);
  const int lane = threadIdx.x & 63;
  const int row = lane & 15, grp = lane >> 4;
  const int r = a.r;
  const int nvox = r * r * r;
  const int vbase = (static_cast<int>(pm::xcd_block()) * 4 + static_cast<int>(threadIdx.x >> 6)) * 64;
  if (vbase >= nvox) return;  // whole wave idle (no barriers in this kernel)

  int vz[4], vy[4], vx[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const int v = vbase + 16 * mb + row;
    if (v < nvox) {
      vx[mb] = v % r;
      vy[mb] = (v / r) % r;
      vz[mb] = v / (r * r);
    } else {
      vz[mb] = -4;  // every neighbour of this row falls outside the box
      vy[mb] = vx[mb] = 0;
    }
  }
  f32x4_t acc[4][NB];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  // this lane's k = kc + 8*grp as (tap, ci); advanced by 32 per chunk
  int tap = (kbeg + 8 * grp) / a.cin;
  int ci = kbeg + 8 * grp - tap * a.cin;
  const uint16_t* wrow[NB];
  bool wvalid[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    wvalid[nb] = co0 + 16 * nb < a.rows;
    wrow[nb] = a.w + static_cast<int64_t>(wvalid[nb] ? co0 + 16 * nb + row : 0) * a.kpad + 8 * grp;
  }

  // G consecutive 32-wide chunks per round: every operand load of the round is issued
  // before its first MFMA (a chunk's few MFMAs cannot hide a load's latency on their own),
  // and the chunks are accumulated in order, as one chunk per round would
  constexpr int G = NB <= 2 ? 2 : 1;
  for (int kc = kbeg; kc < kend; kc += 32 * G) {
    u32x4_t av[G][4];
    u32x4_t bv[G][NB];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int kcg = kc + 32 * g;
      const bool live = kcg < kend;
      const bool kin = live && kcg + 8 * grp < a.kreal;
      const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const uint16_t* p = kin ? conv_src(a.xa, a.xb, a.ca, a.cb, r, a.ra, a.up, vz[mb] + dz,
                                           vy[mb] + dy, vx[mb] + dx, ci)
                                : nullptr;
        av[g][mb] = p ? load16(p) : u32x4_t{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
        bv[g][nb] = (live && wvalid[nb]) ? load16(wrow[nb] + kcg) : u32x4_t{0u, 0u, 0u, 0u};
      ci += 32;
      while (ci >= a.cin) {
        ci -= a.cin;
        ++tap;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (kc + 32 * g >= kend) break;  // uniform
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mb][nb] = mfma_bf16(av[g][mb], bv[g][nb], acc[mb][nb]);
    }
  }

  // D layout: lane holds rows 4*grp + i (voxels), column `row` (output channel)
  if (a.partial) {
    float* out = a.partial + static_cast<int64_t>(blockIdx.z) * nvox * a.rows;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int co = co0 + 16 * nb + row;
      if (co >= a.rows) continue;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int v = vbase + 16 * mb + 4 * grp + i;
          if (v < nvox) out[static_cast<int64_t>(v) * a.rows + co] = acc[mb][nb][i];
        }
    }
    return;
  }
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int co = co0 + 16 * nb + row;
    if (co >= a.cout_st) continue;
    const bool real = co < a.cout;
    const float b = (real && a.bias) ? a.bias[co] : 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int v = vbase + 16 * mb + 4 * grp + i;
        if (v >= nvox) continue;
        float val = 0.0f;
        if (real) {
          val = acc[mb][nb][i] + b;
          if (a.act && !(val > 0.0f)) val = val * a.slope;
        }
        if (a.y) a.y[static_cast<int64_t>(v) * a.cout_st + co] = f2bf(val);
        if (a.y32 && real) a.y32[static_cast<int64_t>(v) * a.cout + co] = val;
      }
    }
  }
}

// Split-K epilogue: y = act(bias + sum over splits in order), one thread per output value.
__global__ void __launch_bounds__(256) conv_epilogue_kernel(const float* __restrict__ partial,
                                                            int splits, int nvox, int rows,
                                                            const float* __restrict__ bias,
                                                            int cout, int cout_st, int act,
                                                            float slope, uint16_t* y, float* y32) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(nvox) * cout_st) return;
  const int co = static_cast<int>(idx % cout_st);
  const int v = static_cast<int>(idx / cout_st);
  float val = 0.0f;
  if (co < cout) {
    float sum = 0.0f;
    for (int z = 0; z < splits; ++z)
      sum += partial[(static_cast<int64_t>(z) * nvox + v) * rows + co];
    val = sum + (bias ? bias[co] : 0.0f);
    if (act && !(val > 0.0f)) val = val * slope;
  }
  if (y) y[idx] = f2bf(val);
  if (y32 && co < cout) y32[static_cast<int64_t>(v) * cout + co] = val;
}

struct ConvWgradArgs {
  const uint16_t* xa;
  const uint16_t* xb;
  int ca, cb, cin, r, ra, up;
  const uint16_t* dz;  // [r^3][dz_c]
  int dz_c, kreal, nvox, chunks_per_split;
  float* partial;      // [splits][16*NR][kcols]
  int kcols;
};

// grid (kcols/64, splits): one workgroup = 64 reduction columns x all output channels over
// one voxel range; wave w owns columns 16w..16w+15.
template <int NR>
__global__ void __launch_bounds__(256) conv_wgrad_kernel(ConvWgradArgs a) {
  constexpr int kLd = 40;  // 32 voxels + 8 pad (80-byte rows keep 16-byte alignment)
  __shared__ __align__(16) uint16_t sdz[NR * 16][kLd];
  __shared__ __align__(16) uint16_t sx[64][kLd];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, grp = lane >> 4;
  const int k0 = blockIdx.x * 64;
  const int split = blockIdx.y;

  for (int p = tid; p < NR * 16 * kLd; p += 256) (&sdz[0][0])[p] = 0;  // rows >= dz_c stay 0

  const int cg = tid >> 5, vs = tid & 31;  // x staging: column group (8 columns), voxel slot
  const int kq = k0 + 8 * cg;
  const bool kvalid = kq < a.kreal;
  int ci = 0, dz = 0, dy = 0, dx = 0;
  if (kvalid) {
    const int tap = kq / a.cin;
    ci = kq - tap * a.cin;
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
  }
  const int c8n = a.dz_c >> 3;
  f32x4_t acc[NR];
#pragma unroll
  for (int rb = 0; rb < NR; ++rb) acc[rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  const int vbeg = split * a.chunks_per_split * 32;
  const int vend = min(a.nvox, vbeg + a.chunks_per_split * 32);
  for (int vc = vbeg; vc < vend; vc += 32) {
    __syncthreads();  // previous chunk's MFMA reads are done (and the zero fill above)
    for (int p = tid; p < 32 * c8n; p += 256) {
      const int v = p / c8n, c8 = p - v * c8n;
      u32x4_t val{0u, 0u, 0u, 0u};
      if (vc + v < a.nvox) val = load16(a.dz + static_cast<int64_t>(vc + v) * a.dz_c + 8 * c8);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sdz[8 * c8 + 2 * q][v] = static_cast<uint16_t>(val[q] & 0xffffu);
        sdz[8 * c8 + 2 * q + 1][v] = static_cast<uint16_t>(val[q] >> 16);
      }
    }
    {
      u32x4_t val{0u, 0u, 0u, 0u};
      const int v = vc + vs;
      if (kvalid && v < a.nvox) {
        const int x = v % a.r, y = (v / a.r) % a.r, z = v / (a.r * a.r);
        const uint16_t* p =
            conv_src(a.xa, a.xb, a.ca, a.cb, a.r, a.ra, a.up, z + dz, y + dy, x + dx, ci);
        if (p) val = load16(p);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sx[8 * cg + 2 * q][vs] = static_cast<uint16_t>(val[q] & 0xffffu);
        sx[8 * cg + 2 * q + 1][vs] = static_cast<uint16_t>(val[q] >> 16);
      }
    }
    __syncthreads();
    const u32x4_t bq = *reinterpret_cast<const u32x4_t*>(&sx[16 * wave + row][8 * grp]);
#pragma unroll
    for (int rb = 0; rb < NR; ++rb) {
      const u32x4_t aq = *reinterpret_cast<const u32x4_t*>(&sdz[16 * rb + row][8 * grp]);
      acc[rb] = mfma_bf16(aq, bq, acc[rb]);
    }
  }
  const int kcol = k0 + 16 * wave + row;
#pragma unroll
  for (int rb = 0; rb < NR; ++rb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = 16 * rb + 4 * grp + i;
      a.partial[(static_cast<int64_t>(split) * (16 * NR) + co) * a.kcols + kcol] = acc[rb][i];
    }
}

// dw[co][ci][tap] (torch layout) = (accumulate ? dw : 0) + sum over splits in order.
__global__ void __launch_bounds__(256) wgrad_fold_kernel(const float* __restrict__ partial,
                                                         int splits, int cout, int rows,
                                                         int kcols, int cin, float* dw,
                                                         int accumulate) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = cout * cin * 27;
  if (idx >= total) return;
  const int co = idx / (cin * 27);
  const int rem = idx - co * cin * 27;
  const int c = rem / 27, tap = rem - c * 27;
  const int k = tap * cin + c;
  float s = 0.0f;
  for (int sp = 0; sp < splits; ++sp) s += partial[(static_cast<int64_t>(sp) * rows + co) * kcols + k];
  dw[idx] = accumulate ? dw[idx] + s : s;
}

// Per-block channel sums of dz[v][c] (c < cout) over a voxel range -> partial[block][c].
__global__ void __launch_bounds__(256) channel_sum_kernel(const uint16_t* __restrict__ dz,
                                                          int nvox, int dz_c, int cout,
                                                          int vox_per_block,
                                                          float* __restrict__ partial) {
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int lanes = 256 / dz_c;  // dz_c is a power of two in [8, 128]
  const int c = tid % dz_c, l = tid / dz_c;
  const int v0 = blockIdx.x * vox_per_block;
  const int v1 = min(nvox, v0 + vox_per_block);
  float s = 0.0f;
  for (int v = v0 + l; v < v1; v += lanes) s += bf2f(dz[static_cast<int64_t>(v) * dz_c + c]);
  red[tid] = s;
  __syncthreads();
  if (l == 0 && c < cout) {
    float t = 0.0f;
    for (int j = 0; j < lanes; ++j) t += red[j * dz_c + c];
    partial[static_cast<int64_t>(blockIdx.x) * cout + c] = t;
  }
}

__global__ void __launch_bounds__(256) channel_fold_kernel(const float* __restrict__ partial,
                                                           int blocks, int cout, float* db,
                                                           int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cout) return;
  float s = 0.0f;
  for (int b = 0; b < blocks; ++b) s += partial[static_cast<int64_t>(b) * cout + c];
  db[c] = accumulate ? db[c] + s : s;
}

// Pack torch-layout float32 weights w[co][ci][27] into bf16 GEMM operands.
//   forward:   wp[co][tap*cin + ci]            rows co < rows_pad, k < kpad
//   transpose: wp[ci][tap'*cout_st + co] = w[co][ci][26 - tap']   (backward-data operand)
__global__ void __launch_bounds__(256) pack_weights_kernel(const float* __restrict__ w, int cout,
                                                           int cin, int transpose, int cin_eff,
                                                           int rows_pad, int kpad,
                                                           uint16_t* __restrict__ wp) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(rows_pad) * kpad) return;
  const int rrow = static_cast<int>(idx / kpad);
  const int k = static_cast<int>(idx - static_cast<int64_t>(rrow) * kpad);
  float v = 0.0f;
  if (k < 27 * cin_eff) {
    const int tap = k / cin_eff, c = k - tap * cin_eff;
    if (!transpose) {
      // rrow = co, c = ci
      if (rrow < cout && c < cin) v = w[(static_cast<int64_t>(rrow) * cin + c) * 27 + tap];
    } else {
      // rrow = ci, c = co (cin_eff = cout_st)
      if (rrow < cin && c < cout) v = w[(static_cast<int64_t>(c) * cin + rrow) * 27 + (26 - tap)];
    }
  }
  wp[idx] = f2bf(v);
}

// 2x2x2 average pooling, bf16 -> bf16 (float32 sum in a fixed order), C % 8 == 0.
__global__ void __launch_bounds__(256) pool2_kernel(const uint16_t* __restrict__ x, int r, int c,
                                                    uint16_t* __restrict__ y) {
  const int rc = r >> 1;
  const int groups = c >> 3;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(rc) * rc * rc * groups) return;
  const int g = static_cast<int>(idx % groups);
  const int vc = static_cast<int>(idx / groups);
  const int X = vc % rc, Y = (vc / rc) % rc, Z = vc / (rc * rc);
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int child = 0; child < 8; ++child) {
    const int z = 2 * Z + (child >> 2), yy = 2 * Y + ((child >> 1) & 1), xx = 2 * X + (child & 1);
    const u32x4_t v = load16(x + static_cast<int64_t>((z * r + yy) * r + xx) * c + 8 * g);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      s[2 * q] += bf2f(static_cast<uint16_t>(v[q] & 0xffffu));
      s[2 * q + 1] += bf2f(static_cast<uint16_t>(v[q] >> 16));
    }
  }
  u32x4_t o;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    o[q] = static_cast<uint32_t>(f2bf(s[2 * q] * 0.125f)) |
           (static_cast<uint32_t>(f2bf(s[2 * q + 1] * 0.125f)) << 16);
  *reinterpret_cast<u32x4_t*>(y + static_cast<int64_t>(vc) * c + 8 * g) = o;
}

// dz[v][c] = bf16((g1[v][c] + g2[parent(v)][c] / 8) * act'(y[v][c])) for c < gc, 0 for the
// pad channels c in [gc, dz_c). g1 has stride g1_stride and may be null; g2 (the gradient
// of a following 2x2x2 average pool, at resolution r/2 with gc channels) may be null;
// y null means no activation (derivative 1).
__global__ void __launch_bounds__(256) act_bwd_kernel(const float* __restrict__ g1,
                                                      int g1_stride, const float* __restrict__ g2,
                                                      const uint16_t* __restrict__ y, float slope,
                                                      int r, int gc, int dz_c,
                                                      uint16_t* __restrict__ dz) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nvox = static_cast<int64_t>(r) * r * r;
  if (idx >= nvox * dz_c) return;
  const int c = static_cast<int>(idx % dz_c);
  const int v = static_cast<int>(idx / dz_c);
  float g = 0.0f;
  if (c < gc) {
    if (g1) g = g1[static_cast<int64_t>(v) * g1_stride + c];
    if (g2) {
      const int rc = r >> 1;
      const int x = v % r, yy = (v / r) % r, z = v / (r * r);
      const int p = ((z >> 1) * rc + (yy >> 1)) * rc + (x >> 1);
      g += g2[static_cast<int64_t>(p) * gc + c] * 0.125f;
    }
    if (y && !(bf2f(y[static_cast<int64_t>(v) * gc + c]) > 0.0f)) g *= slope;
  }
  dz[idx] = f2bf(g);
}

// Gradient of a nearest-upsampled input: ga[vc][c] = sum over the 8 children of
// dx[child][c] (c < ca; dx has stride cstride), in child order.
__global__ void __launch_bounds__(256) upsum_kernel(const float* __restrict__ dx, int r,
                                                    int cstride, int ca, float* __restrict__ ga) {
  const int rc = r >> 1;
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<int64_t>(rc) * rc * rc * ca) return;
  const int c = static_cast<int>(idx % ca);
  const int vc = static_cast<int>(idx / ca);
  const int X = vc % rc, Y = (vc / rc) % rc, Z = vc / (rc * rc);
  float s = 0.0f;
  for (int child = 0; child < 8; ++child) {
    const int z = 2 * Z + (child >> 2), y = 2 * Y + ((child >> 1) & 1), x = 2 * X + (child & 1);
    s += dx[static_cast<int64_t>((z * r + y) * r + x) * cstride + c];
  }
  ga[idx] = s;
}

__global__ void __launch_bounds__(256) f32_to_bf16_kernel(const float* __restrict__ x,
                                                          int64_t count,
                                                          uint16_t* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) y[i] = f2bf(x[i]);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

namespace pm {
namespace conv {

int rows_pad(int cout) { return (cout + 15) / 16 * 16; }
int kpad(int cin) { return (27 * cin + 31) / 32 * 32; }
int store_channels(int c) { return (c + 7) / 8 * 8; }

pm_status check_input(const uint16_t* xa, int ca, const uint16_t* xb, int cb, int r, int up,
                      const char* who) {
  PM_REQUIRE(xa, std::string(who) + ": null input");
  PM_REQUIRE(r >= 1 && r <= 512, std::string(who) + ": resolution must be in [1, 512]");
  PM_REQUIRE(ca > 0 && ca % 8 == 0 && cb >= 0 && cb % 8 == 0 && ca + cb <= 1024,
             std::string(who) + ": channel counts must be multiples of 8 (ca > 0, total <= 1024)");
  PM_REQUIRE(cb == 0 || xb, std::string(who) + ": cb > 0 needs xb");
  PM_REQUIRE(!up || r % 2 == 0, std::string(who) + ": upsampled input needs an even resolution");
  PM_REQUIRE(aligned16(xa) && (!xb || aligned16(xb)),
             std::string(who) + ": inputs must be 16-byte aligned");
  return PM_OK;
}

// Launch shape of a forward convolution: column groups of 16*NB output channels and
// reduction splits so that small grids (the network's coarse levels) still spread over the
// chip. Split-K needs float32 scratch for the raw sums (fwd_scratch_floats).
struct FwdPlan {
  int nb = 8, groups = 1, ksplit = 1, kc_per = 1;
};

FwdPlan fwd_plan(int nvox, int rows, int kpad_, bool allow_split) {
  constexpr int64_t kTargetWaves = 2048;
  FwdPlan p;
  const int64_t vox_waves = (nvox + 63) / 64;
  p.nb = std::min(8, rows / 16);
  while (p.nb > 1 && vox_waves * ((rows + 16 * p.nb - 1) / (16 * p.nb)) < kTargetWaves) p.nb /= 2;
  if (p.nb == 3) p.nb = 2;
  p.groups = (rows + 16 * p.nb - 1) / (16 * p.nb);
  const int kchunks = kpad_ / 32;
  const int64_t waves = vox_waves * p.groups;
  if (allow_split && waves < kTargetWaves) {
    int64_t want = kTargetWaves / std::max<int64_t>(1, waves);
    want = std::min<int64_t>(want, std::max(1, kchunks / 4));
    want = std::min<int64_t>(want, 32);
    p.ksplit = static_cast<int>(std::max<int64_t>(1, want));
  }
  p.kc_per = (kchunks + p.ksplit - 1) / p.ksplit;
  p.ksplit = (kchunks + p.kc_per - 1) / p.kc_per;
  return p;
}

int64_t fwd_scratch_floats(int cin, int cout, int r) {
  const int64_t nvox = static_cast<int64_t>(r) * r * r;
  const FwdPlan p = fwd_plan(static_cast<int>(nvox), rows_pad(cout), kpad(cin), true);
  return p.ksplit > 1 ? p.ksplit * nvox * rows_pad(cout) : 0;
}

pm_status fwd(const uint16_t* xa, int ca, const uint16_t* xb, int cb, int r, int up,
              const uint16_t* wp, const float* bias, int cout, int act, float slope, uint16_t* y,
              float* y32, const char* timer, hipStream_t s, float* scratch,
              int64_t scratch_floats) {
  pm_status st;
  if ((st = check_input(xa, ca, xb, cb, r, up, "pm_conv3d_fwd_bf16"))) return st;
  PM_REQUIRE(wp && aligned16(wp), "pm_conv3d_fwd_bf16: weights null or not 16-byte aligned");
  PM_REQUIRE(cout >= 1 && cout <= 1024, "pm_conv3d_fwd_bf16: cout must be in [1, 1024]");
  PM_REQUIRE(y || y32, "pm_conv3d_fwd_bf16: no output");
  ConvFwdArgs a;
  a.xa = xa;
  a.xb = cb ? xb : nullptr;
  a.ca = ca;
  a.cb = cb;
  a.cin = ca + cb;
  a.r = r;
  a.ra = up ? r / 2 : r;
  a.up = up ? 1 : 0;
  a.w = wp;
  a.kpad = kpad(a.cin);
  a.kreal = 27 * a.cin;
  a.rows = rows_pad(cout);
  a.bias = bias;
  a.cout = cout;
  a.cout_st = store_channels(cout);
  a.act = act ? 1 : 0;
  a.slope = slope;
  a.y = y;
  a.y32 = y32;
  const int nvox = r * r * r;
  FwdPlan p = fwd_plan(nvox, a.rows, a.kpad, scratch != nullptr);
  if (p.ksplit > 1 && static_cast<int64_t>(p.ksplit) * nvox * a.rows > scratch_floats)
    p = fwd_plan(nvox, a.rows, a.kpad, false);
  a.kc_per = p.kc_per;
  a.partial = p.ksplit > 1 ? scratch : nullptr;
  const dim3 grid(static_cast<unsigned>((nvox + 255) / 256), static_cast<unsigned>(p.groups),
                  static_cast<unsigned>(p.ksplit));
  const double flops = 2.0 * 27.0 * a.cin * cout * static_cast<double>(nvox);
  PM_TIMED(timer, flops, s);
  switch (p.nb) {
    case 1: conv_fwd_kernel<1><<<grid, 256, 0, s>>>(a); break;
    case 2: conv_fwd_kernel<2><<<grid, 256, 0, s>>>(a); break;
    case 4: conv_fwd_kernel<4><<<grid, 256, 0, s>>>(a); break;
    default: conv_fwd_kernel<8><<<grid, 256, 0, s>>>(a); break;
  }
  PM_CHECK_LAUNCH("conv_fwd_kernel");
  if (p.ksplit > 1) {
    const int64_t total = static_cast<int64_t>(nvox) * a.cout_st;
    conv_epilogue_kernel<<<pm::blocks_for(total), 256, 0, s>>>(
        scratch, p.ksplit, nvox, a.rows, bias, cout, a.cout_st, a.act, slope, y, y32);
    PM_CHECK_LAUNCH("conv_epilogue_kernel");
  }
  return PM_OK;
}

int64_t wgrad_partial_floats(int cin, int cout, int r) {
  int splits, chunks;
  wgrad_plan(cin, cout, r, &splits, &chunks);
  const int kcols = (kpad(cin) + 63) / 64 * 64;
  const int rows = rows_pad(store_channels(cout));
  const int64_t bias_floats = (static_cast<int64_t>(r) * r * r + 2047) / 2048 * cout;
  return std::max(static_cast<int64_t>(splits) * rows * kcols, bias_floats);
}

void wgrad_plan(int cin, int cout, int r, int* splits, int* chunks_per_split) {
  (void)cout;
  const int64_t nvox = static_cast<int64_t>(r) * r * r;
  const int64_t nchunks = (nvox + 31) / 32;
  const int kgroups = (kpad(cin) + 63) / 64;
  int64_t want = std::max<int64_t>(1, 2048 / kgroups);
  want = std::min<int64_t>(want, nchunks);
  const int64_t per = (nchunks + want - 1) / want;
  *chunks_per_split = static_cast<int>(per);
  *splits = static_cast<int>((nchunks + per - 1) / per);
}

pm_status wgrad(const uint16_t* xa, int ca, const uint16_t* xb, int cb, int r, int up,
                const uint16_t* dz, int cout, float* dw, float* db, int accumulate,
                float* partial, int64_t partial_floats, const char* timer, hipStream_t s) {
  pm_status st;
  if ((st = check_input(xa, ca, xb, cb, r, up, "pm_conv3d_bwd_weight_bf16"))) return st;
  PM_REQUIRE(dz && aligned16(dz) && dw && partial,
             "pm_conv3d_bwd_weight_bf16: null or misaligned pointer");
  PM_REQUIRE(cout >= 1 && cout <= 128, "pm_conv3d_bwd_weight_bf16: cout must be in [1, 128]");
  const int cin = ca + cb;
  const int dz_c = store_channels(cout);
  PM_REQUIRE((dz_c & (dz_c - 1)) == 0,
             "pm_conv3d_bwd_weight_bf16: cout must round up (to a multiple of 8) to a power of two");
  PM_REQUIRE(partial_floats >= wgrad_partial_floats(cin, cout, r),
             "pm_conv3d_bwd_weight_bf16: partial buffer too small");
  int splits, chunks;
  wgrad_plan(cin, cout, r, &splits, &chunks);
  ConvWgradArgs a;
  a.xa = xa;
  a.xb = cb ? xb : nullptr;
  a.ca = ca;
  a.cb = cb;
  a.cin = cin;
  a.r = r;
  a.ra = up ? r / 2 : r;
  a.up = up ? 1 : 0;
  a.dz = dz;
  a.dz_c = dz_c;
  a.kreal = 27 * cin;
  a.nvox = r * r * r;
  a.chunks_per_split = chunks;
  a.partial = partial;
  a.kcols = (kpad(cin) + 63) / 64 * 64;
  const int nr = std::max(1, rows_pad(dz_c) / 16);  // 1, 2, 4 or 8 (dz_c is 8..128)
  const dim3 grid(static_cast<unsigned>(a.kcols / 64), static_cast<unsigned>(splits));
  {
    const double flops = 2.0 * 27.0 * cin * cout * static_cast<double>(a.nvox);
    PM_TIMED(timer, flops, s);
    switch (nr) {
      case 1: conv_wgrad_kernel<1><<<grid, 256, 0, s>>>(a); break;
      case 2: conv_wgrad_kernel<2><<<grid, 256, 0, s>>>(a); break;
      case 4: conv_wgrad_kernel<4><<<grid, 256, 0, s>>>(a); break;
      default: conv_wgrad_kernel<8><<<grid, 256, 0, s>>>(a); break;
    }
    PM_CHECK_LAUNCH("conv_wgrad_kernel");
  }
  const int rows = 16 * nr;
  const int total = cout * cin * 27;
  wgrad_fold_kernel<<<pm::blocks_for(total), 256, 0, s>>>(partial, splits, cout, rows, a.kcols,
                                                           cin, dw, accumulate);
  PM_CHECK_LAUNCH("wgrad_fold_kernel");
  if (db) {
    // bias gradient: channel sums of dz, reusing the front of the partial buffer
    const int64_t nvox = a.nvox;
    const int vpb = 2048;
    const int blocks = static_cast<int>((nvox + vpb - 1) / vpb);
    PM_REQUIRE(static_cast<int64_t>(blocks) * cout <= partial_floats,
               "pm_conv3d_bwd_weight_bf16: partial buffer too small for the bias gradient");
    // the weight fold above reads `partial` on the same stream first, so reuse is ordered
    channel_sum_kernel<<<blocks, 256, 0, s>>>(dz, a.nvox, dz_c, cout, vpb, partial);
    PM_CHECK_LAUNCH("channel_sum_kernel");
    channel_fold_kernel<<<pm::blocks_for(cout), 256, 0, s>>>(partial, blocks, cout, db, accumulate);
    PM_CHECK_LAUNCH("channel_fold_kernel");
  }
  return PM_OK;
}

pm_status pack(const float* w, int cout, int cin, int transpose, uint16_t* wp, hipStream_t s) {
  PM_REQUIRE(w && wp, "pm_conv3d_pack_weights: null pointer");
  PM_REQUIRE(cout >= 1 && cout <= 1024 && cin >= 1 && cin <= 1024,
             "pm_conv3d_pack_weights: channel counts out of range");
  int rows, kp, cin_eff;
  if (!transpose) {
    rows = rows_pad(cout);
    cin_eff = cin;
    kp = kpad(cin);
  } else {
    rows = rows_pad(cin);
    cin_eff = store_channels(cout);
    kp = kpad(cin_eff);
  }
  const int64_t total = static_cast<int64_t>(rows) * kp;
  pack_weights_kernel<<<pm::blocks_for(total), 256, 0, s>>>(w, cout, cin, transpose, cin_eff, rows,
                                                            kp, wp);
  PM_CHECK_LAUNCH("pack_weights_kernel");
  return PM_OK;
}

int64_t packed_elems(int cout, int cin, int transpose) {
  if (!transpose) return static_cast<int64_t>(rows_pad(cout)) * kpad(cin);
  return static_cast<int64_t>(rows_pad(cin)) * kpad(store_channels(cout));
}

pm_status pool2(const uint16_t* x, int r, int c, uint16_t* y, hipStream_t s) {
  PM_REQUIRE(x && y && aligned16(x) && aligned16(y), "pm_pool2_bf16: null or misaligned pointer");
  PM_REQUIRE(r >= 2 && r % 2 == 0 && c % 8 == 0 && c > 0, "pm_pool2_bf16: bad shape");
  const int64_t total = static_cast<int64_t>(r / 2) * (r / 2) * (r / 2) * (c / 8);
  PM_TIMED("net.pool", 2.0 * c * (static_cast<double>(r) * r * r + total * 8.0 / c), s);
  pool2_kernel<<<pm::blocks_for(total), 256, 0, s>>>(x, r, c, y);
  PM_CHECK_LAUNCH("pool2_kernel");
  return PM_OK;
}

pm_status act_bwd(const float* g1, int g1_stride, const float* g2, const uint16_t* y, float slope,
                  int r, int gc, int dz_c, uint16_t* dz, hipStream_t s) {
  PM_REQUIRE(dz, "pm_act_bwd_bf16: null output");
  PM_REQUIRE(gc >= 1 && dz_c >= gc && r >= 1, "pm_act_bwd_bf16: bad shape");
  PM_REQUIRE(!g2 || r % 2 == 0, "pm_act_bwd_bf16: pooled gradient needs an even resolution");
  const int64_t total = static_cast<int64_t>(r) * r * r * dz_c;
  PM_TIMED("net.act_bwd", 6.0 * static_cast<double>(total), s);
  act_bwd_kernel<<<pm::blocks_for(total), 256, 0, s>>>(g1, g1_stride, g2, y, slope, r, gc, dz_c, dz);
  PM_CHECK_LAUNCH("act_bwd_kernel");
  return PM_OK;
}

pm_status upsum(const float* dx, int r, int cstride, int ca, float* ga, hipStream_t s) {
  PM_REQUIRE(dx && ga, "pm_upsum_f32: null pointer");
  PM_REQUIRE(r % 2 == 0 && ca >= 1 && cstride >= ca, "pm_upsum_f32: bad shape");
  const int64_t total = static_cast<int64_t>(r / 2) * (r / 2) * (r / 2) * ca;
  PM_TIMED("net.upsum", 36.0 * static_cast<double>(total), s);
  upsum_kernel<<<pm::blocks_for(total), 256, 0, s>>>(dx, r, cstride, ca, ga);
  PM_CHECK_LAUNCH("upsum_kernel");
  return PM_OK;
}

pm_status to_bf16(const float* x, int64_t count, uint16_t* y, hipStream_t s) {
  PM_REQUIRE(x && y && count >= 0, "pm_f32_to_bf16: bad argument");
  if (count == 0) return PM_OK;
  f32_to_bf16_kernel<<<pm::blocks_for(count), 256, 0, s>>>(x, count, y);
  PM_CHECK_LAUNCH("f32_to_bf16_kernel");
  return PM_OK;
}

}  // namespace conv
}  // namespace pm

extern "C" {

pm_status pm_conv3d_packed_size(int32_t cout, int32_t cin, int32_t transpose, int64_t* elems) {
  PM_REQUIRE(elems && cout >= 1 && cin >= 1, "pm_conv3d_packed_size: bad argument");
  *elems = pm::conv::packed_elems(cout, cin, transpose);
  return PM_OK;
}

pm_status pm_conv3d_pack_weights(const float* w, int32_t cout, int32_t cin, int32_t transpose,
                                 uint16_t* packed, hipStream_t s) {
  pm::clear_error();
  return pm::conv::pack(w, cout, cin, transpose, packed, s);
}

pm_status pm_conv3d_scratch_size(int32_t cin, int32_t cout, int32_t r, int64_t* floats) {
  PM_REQUIRE(floats && cin >= 1 && cout >= 1 && r >= 1, "pm_conv3d_scratch_size: bad argument");
  *floats = std::max(pm::conv::fwd_scratch_floats(cin, cout, r),
                     pm::conv::fwd_scratch_floats(pm::conv::store_channels(cout), cin, r));
  return PM_OK;
}

pm_status pm_conv3d_fwd_bf16(const uint16_t* xa, int32_t ca, const uint16_t* xb, int32_t cb,
                             int32_t r, int32_t up_a, const uint16_t* packed, const float* bias,
                             int32_t cout, int32_t act, float slope, uint16_t* y, float* y32,
                             float* scratch, int64_t scratch_floats, hipStream_t s) {
  pm::clear_error();
  return pm::conv::fwd(xa, ca, xb, cb, r, up_a, packed, bias, cout, act, slope, y, y32,
                       "net.conv_fwd", s, scratch, scratch ? scratch_floats : 0);
}

pm_status pm_conv3d_bwd_data_bf16(const uint16_t* dz, int32_t cout, int32_t r,
                                  const uint16_t* packed_t, int32_t cin, float* dx,
                                  float* scratch, int64_t scratch_floats, hipStream_t s) {
  pm::clear_error();
  PM_REQUIRE(dz && packed_t && dx, "pm_conv3d_bwd_data_bf16: null pointer");
  return pm::conv::fwd(dz, pm::conv::store_channels(cout), nullptr, 0, r, 0, packed_t, nullptr,
                       cin, 0, 0.0f, nullptr, dx, "net.conv_bwd_data", s, scratch,
                       scratch ? scratch_floats : 0);
}

pm_status pm_conv3d_wgrad_partial_size(int32_t cin, int32_t cout, int32_t r, int64_t* floats) {
  PM_REQUIRE(floats && cin >= 1 && cout >= 1 && r >= 1, "pm_conv3d_wgrad_partial_size: bad argument");
  *floats = pm::conv::wgrad_partial_floats(cin, cout, r);
  return PM_OK;
}

pm_status pm_conv3d_bwd_weight_bf16(const uint16_t* xa, int32_t ca, const uint16_t* xb,
                                    int32_t cb, int32_t r, int32_t up_a, const uint16_t* dz,
                                    int32_t cout, float* dw, float* db, int32_t accumulate,
                                    float* partial, int64_t partial_floats, hipStream_t s) {
  pm::clear_error();
  return pm::conv::wgrad(xa, ca, xb, cb, r, up_a, dz, cout, dw, db, accumulate, partial,
                         partial_floats, "net.conv_bwd_weight", s);
}

pm_status pm_pool2_bf16(const uint16_t* x, int32_t r, int32_t c, uint16_t* y, hipStream_t s) {
  pm::clear_error();
  return pm::conv::pool2(x, r, c, y, s);
}

pm_status pm_act_bwd_bf16(const float* g1, int32_t g1_stride, const float* g2, const uint16_t* y,
                          float slope, int32_t r, int32_t gc, int32_t dz_c, uint16_t* dz,
                          hipStream_t s) {
  pm::clear_error();
  return pm::conv::act_bwd(g1, g1_stride, g2, y, slope, r, gc, dz_c, dz, s);
}

pm_status pm_upsum_f32(const float* dx, int32_t r, int32_t cstride, int32_t ca, float* ga,
                       hipStream_t s) {
  pm::clear_error();
  return pm::conv::upsum(dx, r, cstride, ca, ga, s);
}

pm_status pm_f32_to_bf16(const float* x, int64_t count, uint16_t* y, hipStream_t s) {
  pm::clear_error();
  return pm::conv::to_bf16(x, count, y, s);
}

}  // extern "C"
