// Deterministic float64 reductions shared by the multigrid solver and the diagnostics.
//
// Every reduction is two-stage with a fixed tree: stage 1 runs min(kRedBlocks,
// ceil(count/256)) workgroups; each thread sums its grid-stride elements sequentially,
// then the workgroup combines lanes with a shuffle-down tree and the four waves in order.
// Stage 2 is one workgroup that folds the partials the same way. For a given element
// count the result is therefore identical run to run.
#pragma once

#include "pm_common.h"

namespace pm {

constexpr int kRedBlocks = 2048;
constexpr int kRedMaxQ = 16;  // quantities per reduction (partials stride)

inline unsigned red_blocks(int64_t count) {
  int64_t b = (count + kBlock - 1) / kBlock;
  if (b > kRedBlocks) b = kRedBlocks;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

// x from lane + N of the same row of 16 lanes (a DPP row shift; a lane whose source lies
// past its row keeps its own x, which no lane that feeds lane 0 reads).
template <int N>
__device__ __forceinline__ double row_down(double x) {
  static_assert(N >= 1 && N <= 15, "row shift");
  const int lo = __double2loint(x), hi = __double2hiint(x);
  const int l = __builtin_amdgcn_update_dpp(lo, lo, 0x100 + N, 0xF, 0xF, false);
  const int h = __builtin_amdgcn_update_dpp(hi, hi, 0x100 + N, 0xF, 0xF, false);
  return __hiloint2double(h, l);
}

// Sum Q doubles across a 256-thread workgroup; the result is valid on thread 0.
// s_red must hold 4*Q doubles. Contains __syncthreads(): call from uniform control flow.
// Each wave sums with a shuffle-down tree (offsets 32, 16, ..., 1). From offset 8 on, the
// lanes whose sums reach lane 0 read lanes of their own row of 16, so those levels use DPP
// row shifts (no LDS round trip), which hand them the same operands: lane 0's sum is the
// shuffle tree's.
template <int Q>
__device__ __forceinline__ void block_sum(double (&v)[Q], double* s_red) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off >= 16; off >>= 1) {
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] += __shfl_down(v[q], off, 64);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] += row_down<8>(v[q]);
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] += row_down<4>(v[q]);
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] += row_down<2>(v[q]);
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] += row_down<1>(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) s_red[wave * Q + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double t = s_red[q];
      for (int w = 1; w < 4; ++w) t = t + s_red[w * Q + q];
      v[q] = t;
    }
  }
  __syncthreads();
}

// v[q] += p[b*sb + q*sq] for b = threadIdx.x, threadIdx.x + blockDim.x, ... < blocks, in that
// order (the stage-2 loop of every fold), with the loads of fold_unroll<Q>() iterations
// issued before they are added: a fold is a chain of memory round trips, one per batch
// (16 iterations hold 32*Q registers: 16 batches up to 3 quantities, fewer beyond).
template <int Q>
constexpr int fold_unroll() {
  return Q <= 3 ? 16 : (Q <= 6 ? 8 : 4);
}
template <int Q>
__device__ __forceinline__ void fold_loop(const double* __restrict__ p, int blocks, int64_t sb,
                                          int64_t sq, double (&v)[Q]) {
  constexpr int kFoldUnroll = fold_unroll<Q>();
  for (int b0 = threadIdx.x; b0 < blocks; b0 += kFoldUnroll * blockDim.x) {
    double t[kFoldUnroll][Q];
#pragma unroll
    for (int u = 0; u < kFoldUnroll; ++u) {
      const int b = b0 + u * static_cast<int>(blockDim.x);
#pragma unroll
      for (int q = 0; q < Q; ++q) t[u][q] = b < blocks ? p[b * sb + q * sq] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFoldUnroll; ++u) {
      if (b0 + u * static_cast<int>(blockDim.x) < blocks) {
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = v[q] + t[u][q];
      }
    }
  }
}

// Stage 2: fold `blocks` partials of Q quantities (partials[b*kRedMaxQ + q]) into out[q].
template <int Q>
__global__ void __launch_bounds__(256) fold_partials_kernel(const double* __restrict__ partials,
                                                            int blocks,
                                                            double* __restrict__ out) {
  __shared__ double s_red[4 * Q];
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = 0.0;
  fold_loop<Q>(partials, blocks, kRedMaxQ, 1, v);
  block_sum<Q>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q] = v[q];
  }
}

// fold_partials_kernel for partials stored quantity-major, partials[q * blocks + b] (one
// contiguous column per quantity): the same per-thread order and tree, so the same result.
template <int Q>
__global__ void __launch_bounds__(256) fold_columns_kernel(const double* __restrict__ partials,
                                                           int blocks,
                                                           double* __restrict__ out) {
  __shared__ double s_red[4 * Q];
  double v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = 0.0;
  fold_loop<Q>(partials, blocks, 1, blocks, v);
  block_sum<Q>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q] = v[q];
  }
}

// Two fold_columns_kernel folds in one launch: Q1 columns of p1 into out1 and Q2 columns of
// p2 into out2, each quantity summed exactly as its own fold_columns_kernel would.
template <int Q1, int Q2>
__global__ void __launch_bounds__(256) fold_columns_pair_kernel(const double* __restrict__ p1,
                                                                int blocks1,
                                                                double* __restrict__ out1,
                                                                const double* __restrict__ p2,
                                                                int blocks2,
                                                                double* __restrict__ out2) {
  __shared__ double s_red[4 * (Q1 + Q2)];
  double v[Q1 + Q2];
#pragma unroll
  for (int q = 0; q < Q1 + Q2; ++q) v[q] = 0.0;
  double v1[Q1], v2[Q2];
#pragma unroll
  for (int q = 0; q < Q1; ++q) v1[q] = 0.0;
#pragma unroll
  for (int q = 0; q < Q2; ++q) v2[q] = 0.0;
  fold_loop<Q1>(p1, blocks1, 1, blocks1, v1);
  fold_loop<Q2>(p2, blocks2, 1, blocks2, v2);
#pragma unroll
  for (int q = 0; q < Q1; ++q) v[q] = v1[q];
#pragma unroll
  for (int q = 0; q < Q2; ++q) v[Q1 + q] = v2[q];
  block_sum<Q1 + Q2>(v, s_red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q1; ++q) out1[q] = v[q];
#pragma unroll
    for (int q = 0; q < Q2; ++q) out2[q] = v[Q1 + q];
  }
}

// Scratch for reductions issued without a handle: kRedBlocks*kRedMaxQ partial doubles and
// kRedMaxQ result doubles on the device, plus a pinned host mirror of the result.
struct RedScratch {
  double* partials = nullptr;
  double* result = nullptr;
  double* host = nullptr;
};
pm_status red_scratch(RedScratch** out);

// sum of x^2 over `count` float32 values -> out_dev[0] (float64). Enqueued only.
pm_status launch_sumsq_f32(const float* x, int64_t count, double* partials, double* out_dev,
                           hipStream_t stream);
// sum of x^2 over `count` float64 values -> out_dev[0]. Enqueued only.
pm_status launch_sumsq_f64(const double* x, int64_t count, double* partials, double* out_dev,
                           hipStream_t stream);
// Ten raw moments of rho -> sums_dev[0..9]; if face_params is non-null also writes the
// eleven float32 face parameters [G, M, cx, cy, cz, sxx, syy, szz, sxy, sxz, syz]
// computed from them exactly as reference multigrid.py:364-389 does on the host. With
// f_sumsq_dev the same pass sums f^2 for f = factor * rho into it, and with f also writes f.
pm_status launch_moments(const float* rho, int n, float dx, double grav, double* partials,
                         double* sums_dev, float* face_params, hipStream_t stream,
                         float* f = nullptr, float factor = 0.0f, double* f_sumsq_dev = nullptr);

}  // namespace pm
