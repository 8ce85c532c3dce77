// Warm-start kernels used by pm_mg_solve_warm (pm_warm.hip).
#pragma once

#include "pm_common.h"

namespace pm {
namespace warm {

// Interior phi <- 2*phi - hist, hist <- old phi; cand (optional) <- new phi.
pm_status extrapolate(float* phi, float* hist, float* cand, int n, hipStream_t s);
// Network input: channels [r, phi - hist, f] pooled by p^3, each divided by its RMS, bf16
// [R^3][8]; scales[0..2] the RMS values, scales[3] = RMS(r)*(p*h)^2, scales[4] = 1/scales[3].
// pooled: float32 scratch [R^3][3]; partials/red: reduction scratch (red gets 3 doubles).
pm_status make_input(const float* r32, const float* phi, const float* hist, const float* f, int n,
                     int p, double h, float* pooled, uint16_t* input, double* partials,
                     double* red, float* scales, hipStream_t s);
// e = scales[3] * trilinear(out) on the interior, 0 on the faces.
pm_status apply(const float* out, int n, int p, const float* scales, float* e, hipStream_t s);
// rnew = r32 - lap(e) (float64 arithmetic, 0 on the faces); red[0], red[1] = sums over the
// network voxels of the squared mean residual before (r32) and after (rnew); red[2] =
// sum of rnew^2. The pool factor p must divide 64.
pm_status guard(const float* r32, const float* e, int n, int p, double h, float* rnew,
                double* partials, double* red, hipStream_t s);
// tgt[V] = scales[4] * mean over V's fine nodes of (phi - cand).
pm_status target(const float* phi, const float* cand, int n, int p, const float* scales, float* tgt,
                 hipStream_t s);

}  // namespace warm
}  // namespace pm
