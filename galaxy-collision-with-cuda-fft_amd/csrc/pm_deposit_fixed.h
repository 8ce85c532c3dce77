// Fixed-point CIC deposit (pm_deposit_fixed.hip), used by pm_deposit_cic in
// PM_DEPOSIT_FIXED mode.
#pragma once

#include "pm_common.h"

namespace pm {
namespace depfix {

constexpr unsigned kMaxBlocks = 4096;  // blocks of the max reduction

// acc: n^3 zero-initialized 64-bit accumulators (left zeroed on return); maxbits: 1 +
// kMaxBlocks device words of scratch. keys_out (optional): the extended cell key of every
// particle.
pm_status deposit(const float* px, const float* py, const float* pz, const float* mass,
                  int64_t np, int n, float dx, float* rho, uint32_t* keys_out,
                  unsigned long long* acc, unsigned int* maxbits, hipStream_t s);
pm_status zero(unsigned long long* acc, int64_t count, hipStream_t s);
// maxbits[0] = float bits of the largest finite positive mass*inv_vol (0 if none);
// maxbits[1 .. kMaxBlocks] are scratch.
pm_status max_mprime(const float* mass, int64_t np, float inv_vol, unsigned int* maxbits,
                     hipStream_t s);

}  // namespace depfix
}  // namespace pm
