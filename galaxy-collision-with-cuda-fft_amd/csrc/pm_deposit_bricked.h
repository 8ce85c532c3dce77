// Bricked fixed-point CIC deposit (pm_deposit_bricked.hip): PM_DEPOSIT_BRICKED of
// pm_deposit_cic and the step deposit pm_deposit_cic_step.
#pragma once

#include "pm_common.h"

namespace pm {
namespace depbrick {

struct State;

pm_status create(int n, int64_t cap, State** out);
void destroy(State* st);

// Sort buffers borrowed from the deposit plan: keys[2], vals[2] of cap entries, hist of
// radix::hist_words(cap) words, digit of 2^radix::kMaxDigitBits words.
struct SortBuffers {
  uint32_t* keys[2];
  int32_t* vals[2];
  uint32_t* hist;
  uint32_t* digit;
};

// rho = the fixed-point deposit of the particles. bin: bin the particles into bricks
// again; otherwise the binning of the last call is reused when it covers np particles with
// the same dx (particles that left their brick's region take the slow exact path).
// rho_kept: rho is the array of the last call, unchanged since, so the node bricks far
// from every particle (still zero) need not be stored again.
pm_status deposit(State* st, const SortBuffers& sb, const float* px, const float* py,
                  const float* pz, const float* mass, int64_t np, float dx, float* rho, bool bin,
                  bool rho_kept, hipStream_t s);
// The particles now lie in the order of the last binning (order()): later deposits read
// them in place instead of through that order.
pm_status mark_ordered(State* st);
// Particle order of the binning done by the last deposit (binned position -> particle
// index), or null when the last deposit reused an older binning.
const int32_t* order(const State* st);
// Node bricks a gather may read (those near a particle), for pm_grad4_listed.
void node_list(const State* st, const int32_t** list, const uint32_t** count, int64_t* max);
// Forget the binning (its order lives in the plan's sort buffers, which another mode's
// deposit overwrites).
void invalidate(State* st);

}  // namespace depbrick
}  // namespace pm
