// Optional per-kernel timing with HIP events on the launch stream.
//
// When enabled (pm_timing_enable), every instrumented launch records an event pair around
// itself on its own stream, tagged with a kernel name and the kernel's algorithmic
// (compulsory) bytes for that launch. pm_timing_report synchronizes the events and returns
// per-name totals. When disabled, a scope costs one branch.
#pragma once

#include "pm_common.h"

namespace pm {

bool timing_enabled();

class KernelTimer {
 public:
  KernelTimer(const char* name, double bytes, hipStream_t stream);
  ~KernelTimer();
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;

 private:
  int slot_ = -1;
  hipStream_t stream_ = nullptr;
};

}  // namespace pm

#define PM_CAT2_(a, b) a##b
#define PM_CAT_(a, b) PM_CAT2_(a, b)
#define PM_TIMED(name, bytes, stream) \
  pm::KernelTimer PM_CAT_(pm_timer_, __LINE__)((name), (bytes), (stream))
