// Optional per-kernel timing with HIP events on the launch stream.
//
// When enabled (pm_timing_enable), every instrumented launch records an event pair around
// itself on its own stream, tagged with a kernel name and the kernel's algorithmic
// (compulsory) bytes for that launch. pm_timing_report synchronizes the events and returns
// per-name totals. When disabled, a scope costs one branch.
#pragma once

#include <hip/hip_ext.h>

#include "pm_common.h"

namespace pm {

// Event-pair timers (KernelTimer) record: mode 1.
bool timing_enabled();
// While a TimingPause lives on the calling thread no timer records (and no timing event is
// created): for entry points that must not acquire resources.
class TimingPause {
 public:
  TimingPause();
  ~TimingPause();
  TimingPause(const TimingPause&) = delete;
  TimingPause& operator=(const TimingPause&) = delete;
};
// Dispatch-recorded timers (LaunchTimer, PM_LAUNCH_TIMED) record: modes 1 and 2. Mode 2 adds
// no packets to the streams, so the timed work runs as it does untimed (graphs included).
bool launch_timing_enabled();

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

// An event pair that the kernel's own dispatch records (hipExtLaunchKernelGGL): the pair
// then spans the kernel's execution alone, as a kernel trace reports it, without the
// dispatch gaps of separate event packets. Used for the kernels whose bandwidth is quoted.
class LaunchTimer {
 public:
  LaunchTimer(const char* name, double bytes);
  hipEvent_t start() const { return start_; }
  hipEvent_t stop() const { return stop_; }

 private:
  hipEvent_t start_ = nullptr;
  hipEvent_t stop_ = nullptr;
};

}  // namespace pm

// Launch `kernel` with grid/block/shmem on `stream`; when timing is enabled its dispatch
// records a LaunchTimer pair.
#define PM_LAUNCH_TIMED(name, bytes, kernel, grid, block, shmem, stream, ...)               \
  do {                                                                                      \
    if (pm::launch_timing_enabled()) {                                                      \
      pm::LaunchTimer pm_lt_((name), (bytes));                                              \
      hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem), (stream),             \
                            pm_lt_.start(), pm_lt_.stop(), 0u, __VA_ARGS__);                \
    } else {                                                                                \
      kernel<<<(grid), (block), (shmem), (stream)>>>(__VA_ARGS__);                          \
    }                                                                                       \
  } while (0)

#define PM_CAT2_(a, b) a##b
#define PM_CAT_(a, b) PM_CAT2_(a, b)
#define PM_TIMED(name, bytes, stream) \
  pm::KernelTimer PM_CAT_(pm_timer_, __LINE__)((name), (bytes), (stream))
