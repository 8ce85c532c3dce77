// Event-pair kernel timing (see pm_timing.h).
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pm_timing.h"

namespace {

struct Record {
  std::string name;
  double bytes = 0.0;
  hipEvent_t start = nullptr;
  hipEvent_t stop = nullptr;
};

std::atomic<int> g_mode{0};  // 0 off, 1 every timer, 2 dispatch-recorded timers only
std::mutex g_mu;
std::vector<Record> g_records;      // pending event pairs
std::vector<hipEvent_t> g_free;     // recycled events

struct Total {
  int64_t launches = 0;
  double ms = 0.0;
  double bytes = 0.0;
};
std::map<std::string, Total> g_totals;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  // timing-only events: no system-scope fence (and its cache writeback and invalidate,
  // which would also slow the kernel that follows) when they are recorded
  if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
  return e;
}

// Folds completed records into g_totals (caller holds g_mu). Synchronizes the events.
pm_status drain() {
  for (Record& r : g_records) {
    hipError_t e = hipEventSynchronize(r.stop);
    if (e != hipSuccess) return pm::hip_fail("hipEventSynchronize (timing)", e);
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, r.start, r.stop);
    if (e != hipSuccess) return pm::hip_fail("hipEventElapsedTime (timing)", e);
    Total& t = g_totals[r.name];
    t.launches += 1;
    t.ms += ms;
    t.bytes += r.bytes;
    g_free.push_back(r.start);
    g_free.push_back(r.stop);
  }
  g_records.clear();
  return PM_OK;
}

}  // namespace

namespace pm {

thread_local int t_paused = 0;

bool timing_enabled() { return t_paused == 0 && g_mode.load(std::memory_order_relaxed) == 1; }

bool launch_timing_enabled() {
  return t_paused == 0 && g_mode.load(std::memory_order_relaxed) != 0;
}

TimingPause::TimingPause() { ++t_paused; }

TimingPause::~TimingPause() { --t_paused; }

KernelTimer::KernelTimer(const char* name, double bytes, hipStream_t stream) {
  if (!timing_enabled()) return;
  std::lock_guard<std::mutex> lock(g_mu);
  Record r;
  r.name = name;
  r.bytes = bytes;
  r.start = take_event();
  r.stop = take_event();
  if (!r.start || !r.stop) return;
  if (hipEventRecord(r.start, stream) != hipSuccess) return;
  stream_ = stream;
  slot_ = static_cast<int>(g_records.size());
  g_records.push_back(r);
}

KernelTimer::~KernelTimer() {
  if (slot_ < 0) return;
  std::lock_guard<std::mutex> lock(g_mu);
  if (slot_ < static_cast<int>(g_records.size()))
    (void)hipEventRecord(g_records[slot_].stop, stream_);
  // very long timed runs: fold completed pairs to bound the pending list
  if (g_records.size() > 200000) (void)drain();
}

LaunchTimer::LaunchTimer(const char* name, double bytes) {
  if (!launch_timing_enabled()) return;
  std::lock_guard<std::mutex> lock(g_mu);
  Record r;
  r.name = name;
  r.bytes = bytes;
  r.start = take_event();
  r.stop = take_event();
  if (!r.start || !r.stop) return;
  start_ = r.start;
  stop_ = r.stop;
  g_records.push_back(r);
  if (g_records.size() > 200000) (void)drain();
}

}  // namespace pm

extern "C" {

pm_status pm_timing_enable(int32_t on) {
  std::lock_guard<std::mutex> lock(g_mu);
  PM_REQUIRE(on >= 0 && on <= 2, "pm_timing_enable: mode must be 0, 1 or 2");
  g_mode.store(on);
  return PM_OK;
}

pm_status pm_timing_reset(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  pm_status st = drain();
  g_totals.clear();
  return st;
}

pm_status pm_timing_report(int32_t capacity, pm_kernel_time* out, int32_t* count) {
  PM_REQUIRE(count, "pm_timing_report: null count");
  std::lock_guard<std::mutex> lock(g_mu);
  pm_status st = drain();
  if (st != PM_OK) return st;
  int32_t i = 0;
  for (const auto& kv : g_totals) {
    if (out && i < capacity) {
      pm_kernel_time& k = out[i];
      std::memset(&k, 0, sizeof(k));
      std::strncpy(k.name, kv.first.c_str(), sizeof(k.name) - 1);
      k.launches = kv.second.launches;
      k.total_ms = kv.second.ms;
      k.bytes = kv.second.bytes;
    }
    ++i;
  }
  *count = i;
  return PM_OK;
}

}  // extern "C"
