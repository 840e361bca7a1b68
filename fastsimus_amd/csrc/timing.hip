// Optional HIP-event timing of the dominant kernel launches (fs_set_kernel_timing /
// fs_kernel_time_ms in include/fastsimus.h). Process-wide, not thread-safe: meant
// for benchmarks that time one stream at a time.
#include "fs_device.h"

namespace fs {

constexpr int kSlots = 512;
static int g_mask = 0;  // FS_TIMING_* classes being timed
static bool g_init = false;
static hipEvent_t g_start[kSlots], g_stop[kSlots];
static int g_used = 0;

int timing_start(hipStream_t st, int cls) {
  if (!(g_mask & cls) || g_used >= kSlots) return -1;
  const int s = g_used++;
  hipEventRecord(g_start[s], st);
  return s;
}

void timing_stop(int slot, hipStream_t st) {
  if (slot >= 0) hipEventRecord(g_stop[slot], st);
}

}  // namespace fs

extern "C" {

int fs_set_kernel_timing(int enable) {
  using namespace fs;
  if (enable && !g_init) {
    for (int i = 0; i < kSlots; ++i) {
      if (hipEventCreate(&g_start[i]) != hipSuccess || hipEventCreate(&g_stop[i]) != hipSuccess) return FS_ERR_LAUNCH;
    }
    g_init = true;
  }
  g_mask = enable;
  g_used = 0;
  return FS_OK;
}

double fs_kernel_time_ms(int* launches) {
  using namespace fs;
  double total = 0.0;
  for (int i = 0; i < g_used; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(g_stop[i]) == hipSuccess && hipEventElapsedTime(&ms, g_start[i], g_stop[i]) == hipSuccess)
      total += ms;
  }
  if (launches) *launches = g_used;
  g_used = 0;
  return total;
}

}  // extern "C"
