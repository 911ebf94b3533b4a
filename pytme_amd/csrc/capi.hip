// Library-wide C-ABI helpers: error reporting, device queries and the
// optional per-kernel event timing used by the benchmark.
#include <stdarg.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "tm_common.h"

static thread_local char g_err[1024] = {0};

void tm_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {
struct TimedLaunch {
  std::string name;
  hipEvent_t a, b;
  double bytes;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimedLaunch> g_launches;
std::vector<hipEvent_t> g_pool;  // all events ever created
size_t g_pool_used = 0;
}  // namespace

bool tm_timing_on() { return g_timing; }

hipEvent_t tm_timing_event() {
  std::lock_guard<std::mutex> lk(g_tmu);
  if (g_pool_used == g_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_pool.push_back(e);
  }
  return g_pool[g_pool_used++];
}

void tm_timing_push(const char* name, hipEvent_t a, hipEvent_t b, double bytes) {
  if (!a || !b) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  g_launches.push_back({name, a, b, bytes});
}

extern "C" {

#define TM_ABI_VERSION 1

int tm_abi_version(void) { return TM_ABI_VERSION; }

const char* tm_last_error(void) { return g_err; }

int tm_device_count(int* n) {
  TM_REQUIRE(n != nullptr, "null pointer");
  TM_CHECK_HIP(hipGetDeviceCount(n));
  return TM_OK;
}

// Block until all work queued on the stream has finished; reports any
// asynchronous kernel fault as an error status.
int tm_stream_sync(void* stream) {
  TM_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return TM_OK;
}

// Start/stop recording launch events (recording clears earlier records).
int tm_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
  if (on) {
    g_launches.clear();
    g_pool_used = 0;
  }
  return TM_OK;
}

// Names of the kernels recorded so far, '\n'-separated, into buf.
int tm_timing_names(char* buf, int len) {
  TM_REQUIRE(buf && len > 0, "null buffer");
  std::lock_guard<std::mutex> lk(g_tmu);
  std::vector<std::string> names;
  for (auto& l : g_launches) {
    bool seen = false;
    for (auto& n : names) seen = seen || n == l.name;
    if (!seen) names.push_back(l.name);
  }
  std::string out;
  for (auto& n : names) out += n + "\n";
  TM_REQUIRE((int)out.size() < len, "buffer too small");
  memcpy(buf, out.c_str(), out.size() + 1);
  return TM_OK;
}

// Total device time (ms), launch count and algorithmic bytes of one kernel
// name.  Waits for the recorded events.
int tm_timing_query(const char* name, double* total_ms, int64_t* count, double* bytes) {
  TM_REQUIRE(name && total_ms && count && bytes, "null pointer");
  std::lock_guard<std::mutex> lk(g_tmu);
  double ms = 0.0, by = 0.0;
  int64_t n = 0;
  for (auto& l : g_launches) {
    if (l.name != name) continue;
    TM_CHECK_HIP(hipEventSynchronize(l.b));
    float e = 0.f;
    TM_CHECK_HIP(hipEventElapsedTime(&e, l.a, l.b));
    ms += e;
    by += l.bytes;
    ++n;
  }
  *total_ms = ms;
  *count = n;
  *bytes = by;
  return TM_OK;
}

}  // extern "C"
