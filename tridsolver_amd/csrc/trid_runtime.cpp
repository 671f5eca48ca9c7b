// trid_runtime.cpp -- stream, error and workspace state of tridsolver_amd.
#include "trid_runtime.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/trid_hip_ext.h"

namespace tridhip {

namespace {
std::atomic<hipStream_t> g_stream{nullptr};
std::atomic<int> g_autotune{-1};
std::atomic<int> g_wide{0};
std::atomic<int> g_cache_policy{-1};
std::mutex g_mu;
ForcedConfig g_forced;
thread_local std::string t_error;
thread_local std::string t_config;
thread_local LaunchEvents t_events;

struct DevWs {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool captured = false;  // handed to a call during a stream capture
};
// keyed by (device, stream)
std::map<std::pair<int, hipStream_t>, DevWs> g_ws;
// buffers replaced after a capture used them: a caller's graph may still
// launch kernels that read them, so they are kept until the process ends
std::vector<void *> g_retired;
}  // namespace

hipStream_t current_stream() { return g_stream.load(); }

LaunchEvents take_launch_events() {
  const LaunchEvents e = t_events;
  t_events = LaunchEvents{};
  return e;
}

tridStatus_t fail(tridStatus_t st, const std::string &msg) {
  t_error = msg;
  return st;
}

void clear_error() { t_error.clear(); }

void set_last_config(const std::string &cfg) { t_config = cfg; }

tridStatus_t check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return TRID_STATUS_SUCCESS;
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(TRID_STATUS_ALLOC_FAILED, m);
  if (e == hipErrorNoBinaryForGpu || e == hipErrorInvalidDeviceFunction)
    return fail(TRID_STATUS_ARCH_MISMATCH, m);
  return fail(TRID_STATUS_EXECUTION_FAILED, m);
}

bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return cs != hipStreamCaptureStatusNone;
}

void *workspace(size_t bytes, hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  DevWs &w = g_ws[std::make_pair(dev, st)];
  const bool cap = capturing(st);
  if (w.bytes >= bytes && w.ptr) {
    w.captured |= cap;
    return w.ptr;
  }
  if (cap) {
    fail(TRID_STATUS_ALLOC_FAILED,
         "workspace of " + std::to_string(bytes) +
             " bytes cannot be allocated during stream capture: run the same call once "
             "before capturing it");
    return nullptr;
  }
  if (w.ptr) {
    if (w.captured) {
      g_retired.push_back(w.ptr);
    } else {
      // the old buffer may still be read by queued kernels on this stream
      (void)hipStreamSynchronize(st);
      (void)hipFree(w.ptr);
    }
    w.ptr = nullptr;
    w.bytes = 0;
    w.captured = false;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    fail(TRID_STATUS_ALLOC_FAILED, "workspace allocation of " + std::to_string(bytes) +
                                       " bytes failed");
    return nullptr;
  }
  w.ptr = p;
  w.bytes = bytes;
  return p;
}

int autotune_mode() {
  int m = g_autotune.load();
  if (m < 0) {
    const char *e = std::getenv("TRIDHIP_AUTOTUNE");
    m = (e && e[0] == '1') ? 1 : 0;
    g_autotune.store(m);
  }
  return m;
}

bool wide_index() { return g_wide.load() != 0; }

int forced_cache_policy() { return g_cache_policy.load(); }

unsigned resident_blocks(const void *kernel, int threads, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, int, size_t>, unsigned> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const auto key = std::make_tuple(dev, kernel, threads, lds);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    per_cu = 1;
    cus = 256;
  }
  const unsigned n = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
  cache[key] = n;
  return n;
}

int cache_policy(size_t bytes, size_t rhs_bytes) {
  const int f = g_cache_policy.load();
  if (f >= 0) return f;
  if (bytes <= kMallFitBytes) return 1;
  return rhs_bytes <= kMallFitBytes ? 2 : 0;
}

ForcedConfig forced_config() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_forced;
}

}  // namespace tridhip

using namespace tridhip;

EXTERN_C tridStatus_t tridSetLaunchEvents(void *start, void *stop) {
  tridhip::t_events.start = static_cast<hipEvent_t>(start);
  tridhip::t_events.stop = static_cast<hipEvent_t>(stop);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridSetStream(void *stream) {
  g_stream.store(reinterpret_cast<hipStream_t>(stream));
  return TRID_STATUS_SUCCESS;
}

EXTERN_C void *tridGetStream(void) { return g_stream.load(); }

EXTERN_C tridStatus_t tridSetAutotune(int mode) {
  if (mode != 0 && mode != 1)
    return fail(TRID_STATUS_INVALID_VALUE, "tridSetAutotune: mode must be 0 or 1");
  g_autotune.store(mode);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C const char *tridGetLastError(void) { return t_error.c_str(); }

EXTERN_C const char *tridGetLastConfig(void) { return t_config.c_str(); }

EXTERN_C const char *tridGetStatusString(tridStatus_t s) {
  switch (s) {
    case TRID_STATUS_SUCCESS: return "TRID_STATUS_SUCCESS";
    case TRID_STATUS_NOT_INITIALIZED: return "TRID_STATUS_NOT_INITIALIZED";
    case TRID_STATUS_ALLOC_FAILED: return "TRID_STATUS_ALLOC_FAILED";
    case TRID_STATUS_INVALID_VALUE: return "TRID_STATUS_INVALID_VALUE";
    case TRID_STATUS_ARCH_MISMATCH: return "TRID_STATUS_ARCH_MISMATCH";
    case TRID_STATUS_MAPPING_ERROR: return "TRID_STATUS_MAPPING_ERROR";
    case TRID_STATUS_EXECUTION_FAILED: return "TRID_STATUS_EXECUTION_FAILED";
    case TRID_STATUS_INTERNAL_ERROR: return "TRID_STATUS_INTERNAL_ERROR";
    case TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED: return "TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED";
    case TRID_STATUS_ZERO_PIVOT: return "TRID_STATUS_ZERO_PIVOT";
  }
  return "unknown tridStatus_t";
}

EXTERN_C tridStatus_t tridForceConfig(int kind, int rows, int cps, int spb) {
  if (kind < 0 || kind > 4) return fail(TRID_STATUS_INVALID_VALUE, "tridForceConfig: kind");
  if (rows != 0 && rows != 4 && rows != 5 && rows != 6 && rows != 7 && rows != 8 &&
      rows != 12 && rows != 16)
    return fail(TRID_STATUS_INVALID_VALUE, "tridForceConfig: rows must be 0, 4..8, 12 or 16");
  auto pow2 = [](int v) { return v == 0 || (v > 0 && (v & (v - 1)) == 0); };
  if (!pow2(cps) || !pow2(spb))
    return fail(TRID_STATUS_INVALID_VALUE, "tridForceConfig: cps/spb must be powers of two");
  std::lock_guard<std::mutex> lk(g_mu);
  g_forced.kind = kind;
  g_forced.rows = rows;
  g_forced.cps = cps;
  g_forced.spb = spb;
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridForceWideIndex(int on) {
  if (on != 0 && on != 1) return fail(TRID_STATUS_INVALID_VALUE, "tridForceWideIndex: on must be 0 or 1");
  g_wide.store(on);
  return TRID_STATUS_SUCCESS;
}

EXTERN_C tridStatus_t tridForceCachePolicy(int policy) {
  if (policy < -1 || policy > 2)
    return fail(TRID_STATUS_INVALID_VALUE, "tridForceCachePolicy: policy must be -1, 0, 1 or 2");
  g_cache_policy.store(policy);
  return TRID_STATUS_SUCCESS;
}
