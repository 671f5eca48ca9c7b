// trid_runtime.hpp -- host-side runtime shared by the tridsolver_amd entry
// points: current stream, per-thread error text, per-device workspace cache.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/trid_common.h"

namespace tridhip {

hipStream_t current_stream();

// Records `msg` as this thread's last error and returns `st`.
tridStatus_t fail(tridStatus_t st, const std::string &msg);
void clear_error();
void set_last_config(const std::string &cfg);

// Maps a HIP error to a status (and records it); hipSuccess -> SUCCESS.
tridStatus_t check_hip(hipError_t e, const char *what);

// True when work enqueued on `st` would be recorded into a graph rather
// than run: the stream itself is capturing, or the query is refused (the
// null stream while another stream captures in global mode).
bool capturing(hipStream_t st);

// Grow-only device workspace of the current device and stream `st`: calls
// on different streams never share a buffer, calls on one stream run in
// order.  Growing waits for `st` and frees the old buffer, except a buffer
// that a call under capture used: a caller's graph keeps its address, so it
// is kept until the process ends.  During a capture on `st` no allocation
// is made: a big-enough buffer from an earlier call is returned, otherwise
// nullptr with the reason recorded as the last error.  Returns nullptr on
// allocation failure.
void *workspace(size_t bytes, hipStream_t st);

int autotune_mode();

// Events that the calling thread's next kernel launch of the library records
// (tridSetLaunchEvents): hipExtLaunchKernel timestamps the dispatch itself,
// so start/stop bracket exactly the kernel's execution.  Taking them clears
// them.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
LaunchEvents take_launch_events();

// Launches `kernel` on `st`; when launch events are pending (see above) the
// launch records them.  Returns hipGetLastError().
template <typename... Args, typename F = void (*)(Args...)>
inline hipError_t launch_kernel(F kernel, dim3 grid, dim3 block, std::uint32_t lds,
                                hipStream_t st, Args... args) {
  const LaunchEvents ev = take_launch_events();
  if (ev.start || ev.stop)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, st, ev.start, ev.stop, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
  return hipGetLastError();
}

// Cache policy of the streaming accesses (0 non-temporal, 1 ordinary, 2
// non-temporal coefficient arrays with ordinary d and u; the kernels'
// kStream / kCached / kMixed) for a solve, or a chain of solves such as an
// ADI step, whose arrays span `bytes`, of which `rhs_bytes` are d and u:
// ordinary accesses when all of them fit in about 1.5x MI355X's 256 MB
// last-level cache (kMallFitBytes), so that each kernel finds much of what
// the previous one (or call) wrote there; else the mixed policy when d and u
// fit (the coefficients stream past the cache and d stays in it);
// non-temporal past that.  tridForceCachePolicy overrides the rule.
constexpr size_t kMallFitBytes = 384ull << 20;
int cache_policy(size_t bytes, size_t rhs_bytes = ~size_t(0));
// The policy tridForceCachePolicy fixed, -1 when none is.
int forced_cache_policy();

// Work-groups of `threads` threads and `lds` bytes of dynamic LDS of
// `kernel` that the current device holds at once (occupancy x compute
// units; at least 1), cached per device and configuration.
unsigned resident_blocks(const void *kernel, int threads, size_t lds);

// True while tridForceWideIndex(1) is in effect: the kernels then use 64-bit
// element offsets and 64-bit system-index arithmetic whatever the sizes.
bool wide_index();

struct ForcedConfig {
  int kind = 0, rows = 0, cps = 0, spb = 0, stage = 1;
};
ForcedConfig forced_config();

}  // namespace tridhip
