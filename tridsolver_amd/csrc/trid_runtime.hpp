// trid_runtime.hpp -- host-side runtime shared by the tridsolver_amd entry
// points: current stream, per-thread error text, per-device workspace cache.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
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

// Grow-only device workspace of the current device (for the fallback
// kernel and autotuning scratch).  Returns nullptr on allocation failure.
void *workspace(size_t bytes);

int autotune_mode();

// True while tridForceWideIndex(1) is in effect: the kernels then use 64-bit
// element offsets and 64-bit system-index arithmetic whatever the sizes.
bool wide_index();

struct ForcedConfig {
  int kind = 0, rows = 0, cps = 0, spb = 0, stage = 1;
};
ForcedConfig forced_config();

}  // namespace tridhip
