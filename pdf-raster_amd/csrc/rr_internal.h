// rr_internal.h -- context accessors shared by the translation units of the
// HIP backend (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "rr_hip.h"

namespace rrint {
hipStream_t stream(rr_hip_ctx *c);
int device(rr_hip_ctx *c);
bool page_open(rr_hip_ctx *c);
int32_t fail(rr_hip_ctx *c, int32_t code, const std::string &msg);
// Start/stop event pair for the next kernel when per-launch profiling is on,
// nullptr otherwise; *ok is false when the pool could not grow.
hipEvent_t *prof_pair(rr_hip_ctx *c, bool *ok);
void count_launch(rr_hip_ctx *c);
}  // namespace rrint
