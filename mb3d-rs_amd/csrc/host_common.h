// host_common.h -- host-side helpers shared by the backend's sources: the
// thread-local last error (as capi/src/lib.rs:18-26) and the capability
// check (gpu::unsupported, src/gpu.rs:445-477).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/mb3d_hip.h"

namespace mb3d {

// The library's one device context (mb3d_hip_init): its lock, stream (null
// before init) and compute-unit count.
std::mutex &lib_mutex();
hipStream_t lib_stream();
int lib_cus();
int lib_device();  // -1 before init

// This thread's work counters (mb3d_hip_counters_get).
mb3d_hip_counters &thread_counters();

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const char *msg);

// Why this backend cannot march `p` ("" when it can).
std::string unsupported_reason(const mb3d_hip_params *p);

}  // namespace mb3d
