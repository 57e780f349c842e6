// Test-only HIP runtime stand-in (hipstub.cpp): interface for the CPU kernels (kernels_cpu.cpp) and the
// harness driver (driver.cpp).
#pragma once

#include <functional>

#include "hip/hip_runtime_api.h"

namespace hipstub {
// Enqueues `fn` on `stream` (executed in order by the stream's worker thread), or appends it to the graph
// being captured on that stream.
void enqueue(hipStream_t stream, std::function<void()> fn);
}  // namespace hipstub

extern "C" {
// API-usage violations recorded so far (unsafe calls during a capture, device-wide synchronisation while
// another thread captures, copies touching freed or unknown memory, unjoined captures, ...)
int hipstub_violation_count(void);
const char* hipstub_violation(int i);
// number of captures that were active while another thread synchronised the device, and the largest
// number of captures that were active at once
int hipstub_max_concurrent_captures(void);
// bytes of device memory allocated and not freed
long long hipstub_live_device_bytes(void);
}
