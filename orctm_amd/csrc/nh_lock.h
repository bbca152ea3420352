// nh_lock.h -- the process-wide lock shared by every graph capture and by every entry point that
// allocates, frees, copies on the null stream or synchronises the device (defined in nh_api.hip).
// The runtime fails such a call made on one thread while another thread captures a graph
// (hipErrorStreamCaptureImplicit), even in thread-local capture mode.
#pragma once

#include <mutex>

namespace onh {
std::recursive_mutex& device_lock();
}  // namespace onh
