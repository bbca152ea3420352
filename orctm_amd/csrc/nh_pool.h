// nh_pool.h -- the library's device-memory pool.
//
// Handles, solver contexts and serving requests take their device memory from here instead of from
// hipMalloc/hipFree per object. A freed block returns to the pool, never to the driver, so creating
// and destroying handles while other work runs issues no hipFree (which waits for the whole device)
// and no hipMalloc after the pool has grown to the working set. Large blocks (1 MiB or more) start on
// a 2 MiB boundary of their own, so every array of a handle has the same placement relative to pages
// and memory channels whatever was allocated and freed before; small blocks (Sys arrays, tile lists,
// reduction state) are 256-byte aligned in shared slabs.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace onh {

// A block of at least `bytes` on the current device (hipSetDevice done by the caller). Contents are
// undefined: callers clear what must start at zero, on their own stream.
hipError_t pool_alloc(int device, size_t bytes, void** p);
// Returns a block obtained from pool_alloc (nullptr is ignored). The caller guarantees that no work
// still queued on the device uses it.
void pool_free(int device, void* p);
// The same for small pinned host blocks (hipHostMalloc slabs), for device-to-host status reads.
hipError_t pool_alloc_host(size_t bytes, void** p);
void pool_free_host(void* p);
// Bytes held by the pool on `device` that no live block uses (free for new blocks without growth).
size_t pool_idle_bytes(int device);
// Bytes of device memory the pool holds on `device` (live and idle).
size_t pool_held_bytes(int device);

}  // namespace onh
