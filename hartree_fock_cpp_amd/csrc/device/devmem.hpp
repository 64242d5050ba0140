// devmem.hpp -- block cache for the device and pinned host memory of per-job
// objects (bases, Fock contexts, device SCF state, one-electron scratch).
//
// hipFree and hipHostFree synchronize the whole device.  When several SCF jobs
// run side by side (the server), each job's teardown would then wait for the
// work of every other job in flight.  Blocks of up to BLOCK_MAX bytes are
// instead rounded up to a power of two and kept for reuse by later
// allocations of this process; larger blocks go straight to hipMalloc /
// hipFree.  A block must not be released while device work that uses it is
// still pending: callers synchronize the streams that used it first.
#pragma once
#include <cstddef>

#include <hip/hip_runtime.h>

namespace hfx_detail {

// Device memory on the current device.
hipError_t dev_alloc(void** p, size_t bytes);
void dev_free(void* p);
// Page-locked host memory.
hipError_t host_alloc(void** p, size_t bytes);
void host_free(void* p);

}  // namespace hfx_detail
