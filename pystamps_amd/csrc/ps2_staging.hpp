// Runtime services shared by the library's translation units.
//
// Bulk copies between caller-owned (pageable) host arrays and device memory
// through the library's pinned staging pool (ps2_stage2.hip): the DMA of one
// 64 MB chunk overlaps a multi-threaded host copy of the next.  Both return
// after the data has arrived; errors as the C ABI (PS2_E_*, set_error).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include <shared_mutex>

namespace ps2 {

int staging_h2d(void *dst_device, const void *src_host, size_t bytes, hipStream_t stream);
int staging_d2h(void *dst_host, const void *src_device, size_t bytes, hipStream_t stream);

// Graph capture and device allocation exclude each other process-wide: a
// session's one-time capture of its iteration takes this lock exclusively;
// every device or pinned allocation and free in the library takes it shared,
// so no allocation runs while any stream of the process is being captured.
std::shared_mutex &capture_mutex();
using AllocGuard = std::shared_lock<std::shared_mutex>;

}  // namespace ps2
