// Bulk copies between caller-owned (pageable) host arrays and device memory
// through the library's pinned staging pool (ps2_stage2.hip): the DMA of one
// 64 MB chunk overlaps a multi-threaded host copy of the next.  Both return
// after the data has arrived; errors as the C ABI (PS2_E_*, set_error).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ps2 {

int staging_h2d(void *dst_device, const void *src_host, size_t bytes, hipStream_t stream);
int staging_d2h(void *dst_host, const void *src_device, size_t bytes, hipStream_t stream);

}  // namespace ps2
