// Segment-list copies (transfer.hip): one kernel moves many buffers between
// device memory and mapped page-locked host memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icbc_gpu {

constexpr uint32_t kCopyChunkBytes = 4096;

// One buffer to move.  src and dst are device-accessible addresses (device
// memory, or page-locked host memory through its mapped address).
struct CopySeg {
    const void* src;
    void* dst;
    uint64_t bytes;
    uint64_t first_chunk;  // set by number_copy_segments
};

// Numbers the chunks of segs[0..n) in order; returns the total chunk count.
uint64_t number_copy_segments(CopySeg* segs, uint32_t nsegs);

// Copies every segment of `segs` (device memory, numbered) with at most
// `grid` workgroups.
hipError_t launch_copy_segments(const CopySeg* segs, uint32_t nsegs, uint64_t nchunks, uint32_t grid,
                                hipStream_t stream);

}  // namespace icbc_gpu
