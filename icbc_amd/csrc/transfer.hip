// Segment-list copies between device memory and page-locked host memory.
//
// The request server moves many buffers per launch: each request's texels,
// weights and prior codes in, its codes, errors and counters out.  A DMA copy
// per buffer costs a few microseconds of engine time however small it is
// (16 KiB copies move at 6.7 GB/s on MI355X, 128 KiB ones at 11 GB/s), so a
// launch of a hundred small requests spent most of its time in copy
// commands.  Here one kernel moves a whole list: page-locked host memory is
// mapped into the device's address space, so the kernel reads and writes it
// directly over PCIe (54-57 GB/s measured for one large copy either way,
// the DMA engine's large-copy rate).
//
// Work split: the list's bytes are cut into 4 KiB chunks numbered across the
// segments (CopySeg::first_chunk, computed on the host); workgroup g copies a
// contiguous run of chunks, finds its first segment once by binary search and
// walks forward from there.  A chunk is 256 lanes x 16 bytes when source and
// destination are 16-byte aligned, otherwise 256 lanes x 16 single bytes.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "transfer.h"

namespace icbc_gpu {

namespace {

constexpr uint32_t kCopyThreads = 256;
static_assert(kCopyThreads * 16 == kCopyChunkBytes, "a chunk is one 16-byte access per lane");

__global__ __launch_bounds__(kCopyThreads) void copy_segments_kernel(const CopySeg* __restrict__ segs,
                                                                     uint32_t nsegs, uint64_t nchunks) {
    const uint64_t per = (nchunks + gridDim.x - 1) / gridDim.x;
    uint64_t c = (uint64_t)blockIdx.x * per;
    const uint64_t end = min<uint64_t>(c + per, nchunks);
    if (c >= end) return;
    // last segment whose first chunk is at or before c (segments with no
    // bytes own no chunk and are skipped by the walk below)
    uint32_t lo = 0, hi = nsegs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (segs[mid].first_chunk <= c) lo = mid;
        else hi = mid - 1;
    }
    uint32_t k = lo;
    CopySeg s = segs[k];
    for (; c < end; c++) {
        while (c >= s.first_chunk + (s.bytes + kCopyChunkBytes - 1) / kCopyChunkBytes) s = segs[++k];
        const uint64_t off = (c - s.first_chunk) * kCopyChunkBytes;
        const uint64_t len = min<uint64_t>(kCopyChunkBytes, s.bytes - off);
        const uint8_t* src = static_cast<const uint8_t*>(s.src) + off;
        uint8_t* dst = static_cast<uint8_t*>(s.dst) + off;
        const uint32_t t = threadIdx.x;
        if ((((uintptr_t)src | (uintptr_t)dst) & 15u) == 0) {
            const uint32_t nvec = (uint32_t)(len >> 4);
            if (t < nvec) reinterpret_cast<uint4*>(dst)[t] = reinterpret_cast<const uint4*>(src)[t];
            const uint32_t tail = (uint32_t)len - (nvec << 4);
            if (t < tail) dst[(nvec << 4) + t] = src[(nvec << 4) + t];
        } else {
            for (uint32_t b = t; b < len; b += kCopyThreads) dst[b] = src[b];
        }
    }
}

}  // namespace

uint64_t number_copy_segments(CopySeg* segs, uint32_t nsegs) {
    uint64_t chunk = 0;
    for (uint32_t k = 0; k < nsegs; k++) {
        segs[k].first_chunk = chunk;
        chunk += (segs[k].bytes + kCopyChunkBytes - 1) / kCopyChunkBytes;
    }
    return chunk;
}

hipError_t launch_copy_segments(const CopySeg* segs, uint32_t nsegs, uint64_t nchunks, uint32_t grid,
                                hipStream_t stream) {
    if (nsegs == 0 || nchunks == 0) return hipSuccess;
    const uint32_t g = (uint32_t)(nchunks < grid ? nchunks : grid);
    hipLaunchKernelGGL(copy_segments_kernel, dim3(g), dim3(kCopyThreads), 0, stream, segs, nsegs, nchunks);
    return hipGetLastError();
}

}  // namespace icbc_gpu
