// Device texture pre-pass (texture_prepass.hip): image -> block gather and
// box-filtered mip levels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icbc_gpu {

constexpr uint32_t kImageRGBA8 = 0;
constexpr uint32_t kImageRGBA32F = 1;
constexpr uint32_t kEdgeReplicate = 0;
constexpr uint32_t kEdgeRoundDown = 1;

// Validated image descriptor as the kernels take it.
struct ImageDev {
    const uint8_t* data;
    uint64_t pitch;
    uint32_t format, width, height, edge;
};

uint32_t image_blocks_x(const ImageDev& img);
uint64_t image_block_count(const ImageDev& img);
// Size of the next level down (a 1-texel dimension stays 1; 1x1 stays 1x1).
void downsampled_size(uint32_t w, uint32_t h, uint32_t& dw, uint32_t& dh);

// Block-major float RGBA (16 texels per block) and, when `weights` is set,
// each texel's alpha as its weight.
hipError_t launch_image_gather(const ImageDev& img, float4* colors, float* weights, hipStream_t stream);
// One box-filtered level of `src` into `dst` (float RGBA, tightly packed).
hipError_t launch_downsample(const ImageDev& src, float4* dst, hipStream_t stream);

}  // namespace icbc_gpu
