// Texture pre-pass on the device: image -> 4x4 block gather and box-filtered
// mip levels.  The reference's harness does the gather on the CPU
// (icbc_test.cpp:353-380, texel j = 4*y + x of the block, channels read as
// c / 255.0f at 401-405).  Everything here is HBM-bound byte movement: one
// thread per output texel, consecutive threads on consecutive outputs.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "texture_prepass.h"

namespace icbc_gpu {

namespace {

// c / 255.0f, correctly rounded like the host's division.
__device__ __forceinline__ float unit8(uint32_t c) { return (float)c / 255.0f; }

__device__ __forceinline__ float4 load_texel(const uint8_t* data, uint32_t format, uint64_t pitch, uint32_t x,
                                             uint32_t y) {
    const uint8_t* row = data + (uint64_t)y * pitch;
    if (format == kImageRGBA8) {
        const uint32_t v = reinterpret_cast<const uint32_t*>(row)[x];
        return make_float4(unit8(v & 0xffu), unit8((v >> 8) & 0xffu), unit8((v >> 16) & 0xffu), unit8(v >> 24));
    }
    return reinterpret_cast<const float4*>(row)[x];
}

__global__ void image_gather_kernel(ImageDev img, uint32_t blocks_x, uint64_t texels, float4* __restrict__ colors,
                                    float* __restrict__ weights) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= texels) return;
    const uint64_t b = t >> 4;
    const uint32_t j = (uint32_t)t & 15u;
    const uint64_t by = b / blocks_x;
    const uint64_t bx = b - by * blocks_x;
    // replicate mode clamps into the image; round-down mode never leaves it
    const uint32_t x = (uint32_t)min<uint64_t>(4 * bx + (j & 3u), img.width - 1);
    const uint32_t y = (uint32_t)min<uint64_t>(4 * by + (j >> 2), img.height - 1);
    const float4 v = load_texel(img.data, img.format, img.pitch, x, y);
    colors[t] = v;
    if (weights) weights[t] = v.w;
}

// One level of the box-filtered chain (float RGBA out, tightly packed).
__global__ void downsample_kernel(ImageDev src, uint32_t dw, uint32_t dh, float4* __restrict__ dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)dw * dh) return;
    const uint32_t y = (uint32_t)(t / dw);
    const uint32_t x = (uint32_t)(t - (uint64_t)y * dw);
    float4 o;
    if (src.width == 1 && src.height == 1) {
        o = load_texel(src.data, src.format, src.pitch, 0, 0);
    } else if (src.height == 1) {
        const float4 a = load_texel(src.data, src.format, src.pitch, 2 * x, 0);
        const float4 b = load_texel(src.data, src.format, src.pitch, 2 * x + 1, 0);
        o = make_float4((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, (a.w + b.w) * 0.5f);
    } else if (src.width == 1) {
        const float4 a = load_texel(src.data, src.format, src.pitch, 0, 2 * y);
        const float4 b = load_texel(src.data, src.format, src.pitch, 0, 2 * y + 1);
        o = make_float4((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, (a.w + b.w) * 0.5f);
    } else {
        const float4 a = load_texel(src.data, src.format, src.pitch, 2 * x, 2 * y);
        const float4 b = load_texel(src.data, src.format, src.pitch, 2 * x + 1, 2 * y);
        const float4 c = load_texel(src.data, src.format, src.pitch, 2 * x, 2 * y + 1);
        const float4 d = load_texel(src.data, src.format, src.pitch, 2 * x + 1, 2 * y + 1);
        o = make_float4(((a.x + b.x) + (c.x + d.x)) * 0.25f, ((a.y + b.y) + (c.y + d.y)) * 0.25f,
                        ((a.z + b.z) + (c.z + d.z)) * 0.25f, ((a.w + b.w) + (c.w + d.w)) * 0.25f);
    }
    dst[t] = o;
}

constexpr uint32_t kThreads = 256;

uint32_t groups(uint64_t n) { return (uint32_t)((n + kThreads - 1) / kThreads); }

}  // namespace

uint32_t image_blocks_x(const ImageDev& img) {
    return img.edge == kEdgeRoundDown ? img.width / 4 : (img.width + 3) / 4;
}

uint64_t image_block_count(const ImageDev& img) {
    const uint64_t by = img.edge == kEdgeRoundDown ? img.height / 4 : (img.height + 3) / 4;
    return (uint64_t)image_blocks_x(img) * by;
}

void downsampled_size(uint32_t w, uint32_t h, uint32_t& dw, uint32_t& dh) {
    if (w == 1 && h == 1) {
        dw = dh = 1;
    } else if (h == 1) {
        dw = w / 2;
        dh = 1;
    } else if (w == 1) {
        dw = 1;
        dh = h / 2;
    } else {
        dw = w / 2;
        dh = h / 2;
    }
}

hipError_t launch_image_gather(const ImageDev& img, float4* colors, float* weights, hipStream_t stream) {
    const uint64_t blocks = image_block_count(img);
    if (blocks == 0) return hipSuccess;
    const uint64_t texels = blocks * 16;
    hipLaunchKernelGGL(image_gather_kernel, dim3(groups(texels)), dim3(kThreads), 0, stream, img,
                       image_blocks_x(img), texels, colors, weights);
    return hipGetLastError();
}

hipError_t launch_downsample(const ImageDev& src, float4* dst, hipStream_t stream) {
    uint32_t dw, dh;
    downsampled_size(src.width, src.height, dw, dh);
    const uint64_t n = (uint64_t)dw * dh;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(downsample_kernel, dim3(groups(n)), dim3(kThreads), 0, stream, src, dw, dh, dst);
    return hipGetLastError();
}

}  // namespace icbc_gpu
