// DDS and KTX containers for BC1 textures with mip levels (host side).
//
// The reference's harness writes single-level files (icbc_test.cpp:155-211
// DDS, 233-279 KTX) and compares block data against a DDS file for its
// regression mode (icbc_test.cpp:213-231).  The same header fields are
// written here, little-endian, with the mip-level fields of both formats
// filled in when a file holds more than one level.

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/icbc_gpu.h"
#include "icbc_internal.h"

namespace {

uint32_t fourcc(const char* s) {
    return (uint32_t)(uint8_t)s[0] | ((uint32_t)(uint8_t)s[1] << 8) | ((uint32_t)(uint8_t)s[2] << 16) |
           ((uint32_t)(uint8_t)s[3] << 24);
}

uint32_t mip_dim(uint32_t d, uint32_t k) {
    const uint32_t v = k < 32 ? d >> k : 0;
    return v ? v : 1;
}

uint64_t level_bytes(uint32_t w, uint32_t h) { return 8ull * ((w + 3) / 4) * ((h + 3) / 4); }

uint64_t data_bytes(uint32_t w, uint32_t h, uint32_t levels) {
    uint64_t n = 0;
    for (uint32_t k = 0; k < levels; k++) n += level_bytes(mip_dim(w, k), mip_dim(h, k));
    return n;
}

constexpr uint64_t kDdsHeader = 128;
constexpr uint64_t kKtxHeader = 64;

void put32(std::vector<uint8_t>& b, size_t at, uint32_t v) {
    b[at] = (uint8_t)v;
    b[at + 1] = (uint8_t)(v >> 8);
    b[at + 2] = (uint8_t)(v >> 16);
    b[at + 3] = (uint8_t)(v >> 24);
}

int check_args(int kind, uint32_t width, uint32_t height, uint32_t levels) {
    if (kind != ICBC_CONTAINER_DDS && kind != ICBC_CONTAINER_KTX)
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "unknown container kind %d", kind);
    if (width == 0 || height == 0 || levels == 0 || levels > 32)
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "empty texture or bad level count");
    return ICBC_OK;
}

// Whole file in memory.
std::vector<uint8_t> serialise(int kind, uint32_t w, uint32_t h, uint32_t levels, const uint8_t* codes) {
    const uint64_t data = data_bytes(w, h, levels);
    std::vector<uint8_t> b;
    if (kind == ICBC_CONTAINER_DDS) {
        // DDS_HEADER after the magic (icbc_test.cpp:157-196)
        const uint32_t DDSD_CAPS = 0x1, DDSD_HEIGHT = 0x2, DDSD_WIDTH = 0x4, DDSD_PIXELFORMAT = 0x1000,
                       DDSD_MIPMAPCOUNT = 0x20000, DDSD_LINEARSIZE = 0x80000;
        const uint32_t DDPF_FOURCC = 0x4;
        const uint32_t DDSCAPS_COMPLEX = 0x8, DDSCAPS_TEXTURE = 0x1000, DDSCAPS_MIPMAP = 0x400000;
        b.assign(kDdsHeader + data, 0);
        put32(b, 0, fourcc("DDS "));
        put32(b, 4, 124);
        put32(b, 8, DDSD_CAPS | DDSD_PIXELFORMAT | DDSD_WIDTH | DDSD_HEIGHT | DDSD_LINEARSIZE |
                        (levels > 1 ? DDSD_MIPMAPCOUNT : 0));
        put32(b, 12, h);
        put32(b, 16, w);
        put32(b, 20, (uint32_t)level_bytes(w, h));  // linear size of the top level
        put32(b, 24, 0);                             // depth
        put32(b, 28, levels);                        // mipmap count (1 in the reference)
        // 11 reserved words at 32..75
        put32(b, 76, 32);  // pixel format size
        put32(b, 80, DDPF_FOURCC);
        put32(b, 84, fourcc("DXT1"));
        // bit count and masks 88..107 stay 0
        put32(b, 108, DDSCAPS_TEXTURE | (levels > 1 ? DDSCAPS_COMPLEX | DDSCAPS_MIPMAP : 0));
        // caps2..4 and the unused word stay 0
        memcpy(b.data() + kDdsHeader, codes, data);
        return b;
    }
    // KTX 1.1 (icbc_test.cpp:233-279): identifier, endianness, GL fields, then
    // per level imageSize + data (BC1 levels are multiples of 4 bytes, so no
    // mip padding is needed)
    const uint32_t GL_COMPRESSED_RGB_S3TC_DXT1_EXT = 0x83F0, GL_RGBA = 0x1908;
    static const uint8_t ident[12] = {0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31, 0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A};
    b.assign(kKtxHeader + 4ull * levels + data, 0);
    memcpy(b.data(), ident, 12);
    put32(b, 12, 0x04030201);
    put32(b, 16, 0);  // glType
    put32(b, 20, 1);  // glTypeSize
    put32(b, 24, 0);  // glFormat
    put32(b, 28, GL_COMPRESSED_RGB_S3TC_DXT1_EXT);
    put32(b, 32, GL_RGBA);
    put32(b, 36, w);
    put32(b, 40, h);
    put32(b, 44, 0);  // pixelDepth
    put32(b, 48, 0);  // numberOfArrayElements
    put32(b, 52, 1);  // numberOfFaces
    put32(b, 56, levels);
    put32(b, 60, 0);  // bytesOfKeyValueData
    size_t at = kKtxHeader;
    const uint8_t* src = codes;
    for (uint32_t k = 0; k < levels; k++) {
        const uint64_t lb = level_bytes(mip_dim(w, k), mip_dim(h, k));
        put32(b, at, (uint32_t)lb);
        memcpy(b.data() + at + 4, src, lb);
        at += 4 + lb;
        src += lb;
    }
    return b;
}

}  // namespace

extern "C" {

uint64_t icbc_container_size(int kind, uint32_t width, uint32_t height, uint32_t levels) {
    if (check_args(kind, width, height, levels)) return 0;
    const uint64_t data = data_bytes(width, height, levels);
    return kind == ICBC_CONTAINER_DDS ? kDdsHeader + data : kKtxHeader + 4ull * levels + data;
}

uint64_t icbc_container_data_bytes(uint32_t width, uint32_t height, uint32_t levels) {
    if (width == 0 || height == 0 || levels == 0 || levels > 32) return 0;
    return data_bytes(width, height, levels);
}

int icbc_container_encode(int kind, uint32_t width, uint32_t height, uint32_t levels, const void* codes, void* out,
                          uint64_t out_bytes) {
    int r = check_args(kind, width, height, levels);
    if (r) return r;
    if (!codes || !out) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null codes or output");
    const std::vector<uint8_t> b = serialise(kind, width, height, levels, (const uint8_t*)codes);
    if (out_bytes < b.size())
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "output holds %llu bytes, file needs %zu",
                                  (unsigned long long)out_bytes, b.size());
    memcpy(out, b.data(), b.size());
    return ICBC_OK;
}

int icbc_container_write(int kind, const char* path, uint32_t width, uint32_t height, uint32_t levels,
                         const void* codes) {
    int r = check_args(kind, width, height, levels);
    if (r) return r;
    if (!path || !codes) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null path or codes");
    const std::vector<uint8_t> b = serialise(kind, width, height, levels, (const uint8_t*)codes);
    FILE* fp = fopen(path, "wb");
    if (!fp) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "cannot open %s for writing", path);
    const size_t n = fwrite(b.data(), 1, b.size(), fp);
    const int closed = fclose(fp);
    if (n != b.size() || closed != 0) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "short write to %s", path);
    return ICBC_OK;
}

int icbc_container_compare_dds(const char* path, uint32_t width, uint32_t height, const void* codes) {
    if (!path || !codes || width == 0 || height == 0)
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null path/codes or empty texture");
    const uint64_t n = level_bytes(width, height);
    FILE* fp = fopen(path, "rb");
    if (!fp) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "cannot open %s", path);
    std::vector<uint8_t> file(n);
    const bool ok = fseek(fp, (long)kDdsHeader, SEEK_SET) == 0 && fread(file.data(), 1, n, fp) == n;
    fclose(fp);
    if (!ok) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "%s is shorter than the top level", path);
    return memcmp(file.data(), codes, n) == 0 ? 1 : 0;
}

}  // extern "C"
