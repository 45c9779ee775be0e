// Device-side building blocks of the BC1 encoder (CDNA4 / gfx950).
//
// Every function reproduces the arithmetic of the reference encoder's scalar
// build (castano/icbc v1.05, icbc.h with ICBC_SIMD=0) operation by operation,
// so that, compiled with -ffp-contract=off and IEEE fp32 division and
// denormals, the 8-byte block codes and float errors are bit-identical:
//   max(a,b) = (b < a) ? a : b, min(a,b) = (a < b) ? a : b   (icbc.h:207-213)
//   round01(a) = float(int(a + 0.5f))                        (icbc.h:417)
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z                 (icbc.h:305, 1533)
// Sums over texels and colours run left to right, as in the reference.
//
// The functions here are per-lane ("one lane owns one 4x4 block"); the
// wave-cooperative candidate search lives in bc1_encoder.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace icbc_gpu {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ float max_r(float a, float b) { return (b < a) ? a : b; }
__device__ __forceinline__ float min_r(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float sat01(float a) { return min_r(max_r(a, 0.0f), 1.0f); }
__device__ __forceinline__ f3 sat3(f3 a) { return mk(sat01(a.x), sat01(a.y), sat01(a.z)); }
__device__ __forceinline__ float round01(float a) { return (float)(int)(a + 0.5f); }

// ---------------------------------------------------------------------------
// Tables written once by the host at init (icbc.h:1560-1570, 1947-1950,
// 3374-3375, 3495-3515) -- see upload_tables() in bc1_encoder.hip.

struct DecoderWeights {
    float w1_3, w2_3, w1_9, w4_9, w2_9;
};

extern __constant__ float k_mid5[32];          // 565 rounding midpoints
extern __constant__ float k_mid6[64];
extern __constant__ uint32_t k_four[976];      // c0 | c1 << 8 | c2 << 16 | pair(c0, c1) << 24
extern __constant__ uint32_t k_three[160];     // c0 | c1 << 8 | (pair p's c0 | c1 << 8) << 16
extern __constant__ int k_four_total[16];
extern __constant__ int k_three_total[16];
extern __constant__ uint16_t k_match5[4][256]; // max | min << 8
extern __constant__ uint16_t k_match6[4][256];
extern __constant__ DecoderWeights k_dweights[4];

// Encoded block: col0 | col1 << 16, indices.
struct Code {
    uint32_t cols;
    uint32_t idx;
};

__device__ __forceinline__ uint32_t col0(Code c) { return c.cols & 0xffffu; }
__device__ __forceinline__ uint32_t col1(Code c) { return c.cols >> 16; }
__device__ __forceinline__ Code make_code(uint32_t c0, uint32_t c1, uint32_t idx) {
    return Code{c0 | (c1 << 16), idx};
}

// ---------------------------------------------------------------------------
// 565 conversion (icbc.h:1588-1667)

// round01 of a saturated value: x*31 + 0.5 lies in [0.5, 31.5] (x*63 + 0.5 in
// [0.5, 63.5]), where float(int(y)) equals truncf(y); v_trunc_f32 does it in
// one instruction instead of a convert pair.
__device__ __forceinline__ float q5(float x) { return truncf(x * 31.0f + 0.5f) * (1.0f / 31.0f); }
__device__ __forceinline__ float q6(float x) { return truncf(x * 63.0f + 0.5f) * (1.0f / 63.0f); }
__device__ __forceinline__ f3 quantize_endpoint(f3 v) {
    return mk(q5(sat01(v.x)), q6(sat01(v.y)), q5(sat01(v.z)));
}

__device__ __forceinline__ uint32_t to565(f3 v) {
    uint32_t r = (uint32_t)min_r(max_r(v.x * 31.0f, 0.0f), 31.0f);
    uint32_t g = (uint32_t)min_r(max_r(v.y * 63.0f, 0.0f), 63.0f);
    uint32_t b = (uint32_t)min_r(max_r(v.z * 31.0f, 0.0f), 31.0f);
    r += (v.x > k_mid5[r]) ? 1u : 0u;
    g += (v.y > k_mid6[g]) ? 1u : 0u;
    b += (v.z > k_mid5[b]) ? 1u : 0u;
    return (r << 11) | (g << 5) | b;
}

__device__ __forceinline__ uint32_t to_u8(float v) { return (uint32_t)(uint8_t)(sat01(v) * 255 + 0.5f); }

// ---------------------------------------------------------------------------
// Palettes (icbc.h:2751-2900).  Channels are kept as ints 0..255.

struct Pal8 {
    int r[4], g[4], b[4];
};

__device__ __forceinline__ void expand565(uint32_t c, int& r, int& g, int& b) {
    int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    r = (r5 << 3) | (r5 >> 2);
    g = (g6 << 2) | (g6 >> 4);
    b = (b5 << 3) | (b5 >> 2);
}

__device__ __forceinline__ int interp_ch(int decoder, bool four, int p0, int p1) {
    // third palette entry of one channel for D3D10 / AMD / Intel
    if (four) {
        if (decoder == 2) return (uint8_t)((43 * p0 + 21 * p1 + 32) >> 6);
        if (decoder == 3) return (uint8_t)((171 * p0 + 85 * p1 + 128) >> 8);
        return (2 * p0 + p1) / 3;
    }
    if (decoder == 2) return (uint8_t)((p0 + p1 + 1) / 2);
    if (decoder == 3) return (uint8_t)((128 * p0 + 128 * p1 + 128) >> 8);
    return (p0 + p1) / 2;
}

// The four entries of a palette as named scalars (no array, so the compiler
// keeps them in registers on every decoder path).
struct PalInts {
    int r0, g0, b0, r1, g1, b1, r2, g2, b2, r3, g3, b3;
};

// With `force_four` the four-colour formulas apply whatever the endpoint order
// (evaluate_palette4, icbc.h:2883-2891, which BC3's colour block uses).
__device__ __forceinline__ PalInts palette_ints(uint32_t c0, uint32_t c1, int decoder, bool force_four = false) {
    PalInts p;
    expand565(c0, p.r0, p.g0, p.b0);
    expand565(c1, p.r1, p.g1, p.b1);
    const bool four = force_four || c0 > c1;
    p.r3 = p.g3 = p.b3 = 0;
    if (decoder == 1) {
        // NVIDIA: red and blue interpolate the 5-bit values, green the 8-bit ones.
        const int r0 = (c0 >> 11) & 31, b0 = c0 & 31, r1 = (c1 >> 11) & 31, b1 = c1 & 31;
        const int gd = p.g1 - p.g0;
        if (four) {
            p.r2 = (uint8_t)(((2 * r0 + r1) * 22) / 8);
            p.g2 = (uint8_t)((256 * p.g0 + gd / 4 + 128 + gd * 80) / 256);
            p.b2 = (uint8_t)(((2 * b0 + b1) * 22) / 8);
            p.r3 = (uint8_t)(((2 * r1 + r0) * 22) / 8);
            p.g3 = (uint8_t)((256 * p.g1 - gd / 4 + 128 - gd * 80) / 256);
            p.b3 = (uint8_t)(((2 * b1 + b0) * 22) / 8);
        } else {
            p.r2 = (uint8_t)(((r0 + r1) * 33) / 8);
            p.g2 = (uint8_t)((256 * p.g0 + gd / 4 + 128 + gd * 128) / 256);
            p.b2 = (uint8_t)(((b0 + b1) * 33) / 8);
        }
        return p;
    }
    p.r2 = interp_ch(decoder, four, p.r0, p.r1);
    p.g2 = interp_ch(decoder, four, p.g0, p.g1);
    p.b2 = interp_ch(decoder, four, p.b0, p.b1);
    if (four) {
        p.r3 = interp_ch(decoder, true, p.r1, p.r0);
        p.g3 = interp_ch(decoder, true, p.g1, p.g0);
        p.b3 = interp_ch(decoder, true, p.b1, p.b0);
    }
    return p;
}

// Array form for the decode / evaluate kernels, which index entries per texel.
__device__ __forceinline__ void palette8(uint32_t c0, uint32_t c1, int decoder, Pal8& p, bool force_four = false) {
    const PalInts q = palette_ints(c0, c1, decoder, force_four);
    p.r[0] = q.r0; p.g[0] = q.g0; p.b[0] = q.b0;
    p.r[1] = q.r1; p.g[1] = q.g1; p.b[1] = q.b1;
    p.r[2] = q.r2; p.g[2] = q.g2; p.b[2] = q.b2;
    p.r[3] = q.r3; p.g[3] = q.g3; p.b[3] = q.b3;
}

// k / 255.0f for k in 0..255, correctly rounded (the value of the IEEE
// division): the product with the rounded reciprocal, corrected by one
// fused multiply-add residual step.  The residual k - q * 255 is exact, and
// the corrected quotient equals the division for every k (checked for all 256
// values in exact rational arithmetic by the CPU tests).
__device__ __forceinline__ float u8_to_unit(uint32_t k) {
    constexpr float kRecip = 1.0f / 255.0f;
    const float x = (float)k;
    const float q = x * kRecip;
    return __builtin_fmaf(__builtin_fmaf(-q, 255.0f, x), kRecip, q);
}

// palette as floats: channel / 255.0f (icbc.h:1656, 2902-2909).
__device__ __forceinline__ void palette_f(uint32_t c0, uint32_t c1, int decoder, f3 pal[4]) {
    const PalInts p = palette_ints(c0, c1, decoder);
    pal[0] = mk(u8_to_unit(p.r0), u8_to_unit(p.g0), u8_to_unit(p.b0));
    pal[1] = mk(u8_to_unit(p.r1), u8_to_unit(p.g1), u8_to_unit(p.b1));
    pal[2] = mk(u8_to_unit(p.r2), u8_to_unit(p.g2), u8_to_unit(p.b2));
    pal[3] = mk(u8_to_unit(p.r3), u8_to_unit(p.g3), u8_to_unit(p.b3));
}

// ---------------------------------------------------------------------------
// The 16 texels and weights of one block, read from memory on demand (L1/L2
// resident after the first pass) rather than held in registers.

// Global-memory loads and stores.  A pointer read from a job descriptor is
// generic to the compiler, and accesses through it become flat instructions,
// whose loads also wait on the LDS counter; these address global memory
// (global_* instructions, with a scalar base where the address is a wave-
// uniform pointer plus a lane offset).  Vectors move as the compiler's native
// vector types.
template <typename T>
using InGlobal = __attribute__((address_space(1))) T;
typedef float NativeF4 __attribute__((ext_vector_type(4)));
typedef uint32_t NativeU4 __attribute__((ext_vector_type(4)));
typedef uint32_t NativeU2 __attribute__((ext_vector_type(2)));
typedef int NativeI4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float ld_global(const float* p) { return *(const InGlobal<float>*)p; }
__device__ __forceinline__ uint32_t ld_global(const uint32_t* p) { return *(const InGlobal<uint32_t>*)p; }
__device__ __forceinline__ float4 ld_global(const float4* p) {
    const NativeF4 v = *(const InGlobal<NativeF4>*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
// The RGB of a float RGBA texel: a 12-byte load (alpha is never read).
typedef float NativeF3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ f3 ld_global_rgb(const float4* p) {
    const NativeF3 v = *(const InGlobal<NativeF3>*)p;
    return mk(v.x, v.y, v.z);
}
__device__ __forceinline__ uint4 ld_global(const uint4* p) {
    const NativeU4 v = *(const InGlobal<NativeU4>*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_global(const uint2* p) {
    const NativeU2 v = *(const InGlobal<NativeU2>*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void st_global(float* p, float v) { *(InGlobal<float>*)p = v; }
__device__ __forceinline__ void st_global(int* p, int v) { *(InGlobal<int>*)p = v; }
__device__ __forceinline__ void st_global(uint2* p, uint2 v) {
    const NativeU2 n = {v.x, v.y};
    *(InGlobal<NativeU2>*)p = n;
}
__device__ __forceinline__ void st_global(int4* p, int4 v) {
    const NativeI4 n = {v.x, v.y, v.z, v.w};
    *(InGlobal<NativeI4>*)p = n;
}

// A block's 16 texels in HBM: float RGBA (alpha unused), or with kRgba8 16
// RGBA8 words (R in the low byte) read as c / 255.0f, as the reference's
// harness converts 8-bit images (icbc_test.cpp:401-405); 16 weights or null
// for all 1.0.
template <bool kRgba8>
struct TexelsOf {
    const void* c;
    const float* w;
    __device__ __forceinline__ f3 color(int i) const {
        if constexpr (kRgba8) {
            const uint32_t v = ld_global(static_cast<const uint32_t*>(c) + i);
            return mk(u8_to_unit(v & 0xffu), u8_to_unit((v >> 8) & 0xffu), u8_to_unit((v >> 16) & 0xffu));
        } else {
            return ld_global_rgb(static_cast<const float4*>(c) + i);
        }
    }
    __device__ __forceinline__ float weight(int i) const { return w ? ld_global(w + i) : 1.0f; }
};
using Texels = TexelsOf<false>;

// ---------------------------------------------------------------------------
// Error of a block for a palette and index word (icbc.h:2955-2958, 3020-3028).

// Channel weights all 1.0 (kUnit, the reference harness's setting): every
// product by a weight or by the metric w * w is x * 1.0f == x, exact, so the
// kUnit forms below skip those products and give bit-identical results.
template <bool kUnit>
__device__ __forceinline__ f3 wmul(f3 x, f3 w) {
    if constexpr (kUnit) return x;
    else return x * w;
}
// dot(e, metric) (icbc.h:2742): ((e.x m.x + e.y m.y) + e.z m.z)
template <bool kUnit>
__device__ __forceinline__ float metric_dot(f3 e, f3 m) {
    if constexpr (kUnit) return (e.x + e.y) + e.z;
    else return dot3(e, m);
}

template <bool kUnit = false>
__device__ __forceinline__ float texel_error(f3 p, f3 c, f3 w) {
    f3 d = wmul<kUnit>(p - c, w) * 255.0f;
    return dot3(d, d);
}

// The block's texels held in registers, for stages that read them many
// times (box fit + least squares, the refinement walk); unweighted blocks keep
// no weights.
template <bool kWeighted>
struct TexelRegs {
    f3 c[16];
    float w[kWeighted ? 16 : 1];
    __device__ __forceinline__ TexelRegs() {}
    template <typename Tx>
    __device__ __forceinline__ explicit TexelRegs(const Tx& t) {
        load(t);
    }
    // With kWeighted, a block without weights holds 1.0 (w * x == x exactly).
    template <typename Tx>
    __device__ __forceinline__ void load(const Tx& t) {
#pragma unroll
        for (int i = 0; i < 16; i++) c[i] = t.color(i);
        if (kWeighted) {
#pragma unroll
            for (int i = 0; i < 16; i++) w[kWeighted ? i : 0] = t.weight(i);
        }
    }
    __device__ __forceinline__ f3 color(int i) const { return c[i]; }
    __device__ __forceinline__ float weight(int i) const { return kWeighted ? w[kWeighted ? i : 0] : 1.0f; }
};

// Loops over the 16 texels unroll fully for register-held texels (static
// indices keep them in registers) and partly for texels read from memory.
template <typename Src>
struct TexelUnroll {
    static constexpr int value = 4;
};
template <bool kWeighted>
struct TexelUnroll<TexelRegs<kWeighted>> {
    static constexpr int value = 16;
};

__device__ __forceinline__ f3 pick(const f3 pal[4], uint32_t k) {
    f3 lo = (k & 1) ? pal[1] : pal[0];
    f3 hi = (k & 1) ? pal[3] : pal[2];
    return (k & 2) ? hi : lo;
}

template <bool kUnit, typename Tx>
__device__ __forceinline__ float palette_error(const Tx& t, f3 cw, const f3 pal[4], uint32_t indices) {
    float e = 0.0f;
#pragma unroll 4
    for (int i = 0; i < 16; i++)
        e += t.weight(i) * texel_error<kUnit>(pick(pal, (indices >> (2 * i)) & 3), t.color(i), cw);
    return e;
}

template <bool kUnit, typename Tx>
__device__ __forceinline__ float code_error(const Tx& t, f3 cw, int decoder, Code c) {
    f3 pal[4];
    palette_f(col0(c), col1(c), decoder, pal);
    return palette_error<kUnit>(t, cw, pal, c.idx);
}

// ---------------------------------------------------------------------------
// Index selection (icbc.h:3097-3235) fused with the error of the chosen
// indices: texel i's index goes to bits 2i..2i+1 and its weighted error is
// added in texel order, exactly as selecting all indices first and then
// summing the errors would.

// compute_indices4 rule (icbc.h:3097-3131)
template <bool kUnit, typename Src>
__device__ __forceinline__ float select4(const Src& t, f3 w, const f3 pal[4], uint32_t& indices) {
    const f3 p0 = wmul<kUnit>(pal[0], w), p1 = wmul<kUnit>(pal[1], w), p2 = wmul<kUnit>(pal[2], w),
             p3 = wmul<kUnit>(pal[3], w);
    uint32_t out = 0;
    float e = 0.0f;
#pragma unroll TexelUnroll<Src>::value
    for (int i = 0; i < 16; i++) {
        const f3 x = t.color(i);
        const f3 c = wmul<kUnit>(x, w);
        const f3 e0 = c - p0, e1 = c - p1, e2 = c - p2, e3 = c - p3;
        const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(e3, e3);
        const uint32_t hi = ((d1 > d2) & (d0 > d2)) | ((d0 > d3) & (d1 > d3));
        const uint32_t lo = (d0 > d3) & (d2 > d3);
        const uint32_t k = lo | (hi << 1);
        out |= k << (2 * i);
        if constexpr (kUnit) {
            // with unit channel weights c = x, and the error's difference
            // p - x is the chosen entry's c - p negated, exactly (IEEE
            // subtraction is antisymmetric): the error squares the same values
            const f3 lo3 = (k & 1) ? e1 : e0, hi3 = (k & 1) ? e3 : e2;
            const f3 d = ((k & 2) ? hi3 : lo3) * 255.0f;
            e += t.weight(i) * dot3(d, d);
        } else {
            e += t.weight(i) * texel_error<kUnit>(pick(pal, k), x, w);
        }
    }
    indices = out;
    return e;
}

// compute_indices3 rule (icbc.h:3133-3176)
template <bool kUnit, typename Tx>
__device__ __forceinline__ float select3(const Tx& t, f3 w, bool allow_black, const f3 pal[4],
                                         uint32_t& indices) {
    const f3 p0 = wmul<kUnit>(pal[0], w), p1 = wmul<kUnit>(pal[1], w), p2 = wmul<kUnit>(pal[2], w);
    uint32_t out = 0;
    float e = 0.0f;
#pragma unroll 4
    for (int i = 0; i < 16; i++) {
        const f3 x = t.color(i);
        const f3 c = wmul<kUnit>(x, w);
        uint32_t k;
        if (allow_black) {
            const f3 e0 = p0 - c, e1 = p1 - c, e2 = p2 - c;
            const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(c, c);
            const uint32_t i1 = d1 < d2;
            const uint32_t i2 = (d2 <= d0) & (d2 <= d1);
            const uint32_t i3 = (d3 <= d0) & (d3 <= d1) & (d3 <= d2);
            k = (i1 | i3) | ((i2 | i3) << 1);
        } else {
            const f3 e0 = c - p0, e1 = c - p1, e2 = c - p2;
            const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2);
            const uint32_t i1 = d1 < d2;
            const uint32_t i2 = (d2 <= d0) & (d2 <= d1);
            k = i1 | (i2 << 1);
        }
        out |= k << (2 * i);
        e += t.weight(i) * texel_error<kUnit>(pick(pal, k), x, w);
    }
    indices = out;
    return e;
}

// compute_indices rule used by refinement (icbc.h:3179-3235): indices chosen
// against `sel` (the current best block's palette), error measured against
// `pal` (the candidate block's palette).
// Indices of the 16 texels against `sel` (palette of the current best block):
// nearest entry under the colour weights, lowest index on ties.
// Index of colour x against the weighted palette entries p0..p3 (the
// entries already multiplied by the colour weights w).
template <bool kUnit>
__device__ __forceinline__ uint32_t nearest_index(f3 x, f3 w, f3 p0, f3 p1, f3 p2, f3 p3) {
    const f3 c = wmul<kUnit>(x, w);
    const f3 e0 = c - p0, e1 = c - p1, e2 = c - p2, e3 = c - p3;
    const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(e3, e3);
    const uint32_t i1 = (d1 <= d0) & (d1 < d2) & (d1 < d3);
    const uint32_t i2 = (d2 <= d0) & (d2 <= d1) & (d2 < d3);
    const uint32_t i3 = (d3 <= d0) & (d3 <= d1) & (d3 <= d2);
    return (i1 | i3) | ((i2 | i3) << 1);
}

template <bool kUnit, typename Src>
__device__ __forceinline__ uint32_t select_indices(const Src& t, f3 w, const f3 sel[4]) {
    const f3 p0 = wmul<kUnit>(sel[0], w), p1 = wmul<kUnit>(sel[1], w), p2 = wmul<kUnit>(sel[2], w),
             p3 = wmul<kUnit>(sel[3], w);
    uint32_t out = 0;
#pragma unroll TexelUnroll<Src>::value
    for (int i = 0; i < 16; i++) out |= nearest_index<kUnit>(t.color(i), w, p0, p1, p2, p3) << (2 * i);
    return out;
}

// The indices of select_indices computed by a group of K lanes (K a power of
// two, wavefront-uniform) that hold the same texels: with K = 2 each lane
// takes 8 texels, with K >= 4 lanes 0-3 of the group take 4 each, and the
// lanes' index bits are combined over the group.  Every lane of the group
// must be active.
template <bool kUnit, bool kWeighted>
__device__ __forceinline__ uint32_t group_select_indices(const TexelRegs<kWeighted>& r, f3 w, const f3 sel[4], int K,
                                                         int sub) {
    const f3 p0 = wmul<kUnit>(sel[0], w), p1 = wmul<kUnit>(sel[1], w), p2 = wmul<kUnit>(sel[2], w),
             p3 = wmul<kUnit>(sel[3], w);
    uint32_t out = 0;
    if (K == 1) {
#pragma unroll
        for (int i = 0; i < 16; i++) out |= nearest_index<kUnit>(r.c[i], w, p0, p1, p2, p3) << (2 * i);
        return out;
    }
    if (K == 2) {
        const bool hi = sub != 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const f3 a = r.c[j], b = r.c[8 + j];
            const f3 x = mk(hi ? b.x : a.x, hi ? b.y : a.y, hi ? b.z : a.z);
            out |= nearest_index<kUnit>(x, w, p0, p1, p2, p3) << (2 * j);
        }
        out <<= hi ? 16 : 0;
    } else if (sub < 4) {
        const bool b0 = (sub & 1) != 0, b1 = (sub & 2) != 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const f3 a = b0 ? r.c[4 + j] : r.c[j], b = b0 ? r.c[12 + j] : r.c[8 + j];
            const f3 x = b1 ? b : a;
            out |= nearest_index<kUnit>(x, w, p0, p1, p2, p3) << (2 * j);
        }
        out <<= 8 * sub;
    }
    for (int o = 1; o < K; o <<= 1) out |= (uint32_t)__shfl_xor((int)out, o, 64);
    return out;
}



// ---------------------------------------------------------------------------
// Block output (icbc.h:3238-3274)

template <bool kUnit, typename Src>
__device__ __forceinline__ float emit4(const Src& t, f3 cw, int decoder, f3 a, f3 b, Code& out) {
    uint32_t c0 = to565(a), c1 = to565(b);
    if (c0 < c1) {
        uint32_t tmp = c0;
        c0 = c1;
        c1 = tmp;
    }
    f3 pal[4];
    palette_f(c0, c1, decoder, pal);
    uint32_t idx;
    const float e = select4<kUnit>(t, cw, pal, idx);
    out = make_code(c0, c1, idx);
    return e;
}

template <bool kUnit, typename Tx>
__device__ __forceinline__ float emit3(const Tx& t, f3 cw, int decoder, bool allow_black,
                                       f3 a, f3 b, Code& out) {
    uint32_t c0 = to565(a), c1 = to565(b);
    if (c0 > c1) {
        uint32_t tmp = c0;
        c0 = c1;
        c1 = tmp;
    }
    f3 pal[4];
    palette_f(c0, c1, decoder, pal);
    uint32_t idx;
    const float e = select3<kUnit>(t, cw, allow_black, pal, idx);
    out = make_code(c0, c1, idx);
    return e;
}

// output_block4 and, when `with3`, output_block3 (icbc.h:3238-3274) sharing
// one pass over the texels: per texel both index rules run and each error sum
// accumulates in texel order, as two separate passes would.
template <bool kUnit, typename Tx>
__device__ __forceinline__ float emit43(const Tx& t, f3 w, int decoder, bool allow_black,
                                        bool with3, f3 a4, f3 b4, f3 a3, f3 b3, Code& out4, Code& out3, float& err3) {
    uint32_t c40 = to565(a4), c41 = to565(b4);
    if (c40 < c41) {
        uint32_t tmp = c40;
        c40 = c41;
        c41 = tmp;
    }
    uint32_t c30 = to565(a3), c31 = to565(b3);
    if (c30 > c31) {
        uint32_t tmp = c30;
        c30 = c31;
        c31 = tmp;
    }
    f3 pal4[4], pal3[4];
    palette_f(c40, c41, decoder, pal4);
    palette_f(c30, c31, decoder, pal3);
    const f3 p0 = wmul<kUnit>(pal4[0], w), p1 = wmul<kUnit>(pal4[1], w), p2 = wmul<kUnit>(pal4[2], w),
             p3 = wmul<kUnit>(pal4[3], w);
    const f3 q0 = wmul<kUnit>(pal3[0], w), q1 = wmul<kUnit>(pal3[1], w), q2 = wmul<kUnit>(pal3[2], w);
    uint32_t idx4 = 0, idx3 = 0;
    float e4 = 0.0f, e3 = 0.0f;
#pragma unroll 2
    for (int i = 0; i < 16; i++) {
        const f3 x = t.color(i);
        const float wt = t.weight(i);
        const f3 c = wmul<kUnit>(x, w);
        {  // compute_indices4 rule (icbc.h:3097-3131)
            const f3 e0 = c - p0, e1 = c - p1, e2 = c - p2, ee3 = c - p3;
            const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(ee3, ee3);
            const uint32_t hi = ((d1 > d2) & (d0 > d2)) | ((d0 > d3) & (d1 > d3));
            const uint32_t lo = (d0 > d3) & (d2 > d3);
            const uint32_t k = lo | (hi << 1);
            idx4 |= k << (2 * i);
            e4 += wt * texel_error<kUnit>(pick(pal4, k), x, w);
        }
        if (with3) {  // compute_indices3 rule (icbc.h:3133-3176)
            uint32_t k;
            if (allow_black) {
                const f3 e0 = q0 - c, e1 = q1 - c, e2 = q2 - c;
                const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2), d3 = dot3(c, c);
                const uint32_t i1 = d1 < d2;
                const uint32_t i2 = (d2 <= d0) & (d2 <= d1);
                const uint32_t i3 = (d3 <= d0) & (d3 <= d1) & (d3 <= d2);
                k = (i1 | i3) | ((i2 | i3) << 1);
            } else {
                const f3 e0 = c - q0, e1 = c - q1, e2 = c - q2;
                const float d0 = dot3(e0, e0), d1 = dot3(e1, e1), d2 = dot3(e2, e2);
                const uint32_t i1 = d1 < d2;
                const uint32_t i2 = (d2 <= d0) & (d2 <= d1);
                k = i1 | (i2 << 1);
            }
            idx3 |= k << (2 * i);
            e3 += wt * texel_error<kUnit>(pick(pal3, k), x, w);
        }
    }
    out4 = make_code(c40, c41, idx4);
    out3 = make_code(c30, c31, idx3);
    err3 = e3;
    return e4;
}

// ---------------------------------------------------------------------------
// Colour reduction (icbc.h:1673-1748)

__device__ __forceinline__ bool is_black(f3 c) {
    return c.x < 1.0f / 8 && c.y < 1.0f / 8 && c.z < 1.0f / 8;
}

__device__ __forceinline__ bool near3(f3 a, f3 b, float eps) {
    return fabsf(a.x - b.x) < eps && fabsf(a.y - b.y) < eps && fabsf(a.z - b.z) < eps;
}

// The colour-reduction, principal-axis and sort stages below keep their
// arrays in registers: every loop is fully unrolled so each array index is a
// compile-time constant, and data-dependent choices (which slot a texel merges
// into, where an appended colour goes, whether a sort step swaps) become
// predicated register moves.  This gives the same results as the reference's
// index-driven loops without per-lane memory traffic.

// Merge each weighted texel into the first earlier colour within `threshold`
// (running weighted average), else append it; texels with weight <= 0 (or
// NaN) are dropped (icbc.h:1683-1728).  Returns the colour count.
template <typename Tx>
__device__ __forceinline__ int reduce_colors(const Tx& t, float threshold, f3 (&c)[16], float (&w)[16],
                                             bool& any_black) {
    int n = 0;
    any_black = false;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        c[j] = mk(0.0f, 0.0f, 0.0f);
        w[j] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float wi = t.weight(i);
        const f3 ci = t.color(i);
        const bool valid = wi > 0;
        bool found = false;
#pragma unroll
        for (int j = 0; j < i; j++) {  // slot j holds a colour only when j < n (n <= i here)
            const bool m = valid && !found && j < n && near3(c[j], ci, threshold);
            if (m) {
                const float wj = w[j];
                const f3 s = c[j] * wj + ci * wi;
                const float d = wj + wi;
                c[j] = mk(s.x / d, s.y / d, s.z / d);
                w[j] = wj + wi;
            }
            found = found || m;
        }
        const bool add = valid && !found;
#pragma unroll
        for (int j = 0; j <= i; j++) {
            if (add && j == n) {
                c[j] = ci;
                w[j] = wi;
            }
        }
        n += add ? 1 : 0;
        if (valid && is_black(ci)) any_black = true;
    }
    return n;
}

// Removes black colours, keeping the order of the rest (icbc.h:1730-1748).
__device__ __forceinline__ int drop_blacks(f3 (&c)[16], float (&w)[16], int n) {
    int m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const bool keep = i < n && !is_black(c[i]);
#pragma unroll
        for (int k = 0; k < i; k++) {  // m <= i: move colour i down to slot m
            if (keep && k == m) {
                c[k] = c[i];
                w[k] = w[i];
            }
        }
        m += keep ? 1 : 0;
    }
    return m;
}

// ---------------------------------------------------------------------------
// Principal axis: weighted centroid, covariance, 6 power-method steps
// (icbc.h:1752-1844).

__device__ __forceinline__ f3 principal_axis(const f3 (&pts)[16], const float (&ws)[16], int n) {
    float total = 0.0f;
    f3 c = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i < n) {
            total += ws[i];
            c = c + pts[i] * ws[i];
        }
    }
    c = c * (1.0f / total);

    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f, m5 = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i < n) {
            f3 a = pts[i] - c;
            f3 b = a * ws[i];
            m0 += a.x * b.x;
            m1 += a.x * b.y;
            m2 += a.x * b.z;
            m3 += a.y * b.y;
            m4 += a.y * b.z;
            m5 += a.z * b.z;
        }
    }
    if (m0 == 0 && m3 == 0 && m5 == 0) return mk(0.0f, 0.0f, 0.0f);

    f3 r0 = mk(m0, m1, m2), r1 = mk(m1, m3, m4), r2 = mk(m2, m4, m5);
    float l0 = dot3(r0, r0), l1 = dot3(r1, r1), l2 = dot3(r2, r2);
    f3 v = (l0 > l1 && l0 > l2) ? r0 : ((l1 > l2) ? r1 : r2);
#pragma unroll
    for (int it = 0; it < 6; it++) {
        float x = v.x * m0 + v.y * m1 + v.z * m2;
        float y = v.x * m1 + v.y * m3 + v.z * m4;
        float z = v.x * m2 + v.y * m4 + v.z * m5;
        float norm = max_r(max_r(x, y), z);
        v = mk(x, y, z) * (1.0f / norm);
    }
    return v;
}

// Summed-area table of the colours sorted along the principal axis
// (icbc.h:1857-1937).  The colours are sorted in place (a stable insertion
// sort with strict <, the reference's loop, which behaves the same with NaN
// keys).  Row k+1 of `sat` (float4 r,g,b,w) holds the prefix over the first
// k+1 sorted colours; row 0 holds zeros so that a cluster boundary of 0 reads
// a zero sum.  Rows 0..n are written.
// With `copy` non-null the rows are written there as well.
__device__ __forceinline__ void build_sat(f3 (&c)[16], float (&w)[16], int n, float4* sat, float4* copy) {
    const f3 axis = principal_axis(c, w, n);
    float key[16];
#pragma unroll
    for (int i = 0; i < 16; i++) key[i] = dot3(c[i], axis);
#pragma unroll
    for (int i = 1; i < 16; i++) {
        bool moving = i < n;
#pragma unroll
        for (int j = i; j > 0; j--) {
            moving = moving && (key[j] < key[j - 1]);
            if (moving) {
                const float tk = key[j];
                key[j] = key[j - 1];
                key[j - 1] = tk;
                const f3 tc = c[j];
                c[j] = c[j - 1];
                c[j - 1] = tc;
                const float tw = w[j];
                w[j] = w[j - 1];
                w[j - 1] = tw;
            }
        }
    }
    sat[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (copy) copy[0] = sat[0];
    float r = 0.0f, g = 0.0f, b = 0.0f, s = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i < n) {
            const float wi = w[i];
            if (i == 0) {
                r = c[i].x * wi;
                g = c[i].y * wi;
                b = c[i].z * wi;
                s = wi;
            } else {
                r = r + c[i].x * wi;
                g = g + c[i].y * wi;
                b = b + c[i].z * wi;
                s = s + wi;
            }
            sat[i + 1] = make_float4(r, g, b, s);
            if (copy) copy[i + 1] = sat[i + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// One cluster-fit candidate (icbc.h:2685-2742 four clusters, 2299-2344 three).
// s0..s2 are SAT rows at the cluster boundaries, tot the full-sum row.
// Returns the candidate's error (without the constant sum(w x^2) term) and the
// snapped endpoints.

// Correctly rounded 1/b, the value of IEEE-754 division 1.0f / b, in three
// instructions where b and 1/b are both normal (biased exponent 1..252): the
// hardware reciprocal estimate (within 1 ulp) and one Newton step with fused
// multiply-adds, e = 1 - b*y0 (exact) and y0 + e*y0.  Zero, denormal, huge,
// infinite and NaN b take the division itself, behind a wavefront-uniform
// branch: left to itself the compiler computes the division on every call and
// selects it, about ten instructions per candidate evaluation for a case
// that almost never occurs.  bc1_check_reciprocal_kernel compares the two for
// all 2^32 bit patterns on the device.
__device__ __forceinline__ float recip_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    float y1 = __builtin_fmaf(e, y0, y0);
    const uint32_t ex = (__float_as_uint(b) >> 23) & 0xffu;
    const bool slow = ex - 1u >= 252u;
    if (__builtin_expect(__ballot(slow) != 0, 0)) {
        if (slow) y1 = 1.0f / b;
    }
    return y1;
}

template <bool kUnit>
__device__ __forceinline__ float candidate4(float4 s0, float4 s1, float4 s2, float4 tot, f3 metric,
                                            const DecoderWeights& dw, f3& a, f3& b) {
    f3 x0 = mk(s0.x, s0.y, s0.z), x1 = mk(s1.x, s1.y, s1.z), x2 = mk(s2.x, s2.y, s2.z);
    float w0 = s0.w, w1 = s1.w, w2 = s2.w;
    const float w3 = tot.w - w2;
    x2 = x2 - x1;
    x1 = x1 - x0;
    w2 = w2 - w1;
    w1 = w1 - w0;

    const float alpha2 = w2 * dw.w1_9 + (w1 * dw.w4_9 + w0);
    const float beta2 = w1 * dw.w1_9 + (w2 * dw.w4_9 + w3);
    const float ab = (w1 + w2) * dw.w2_9;
    const float factor = recip_rn(alpha2 * beta2 - ab * ab);

    const f3 ax = x2 * dw.w1_3 + (x1 * dw.w2_3 + x0);
    const f3 bx = mk(tot.x, tot.y, tot.z) - ax;

    a = quantize_endpoint((ax * beta2 - bx * ab) * factor);
    b = quantize_endpoint((bx * alpha2 - ax * ab) * factor);

    const f3 e2 = (a * (b * ab - ax) - b * bx) * 2.0f;
    const f3 e1 = (a * a) * alpha2 + ((b * b) * beta2 + e2);
    return metric_dot<kUnit>(e1, metric);
}

// ---------------------------------------------------------------------------
// Lower bound for pruning the four-cluster search.  A candidate's error, as
// candidate4 computes it, is the quadratic
//   E(a, b) = sum_c m_c (a^2 A + b^2 B + 2 a b C - 2 a X_c - 2 b Y_c)
// at its snapped endpoints, with A + B + 2C = W (total weight: the palette
// weights of each cluster sum to 1), A + C = Wa (sum of w * alpha) and
// Y = t - X (t = block totals).  Over all real (a, b) its minimum is
//   -K0 - W * sum_c m_c (X_c - (Wa / W) t_c)^2 / (A W - Wa^2),  K0 = sum_c m_c t_c^2 / W,
// so no candidate can score below that.  X, Wa and A are fixed combinations
// of the candidate's three SAT rows S0, S1, S2 (X = w1_3 S2 + (w2_3 - w1_3) S1
// + (1 - w2_3) S0, likewise Wa on the weights, A with the squared palette
// weights), so X - (Wa / W) t is the same combination of the centred rows
// R = (S.xyz - (S.w / W) t, S.w), staged once per block with their colour
// part scaled by the channel weights (then sum_c m_c d_c^2 = |d|^2, m = cw^2).
// The terms of the first two rows depend only on the pair (c0, c1), of which
// a block has at most 153: they are summed once per block (pair_sums), and a
// candidate adds its third row (candidate4_bound_q).  The search prunes on
// q = |d|^2 / (A W - Wa^2), the bound being -K0 - W q.  Fused multiply-adds
// and the hardware reciprocal are used (the bound only decides which
// candidates are evaluated, never a result); the rounding of this form and
// of candidate4's own arithmetic is covered by the block's margin
// (bound_row).  Candidates with a poorly conditioned system
// (A W - Wa^2 < A W / 16) get q = NaN: never pruned.
struct BoundCoefs {
    float c0, c1, c2;  // rows' weights in X and Wa
    float a0, a1, a2;  // rows' weights in A
};

__device__ __forceinline__ BoundCoefs bound_coefs(const DecoderWeights& dw) {
    return BoundCoefs{1.0f - dw.w2_3, dw.w2_3 - dw.w1_3, dw.w1_3, 1.0f - dw.w4_9, dw.w4_9 - dw.w1_9, dw.w1_9};
}

// Centred SAT row (`invW` = 1 / total weight, `t` = colour totals), colour
// part scaled by the channel weights `cw`.
template <bool kUnit>
__device__ __forceinline__ float4 centred_row(float4 s, float invW, f3 t, f3 cw) {
    const float f = s.w * invW;
    const f3 d = wmul<kUnit>(mk(__builtin_fmaf(-f, t.x, s.x), __builtin_fmaf(-f, t.y, s.y), __builtin_fmaf(-f, t.z, s.z)),
                             cw);
    return make_float4(d.x, d.y, d.z, s.w);
}

// Sums of the bound's terms over the first two rows of the candidates with
// boundaries (c0, c1, *): `c` for d (xyz) and Wa (w), `a` for A.
__device__ __forceinline__ void pair_sums(float4 r0, float4 r1, const BoundCoefs& k, float4& c, float& a) {
    c = make_float4(__builtin_fmaf(k.c1, r1.x, k.c0 * r0.x), __builtin_fmaf(k.c1, r1.y, k.c0 * r0.y),
                    __builtin_fmaf(k.c1, r1.z, k.c0 * r0.z), __builtin_fmaf(k.c1, r1.w, k.c0 * r0.w));
    a = __builtin_fmaf(k.a1, r1.w, k.a0 * r0.w);
}

// q of the candidate with pair sums (pc, pa) and third centred row r2.
__device__ __forceinline__ float candidate4_bound_q(float4 pc, float pa, float4 r2, float W, const BoundCoefs& k) {
    const float A = __builtin_fmaf(k.a2, r2.w, pa);
    const float Wa = __builtin_fmaf(k.c2, r2.w, pc.w);
    const float AW = A * W;
    const float D = __builtin_fmaf(-Wa, Wa, AW);
    const float dx = __builtin_fmaf(k.c2, r2.x, pc.x);
    const float dy = __builtin_fmaf(k.c2, r2.y, pc.y);
    const float dz = __builtin_fmaf(k.c2, r2.z, pc.z);
    const float sd = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float q = sd * __builtin_amdgcn_rcpf(D);
    return (D >= AW * 0.0625f) ? q : __builtin_nanf("");
}

// Margin of a block's four-cluster pruning: 2^-13 sum_c m_c P_c (P = sum of
// w x^2, the scale of the errors and of the bound's rounding) + 2^-18 sum_c m_c
// (4 W + 4 sum w |x_c|) (the rounding of candidate4's error, whose terms are
// bounded by the weights and weighted colour sums since the snapped endpoints
// lie in [0, 1]).  Blocks whose sums are not finite are never pruned.
__device__ __forceinline__ float4 bound_row(const f3 (&c)[16], const float (&w)[16], int n, f3 m) {
    float px = 0.0f, py = 0.0f, pz = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f, W = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i < n) {
            px = __builtin_fmaf(w[i] * c[i].x, c[i].x, px);
            py = __builtin_fmaf(w[i] * c[i].y, c[i].y, py);
            pz = __builtin_fmaf(w[i] * c[i].z, c[i].z, pz);
            ax = __builtin_fmaf(w[i], fabsf(c[i].x), ax);
            ay = __builtin_fmaf(w[i], fabsf(c[i].y), ay);
            az = __builtin_fmaf(w[i], fabsf(c[i].z), az);
            W += w[i];
        }
    }
    const float scale = (m.x * px + m.y * py) + m.z * pz;
    const float terms = 4.0f * W * ((m.x + m.y) + m.z) + 4.0f * ((m.x * ax + m.y * ay) + m.z * az);
    const float margin = 0x1p-13f * scale + 0x1p-18f * terms;
    const bool ok = margin == margin && margin < 1e30f && W > 0.0f;
    return make_float4(margin, ok ? 1.0f : 0.0f, 0.0f, 0.0f);
}

// Lower bound of a four-cluster candidate's error (as candidate4 computes it,
// without the constant sum(w x^2) term) from its clusters alone.  All texels
// of a cluster share one palette point p, and a cluster's share of the
// quadratic, sum_c m_c (W p_c^2 - 2 p_c X_c) (X, W its colour and weight
// sums), is at least -|X|_m^2 / W; the sum over the four clusters bounds the
// candidate's error whatever its endpoints.  No system is solved, so it also
// decides candidates whose system is nearly singular.  Each cluster's sums are
// one subtraction of SAT rows (relative error u) and the terms are non-
// negative with a sum at most sum_c m_c P_c, so the rounding is far inside
// the block's margin, which also covers the deviation of candidate4's
// arithmetic (rounded palette constants, its own cluster sums) from the
// quadratic.
template <bool kUnit>
__device__ __forceinline__ float cluster_bound(float4 s0, float4 s1, float4 s2, float4 tot, f3 metric) {
    auto term = [&](float4 lo, float4 hi) {
        const f3 x = mk(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
        const float w = hi.w - lo.w;
        const float nrm = metric_dot<kUnit>(x * x, metric);
        return w > 0.0f ? nrm * __builtin_amdgcn_rcpf(w) : 0.0f;  // an empty cluster has X = 0
    };
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return -(((term(zero, s0) + term(s0, s1)) + term(s1, s2)) + term(s2, tot));
}

template <bool kUnit>
__device__ __forceinline__ float candidate3(float4 s0, float4 s1, float4 tot, f3 metric, f3& a, f3& b) {
    f3 x0 = mk(s0.x, s0.y, s0.z), x1 = mk(s1.x, s1.y, s1.z);
    float w0 = s0.w, w1 = s1.w;
    const float w2 = tot.w - w1;
    x1 = x1 - x0;
    w1 = w1 - w0;

    const float ab = w1 * 0.25f;
    const float alpha2 = w0 + ab;
    const float beta2 = w2 + ab;
    const float factor = recip_rn(alpha2 * beta2 - ab * ab);

    const f3 ax = x0 + x1 * 0.5f;
    const f3 bx = mk(tot.x, tot.y, tot.z) - ax;

    a = quantize_endpoint((ax * beta2 - bx * ab) * factor);
    b = quantize_endpoint((bx * alpha2 - ax * ab) * factor);

    const f3 e2 = (a * (b * ab - ax) - b * bx) * 2.0f;
    const f3 e1 = (a * a) * alpha2 + ((b * b) * beta2 + e2);
    return metric_dot<kUnit>(e1, metric);
}

// ---------------------------------------------------------------------------
// Least-squares endpoints for fixed indices (icbc.h:3278-3310) and box fit
// (icbc.h:3313-3368).

template <typename Src>
__device__ __forceinline__ bool lsq_endpoints(uint32_t indices, const Src& t, f3& a, f3& b) {
    float a2 = 0.0f, b2 = 0.0f, ab = 0.0f;
    f3 ax = mk(0.0f, 0.0f, 0.0f), bx = mk(0.0f, 0.0f, 0.0f);
#pragma unroll TexelUnroll<Src>::value
    for (int i = 0; i < 16; i++) {
        const uint32_t bits = indices >> (2 * i);
        // beta = bits & 1, or (1 + beta) / 3 when bit 1 is set: the quotients
        // are the correctly rounded constants 1/3 and 2/3 (no division per texel)
        float beta = (float)(bits & 1);
        if (bits & 2) beta = (bits & 1) ? 2.0f / 3.0f : 1.0f / 3.0f;
        const float alpha = 1.0f - beta;
        a2 += alpha * alpha;
        b2 += beta * beta;
        ab += alpha * beta;
        const f3 x = t.color(i);
        ax = ax + x * alpha;
        bx = bx + x * beta;
    }
    const float denom = a2 * b2 - ab * ab;
    if (fabsf(denom - 0.0f) < (float)0.0001) return false;
    const float factor = 1.0f / denom;
    a = sat3((ax * b2 - bx * ab) * factor);
    b = sat3((bx * a2 - ax * ab) * factor);
    return true;
}

// `get(i)` returns colour i of the n colours to fit.
template <typename Get>
__device__ __forceinline__ void box_fit(Get get, int n, f3& c0, f3& c1) {
    f3 hi = mk(0.0f, 0.0f, 0.0f), lo = mk(1.0f, 1.0f, 1.0f);
#pragma unroll
    for (int i = 0; i < 16; i++) {  // unrolled with a guard: colour arrays stay in registers
        if (i < n) {
            const f3 c = get(i);
            hi = mk(max_r(hi.x, c.x), max_r(hi.y, c.y), max_r(hi.z, c.z));
            lo = mk(min_r(lo.x, c.x), min_r(lo.y, c.y), min_r(lo.z, c.z));
        }
    }
    const float bias = (8.0f / 255.0f) / 16.0f;
    const f3 d = hi - lo;
    const f3 inset = sat3(mk(d.x / 16.0f - bias, d.y / 16.0f - bias, d.z / 16.0f - bias));
    hi = sat3(hi - inset);
    lo = sat3(lo + inset);
    const f3 center = (hi + lo) * 0.5f;
    float cxz = 0.0f, cyz = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (i < n) {
            const f3 t = get(i) - center;
            cxz += t.x * t.z;
            cyz += t.y * t.z;
        }
    }
    float x0 = hi.x, y0 = hi.y, x1 = lo.x, y1 = lo.y;
    if (cxz < 0) {
        float t = x0;
        x0 = x1;
        x1 = t;
    }
    if (cyz < 0) {
        float t = y0;
        y0 = y1;
        y1 = t;
    }
    c0 = mk(x0, y0, hi.z);
    c1 = mk(x1, y1, lo.z);
}

// ---------------------------------------------------------------------------
// Single-colour block (icbc.h:3519-3534)

__device__ __forceinline__ Code single_color(uint32_t r, uint32_t g, uint32_t b, int decoder) {
    const uint32_t mr = k_match5[decoder][r], mg = k_match6[decoder][g], mb = k_match5[decoder][b];
    uint32_t c0 = ((mr & 0xff) << 11) | ((mg & 0xff) << 5) | (mb & 0xff);
    uint32_t c1 = ((mr >> 8) << 11) | ((mg >> 8) << 5) | (mb >> 8);
    uint32_t idx = 0xaaaaaaaau;
    if (c0 < c1) {
        uint32_t t = c0;
        c0 = c1;
        c1 = t;
        idx ^= 0x55555555u;
    }
    return make_code(c0, c1, idx);
}

// ---------------------------------------------------------------------------
// Endpoint refinement (icbc.h:3574-3651)

// add a signed step to each 565 field, wrapping inside the field
__device__ __forceinline__ uint32_t nudge565(uint32_t c, int dr, int dg, int db) {
    uint32_t r = (((c >> 11) & 31) + dr) & 31;
    uint32_t g = (((c >> 5) & 63) + dg) & 63;
    uint32_t b = ((c & 31) + db) & 31;
    return (r << 11) | (g << 5) | b;
}

// The 16 step directions of the refinement walk, in the reference's order
// (icbc.h:3584-3609): (1,0,0) (0,1,0) (0,0,1) (-1,0,0) (0,-1,0) (0,0,-1)
// (1,1,0) (1,0,1) (0,1,1) (-1,-1,0) (-1,0,-1) (0,-1,-1) (-1,1,0) (1,-1,0)
// (0,-1,1) (0,1,-1).  Each channel's 16 steps are packed as 2-bit fields
// holding d + 1, so a step is decoded with shifts instead of a memory table.
constexpr int8_t kStepR[16] = {1, 0, 0, -1, 0, 0, 1, 1, 0, -1, -1, 0, -1, 1, 0, 0};
constexpr int8_t kStepG[16] = {0, 1, 0, 0, -1, 0, 1, 0, 1, -1, 0, -1, 1, -1, -1, 1};
constexpr int8_t kStepB[16] = {0, 0, 1, 0, 0, -1, 0, 1, 1, 0, -1, -1, 0, 0, 1, -1};
constexpr uint32_t pack_steps(const int8_t (&d)[16]) {
    uint32_t v = 0;
    for (int k = 0; k < 16; k++) v |= uint32_t(d[k] + 1) << (2 * k);
    return v;
}
constexpr uint32_t kPackR = pack_steps(kStepR);
constexpr uint32_t kPackG = pack_steps(kStepG);
constexpr uint32_t kPackB = pack_steps(kStepB);

__device__ __forceinline__ void refine_step(int k, int& dr, int& dg, int& db) {
    dr = (int)((kPackR >> (2 * k)) & 3) - 1;
    dg = (int)((kPackG >> (2 * k)) & 3) - 1;
    db = (int)((kPackB >> (2 * k)) & 3) - 1;
}

// ---------------------------------------------------------------------------
// Per-level stage selection (icbc.h:3653-3731)

struct Plan {
    float threshold;
    bool box, lsq, cluster, fit3_always, fit3_black_only, refine;
};

__device__ __forceinline__ Plan make_plan(int level, bool tcm, bool tcb) {
    Plan p{0.0f, false, false, false, false, false, false};
    if (level < 0 || level > 9) return p;
    p.box = level <= 3;
    p.lsq = level == 1 || level == 2;
    p.cluster = level >= 2;
    p.fit3_always = level >= 8 && tcm;
    p.fit3_black_only = level >= 2 && level <= 7 && tcm && tcb;
    p.refine = level == 9;
    switch (level) {
        case 2: p.threshold = 1.0f / 24; break;
        case 3: p.threshold = 1.0f / 32; break;
        case 4: p.threshold = 1.0f / 48; break;
        case 5: p.threshold = 1.0f / 64; break;
        case 6: p.threshold = 1.0f / 96; break;
        case 7: p.threshold = 1.0f / 128; break;
        case 8:
        case 9: p.threshold = 1.0f / 256; break;
        default: break;
    }
    return p;
}

}  // namespace icbc_gpu
