// C entry points over the reference encoder, for checking only.
//
// This translation unit is the whole "reference build": it includes the
// unmodified reference header (found through -I on the reference checkout,
// never copied here) with ICBC_IMPLEMENTATION, the same way the reference's
// own harness does (icbc_test.cpp:15-16), and exposes plain C functions so
// that tests can call the reference's public API and a few of its internal
// stages through ctypes.  The SIMD level is chosen by -DICBC_SIMD on the
// compiler line (oracle/Makefile): ICBC_SIMD=0 is the scalar build whose
// results define parity; the AVX2 / AVX-512 builds are timed as the CPU
// baseline only.
//
// Nothing in the product links this file.

#define ICBC_IMPLEMENTATION
#include "icbc.h"

#include <atomic>
#include <chrono>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <thread>
#include <vector>

using namespace icbc;

extern "C" {

int ref_simd_level(void) { return ICBC_SIMD; }

void ref_init(int decoder) { icbc::init((Decoder)decoder); }

// icbc.h:3846 public compress_bc1.
float ref_compress_bc1(int level, const float* colors, const float* weights, const float* cw,
                       int three_color_mode, int three_color_black, void* out) {
    return icbc::compress_bc1((Quality)level, colors, weights, cw, three_color_mode != 0,
                              three_color_black != 0, out);
}

// Serial loop over block-major input: colors 64 floats/block, weights 16 floats/block.
void ref_compress_bc1_many(int level, const float* colors, const float* weights, const float* cw,
                           int three_color_mode, int three_color_black, uint8_t* out, float* err,
                           size_t nblocks) {
    for (size_t b = 0; b < nblocks; b++) {
        alignas(64) float c[64];
        alignas(64) float w[16];
        memcpy(c, colors + 64 * b, sizeof(c));
        memcpy(w, weights + 16 * b, sizeof(w));
        float e = icbc::compress_bc1((Quality)level, c, w, cw, three_color_mode != 0,
                                     three_color_black != 0, out + 8 * b);
        if (err) err[b] = e;
    }
}

// Multithreaded loop with the harness's scheduling: blocks handed out in
// chunks of 32 from an atomic counter (icbc_test.cpp:397, ic_pfor.h:669-681).
// Returns the wall time of the encode loop in seconds (threads are started
// and parked before the clock starts).
double ref_compress_bc1_parallel(int level, const float* colors, const float* weights,
                                 const float* cw, int three_color_mode, int three_color_black,
                                 uint8_t* out, float* err, size_t nblocks, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    std::atomic<size_t> next(0);
    std::atomic<int> ready(0);
    std::atomic<bool> go(false);
    const size_t chunk = 32;
    auto worker = [&]() {
        ready.fetch_add(1);
        while (!go.load(std::memory_order_acquire)) std::this_thread::yield();
        for (;;) {
            size_t b0 = next.fetch_add(chunk);
            if (b0 >= nblocks) break;
            size_t b1 = b0 + chunk < nblocks ? b0 + chunk : nblocks;
            for (size_t b = b0; b < b1; b++) {
                alignas(64) float c[64];
                alignas(64) float w[16];
                memcpy(c, colors + 64 * b, sizeof(c));
                memcpy(w, weights + 16 * b, sizeof(w));
                float e = icbc::compress_bc1((Quality)level, c, w, cw, three_color_mode != 0,
                                             three_color_black != 0, out + 8 * b);
                if (err) err[b] = e;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
    while (ready.load() < nthreads) std::this_thread::yield();
    auto t0 = std::chrono::steady_clock::now();
    go.store(true, std::memory_order_release);
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void ref_decode_bc1(const void* block, uint8_t* rgba, int decoder) {
    icbc::decode_bc1(block, rgba, (Decoder)decoder);
}

float ref_evaluate_bc1_error(const uint8_t* rgba, const void* block, int decoder) {
    return icbc::evaluate_bc1_error(rgba, block, (Decoder)decoder);
}

void ref_decode_bc3(const void* block, uint8_t* rgba, int decoder) {
    icbc::decode_bc3(block, rgba, (Decoder)decoder);
}

float ref_evaluate_bc3_error(const uint8_t* rgba, const void* block, int alpha_blend, int decoder) {
    return icbc::evaluate_bc3_error(rgba, block, alpha_blend != 0, (Decoder)decoder);
}

// ---- internal stages (static functions of the header, reachable from this TU) ----

// icbc.h:1683 reduce_colors over the 16 texels of one block.
int ref_reduce_colors(const float* colors64, const float* weights16, float threshold,
                      float* out_rgb48, float* out_w16, int* any_black) {
    Vector3 cols[16];
    float w[16];
    bool ab = false;
    int n = reduce_colors((const Vector4*)colors64, weights16, 16, threshold, cols, w, &ab);
    for (int i = 0; i < n; i++) {
        out_rgb48[3 * i + 0] = cols[i].x;
        out_rgb48[3 * i + 1] = cols[i].y;
        out_rgb48[3 * i + 2] = cols[i].z;
        out_w16[i] = w[i];
    }
    *any_black = ab ? 1 : 0;
    return n;
}

// icbc.h:1857 compute_sat: out is r[16], g[16], b[16], w[16].
int ref_compute_sat(const float* rgb48, const float* w16, int n, float* out64) {
    SummedAreaTable sat;
    int c = compute_sat((const Vector3*)rgb48, w16, n, &sat);
    memcpy(out64 + 0, sat.r, 64);
    memcpy(out64 + 16, sat.g, 64);
    memcpy(out64 + 32, sat.b, 64);
    memcpy(out64 + 48, sat.w, 64);
    return c;
}

// icbc.h:1798-1844 principal axis of n weighted colours.
void ref_principal_component(const float* rgb48, const float* w16, int n, float* out3) {
    Vector3 v = computePrincipalComponent_PowerMethod(n, (const Vector3*)rgb48, w16);
    out3[0] = v.x;
    out3[1] = v.y;
    out3[2] = v.z;
}

static void load_sat(const float* sat64, SummedAreaTable* sat) {
    memcpy(sat->r, sat64 + 0, 64);
    memcpy(sat->g, sat64 + 16, 64);
    memcpy(sat->b, sat64 + 32, 64);
    memcpy(sat->w, sat64 + 48, 64);
}

// icbc.h:2348 cluster_fit_four.
void ref_cluster_fit_four(const float* sat64, int n, const float* metric3, float* start3,
                          float* end3) {
    SummedAreaTable sat;
    load_sat(sat64, &sat);
    Vector3 s, e;
    cluster_fit_four(sat, n, {metric3[0], metric3[1], metric3[2]}, &s, &e);
    start3[0] = s.x; start3[1] = s.y; start3[2] = s.z;
    end3[0] = e.x; end3[1] = e.y; end3[2] = e.z;
}

// icbc.h:2075 cluster_fit_three.
void ref_cluster_fit_three(const float* sat64, int n, const float* metric3, float* start3,
                           float* end3) {
    SummedAreaTable sat;
    load_sat(sat64, &sat);
    Vector3 s, e;
    cluster_fit_three(sat, n, {metric3[0], metric3[1], metric3[2]}, &s, &e);
    start3[0] = s.x; start3[1] = s.y; start3[2] = s.z;
    end3[0] = e.x; end3[1] = e.y; end3[2] = e.z;
}

// icbc.h:3257 output_block4.
float ref_output_block4(const float* colors64, const float* weights16, const float* cw,
                        const float* v0, const float* v1, void* block) {
    return output_block4((const Vector4*)colors64, weights16, {cw[0], cw[1], cw[2]},
                         {v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]}, (BlockBC1*)block);
}

// icbc.h:3238 output_block3.
float ref_output_block3(const float* colors64, const float* weights16, const float* cw,
                        int allow_black, const float* v0, const float* v1, void* block) {
    return output_block3((const Vector4*)colors64, weights16, {cw[0], cw[1], cw[2]},
                         allow_black != 0, {v0[0], v0[1], v0[2]}, {v1[0], v1[1], v1[2]},
                         (BlockBC1*)block);
}

// icbc.h:3278 optimize_end_points4.
int ref_optimize_end_points4(uint32_t indices, const float* colors64, float* a3, float* b3) {
    Vector3 a, b;
    bool ok = optimize_end_points4(indices, (const Vector4*)colors64, 16, &a, &b);
    if (ok) {
        a3[0] = a.x; a3[1] = a.y; a3[2] = a.z;
        b3[0] = b.x; b3[1] = b.y; b3[2] = b.z;
    }
    return ok ? 1 : 0;
}

// icbc.h:3771-3773 box fit: bbox, inset, diagonal.
void ref_box_fit(const float* rgb48, int n, float* c0, float* c1) {
    Vector3 a, b;
    fit_colors_bbox((const Vector3*)rgb48, n, &a, &b);
    inset_bbox(&a, &b);
    select_diagonal((const Vector3*)rgb48, n, &a, &b);
    c0[0] = a.x; c0[1] = a.y; c0[2] = a.z;
    c1[0] = b.x; c1[1] = b.y; c1[2] = b.z;
}

// icbc.h:3574 refine_endpoints (block is updated in place).
float ref_refine_endpoints(const float* colors64, const float* weights16, const float* cw,
                           int three_color_mode, float input_error, void* block) {
    return refine_endpoints((const Vector4*)colors64, weights16, {cw[0], cw[1], cw[2]},
                            three_color_mode != 0, input_error, (BlockBC1*)block);
}

// icbc.h:3519 single colour block from an 8-bit colour.
void ref_single_color(const uint8_t* rgb, void* block) {
    Color32 c;
    c.r = rgb[0];
    c.g = rgb[1];
    c.b = rgb[2];
    c.a = 255;
    compress_dxt1_single_color_optimal(c, (BlockBC1*)block);
}

// icbc.h:1624 vector3_to_color16.
uint32_t ref_vector3_to_color16(const float* v3) {
    return vector3_to_color16({v3[0], v3[1], v3[2]}).u;
}

// icbc.h:2876 palette of the current decoder, as 4 x RGBA8.
void ref_evaluate_palette(uint32_t c0, uint32_t c1, uint8_t* rgba16) {
    Color16 a, b;
    a.u = (uint16)c0;
    b.u = (uint16)c1;
    Color32 pal[4];
    evaluate_palette(a, b, pal);
    for (int i = 0; i < 4; i++) {
        rgba16[4 * i + 0] = pal[i].r;
        rgba16[4 * i + 1] = pal[i].g;
        rgba16[4 * i + 2] = pal[i].b;
        rgba16[4 * i + 3] = pal[i].a;
    }
}

// icbc.h:3097 / 3133 / 3179 index selection against a float palette (12 floats).
uint32_t ref_compute_indices4(const float* colors64, const float* cw, const float* pal12) {
    return compute_indices4((const Vector4*)colors64, {cw[0], cw[1], cw[2]}, (const Vector3*)pal12);
}
uint32_t ref_compute_indices3(const float* colors64, const float* cw, int allow_black,
                              const float* pal12) {
    return compute_indices3((const Vector4*)colors64, {cw[0], cw[1], cw[2]}, allow_black != 0,
                            (const Vector3*)pal12);
}
uint32_t ref_compute_indices(const float* colors64, const float* cw, const float* pal12) {
    return compute_indices((const Vector4*)colors64, {cw[0], cw[1], cw[2]}, (const Vector3*)pal12);
}

// icbc.h:2997 evaluate_mse of a finished block under the current decoder.
float ref_evaluate_mse_block(const float* colors64, const float* weights16, const float* cw,
                             const void* block) {
    return evaluate_mse((const Vector4*)colors64, weights16, {cw[0], cw[1], cw[2]},
                        (const BlockBC1*)block);
}

// Tables built by init (icbc.h:1947-1950, 3374-3375, 3495-3515).
void ref_tables(uint8_t* four976x4, uint8_t* three160x4, int* four_total16, int* three_total16,
                uint8_t* match5_512, uint8_t* match6_512, float* weights5) {
    memcpy(four976x4, s_fourCluster, sizeof(s_fourCluster));
    memcpy(three160x4, s_threeCluster, sizeof(s_threeCluster));
    memcpy(four_total16, s_fourClusterTotal, sizeof(s_fourClusterTotal));
    memcpy(three_total16, s_threeClusterTotal, sizeof(s_threeClusterTotal));
    memcpy(match5_512, s_match5, sizeof(s_match5));
    memcpy(match6_512, s_match6, sizeof(s_match6));
    weights5[0] = w1_3;
    weights5[1] = w2_3;
    weights5[2] = w1_9;
    weights5[3] = w4_9;
    weights5[4] = w2_9;
}

// icbc.h:1560-1570 rounding midpoints.
void ref_midpoints(float* m5_32, float* m6_64) {
    memcpy(m5_32, midpoints5, sizeof(midpoints5));
    memcpy(m6_64, midpoints6, sizeof(midpoints6));
}

}  // extern "C"
