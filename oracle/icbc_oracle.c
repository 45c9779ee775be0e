/* CPU restatement of the reference BC1 encoder, ICBC_SCALAR semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the GPU encoder is
 * checked against (tests, smoke check, CPU baseline).  It is never linked into
 * the product library.
 *
 * Every function follows the operation order of the reference's scalar build
 * (icbc.h v1.05 with ICBC_SIMD=0) so results are bit-identical when compiled
 * without FMA contraction (-ffp-contract=off) and with IEEE float:
 *   - max(a,b) = (b < a) ? a : b, min(a,b) = (a < b) ? a : b  (icbc.h:207-213)
 *   - round01(a) = float(int(a + 0.5f))                         (icbc.h:417)
 *   - a*b+c and a*b-c*d evaluated as written, no fused ops      (icbc.h:413-415)
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z                  (icbc.h:305)
 * Sums run left to right over texels / colours exactly as the reference does.
 *
 * Parity status (checked by tests/test_oracle.py against golden vectors written
 * by the compiled reference and, where the reference build is present, against
 * the reference's own stage functions): every function here is pinned. */
#define _POSIX_C_SOURCE 200809L
#include "icbc_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct { float x, y, z; } v3;

/* ------------------------------------------------------------------ */
/* scalar primitives (icbc.h:201-221, 401-425)                         */

static inline float fmax_r(float a, float b) { return (b < a) ? a : b; }
static inline float fmin_r(float a, float b) { return (a < b) ? a : b; }
static inline float sat01(float a) { return fmin_r(fmax_r(a, 0.0f), 1.0f); }
static inline float round01(float a) { return (float)(int)(a + 0.5f); }

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vsat(v3 a) { return mk(sat01(a.x), sat01(a.y), sat01(a.z)); }

/* ------------------------------------------------------------------ */
/* global state built by oracle_init (icbc.h:3813-3818)                */

static int g_decoder = 0;
static float g_w1_3, g_w2_3, g_w1_9, g_w4_9, g_w2_9;
static uint8_t g_match5[256][2], g_match6[256][2];
static int g_four_total[16], g_three_total[16];
static uint8_t g_four[976][4], g_three[160][4];
static float g_mid5[32], g_mid6[64];

/* 565 rounding midpoints (icbc.h:1560-1570): the reference stores the
 * midpoint between adjacent expanded levels as a 6-decimal literal.  Rebuild
 * each literal's text from the exact midpoint k/510 and let strtof parse it,
 * which is what the compiler did with the literal. */
static void build_midpoints(float* out, int bits) {
    int levels = 1 << bits;
    for (int i = 0; i < levels - 1; i++) {
        int e0 = bits == 5 ? ((i << 3) | (i >> 2)) : ((i << 2) | (i >> 4));
        int e1 = bits == 5 ? (((i + 1) << 3) | ((i + 1) >> 2)) : (((i + 1) << 2) | ((i + 1) >> 4));
        long k = e0 + e1; /* midpoint = k / 510 */
        long micro = (k * 2000000L + 510L) / 1020L; /* round-half-up of k/510 * 1e6 */
        char buf[32];
        snprintf(buf, sizeof buf, "%ld.%06ld", micro / 1000000L, micro % 1000000L);
        out[i] = strtof(buf, NULL);
    }
    out[levels - 1] = FLT_MAX;
}

/* Ordered cluster boundaries (icbc.h:1965-2034).  In closed form, the
 * partitions first reachable with t colours are (c0, c1, t) for
 * 0 <= c0 <= c1 <= t, listed with c0 then c1 ascending; the 3-cluster
 * table lists (c0, t) for 0 <= c0 <= t.  Totals C(t+3,3)-1 and C(t+2,2)-1.
 * The last entry is repeated 8 times as padding. */
static void build_cluster_tables(void) {
    int i = 0;
    for (int t = 1; t <= 16; t++) {
        for (int c0 = 0; c0 <= t; c0++)
            for (int c1 = c0; c1 <= t; c1++) {
                g_four[i][0] = (uint8_t)c0;
                g_four[i][1] = (uint8_t)c1;
                g_four[i][2] = (uint8_t)t;
                g_four[i][3] = 0;
                i++;
            }
        g_four_total[t - 1] = i;
    }
    for (int k = 0; k < 8; k++) memcpy(g_four[968 + k], g_four[967], 4);
    i = 0;
    for (int t = 1; t <= 16; t++) {
        for (int c0 = 0; c0 <= t; c0++) {
            g_three[i][0] = (uint8_t)c0;
            g_three[i][1] = (uint8_t)t;
            g_three[i][2] = 0;
            g_three[i][3] = 0;
            i++;
        }
        g_three_total[t - 1] = i;
    }
    for (int k = 0; k < 8; k++) memcpy(g_three[152 + k], g_three[151], 4);
}

/* Optimal single-colour endpoint tables (icbc.h:3377-3485): brute force
 * over all endpoint pairs, keeping the first pair with the smallest error
 * (min endpoint outer loop, max endpoint inner loop). */
static int expand_bits(int v, int bits) {
    return bits == 5 ? ((v << 3) | (v >> 2)) : ((v << 2) | (v >> 4));
}

static void build_match_table(uint8_t table[256][2], int bits, int decoder) {
    int levels = 1 << bits;
    for (int target = 0; target < 256; target++) {
        int best = 256 * 100;
        for (int lo = 0; lo < levels; lo++) {
            for (int hi = 0; hi < levels; hi++) {
                int elo = expand_bits(lo, bits), ehi = expand_bits(hi, bits);
                int amd = abs(((43 * ehi + 21 * elo + 32) >> 6) - target);
                int nv_val;
                if (bits == 5) nv_val = ((2 * hi + lo) * 22) / 8;
                else nv_val = (256 * elo + (ehi - elo) / 4 + 128 + (ehi - elo) * 80) / 256;
                int nv = abs(nv_val - target);
                int intel = abs(((171 * ehi + 85 * elo + 128) >> 8) - target);
                int err;
                if (decoder == 0) { /* D3D10: worst of the three hardware decoders */
                    err = amd;
                    if (nv > err) err = nv;
                    if (intel > err) err = intel;
                } else if (decoder == 1) err = nv;
                else if (decoder == 3) err = intel;
                else err = amd;
                if (err < best) {
                    best = err;
                    table[target][0] = (uint8_t)hi;
                    table[target][1] = (uint8_t)lo;
                }
            }
        }
    }
}

void oracle_init(int decoder) {
    g_decoder = decoder;
    build_match_table(g_match5, 5, decoder);
    build_match_table(g_match6, 6, decoder);
    build_cluster_tables();
    /* icbc.h:3495-3515 */
    g_w1_3 = 1.0f / 3;
    g_w2_3 = 2.0f / 3;
    if (decoder == 2) { g_w1_3 = 21.0f / 64; g_w2_3 = 43.0f / 64; }
    else if (decoder == 3) { g_w1_3 = 85.0f / 256; g_w2_3 = 171.0f / 256; }
    g_w1_9 = g_w1_3 * g_w1_3;
    g_w4_9 = g_w2_3 * g_w2_3;
    g_w2_9 = g_w1_3 * g_w2_3;
    build_midpoints(g_mid5, 5);
    build_midpoints(g_mid6, 6);
}

/* ------------------------------------------------------------------ */
/* colour conversion (icbc.h:1588-1667)                                */

static inline float q5(float x) { return round01(x * 31.0f) * (1.0f / 31.0f); }
static inline float q6(float x) { return round01(x * 63.0f) * (1.0f / 63.0f); }
static inline v3 quantize_ept(v3 v) { return mk(q5(sat01(v.x)), q6(sat01(v.y)), q5(sat01(v.z))); }

static uint16_t to565(v3 v) {
    unsigned r = (unsigned)fmin_r(fmax_r(v.x * 31.0f, 0.0f), 31.0f);
    unsigned g = (unsigned)fmin_r(fmax_r(v.y * 63.0f, 0.0f), 63.0f);
    unsigned b = (unsigned)fmin_r(fmax_r(v.z * 31.0f, 0.0f), 31.0f);
    r += (v.x > g_mid5[r]);
    g += (v.y > g_mid6[g]);
    b += (v.z > g_mid5[b]);
    return (uint16_t)((r << 11) | (g << 5) | b);
}

typedef struct { int r, g, b, a; } rgba_i;

static rgba_i expand565(uint16_t c) {
    int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    rgba_i o = {(r5 << 3) | (r5 >> 2), (g6 << 2) | (g6 >> 4), (b5 << 3) | (b5 >> 2), 255};
    return o;
}

static inline uint8_t to_u8(float v) { return (uint8_t)(sat01(v) * 255 + 0.5f); }

/* ------------------------------------------------------------------ */
/* palettes (icbc.h:2751-2900)                                         */

/* With force_four the four-colour formulas apply whatever the endpoint order
   (evaluate_palette4, icbc.h:2883-2891). */
static void palette_rgba_forced(uint16_t c0, uint16_t c1, int decoder, int force_four, rgba_i pal[4]) {
    pal[0] = expand565(c0);
    pal[1] = expand565(c1);
    int p0[3] = {pal[0].r, pal[0].g, pal[0].b};
    int p1[3] = {pal[1].r, pal[1].g, pal[1].b};
    int q2[3], q3[3];
    int four = force_four || c0 > c1;
    if (decoder == 1) { /* NVIDIA: red/blue from the 5-bit values, green from 8-bit */
        int r0 = (c0 >> 11) & 31, b0 = c0 & 31, r1 = (c1 >> 11) & 31, b1 = c1 & 31;
        int gd = p1[1] - p0[1];
        if (four) {
            q2[0] = (uint8_t)(((2 * r0 + r1) * 22) / 8);
            q2[1] = (uint8_t)((256 * p0[1] + gd / 4 + 128 + gd * 80) / 256);
            q2[2] = (uint8_t)(((2 * b0 + b1) * 22) / 8);
            q3[0] = (uint8_t)(((2 * r1 + r0) * 22) / 8);
            q3[1] = (uint8_t)((256 * p1[1] - gd / 4 + 128 - gd * 80) / 256);
            q3[2] = (uint8_t)(((2 * b1 + b0) * 22) / 8);
        } else {
            q2[0] = (uint8_t)(((r0 + r1) * 33) / 8);
            q2[1] = (uint8_t)((256 * p0[1] + gd / 4 + 128 + gd * 128) / 256);
            q2[2] = (uint8_t)(((b0 + b1) * 33) / 8);
        }
    } else {
        for (int k = 0; k < 3; k++) {
            if (four) {
                if (decoder == 2) {
                    q2[k] = (uint8_t)((43 * p0[k] + 21 * p1[k] + 32) >> 6);
                    q3[k] = (uint8_t)((43 * p1[k] + 21 * p0[k] + 32) >> 6);
                } else if (decoder == 3) {
                    q2[k] = (uint8_t)((171 * p0[k] + 85 * p1[k] + 128) >> 8);
                    q3[k] = (uint8_t)((171 * p1[k] + 85 * p0[k] + 128) >> 8);
                } else {
                    q2[k] = (2 * p0[k] + p1[k]) / 3;
                    q3[k] = (2 * p1[k] + p0[k]) / 3;
                }
            } else {
                if (decoder == 2) q2[k] = (uint8_t)((p0[k] + p1[k] + 1) / 2);
                else if (decoder == 3) q2[k] = (uint8_t)((128 * p0[k] + 128 * p1[k] + 128) >> 8);
                else q2[k] = (p0[k] + p1[k]) / 2;
            }
        }
    }
    pal[2].r = q2[0]; pal[2].g = q2[1]; pal[2].b = q2[2]; pal[2].a = 255;
    if (four) {
        pal[3].r = q3[0]; pal[3].g = q3[1]; pal[3].b = q3[2]; pal[3].a = 255;
    } else {
        pal[3].r = pal[3].g = pal[3].b = pal[3].a = 0;
    }
}

static void palette_rgba(uint16_t c0, uint16_t c1, int decoder, rgba_i pal[4]) {
    palette_rgba_forced(c0, c1, decoder, 0, pal);
}

/* float palette = colour / 255.0f per channel (icbc.h:1656, 2902-2909) */
static void palette_float(uint16_t c0, uint16_t c1, v3 pal[4]) {
    rgba_i p[4];
    palette_rgba(c0, c1, g_decoder, p);
    for (int i = 0; i < 4; i++) pal[i] = mk(p[i].r / 255.0f, p[i].g / 255.0f, p[i].b / 255.0f);
}

/* ------------------------------------------------------------------ */
/* error evaluation (icbc.h:2955-3028)                                 */

static float texel_error(v3 p, v3 c, v3 w) {
    v3 d = vscale(vmul(vsub(p, c), w), 255);
    return vdot(d, d);
}

static float palette_error(const float* colors64, const float* weights16, v3 cw, const v3 pal[4],
                           uint32_t indices) {
    float e = 0.0f;
    for (int i = 0; i < 16; i++) {
        v3 c = mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]);
        e += weights16[i] * texel_error(pal[(indices >> (2 * i)) & 3], c, cw);
    }
    return e;
}

static float block_error(const float* colors64, const float* weights16, v3 cw, uint16_t c0,
                         uint16_t c1, uint32_t indices) {
    v3 pal[4];
    palette_float(c0, c1, pal);
    return palette_error(colors64, weights16, cw, pal, indices);
}

/* ------------------------------------------------------------------ */
/* index selection (icbc.h:3079-3235)                                  */

static uint32_t indices4(const float* colors64, v3 w, const v3 pal[4]) {
    v3 p0 = vmul(pal[0], w), p1 = vmul(pal[1], w), p2 = vmul(pal[2], w), p3 = vmul(pal[3], w);
    uint32_t out = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = vmul(mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]), w);
        v3 e0 = vsub(c, p0), e1 = vsub(c, p1), e2 = vsub(c, p2), e3 = vsub(c, p3);
        float d0 = vdot(e0, e0), d1 = vdot(e1, e1), d2 = vdot(e2, e2), d3 = vdot(e3, e3);
        unsigned hi = ((d1 > d2) & (d0 > d2)) | ((d0 > d3) & (d1 > d3));
        unsigned lo = (d0 > d3) & (d2 > d3);
        out |= (lo | (hi << 1)) << (2 * i);
    }
    return out;
}

static uint32_t indices3(const float* colors64, v3 w, int allow_black, const v3 pal[4]) {
    v3 p0 = vmul(pal[0], w), p1 = vmul(pal[1], w), p2 = vmul(pal[2], w);
    uint32_t out = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = vmul(mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]), w);
        unsigned lo, hi;
        if (allow_black) {
            v3 e0 = vsub(p0, c), e1 = vsub(p1, c), e2 = vsub(p2, c);
            float d0 = vdot(e0, e0), d1 = vdot(e1, e1), d2 = vdot(e2, e2), d3 = vdot(c, c);
            unsigned i1 = d1 < d2;
            unsigned i2 = (d2 <= d0) & (d2 <= d1);
            unsigned i3 = (d3 <= d0) & (d3 <= d1) & (d3 <= d2);
            lo = i1 | i3;
            hi = i2 | i3;
        } else {
            v3 e0 = vsub(c, p0), e1 = vsub(c, p1), e2 = vsub(c, p2);
            float d0 = vdot(e0, e0), d1 = vdot(e1, e1), d2 = vdot(e2, e2);
            lo = d1 < d2;
            hi = (d2 <= d0) & (d2 <= d1);
        }
        out |= (lo | (hi << 1)) << (2 * i);
    }
    return out;
}

static uint32_t indices_refine(const float* colors64, v3 w, const v3 pal[4]) {
    v3 p0 = vmul(pal[0], w), p1 = vmul(pal[1], w), p2 = vmul(pal[2], w), p3 = vmul(pal[3], w);
    uint32_t out = 0;
    for (int i = 0; i < 16; i++) {
        v3 c = vmul(mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]), w);
        v3 e0 = vsub(c, p0), e1 = vsub(c, p1), e2 = vsub(c, p2), e3 = vsub(c, p3);
        float d0 = vdot(e0, e0), d1 = vdot(e1, e1), d2 = vdot(e2, e2), d3 = vdot(e3, e3);
        unsigned i1 = (d1 <= d0) & (d1 < d2) & (d1 < d3);
        unsigned i2 = (d2 <= d0) & (d2 <= d1) & (d2 < d3);
        unsigned i3 = (d3 <= d0) & (d3 <= d1) & (d3 <= d2);
        out |= ((i1 | i3) | ((i2 | i3) << 1)) << (2 * i);
    }
    return out;
}

/* ------------------------------------------------------------------ */
/* block output (icbc.h:3238-3274)                                     */

typedef struct { uint16_t c0, c1; uint32_t idx; } blk;

static float emit4(const float* colors64, const float* weights16, v3 cw, v3 a, v3 b, blk* out) {
    uint16_t c0 = to565(a), c1 = to565(b);
    if (c0 < c1) { uint16_t t = c0; c0 = c1; c1 = t; }
    v3 pal[4];
    palette_float(c0, c1, pal);
    out->c0 = c0;
    out->c1 = c1;
    out->idx = indices4(colors64, cw, pal);
    return palette_error(colors64, weights16, cw, pal, out->idx);
}

static float emit3(const float* colors64, const float* weights16, v3 cw, int allow_black, v3 a,
                   v3 b, blk* out) {
    uint16_t c0 = to565(a), c1 = to565(b);
    if (c0 > c1) { uint16_t t = c0; c0 = c1; c1 = t; }
    v3 pal[4];
    palette_float(c0, c1, pal);
    out->c0 = c0;
    out->c1 = c1;
    out->idx = indices3(colors64, cw, allow_black, pal);
    return palette_error(colors64, weights16, cw, pal, out->idx);
}

/* ------------------------------------------------------------------ */
/* colour reduction (icbc.h:1673-1748)                                 */

static inline int is_black(v3 c) { return c.x < 1.0f / 8 && c.y < 1.0f / 8 && c.z < 1.0f / 8; }

static inline int near3(v3 a, v3 b, float eps) {
    return fabsf(a.x - b.x) < eps && fabsf(a.y - b.y) < eps && fabsf(a.z - b.z) < eps;
}

static int reduce(const float* colors64, const float* weights16, float threshold, v3* cols,
                  float* ws, int* any_black) {
    int n = 0;
    *any_black = 0;
    for (int i = 0; i < 16; i++) {
        float wi = weights16[i];
        if (!(wi > 0)) continue;
        v3 ci = mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]);
        int j = 0;
        while (j < n && !near3(cols[j], ci, threshold)) j++;
        if (j < n) {
            float wj = ws[j];
            v3 s = vadd(vscale(cols[j], wj), vscale(ci, wi));
            float d = wj + wi;
            cols[j] = mk(s.x / d, s.y / d, s.z / d);
            ws[j] = wj + wi;
        } else {
            cols[n] = ci;
            ws[n] = wi;
            n++;
        }
        if (is_black(ci)) *any_black = 1;
    }
    return n;
}

static int drop_blacks(const v3* cols, const float* ws, int n, v3* out, float* wout) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        if (is_black(cols[i])) continue;
        out[m] = cols[i];
        wout[m] = ws[i];
        m++;
    }
    return m;
}

/* ------------------------------------------------------------------ */
/* principal axis (icbc.h:1752-1844)                                   */

static v3 principal_axis(const v3* pts, const float* ws, int n) {
    float total = 0.0f;
    v3 c = mk(0, 0, 0);
    for (int i = 0; i < n; i++) {
        total += ws[i];
        c = vadd(c, vscale(pts[i], ws[i]));
    }
    c = vscale(c, 1.0f / total);

    float m[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        v3 a = vsub(pts[i], c);
        v3 b = vscale(a, ws[i]);
        m[0] += a.x * b.x;
        m[1] += a.x * b.y;
        m[2] += a.x * b.z;
        m[3] += a.y * b.y;
        m[4] += a.y * b.z;
        m[5] += a.z * b.z;
    }

    if (m[0] == 0 && m[3] == 0 && m[5] == 0) return mk(0, 0, 0);

    v3 r0 = mk(m[0], m[1], m[2]), r1 = mk(m[1], m[3], m[4]), r2 = mk(m[2], m[4], m[5]);
    float l0 = vdot(r0, r0), l1 = vdot(r1, r1), l2 = vdot(r2, r2);
    v3 v = (l0 > l1 && l0 > l2) ? r0 : (l1 > l2) ? r1 : r2;

    for (int it = 0; it < 6; it++) {
        float x = v.x * m[0] + v.y * m[1] + v.z * m[2];
        float y = v.x * m[1] + v.y * m[3] + v.z * m[4];
        float z = v.x * m[2] + v.y * m[4] + v.z * m[5];
        float norm = fmax_r(fmax_r(x, y), z);
        v = vscale(mk(x, y, z), 1.0f / norm);
    }
    return v;
}

/* Summed-area table over colours sorted along the principal axis
 * (icbc.h:1857-1937); rows >= n hold FLT_MAX. sat layout: r[16] g[16] b[16] w[16]. */
static void build_sat(const v3* cols, const float* ws, int n, float sat[64]) {
    v3 axis = principal_axis(cols, ws, n);
    int order[16];
    float key[16];
    for (int i = 0; i < n; i++) {
        order[i] = i;
        key[i] = vdot(cols[i], axis);
    }
    for (int i = 1; i < n; i++) {
        for (int j = i; j > 0 && key[j] < key[j - 1]; j--) {
            float tk = key[j]; key[j] = key[j - 1]; key[j - 1] = tk;
            int to = order[j]; order[j] = order[j - 1]; order[j - 1] = to;
        }
    }
    float r = 0, g = 0, b = 0, w = 0;
    for (int i = 0; i < n; i++) {
        float wi = ws[order[i]];
        v3 c = cols[order[i]];
        if (i == 0) {
            r = c.x * wi; g = c.y * wi; b = c.z * wi; w = wi;
        } else {
            r = r + c.x * wi; g = g + c.y * wi; b = b + c.z * wi; w = w + wi;
        }
        sat[i] = r; sat[16 + i] = g; sat[32 + i] = b; sat[48 + i] = w;
    }
    for (int i = n; i < 16; i++) sat[i] = sat[16 + i] = sat[32 + i] = sat[48 + i] = FLT_MAX;
}

/* ------------------------------------------------------------------ */
/* cluster fits (icbc.h:2075-2743, scalar branches)                    */

static void fit4(const float sat[64], int n, v3 metric, v3* start, v3* end) {
    const float rs = sat[n - 1], gs = sat[16 + n - 1], bs = sat[32 + n - 1], ws = sat[48 + n - 1];
    float best = FLT_MAX;
    v3 bs_a = mk(0, 0, 0), bs_b = mk(0, 0, 0);
    const int total = g_four_total[n - 1];
    for (int i = 0; i < total; i++) {
        int k0 = g_four[i][0] - 1, k1 = g_four[i][1] - 1, k2 = g_four[i][2] - 1;
        v3 x0 = mk(0, 0, 0), x1 = mk(0, 0, 0), x2 = mk(0, 0, 0);
        float w0 = 0, w1 = 0, w2 = 0;
        if (k0 >= 0) { x0 = mk(sat[k0], sat[16 + k0], sat[32 + k0]); w0 = sat[48 + k0]; }
        if (k1 >= 0) { x1 = mk(sat[k1], sat[16 + k1], sat[32 + k1]); w1 = sat[48 + k1]; }
        if (k2 >= 0) { x2 = mk(sat[k2], sat[16 + k2], sat[32 + k2]); w2 = sat[48 + k2]; }

        float w3 = ws - w2;
        x2 = vsub(x2, x1);
        x1 = vsub(x1, x0);
        w2 = w2 - w1;
        w1 = w1 - w0;

        float alpha2 = w2 * g_w1_9 + (w1 * g_w4_9 + w0);
        float beta2 = w1 * g_w1_9 + (w2 * g_w4_9 + w3);
        float ab = (w1 + w2) * g_w2_9;
        float factor = 1.0f / (alpha2 * beta2 - ab * ab);

        v3 ax = vadd(vscale(x2, g_w1_3), vadd(vscale(x1, g_w2_3), x0));
        v3 bx = vsub(mk(rs, gs, bs), ax);

        v3 a = vscale(vsub(vscale(ax, beta2), vscale(bx, ab)), factor);
        v3 b = vscale(vsub(vscale(bx, alpha2), vscale(ax, ab)), factor);
        a = quantize_ept(a);
        b = quantize_ept(b);

        v3 e2 = vscale(vsub(vmul(a, vsub(vscale(b, ab), ax)), vmul(b, bx)), 2.0f);
        v3 e1 = vadd(vscale(vmul(a, a), alpha2), vadd(vscale(vmul(b, b), beta2), e2));
        float err = vdot(e1, metric);
        if (err < best) {
            best = err;
            bs_a = a;
            bs_b = b;
        }
    }
    *start = bs_a;
    *end = bs_b;
}

static void fit3(const float sat[64], int n, v3 metric, v3* start, v3* end) {
    const float rs = sat[n - 1], gs = sat[16 + n - 1], bs = sat[32 + n - 1], ws = sat[48 + n - 1];
    float best = FLT_MAX;
    v3 bs_a = mk(0, 0, 0), bs_b = mk(0, 0, 0);
    const int total = g_three_total[n - 1];
    for (int i = 0; i < total; i++) {
        int k0 = g_three[i][0] - 1, k1 = g_three[i][1] - 1;
        v3 x0 = mk(0, 0, 0), x1 = mk(0, 0, 0);
        float w0 = 0, w1 = 0;
        if (k0 >= 0) { x0 = mk(sat[k0], sat[16 + k0], sat[32 + k0]); w0 = sat[48 + k0]; }
        if (k1 >= 0) { x1 = mk(sat[k1], sat[16 + k1], sat[32 + k1]); w1 = sat[48 + k1]; }

        float w2 = ws - w1;
        x1 = vsub(x1, x0);
        w1 = w1 - w0;

        float ab = w1 * 0.25f;
        float alpha2 = w0 + ab;
        float beta2 = w2 + ab;
        float factor = 1.0f / (alpha2 * beta2 - ab * ab);

        v3 ax = vadd(x0, vscale(x1, 0.5f));
        v3 bx = vsub(mk(rs, gs, bs), ax);

        v3 a = vscale(vsub(vscale(ax, beta2), vscale(bx, ab)), factor);
        v3 b = vscale(vsub(vscale(bx, alpha2), vscale(ax, ab)), factor);
        a = quantize_ept(a);
        b = quantize_ept(b);

        v3 e2 = vscale(vsub(vmul(a, vsub(vscale(b, ab), ax)), vmul(b, bx)), 2.0f);
        v3 e1 = vadd(vscale(vmul(a, a), alpha2), vadd(vscale(vmul(b, b), beta2), e2));
        float err = vdot(e1, metric);
        if (err < best) {
            best = err;
            bs_a = a;
            bs_b = b;
        }
    }
    *start = bs_a;
    *end = bs_b;
}

/* ------------------------------------------------------------------ */
/* least squares and box fit (icbc.h:3278-3368)                        */

static int lsq_endpoints(uint32_t indices, const float* colors64, v3* a, v3* b) {
    float a2 = 0.0f, b2 = 0.0f, ab = 0.0f;
    v3 ax = mk(0, 0, 0), bx = mk(0, 0, 0);
    for (int i = 0; i < 16; i++) {
        uint32_t bits = indices >> (2 * i);
        float beta = (float)(bits & 1);
        if (bits & 2) beta = (1 + beta) / 3.0f;
        float alpha = 1.0f - beta;
        v3 c = mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]);
        a2 += alpha * alpha;
        b2 += beta * beta;
        ab += alpha * beta;
        ax = vadd(ax, vscale(c, alpha));
        bx = vadd(bx, vscale(c, beta));
    }
    float denom = a2 * b2 - ab * ab;
    if (fabsf(denom - 0.0f) < (float)0.0001) return 0;
    float factor = 1.0f / denom;
    *a = vsat(vscale(vsub(vscale(ax, b2), vscale(bx, ab)), factor));
    *b = vsat(vscale(vsub(vscale(bx, a2), vscale(ax, ab)), factor));
    return 1;
}

static void box_fit(const v3* cols, int n, v3* c0, v3* c1) {
    v3 hi = mk(0, 0, 0), lo = mk(1, 1, 1);
    for (int i = 0; i < n; i++) {
        hi = mk(fmax_r(hi.x, cols[i].x), fmax_r(hi.y, cols[i].y), fmax_r(hi.z, cols[i].z));
        lo = mk(fmin_r(lo.x, cols[i].x), fmin_r(lo.y, cols[i].y), fmin_r(lo.z, cols[i].z));
    }
    /* inset (icbc.h:3362-3368) */
    float bias = (8.0f / 255.0f) / 16.0f;
    v3 d = vsub(hi, lo);
    v3 inset = vsat(mk(d.x / 16.0f - bias, d.y / 16.0f - bias, d.z / 16.0f - bias));
    hi = vsat(vsub(hi, inset));
    lo = vsat(vadd(lo, inset));
    /* diagonal (icbc.h:3324-3360) */
    v3 center = vscale(vadd(hi, lo), 0.5f);
    float cxz = 0.0f, cyz = 0.0f;
    for (int i = 0; i < n; i++) {
        v3 t = vsub(cols[i], center);
        cxz += t.x * t.z;
        cyz += t.y * t.z;
    }
    float x0 = hi.x, y0 = hi.y, x1 = lo.x, y1 = lo.y;
    if (cxz < 0) { float t = x0; x0 = x1; x1 = t; }
    if (cyz < 0) { float t = y0; y0 = y1; y1 = t; }
    *c0 = mk(x0, y0, hi.z);
    *c1 = mk(x1, y1, lo.z);
}

/* ------------------------------------------------------------------ */
/* single colour (icbc.h:3519-3534)                                    */

static void single_color(int r, int g, int b, blk* out) {
    uint16_t c0 = (uint16_t)((g_match5[r][0] << 11) | (g_match6[g][0] << 5) | g_match5[b][0]);
    uint16_t c1 = (uint16_t)((g_match5[r][1] << 11) | (g_match6[g][1] << 5) | g_match5[b][1]);
    uint32_t idx = 0xaaaaaaaau;
    if (c0 < c1) {
        uint16_t t = c0; c0 = c1; c1 = t;
        idx ^= 0x55555555u;
    }
    out->c0 = c0;
    out->c1 = c1;
    out->idx = idx;
}

/* ------------------------------------------------------------------ */
/* endpoint refinement (icbc.h:3574-3651)                              */

static const int8_t k_steps[16][3] = {
    {1, 0, 0},   {0, 1, 0},   {0, 0, 1},  {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
    {1, 1, 0},   {1, 0, 1},   {0, 1, 1},  {-1, -1, 0}, {-1, 0, -1}, {0, -1, -1},
    {-1, 1, 0},  {1, -1, 0},  {0, -1, 1}, {0, 1, -1},
};

/* add a signed step to each 565 field, wrapping within the field width */
static uint16_t nudge565(uint16_t c, int dr, int dg, int db) {
    int r = (((c >> 11) & 31) + dr) & 31;
    int g = (((c >> 5) & 63) + dg) & 63;
    int b = ((c & 31) + db) & 31;
    return (uint16_t)((r << 11) | (g << 5) | b);
}

static float refine(const float* colors64, const float* weights16, v3 cw, int three_color_mode,
                    float input_error, blk* out, int* steps) {
    float best = input_error;
    int last = 0;
    int i;
    for (i = 0; i < 256; i++) {
        blk r = *out;
        const int8_t* d = k_steps[i % 16];
        if ((i / 16) & 1) r.c0 = nudge565(r.c0, d[0], d[1], d[2]);
        else r.c1 = nudge565(r.c1, d[0], d[1], d[2]);
        if (!three_color_mode) {
            if (r.c0 == r.c1) r.c1 = nudge565(r.c1, 0, 1, 0);
            if (r.c0 < r.c1) { uint16_t t = r.c0; r.c0 = r.c1; r.c1 = t; }
        }
        v3 pal[4];
        palette_float(out->c0, out->c1, pal); /* palette of the current best block */
        r.idx = indices_refine(colors64, cw, pal);
        float e = block_error(colors64, weights16, cw, r.c0, r.c1, r.idx);
        if (e < best) {
            best = e;
            *out = r;
            last = i;
        }
        if (i - last > 32) break;
    }
    if (steps) *steps = (i < 256) ? i + 1 : 256;
    return best;
}

/* ------------------------------------------------------------------ */
/* level schedule and driver (icbc.h:3653-3808)                        */

typedef struct {
    float threshold;
    int box, lsq, cluster, fit3_always, fit3_black_only, refine;
} plan;

static plan make_plan(int level, int tcm, int tcb) {
    plan p = {0.0f, 0, 0, 0, 0, 0, 0};
    static const int denom[10] = {0, 0, 24, 32, 48, 64, 96, 128, 256, 256};
    if (level < 0 || level > 9) return p;
    p.box = level <= 3;
    p.lsq = level == 1 || level == 2;
    p.cluster = level >= 2;
    p.fit3_always = level >= 8 && tcm;
    p.fit3_black_only = level >= 2 && level <= 7 && tcm && tcb;
    p.refine = level == 9;
    if (level >= 2) p.threshold = 1.0f / denom[level];
    return p;
}

static float cluster_fit(const float* colors64, const float* weights16, const v3* cols,
                         const float* ws, int n, v3 cw, int use3, int try_black, int allow_black,
                         blk* out, int* n3) {
    v3 metric = vmul(cw, cw);
    float sat[64];
    build_sat(cols, ws, n, sat);
    v3 a, b;
    fit4(sat, n, metric, &a, &b);
    float best = emit4(colors64, weights16, cw, a, b, out);
    *n3 = -1;
    if (use3) {
        int m = n;
        if (try_black) {
            v3 c2[16];
            float w2[16];
            m = drop_blacks(cols, ws, n, c2, w2);
            if (!m) return best;
            build_sat(c2, w2, m, sat);
        }
        *n3 = m;
        fit3(sat, m, metric, &a, &b);
        blk t;
        float e3 = emit3(colors64, weights16, cw, allow_black, a, b, &t);
        if (e3 < best) {
            best = e3;
            *out = t;
        }
    }
    return best;
}

float oracle_compress_bc1(int level, const float* colors64, const float* weights16,
                          const float* cw3, int three_color_mode, int three_color_black,
                          void* out8, oracle_stats* stats) {
    plan p = make_plan(level, three_color_mode, three_color_black);
    v3 cw = mk(cw3[0], cw3[1], cw3[2]);
    v3 cols[16];
    float ws[16];
    int any_black = 0, n;
    blk out;
    memcpy(&out, out8, 8);
    if (stats) { stats->count = 0; stats->count3 = -1; stats->refine_steps = 0; stats->pad = 0; }

    if (p.cluster) {
        n = reduce(colors64, weights16, p.threshold, cols, ws, &any_black);
    } else {
        for (int i = 0; i < 16; i++) cols[i] = mk(colors64[4 * i], colors64[4 * i + 1], colors64[4 * i + 2]);
        n = 16;
    }
    if (stats) stats->count = n;

    if (n == 0) {
        memset(out8, 0, 8);
        return 0;
    }
    if (n == 1) {
        single_color(to_u8(cols[0].x), to_u8(cols[0].y), to_u8(cols[0].z), &out);
        memcpy(out8, &out, 8);
        return block_error(colors64, weights16, cw, out.c0, out.c1, out.idx);
    }

    float error = FLT_MAX;
    if (p.box) {
        v3 c0, c1;
        box_fit(cols, n, &c0, &c1);
        error = emit4(colors64, weights16, cw, c0, c1, &out);
        if (p.lsq && lsq_endpoints(out.idx, colors64, &c0, &c1)) {
            blk o2;
            float e2 = emit4(colors64, weights16, cw, c0, c1, &o2);
            if (e2 < error) {
                error = e2;
                out = o2;
            }
        }
    }
    if (p.cluster) {
        int use_black = any_black && three_color_black;
        int use3 = p.fit3_always || (use_black && p.fit3_black_only);
        blk o2;
        int n3;
        float e2 = cluster_fit(colors64, weights16, cols, ws, n, cw, use3, use_black,
                               three_color_black, &o2, &n3);
        if (stats) stats->count3 = n3;
        if (e2 < error) {
            out = o2;
            error = e2;
        }
    }
    if (p.refine) {
        int steps = 0;
        error = refine(colors64, weights16, cw, three_color_mode, error, &out, &steps);
        if (stats) stats->refine_steps = steps;
    }
    memcpy(out8, &out, 8);
    return error;
}

/* ------------------------------------------------------------------ */
/* batch driver                                                        */

typedef struct {
    int level, tcm, tcb;
    const float *colors, *weights, *cw;
    uint8_t* out;
    float* err;
    oracle_stats* stats;
    size_t b0, b1;
} job_t;

static void* run_job(void* arg) {
    job_t* j = (job_t*)arg;
    float ones[16];
    for (int i = 0; i < 16; i++) ones[i] = 1.0f;
    for (size_t b = j->b0; b < j->b1; b++) {
        const float* w = j->weights ? j->weights + 16 * b : ones;
        float e = oracle_compress_bc1(j->level, j->colors + 64 * b, w, j->cw, j->tcm, j->tcb,
                                      j->out + 8 * b, j->stats ? j->stats + b : NULL);
        if (j->err) j->err[b] = e;
    }
    return NULL;
}

double oracle_compress_bc1_many(int level, const float* colors, const float* weights,
                                const float* cw3, int three_color_mode, int three_color_black,
                                uint8_t* out, float* err, oracle_stats* stats, size_t nblocks,
                                int nthreads) {
    struct timespec t0, t1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((size_t)nthreads > nblocks) nthreads = nblocks ? (int)nblocks : 1;
    job_t jobs[256];
    pthread_t th[256];
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < nthreads; t++) {
        job_t j = {level, three_color_mode, three_color_black, colors, weights, cw3, out, err, stats,
                   nblocks * t / nthreads, nblocks * (t + 1) / nthreads};
        jobs[t] = j;
    }
    if (nthreads == 1) {
        run_job(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, run_job, &jobs[t]);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------ */
/* decode / evaluate (icbc.h:2915-2927, 3030-3046)                     */

void oracle_decode_bc1(const void* block8, uint8_t* rgba64, int decoder) {
    blk b;
    memcpy(&b, block8, 8);
    rgba_i pal[4];
    palette_rgba(b.c0, b.c1, decoder, pal);
    for (int i = 0; i < 16; i++) {
        rgba_i c = pal[(b.idx >> (2 * i)) & 3];
        rgba64[4 * i + 0] = (uint8_t)c.r;
        rgba64[4 * i + 1] = (uint8_t)c.g;
        rgba64[4 * i + 2] = (uint8_t)c.b;
        rgba64[4 * i + 3] = (uint8_t)c.a;
    }
}

float oracle_evaluate_bc1_error(const uint8_t* rgba64, const void* block8, int decoder) {
    blk b;
    memcpy(&b, block8, 8);
    rgba_i pal[4];
    palette_rgba(b.c0, b.c1, decoder, pal);
    float e = 0.0f;
    for (int i = 0; i < 16; i++) {
        rgba_i p = pal[(b.idx >> (2 * i)) & 3];
        int dr = p.r - rgba64[4 * i], dg = p.g - rgba64[4 * i + 1], db = p.b - rgba64[4 * i + 2];
        e += (float)(dr * dr + dg * dg + db * db);
    }
    return e;
}

/* decode_bc3 (icbc.h:2929-2946): the colour block (bytes 8-15) with the
   four-colour palette whatever the endpoint order (evaluate_palette4,
   icbc.h:2883-2891), alpha = alpha0 for every texel (alpha_palette[0],
   icbc.h:2902-2911; the reference's alpha index is always 0). */
static void palette4_rgba(uint16_t c0, uint16_t c1, int decoder, rgba_i pal[4]) {
    palette_rgba_forced(c0, c1, decoder, 1, pal);
}

void oracle_decode_bc3(const void* block16, uint8_t* rgba64, int decoder) {
    const uint8_t* b = (const uint8_t*)block16;
    blk rgb;
    memcpy(&rgb, b + 8, 8);
    rgba_i pal[4];
    palette4_rgba(rgb.c0, rgb.c1, decoder, pal);
    for (int i = 0; i < 16; i++) {
        rgba_i c = pal[(rgb.idx >> (2 * i)) & 3];
        rgba64[4 * i + 0] = (uint8_t)c.r;
        rgba64[4 * i + 1] = (uint8_t)c.g;
        rgba64[4 * i + 2] = (uint8_t)c.b;
        rgba64[4 * i + 3] = b[0];
    }
}

/* evaluate_bc3_error (icbc.h:3048-3071): per texel, in order, the integer
   colour error against the indexed four-colour entry, times alpha * 1.0f /
   255.0f when alpha_blend, then the squared alpha0 difference. */
float oracle_evaluate_bc3_error(const uint8_t* rgba64, const void* block16, int alpha_blend, int decoder) {
    const uint8_t* b = (const uint8_t*)block16;
    blk rgb;
    memcpy(&rgb, b + 8, 8);
    rgba_i pal[4];
    palette4_rgba(rgb.c0, rgb.c1, decoder, pal);
    float e = 0.0f;
    for (int i = 0; i < 16; i++) {
        rgba_i p = pal[(rgb.idx >> (2 * i)) & 3];
        int dr = p.r - rgba64[4 * i], dg = p.g - rgba64[4 * i + 1], db = p.b - rgba64[4 * i + 2];
        int a = rgba64[4 * i + 3];
        e += (alpha_blend ? (a * 1.0f / 255.0f) : 1.0f) * (float)(dr * dr + dg * dg + db * db);
        e += (float)((b[0] - a) * (b[0] - a));
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* stage entry points                                                  */

int oracle_reduce_colors(const float* colors64, const float* weights16, float threshold,
                         float* out_rgb48, float* out_w16, int* any_black) {
    v3 cols[16];
    float ws[16];
    int n = reduce(colors64, weights16, threshold, cols, ws, any_black);
    for (int i = 0; i < n; i++) {
        out_rgb48[3 * i] = cols[i].x;
        out_rgb48[3 * i + 1] = cols[i].y;
        out_rgb48[3 * i + 2] = cols[i].z;
        out_w16[i] = ws[i];
    }
    return n;
}

int oracle_compute_sat(const float* rgb48, const float* w16, int n, float* out64) {
    build_sat((const v3*)rgb48, w16, n, out64);
    return n;
}

void oracle_principal_component(const float* rgb48, const float* w16, int n, float* out3) {
    v3 v = principal_axis((const v3*)rgb48, w16, n);
    out3[0] = v.x; out3[1] = v.y; out3[2] = v.z;
}

void oracle_cluster_fit_four(const float* sat64, int n, const float* m3, float* s3, float* e3) {
    v3 a, b;
    fit4(sat64, n, mk(m3[0], m3[1], m3[2]), &a, &b);
    s3[0] = a.x; s3[1] = a.y; s3[2] = a.z;
    e3[0] = b.x; e3[1] = b.y; e3[2] = b.z;
}

void oracle_cluster_fit_three(const float* sat64, int n, const float* m3, float* s3, float* e3) {
    v3 a, b;
    fit3(sat64, n, mk(m3[0], m3[1], m3[2]), &a, &b);
    s3[0] = a.x; s3[1] = a.y; s3[2] = a.z;
    e3[0] = b.x; e3[1] = b.y; e3[2] = b.z;
}

float oracle_output_block4(const float* colors64, const float* weights16, const float* cw3,
                           const float* v0, const float* v1, void* block8) {
    blk o;
    float e = emit4(colors64, weights16, mk(cw3[0], cw3[1], cw3[2]), mk(v0[0], v0[1], v0[2]),
                    mk(v1[0], v1[1], v1[2]), &o);
    memcpy(block8, &o, 8);
    return e;
}

float oracle_output_block3(const float* colors64, const float* weights16, const float* cw3,
                           int allow_black, const float* v0, const float* v1, void* block8) {
    blk o;
    float e = emit3(colors64, weights16, mk(cw3[0], cw3[1], cw3[2]), allow_black,
                    mk(v0[0], v0[1], v0[2]), mk(v1[0], v1[1], v1[2]), &o);
    memcpy(block8, &o, 8);
    return e;
}

int oracle_optimize_end_points4(uint32_t indices, const float* colors64, float* a3, float* b3) {
    v3 a, b;
    if (!lsq_endpoints(indices, colors64, &a, &b)) return 0;
    a3[0] = a.x; a3[1] = a.y; a3[2] = a.z;
    b3[0] = b.x; b3[1] = b.y; b3[2] = b.z;
    return 1;
}

void oracle_box_fit(const float* rgb48, int n, float* c0, float* c1) {
    v3 a, b;
    box_fit((const v3*)rgb48, n, &a, &b);
    c0[0] = a.x; c0[1] = a.y; c0[2] = a.z;
    c1[0] = b.x; c1[1] = b.y; c1[2] = b.z;
}

float oracle_refine_endpoints(const float* colors64, const float* weights16, const float* cw3,
                              int three_color_mode, float input_error, void* block8,
                              int* steps) {
    blk o;
    memcpy(&o, block8, 8);
    float e = refine(colors64, weights16, mk(cw3[0], cw3[1], cw3[2]), three_color_mode,
                     input_error, &o, steps);
    memcpy(block8, &o, 8);
    return e;
}

void oracle_single_color(const uint8_t* rgb, void* block8) {
    blk o;
    single_color(rgb[0], rgb[1], rgb[2], &o);
    memcpy(block8, &o, 8);
}

uint32_t oracle_vector3_to_color16(const float* v3p) { return to565(mk(v3p[0], v3p[1], v3p[2])); }

void oracle_evaluate_palette(uint32_t c0, uint32_t c1, uint8_t* rgba16) {
    rgba_i pal[4];
    palette_rgba((uint16_t)c0, (uint16_t)c1, g_decoder, pal);
    for (int i = 0; i < 4; i++) {
        rgba16[4 * i] = (uint8_t)pal[i].r;
        rgba16[4 * i + 1] = (uint8_t)pal[i].g;
        rgba16[4 * i + 2] = (uint8_t)pal[i].b;
        rgba16[4 * i + 3] = (uint8_t)pal[i].a;
    }
}

uint32_t oracle_compute_indices4(const float* colors64, const float* cw3, const float* pal12) {
    return indices4(colors64, mk(cw3[0], cw3[1], cw3[2]), (const v3*)pal12);
}

uint32_t oracle_compute_indices3(const float* colors64, const float* cw3, int allow_black,
                                 const float* pal12) {
    return indices3(colors64, mk(cw3[0], cw3[1], cw3[2]), allow_black, (const v3*)pal12);
}

uint32_t oracle_compute_indices(const float* colors64, const float* cw3, const float* pal12) {
    return indices_refine(colors64, mk(cw3[0], cw3[1], cw3[2]), (const v3*)pal12);
}

float oracle_evaluate_mse_block(const float* colors64, const float* weights16, const float* cw3,
                                const void* block8) {
    blk b;
    memcpy(&b, block8, 8);
    return block_error(colors64, weights16, mk(cw3[0], cw3[1], cw3[2]), b.c0, b.c1, b.idx);
}

void oracle_tables(uint8_t* four976x4, uint8_t* three160x4, int* four_total16,
                   int* three_total16, uint8_t* match5_512, uint8_t* match6_512,
                   float* weights5) {
    memcpy(four976x4, g_four, sizeof g_four);
    memcpy(three160x4, g_three, sizeof g_three);
    memcpy(four_total16, g_four_total, sizeof g_four_total);
    memcpy(three_total16, g_three_total, sizeof g_three_total);
    memcpy(match5_512, g_match5, sizeof g_match5);
    memcpy(match6_512, g_match6, sizeof g_match6);
    weights5[0] = g_w1_3; weights5[1] = g_w2_3; weights5[2] = g_w1_9;
    weights5[3] = g_w4_9; weights5[4] = g_w2_9;
}

void oracle_midpoints(float* m5_32, float* m6_64) {
    memcpy(m5_32, g_mid5, sizeof g_mid5);
    memcpy(m6_64, g_mid6, sizeof g_mid6);
}
