/* CPU restatement of the reference BC1 encoder (icbc.h v1.05, ICBC_SCALAR
 * semantics).  TEST INFRASTRUCTURE ONLY: the tests and the CPU-baseline leg
 * of bench.py use it as a checker; the product never links it.
 *
 * Block layout is the reference's: 16 texels x RGBA float (alpha ignored),
 * 16 float weights, 3 float channel weights; output 8 bytes
 * {u16 col0, u16 col1, u32 indices} (icbc.h:254-258). */
#ifndef ICBC_ORACLE_H
#define ICBC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-block work counters, used for the algorithmic-FLOP roofline. */
typedef struct oracle_stats {
    int32_t count;        /* reduced colour count n (16 at box-fit-only levels) */
    int32_t count3;       /* colours given to the 3-cluster fit, -1 if not tried */
    int32_t refine_steps; /* iterations run by endpoint refinement (L9), else 0 */
    int32_t pad;
} oracle_stats;

void oracle_init(int decoder);

float oracle_compress_bc1(int level, const float* colors64, const float* weights16,
                          const float* cw3, int three_color_mode, int three_color_black,
                          void* out8, oracle_stats* stats);

/* Block-major batch; weights may be NULL (all 1.0); err and stats may be NULL.
 * nthreads > 1 splits the blocks across POSIX threads. Returns seconds. */
double oracle_compress_bc1_many(int level, const float* colors, const float* weights,
                                const float* cw3, int three_color_mode, int three_color_black,
                                uint8_t* out, float* err, oracle_stats* stats, size_t nblocks,
                                int nthreads);

void oracle_decode_bc1(const void* block8, uint8_t* rgba64, int decoder);
float oracle_evaluate_bc1_error(const uint8_t* rgba64, const void* block8, int decoder);
void oracle_decode_bc3(const void* block16, uint8_t* rgba64, int decoder);
float oracle_evaluate_bc3_error(const uint8_t* rgba64, const void* block16, int alpha_blend, int decoder);

/* Stage entry points (for per-stage pinning against the reference build). */
int oracle_reduce_colors(const float* colors64, const float* weights16, float threshold,
                         float* out_rgb48, float* out_w16, int* any_black);
int oracle_compute_sat(const float* rgb48, const float* w16, int n, float* out64);
void oracle_principal_component(const float* rgb48, const float* w16, int n, float* out3);
void oracle_cluster_fit_four(const float* sat64, int n, const float* metric3, float* start3,
                             float* end3);
void oracle_cluster_fit_three(const float* sat64, int n, const float* metric3, float* start3,
                              float* end3);
float oracle_output_block4(const float* colors64, const float* weights16, const float* cw3,
                           const float* v0, const float* v1, void* block8);
float oracle_output_block3(const float* colors64, const float* weights16, const float* cw3,
                           int allow_black, const float* v0, const float* v1, void* block8);
int oracle_optimize_end_points4(uint32_t indices, const float* colors64, float* a3, float* b3);
void oracle_box_fit(const float* rgb48, int n, float* c0, float* c1);
float oracle_refine_endpoints(const float* colors64, const float* weights16, const float* cw3,
                              int three_color_mode, float input_error, void* block8,
                              int* steps);
void oracle_single_color(const uint8_t* rgb, void* block8);
uint32_t oracle_vector3_to_color16(const float* v3);
void oracle_evaluate_palette(uint32_t c0, uint32_t c1, uint8_t* rgba16);
uint32_t oracle_compute_indices4(const float* colors64, const float* cw3, const float* pal12);
uint32_t oracle_compute_indices3(const float* colors64, const float* cw3, int allow_black,
                                 const float* pal12);
uint32_t oracle_compute_indices(const float* colors64, const float* cw3, const float* pal12);
float oracle_evaluate_mse_block(const float* colors64, const float* weights16, const float* cw3,
                                const void* block8);
void oracle_tables(uint8_t* four976x4, uint8_t* three160x4, int* four_total16,
                   int* three_total16, uint8_t* match5_512, uint8_t* match6_512,
                   float* weights5);
void oracle_midpoints(float* m5_32, float* m6_64);

#ifdef __cplusplus
}
#endif

#endif
