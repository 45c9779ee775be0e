/* icbc_gpu.h -- C ABI of the MI355X (gfx950) BC1 encoder.
 *
 * Plain C types only (no HIP or torch types in the signatures), so the library
 * can be bound from C, ctypes, cgo or JNI.  HIP streams are passed as
 * `void*` (a hipStream_t, NULL = the default stream).
 *
 * The encoder reproduces castano/icbc v1.05 (icbc.h) bit for bit: the 8-byte
 * block codes equal those of the reference's scalar build (ICBC_SIMD=0) and the
 * returned per-block errors are the same floats.
 */
#ifndef ICBC_GPU_H
#define ICBC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ICBC_OK 0
#define ICBC_ERR_INVALID_ARGUMENT (-1)
#define ICBC_ERR_NOT_INITIALIZED (-2)
#define ICBC_ERR_OUT_OF_MEMORY (-3)
#define ICBC_ERR_DEVICE (-4)     /* a HIP runtime call failed */
#define ICBC_ERR_NO_DEVICE (-5)  /* no usable gfx950 device */
#define ICBC_ERR_BUDGET (-6)     /* serving: request larger than the memory budget */
#define ICBC_ERR_CANCELLED (-7)  /* serving: request cancelled before it ran */
#define ICBC_ERR_TIMEOUT (-8)    /* serving: wait timed out, request still running */

/* ---- decoders and quality levels (icbc.h:12-17, 21-36) ----------------- */
#define ICBC_DECODER_D3D10 0
#define ICBC_DECODER_NVIDIA 1
#define ICBC_DECODER_AMD 2
#define ICBC_DECODER_INTEL 3

/* ---- one encode job: a run of blocks sharing level and options ---------- */
#define ICBC_JOB_HOST_MEMORY 1u /* colors/weights/codes/errors/stats are host pointers */
#define ICBC_JOB_COLORS_RGBA8 2u /* colors are 16 x RGBA8 (64 bytes) per block, each channel
                                    read as c / 255.0f like the reference's harness
                                    (icbc_test.cpp:401-405) */
#define ICBC_JOB_IMAGE 4u /* texels come from `image` (device memory), gathered into 4x4
                               blocks on the device; `colors` must be NULL and block_count
                               equal icbc_gpu_image_block_count(image) */
#define ICBC_JOB_WEIGHTS_FROM_ALPHA 8u /* with ICBC_JOB_IMAGE: per-texel weight = the texel's
                                          alpha (`weights` must be NULL) */

/* ---- texture pre-pass: images on the device ------------------------------ */
#define ICBC_IMAGE_RGBA8 0u   /* 4 x uint8 per texel, read as c / 255.0f (icbc_test.cpp:401-405) */
#define ICBC_IMAGE_RGBA32F 1u /* 4 x float per texel */
#define ICBC_EDGE_REPLICATE 0u  /* ceil(w/4) x ceil(h/4) blocks; the last row and column are
                                   repeated into the partial blocks */
#define ICBC_EDGE_ROUND_DOWN 1u /* floor(w/4) x floor(h/4) blocks, partial blocks dropped, as the
                                   reference's harness does (icbc_test.cpp:354-380) */

typedef struct icbc_image {
    const void* data; /* device memory, first texel of the top row */
    uint32_t format;  /* ICBC_IMAGE_* */
    uint32_t width;
    uint32_t height;
    uint32_t edge;    /* ICBC_EDGE_* */
    uint64_t pitch;   /* bytes from one row to the next: >= width * texel size, a multiple of
                         the texel size (RGBA8: 4 B, RGBA32F: 16 B, which is also the alignment) */
} icbc_image;

typedef struct icbc_bc1_job {
    int32_t level;             /* icbc::Quality 0..9; others give error FLT_MAX, codes untouched */
    int32_t three_color_mode;  /* bool */
    int32_t three_color_black; /* bool */
    uint32_t flags;            /* ICBC_JOB_* */
    float color_weights[3];    /* per-channel metric weights */
    uint32_t tile_blocks;      /* blocks per wavefront tile: 8, 16, 32 or 64; 0 = chosen by
                                  the library (smaller tiles finish a small job sooner) */
    uint64_t block_count;
    const float* colors;       /* 64 floats per block: 16 texels x RGBA, texel = 4*y + x;
                                  16-byte aligned (64 bytes with ICBC_JOB_COLORS_RGBA8) */
    const float* weights;      /* 16 floats per block (16-byte aligned), NULL = all 1.0 */
    void* codes;               /* 8 bytes per block: u16 col0, u16 col1, u32 indices */
    float* errors;             /* one float per block (the reference's return value), or NULL */
    int32_t* stats;            /* optional, 4 int32 per block: reduced colour count n,
                                  colours given to the 3-cluster fit (-1 = not run),
                                  refinement steps, 0 */
    const icbc_image* image;   /* texel source with ICBC_JOB_IMAGE, else NULL; read at the call
                                  (and at every execute of a plan holding the job) */
} icbc_bc1_job;

/* Replaces icbc::init (icbc.h:19, 3813): selects the hardware decoder whose
 * palette the encoder targets and uploads the encoder's tables to the current
 * device.  Not thread-safe against concurrent compress calls, like the
 * reference. */
int icbc_gpu_init(int decoder);

/* Replaces a loop of icbc::compress_bc1 (icbc.h:48, 3846) over many blocks.
 * All jobs are encoded by one kernel launch: resident workgroups, one
 * wavefront each, take tiles of 8-64 blocks from a queue.  Device-memory jobs are enqueued on `stream` and the call
 * returns without waiting; host-memory jobs are staged, and the call waits
 * for their results. */
int icbc_gpu_compress_bc1_batch(const icbc_bc1_job* jobs, size_t n_jobs, void* stream);

/* Same results as icbc_gpu_compress_bc1_batch, one kernel launch per job
 * (the plain comparison path). */
int icbc_gpu_compress_bc1_plain(const icbc_bc1_job* jobs, size_t n_jobs, void* stream);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same results as icbc_gpu_compress_bc1_batch, but the call
 * allocates nothing -- no device, pinned or host heap memory.  Everything a
 * launch needs besides the caller's inputs and outputs (job table, tile list,
 * tile-queue counters, level-9 handoff list, workgroup workspaces) lives in a
 * fixed scratch array (icbc_gpu_noalloc_scratch_bytes(), 56.5 MiB) that the
 * library holds on every device it runs on until the process ends.  Limits per call:
 *   - device-memory block jobs only (no ICBC_JOB_HOST_MEMORY, no ICBC_JOB_IMAGE);
 *   - at most ICBC_NOALLOC_MAX_JOBS jobs, ICBC_NOALLOC_MAX_TILES tiles and
 *     ICBC_NOALLOC_MAX_LEVEL9_BLOCKS level-9 blocks (larger calls get
 *     ICBC_ERR_INVALID_ARGUMENT and launch nothing);
 *   - the device must be set up first (icbc_gpu_init with it current, or any
 *     other encode call on it), else ICBC_ERR_NOT_INITIALIZED;
 *   - launches run on at most ICBC_NOALLOC_MAX_GRID resident workgroups.
 * Every call on a device uses that device's one scratch array: calls that
 * overlap on one device (other streams, other threads) race and must be
 * serialised by the caller; calls on one stream are ordered and safe. */
#define ICBC_NOALLOC_MAX_JOBS 32
#define ICBC_NOALLOC_MAX_TILES 65536
#define ICBC_NOALLOC_MAX_LEVEL9_BLOCKS 262144
#define ICBC_NOALLOC_MAX_GRID 1024
int icbc_gpu_compress_bc1_batch_noalloc(const icbc_bc1_job* jobs, size_t n_jobs, void* stream);
/* Bytes of the scratch array held per device. */
uint64_t icbc_gpu_noalloc_scratch_bytes(void);

/* One block, host memory: the reference's compress_bc1 signature over C types.
 * Returns the block error (FLT_MAX for an invalid level, output untouched);
 * returns NaN and sets icbc_gpu_last_error() if the device call fails. */
float icbc_gpu_compress_bc1(int level, const float* input_colors, const float* input_weights,
                            const float color_weights[3], int three_color_mode,
                            int three_color_black, void* output);

/* ---- persistent batched plans (many textures, one launch) -------------- */
typedef struct icbc_bc1_plan icbc_bc1_plan;

#define ICBC_SCHEDULE_GRID 0       /* one workgroup per tile */
#define ICBC_SCHEDULE_PERSISTENT 1 /* resident workgroups pull tiles from a queue */

typedef struct icbc_plan_options {
    int32_t schedule;          /* ICBC_SCHEDULE_* */
    int32_t use_graph;         /* capture the launch in a hipGraph, replay on execute */
    int32_t workgroups_per_cu; /* persistent grid size per compute unit, 0 = automatic */
    int32_t order_by_cost;     /* issue expensive tiles first (estimated from level) */
} icbc_plan_options;

/* Builds a plan over device-memory jobs: job table, tile list and (optionally)
 * a captured graph, all resident on the device.  Buffers named by the jobs
 * must stay valid while the plan exists. */
int icbc_gpu_plan_create(const icbc_bc1_job* jobs, size_t n_jobs, const icbc_plan_options* opt,
                         void* stream, icbc_bc1_plan** plan);
/* Runs the plan on `stream` (asynchronous).  Executions of one plan must not
 * overlap on the device (issue them on one stream, or order them with
 * events): they share the plan's tile queues, workspace and hand-off list,
 * whose counters each execution leaves at zero for the next, and the jobs'
 * outputs.  After a failed execution (a device error) destroy the plan. */
int icbc_gpu_plan_execute(icbc_bc1_plan* plan, void* stream);
int icbc_gpu_plan_destroy(icbc_bc1_plan* plan);
/* Number of tiles in the plan. */
uint64_t icbc_gpu_plan_tiles(const icbc_bc1_plan* plan);

/* ---- request server: continuous batching under a memory budget -------- */
/* No counterpart in the reference (which encodes one block per call on the
 * caller's thread); for services that encode textures as requests arrive. */
typedef struct icbc_server icbc_server;

typedef struct icbc_server_options {
    uint64_t memory_budget; /* every device byte the server holds: request buffers, launch
                               tables and level-9 lists, the launch slots' workspaces and an
                               allowance for each slot's stream and the two upload streams (the
                               slots take at most half of it; at least 20 MiB per slot) */
    uint32_t max_inflight;  /* launches overlapping on the device (0 = 3, at most 8) */
    uint32_t max_batch;     /* requests merged into one launch (0 = no limit); a launch also
                               takes no more requests once it holds 2^20 blocks */
} icbc_server_options;

typedef struct icbc_request_info {
    int32_t status;       /* ICBC_OK or an error code */
    uint32_t batch_size;  /* requests in the launch this one ran in */
    uint64_t block_count;
    double submit_s, admit_s, launch_s, done_s; /* steady-clock seconds */
} icbc_request_info;

typedef struct icbc_server_stats_t {
    uint64_t launches, completed, blocks, pending, inflight;
    uint64_t memory_in_use;    /* bytes of pool blocks held by requests and launches */
    uint64_t memory_committed; /* device bytes the server holds in all (<= memory_budget) */
} icbc_server_stats_t;

int icbc_server_create(const icbc_server_options* opt, icbc_server** server);
/* Queues one texture (a job with ICBC_JOB_HOST_MEMORY; its buffers must stay
 * valid until the request completes).  ICBC_ERR_BUDGET if the request can
 * never fit the memory budget. */
int icbc_server_submit(icbc_server* server, const icbc_bc1_job* job, uint64_t* ticket);
/* Waits for a request (timeout < 0: forever); returns its status (or
 * ICBC_ERR_TIMEOUT) and releases the ticket unless it timed out.  A ticket
 * is collected once: other waits on it, also ones already blocked, return
 * ICBC_ERR_INVALID_ARGUMENT.  After a device failure, requests not yet
 * launched complete with that failure's status. */
int icbc_server_wait(icbc_server* server, uint64_t ticket, double timeout_s, icbc_request_info* info);
/* Cancels a request that has not been admitted yet. */
int icbc_server_cancel(icbc_server* server, uint64_t ticket);
int icbc_server_stats(icbc_server* server, icbc_server_stats_t* stats);
/* Drains outstanding requests and frees the server. */
int icbc_server_destroy(icbc_server* server);

/* ---- decode / evaluate (icbc.h:38, 43, 2915-2927, 3030-3046) ----------- */
/* Device memory: codes 8 B/block, rgba 64 B/block (RGBA8), errors 1 float/block. */
int icbc_gpu_decode_bc1_batch(const void* codes, uint64_t block_count, uint8_t* rgba, int decoder,
                              void* stream);
int icbc_gpu_evaluate_bc1_error_batch(const uint8_t* rgba, const void* codes,
                                      uint64_t block_count, float* errors, int decoder,
                                      void* stream);
/* decode_bc3 (icbc.h:2929-2946, 3824) and evaluate_bc3_error (icbc.h:3048-3071,
 * 3841) over 16-byte BC3 blocks (16-byte aligned, device memory): the colour
 * part decoded with the four-colour palette whatever its endpoint order, every
 * texel's alpha = alpha0 (the reference leaves the alpha indices unimplemented);
 * the error adds, per texel in order, the colour error (times alpha / 255 when
 * alpha_blend) and (alpha0 - alpha)^2. */
int icbc_gpu_decode_bc3_batch(const void* blocks, uint64_t block_count, uint8_t* rgba, int decoder,
                              void* stream);
int icbc_gpu_evaluate_bc3_error_batch(const uint8_t* rgba, const void* blocks, uint64_t block_count,
                                      float* errors, int alpha_blend, int decoder, void* stream);

/* ---- autotuning -----------------------------------------------------------
 * Per quality level: blocks per wavefront tile, resident workgroups per CU and
 * the schedule, measured on the caller's own sample.  The current device's
 * tuning then drives jobs with tile_blocks = 0 in large launches (smaller
 * launches keep spreading their blocks over every workgroup), the defaults of
 * plans created with opt = NULL, and the cost order in which mixed-level
 * launches hand out tiles (dearest first). */
typedef struct icbc_tuning {
    uint32_t tile_blocks[10];      /* 8, 16, 32 or 64; 0 = level not tuned */
    int32_t workgroups_per_cu[10]; /* resident workgroups per CU (persistent schedule) */
    int32_t schedule[10];          /* ICBC_SCHEDULE_* */
    float ns_per_block[10];        /* measured time of the chosen setting */
    int32_t prune_min_colors[10];  /* four-cluster search schedule: blocks with at least this
                                      many reduced colours run the pruning bound pass before the
                                      exact evaluations (9..16, 17 = never; 0 = the default, 13) */
    float default_ns_per_block[10]; /* measured time of the untuned default setting (tile 64,
                                       persistent, full occupancy, prune from 13 colours) */
} icbc_tuning;

/* Times, for each level in `level_mask` (bit L = level L) on `sample`
 * (device-memory jobs; their `level` and `tile_blocks` are overridden, their
 * outputs are overwritten): first every launch setting (tile 8/16/32/64 x
 * {persistent at full and half occupancy, one workgroup per tile}), then the
 * four-cluster search schedule (pruning from 9, 11 or 15 colours, or never) on
 * the fastest one.  Keeps the fastest per level as the current device's tuning
 * (merged with levels tuned before) and returns the whole table in `out` (may
 * be NULL). */
int icbc_gpu_autotune(const icbc_bc1_job* sample, size_t n_jobs, uint32_t level_mask, icbc_tuning* out);
/* The current device's tuning (all zero when none). */
int icbc_gpu_get_tuning(icbc_tuning* out);
/* Replaces it (e.g. with a table saved from an earlier run); NULL clears it. */
int icbc_gpu_set_tuning(const icbc_tuning* tuning);

/* ---- texture pre-pass (device) --------------------------------------------
 * The reference's harness turns an RGBA8 image into blocks on the CPU
 * (icbc_test.cpp:353-380: 4x4 gather, texel j = 4*y + x, c / 255.0f at
 * 401-405).  These run that step, and box-filtered mip chains, on the device.
 * All pointers are device memory; work is enqueued on `stream`. */

/* Blocks of an image under its edge mode (0 for an invalid descriptor). */
uint64_t icbc_gpu_image_block_count(const icbc_image* image);

/* Gathers the image into the encoder's block layout: 64 floats per block
 * (RGBA, alpha carried through) and, when `weights` is not NULL, 16 floats per
 * block holding each texel's alpha. */
int icbc_gpu_image_to_blocks(const icbc_image* image, float* colors, float* weights, void* stream);

/* Bytes of float RGBA needed for every level below `width` x `height` down to
 * the first that fits one 4x4 block, and that level count. */
uint64_t icbc_gpu_mip_chain_bytes(uint32_t width, uint32_t height, uint32_t* levels);

/* Builds those levels from `source` into `dst` (tightly packed, level after
 * level, RGBA32F), each a 2x2 box filter of the one above: ((a + b) + (c + d))
 * * 0.25 per channel in float, an odd last row or column dropped, a 1-texel
 * dimension filtered along the other only ((a + b) * 0.5).  `levels`
 * receives up to `max_levels` descriptors (edge mode copied from `source`),
 * `*n_levels` the count. */
int icbc_gpu_build_mip_chain(const icbc_image* source, float* dst, uint64_t dst_bytes, icbc_image* levels,
                             uint32_t max_levels, uint32_t* n_levels, void* stream);

/* ---- containers (host memory) ---------------------------------------------
 * BC1 files with mip levels.  Level k is max(1, width >> k) x max(1, height >>
 * k) texels in ceil(w/4) x ceil(h/4) blocks; `codes` holds the levels' blocks
 * back to back, level 0 first (8 bytes per block, as the encoder writes them). */
#define ICBC_CONTAINER_DDS 0 /* header of the reference's output_dxt_dds (icbc_test.cpp:155-211);
                                mip count and mipmap caps set when levels > 1 */
#define ICBC_CONTAINER_KTX 1 /* KTX 1.1 as the reference's output_dxt_ktx (icbc_test.cpp:233-279),
                                with one imageSize word per level */

/* Bytes of the file (0 for invalid arguments) / of the block data alone. */
uint64_t icbc_container_size(int kind, uint32_t width, uint32_t height, uint32_t levels);
uint64_t icbc_container_data_bytes(uint32_t width, uint32_t height, uint32_t levels);
/* Serialises into `out` (at least icbc_container_size bytes). */
int icbc_container_encode(int kind, uint32_t width, uint32_t height, uint32_t levels, const void* codes, void* out,
                          uint64_t out_bytes);
/* Writes the file to `path`. */
int icbc_container_write(int kind, const char* path, uint32_t width, uint32_t height, uint32_t levels,
                         const void* codes);
/* The reference's regression check (icbc_test.cpp:213-231): 1 if the top-level
 * block data of the DDS file at `path` equals `codes`, 0 if not, < 0 on error. */
int icbc_container_compare_dds(const char* path, uint32_t width, uint32_t height, const void* codes);

/* ---- diagnostics -------------------------------------------------------- */
const char* icbc_gpu_last_error(void); /* text of this thread's last failure */
int icbc_gpu_device_name(char* buf, size_t len);
int icbc_gpu_version(void); /* 10500 + build revision, mirrors icbc v1.05 */
/* Self-check of this device: counts the 2^32 float32 inputs where the
 * encoder's three-instruction reciprocal differs from IEEE-754 division
 * (no counterpart in the reference; bit-exact encoding needs 0). */
int icbc_gpu_check_reciprocal(uint64_t* mismatches);

/* ---- device plumbing for callers without their own HIP runtime ---------- */
void* icbc_gpu_malloc(size_t bytes);
int icbc_gpu_free(void* ptr);
void* icbc_gpu_host_alloc(size_t bytes); /* page-locked host memory (fast DMA) */
int icbc_gpu_host_free(void* ptr);
int icbc_gpu_memcpy_htod(void* dst, const void* src, size_t bytes);
int icbc_gpu_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int icbc_gpu_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int icbc_gpu_memset(void* dst, int value, size_t bytes);
void* icbc_gpu_stream_create(void);
int icbc_gpu_stream_destroy(void* stream);
int icbc_gpu_stream_synchronize(void* stream);
void* icbc_gpu_event_create(void);
int icbc_gpu_event_destroy(void* event);
int icbc_gpu_event_record(void* event, void* stream);
int icbc_gpu_event_synchronize(void* event);
float icbc_gpu_event_elapsed_ms(void* start, void* stop);

#ifdef __cplusplus
}
#endif

#endif /* ICBC_GPU_H */
