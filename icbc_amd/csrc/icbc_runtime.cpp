// Host runtime of the MI355X BC1 encoder: the C ABI of include/icbc_gpu.h and
// the drop-in C++ API of include/icbc.h.
//
// Responsibilities: build the encoder's lookup tables at init and upload them
// to each device that is used; validate jobs; stage host-memory jobs through
// device buffers; turn jobs into 64-block tiles; launch the kernels of
// bc1_encoder.hip (one launch per batch, one per job, or a persistent /
// graph-captured plan).  There is no CPU encoding path: every encode, decode
// and evaluate runs on the GPU, and a missing or failing device is reported.

#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icbc.h"
#include "../../include/icbc_gpu.h"
#include "bc1_kernels.h"
#include "block_context.h"
#include "icbc_internal.h"
#include "texture_prepass.h"

using namespace icbc_gpu;

namespace {

// ---------------------------------------------------------------------------
// error reporting

// A fixed buffer: recording an error allocates nothing.
thread_local char t_last_error[512];

int vfail(int code, const char* fmt, va_list ap) {
    vsnprintf(t_last_error, sizeof t_last_error, fmt, ap);
    return code;
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfail(code, fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? ICBC_ERR_OUT_OF_MEMORY : ICBC_ERR_DEVICE, "%s: %s (%d)",
                what, hipGetErrorString(e), (int)e);
}

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// ---------------------------------------------------------------------------
// tables (built once per process, uploaded once per device)

std::mutex g_mutex;
bool g_tables_built = false;
HostTables g_tables;
int g_decoder = -1;  // -1 = init not called
std::vector<int> g_uploaded;  // device ordinals holding the tables

// 565 rounding midpoints, icbc.h:1560-1570: the reference stores the midpoint
// of adjacent expanded levels as a 6-decimal literal.  The literal's text is
// rebuilt from the exact midpoint k/510 and parsed with strtof, as the
// compiler parsed the literal.
void build_midpoints(float* out, int bits) {
    const int levels = 1 << bits;
    for (int i = 0; i < levels - 1; i++) {
        const int e0 = bits == 5 ? ((i << 3) | (i >> 2)) : ((i << 2) | (i >> 4));
        const int e1 = bits == 5 ? (((i + 1) << 3) | ((i + 1) >> 2)) : (((i + 1) << 2) | ((i + 1) >> 4));
        const long k = e0 + e1;
        const long micro = (k * 2000000L + 510L) / 1020L;  // round-half-up of k/510 * 1e6
        char buf[32];
        snprintf(buf, sizeof buf, "%ld.%06ld", micro / 1000000L, micro % 1000000L);
        out[i] = strtof(buf, nullptr);
    }
    out[levels - 1] = FLT_MAX;
}

int expand_bits(int v, int bits) { return bits == 5 ? ((v << 3) | (v >> 2)) : ((v << 2) | (v >> 4)); }

// Single-colour endpoint tables, icbc.h:3377-3485: exhaustive search over
// (min, max) endpoint pairs, first pair with the smallest decoder error wins.
// Stored as max | min << 8.
void build_match(uint16_t* table, int bits, int decoder) {
    const int levels = 1 << bits;
    for (int target = 0; target < 256; target++) {
        int best = 256 * 100;
        uint16_t pick = 0;
        for (int lo = 0; lo < levels; lo++) {
            for (int hi = 0; hi < levels; hi++) {
                const int elo = expand_bits(lo, bits), ehi = expand_bits(hi, bits);
                const int amd = abs(((43 * ehi + 21 * elo + 32) >> 6) - target);
                const int nv_val = bits == 5 ? ((2 * hi + lo) * 22) / 8
                                             : (256 * elo + (ehi - elo) / 4 + 128 + (ehi - elo) * 80) / 256;
                const int nv = abs(nv_val - target);
                const int intel = abs(((171 * ehi + 85 * elo + 128) >> 8) - target);
                int err;
                if (decoder == ICBC_DECODER_D3D10) err = std::max(std::max(amd, nv), intel);
                else if (decoder == ICBC_DECODER_NVIDIA) err = nv;
                else if (decoder == ICBC_DECODER_INTEL) err = intel;
                else err = amd;
                if (err < best) {
                    best = err;
                    pick = (uint16_t)(hi | (lo << 8));
                }
            }
        }
        table[target] = pick;
    }
}

void build_tables(HostTables& t) {
    build_midpoints(t.mid5, 5);
    build_midpoints(t.mid6, 6);
    // Cluster partitions in the reference's total order (icbc.h:1965-2034):
    // the partitions first reachable with t colours are (c0, c1, t) for
    // 0 <= c0 <= c1 <= t, c0 then c1 ascending; the last entry pads to 976.
    // Four-cluster entries (four_entry) also carry the index of their first
    // two boundaries as a pair (0..152), which the pruning bound uses; bits
    // 16-31 of three-cluster entry p hold that pair's c0 | c1 << 8.
    // Pairs are numbered c0-major (c0 <= c1 <= 16): consecutive candidates
    // of the table, which differ in c1 only, read consecutive pair entries.
    auto pair_index = [](int c0, int c1) { return (uint32_t)(17 * c0 - c0 * (c0 - 1) / 2 + (c1 - c0)); };
    int i = 0;
    for (int tt = 1; tt <= 16; tt++) {
        for (int c0 = 0; c0 <= tt; c0++)
            for (int c1 = c0; c1 <= tt; c1++)
                t.four[i++] = four_entry((uint32_t)c0, (uint32_t)c1, (uint32_t)tt, pair_index(c0, c1));
        t.four_total[tt - 1] = i;
    }
    for (; i < 976; i++) t.four[i] = t.four[967];
    i = 0;
    for (int tt = 1; tt <= 16; tt++) {
        for (int c0 = 0; c0 <= tt; c0++) t.three[i++] = (uint32_t)c0 | ((uint32_t)tt << 8);
        t.three_total[tt - 1] = i;
    }
    for (; i < 160; i++) t.three[i] = t.three[151];
    for (int c0 = 0; c0 <= 16; c0++)
        for (int c1 = c0; c1 <= 16; c1++) t.three[pair_index(c0, c1)] |= ((uint32_t)c0 | ((uint32_t)c1 << 8)) << 16;
    for (uint32_t p = kPairs; p < 160; p++) t.three[p] |= t.three[kPairs - 1] & 0xffff0000u;
    for (int d = 0; d < 4; d++) {
        build_match(t.match5[d], 5, d);
        build_match(t.match6[d], 6, d);
        // decoder interpolation weights, icbc.h:3495-3515
        float w1_3 = 1.0f / 3, w2_3 = 2.0f / 3;
        if (d == ICBC_DECODER_AMD) {
            w1_3 = 21.0f / 64;
            w2_3 = 43.0f / 64;
        } else if (d == ICBC_DECODER_INTEL) {
            w1_3 = 85.0f / 256;
            w2_3 = 171.0f / 256;
        }
        t.dweights[d][0] = w1_3;
        t.dweights[d][1] = w2_3;
        t.dweights[d][2] = w1_3 * w1_3;
        t.dweights[d][3] = w2_3 * w2_3;
        t.dweights[d][4] = w1_3 * w2_3;
        // Nearly singular systems of 16 unit-weight colours: with cumulative
        // weights c0 <= c1 <= c2 at the boundaries, A W - Wa^2 < A W / 16
        // (A = sum w alpha^2, Wa = sum w alpha, W = 16; the device's guard).
        // Which candidates are listed only affects speed: a listed candidate is
        // decided by the cluster bound, an unlisted singular one is evaluated.
        const double a0 = 1.0 - t.dweights[d][3], a1 = (double)t.dweights[d][3] - t.dweights[d][2],
                     a2 = t.dweights[d][2];
        const double k0 = 1.0 - w2_3, k1 = (double)w2_3 - w1_3, k2 = w1_3;
        int count = 0;
        for (int k = 0; k < 64; k++) t.singular_mask[d][k] = 0;
        for (int k = 0; k < 64; k++) t.singular[d][k] = 0;
        for (int i = 0; i < t.four_total[15] && count < 64; i++) {
            const uint32_t e = t.four[i];
            const double s0 = (e >> 21) & 31u, s1 = (e >> 26) & 31u, s2 = (e >> 16) & 31u;
            const double A = a0 * s0 + a1 * s1 + a2 * s2, Wa = k0 * s0 + k1 * s1 + k2 * s2, W = 16.0;
            if (A * W - Wa * Wa < A * W / 16.0 * 1.001) {
                t.singular[d][count++] = (uint16_t)i;
                t.singular_mask[d][i & 63] |= (uint16_t)(1u << (i >> 6));
            }
        }
        t.singular_count[d] = count;
    }
}

// Makes sure the current device holds the tables; returns the decoder.
// Per-device memory pool for the library's stream-ordered scratch (workspace,
// job tables, staging): it keeps up to kPoolKeepBytes reserved across stream
// synchronisations, so repeated calls reuse their memory instead of mapping
// it again (the default pool hands everything back at each synchronisation).
constexpr uint64_t kPoolKeepBytes = 8ull << 30;
std::vector<std::pair<int, hipMemPool_t>> g_pools;

hipMemPool_t pool_for(int dev) {
    for (auto& p : g_pools)
        if (p.first == dev) return p.second;
    return nullptr;
}

int resident_workgroups(uint32_t* grid);

int ensure_device_ready(int* decoder) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_decoder < 0) return fail(ICBC_ERR_NOT_INITIALIZED, "icbc_gpu_init has not been called");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (std::find(g_uploaded.begin(), g_uploaded.end(), dev) == g_uploaded.end()) {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return fail(ICBC_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950", dev,
                        prop.gcnArchName);
        hipError_t e = upload_tables(g_tables, nullptr);
        if (e != hipSuccess) return hip_fail(e, "upload_tables");
        // the no-allocation entry point's scratch address on this device, and
        // the launch caches it relies on, are settled here, once per device
        e = noalloc_scratch_setup(dev);
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress (no-allocation scratch)");
        e = warm_launch_caches();
        if (e != hipSuccess) return hip_fail(e, "launch caches");
        uint32_t grid;
        const int r = resident_workgroups(&grid);  // fills this device's residency cache
        if (r) return r;
        if (!pool_for(dev)) {
            hipMemPoolProps props;
            memset(&props, 0, sizeof props);
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t pool = nullptr;
            e = hipMemPoolCreate(&pool, &props);
            if (e != hipSuccess) return hip_fail(e, "hipMemPoolCreate");
            uint64_t keep = kPoolKeepBytes;
            e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            if (e != hipSuccess) return hip_fail(e, "hipMemPoolSetAttribute");
            g_pools.push_back({dev, pool});
        }
        g_uploaded.push_back(dev);
    }
    *decoder = g_decoder;
    return ICBC_OK;
}

// Like ensure_device_ready, but for a device that must already be set up (by
// icbc_gpu_init with it current, or by an earlier call on it): never takes
// the set-up path, which allocates.
int device_ready_no_setup(int* decoder, int* device) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_decoder < 0) return fail(ICBC_ERR_NOT_INITIALIZED, "icbc_gpu_init has not been called");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (std::find(g_uploaded.begin(), g_uploaded.end(), dev) == g_uploaded.end())
        return fail(ICBC_ERR_NOT_INITIALIZED, "device %d is not set up: call icbc_gpu_init (or any other encode "
                    "entry point) with it current first", dev);
    *decoder = g_decoder;
    *device = dev;
    return ICBC_OK;
}

// Resident encode workgroups the current device holds at once.
int resident_workgroups(uint32_t* grid) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    static std::mutex m;
    static std::vector<std::pair<int, uint32_t>> cache;
    std::lock_guard<std::mutex> lock(m);
    for (auto& c : cache)
        if (c.first == dev) {
            *grid = c.second;
            return ICBC_OK;
        }
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    const uint32_t g = (uint32_t)prop.multiProcessorCount * (uint32_t)encode_workgroups_per_cu();
    cache.push_back({dev, g});
    *grid = g;
    return ICBC_OK;
}

// Stream-ordered allocation from the current device's pool.
hipError_t scratch_alloc(void** p, size_t bytes, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipMemPool_t pool = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        pool = pool_for(dev);
    }
    return pool ? hipMallocFromPoolAsync(p, bytes, pool, s) : hipMallocAsync(p, bytes, s);
}

// The level-9 handoff list laid out at `p` (refine_list_bytes() bytes).
RefineList refine_list_at(void* p) {
    return RefineList{(RefineItem*)((char*)p + kRefineCounterBytes), (uint32_t*)p};
}

// Stream-ordered workspace for one encode launch: the tile-queue counters,
// `grid` workgroup slices and room to hand `level9` blocks to refinement.
struct Workspace {
    void* base = nullptr;
    uint32_t grid = 0;
    float4* slices() const { return (float4*)((char*)base + kQueueBytes); }
    uint32_t* counter() const { return (uint32_t*)base; }
    RefineList refine() const {
        return refine_list_at((char*)base + kQueueBytes + (size_t)grid * kWorkspaceBytesPerWorkgroup);
    }
};

int workspace_alloc(uint32_t grid, uint64_t level9, hipStream_t s, Workspace& w) {
    w.grid = grid;
    HIP_TRY(scratch_alloc((void**)&w.base, kQueueBytes + (size_t)grid * kWorkspaceBytesPerWorkgroup + refine_list_bytes(level9),
                           s));
    return ICBC_OK;
}

// ---------------------------------------------------------------------------
// jobs

// Checks an image descriptor and converts it for the kernels.
int image_desc(const icbc_image* im, ImageDev& out, const char* what) {
    if (!im) return fail(ICBC_ERR_INVALID_ARGUMENT, "%s: null image", what);
    if (im->format != ICBC_IMAGE_RGBA8 && im->format != ICBC_IMAGE_RGBA32F)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%s: unknown image format %u", what, im->format);
    if (im->edge != ICBC_EDGE_REPLICATE && im->edge != ICBC_EDGE_ROUND_DOWN)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%s: unknown edge mode %u", what, im->edge);
    if (!im->data || im->width == 0 || im->height == 0)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%s: empty image", what);
    const uint64_t tb = im->format == ICBC_IMAGE_RGBA8 ? 4 : 16;
    if (im->pitch < (uint64_t)im->width * tb || im->pitch % tb != 0 || ((uintptr_t)im->data % tb) != 0)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%s: pitch must cover a row and be a multiple of %u bytes, "
                    "data aligned to it", what, (unsigned)tb);
    out.data = (const uint8_t*)im->data;
    out.pitch = im->pitch;
    out.format = im->format;
    out.width = im->width;
    out.height = im->height;
    out.edge = im->edge;
    return ICBC_OK;
}

int validate(const icbc_bc1_job& j, size_t k) {
    if ((j.flags & ~(ICBC_JOB_HOST_MEMORY | ICBC_JOB_COLORS_RGBA8 | ICBC_JOB_IMAGE | ICBC_JOB_WEIGHTS_FROM_ALPHA)) != 0)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: unknown flags", k);
    if ((j.flags & ICBC_JOB_WEIGHTS_FROM_ALPHA) && (!(j.flags & ICBC_JOB_IMAGE) || j.weights))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: weights from alpha need an image source and null weights", k);
    if (j.flags & ICBC_JOB_IMAGE) {
        if (j.flags & (ICBC_JOB_HOST_MEMORY | ICBC_JOB_COLORS_RGBA8))
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: image sources are device images, not host or RGBA8 blocks", k);
        if (j.colors) return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: image job with colors set", k);
        ImageDev im;
        char what[32];
        snprintf(what, sizeof what, "job %zu", k);
        int r = image_desc(j.image, im, what);
        if (r) return r;
        if (image_block_count(im) != j.block_count)
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: block_count %llu, image has %llu blocks", k,
                        (unsigned long long)j.block_count, (unsigned long long)image_block_count(im));
    } else if (j.image) {
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: image set without ICBC_JOB_IMAGE", k);
    }
    if (j.tile_blocks != 0 && j.tile_blocks != 8 && j.tile_blocks != 16 && j.tile_blocks != 32 &&
        j.tile_blocks != 64)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: tile_blocks must be 0, 8, 16, 32 or 64", k);
    if (j.block_count == 0) return ICBC_OK;
    if ((!j.colors && !(j.flags & ICBC_JOB_IMAGE)) || !j.codes)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: null colors or codes", k);
    if (!(j.flags & ICBC_JOB_HOST_MEMORY)) {
        if (((uintptr_t)j.colors & 15) || ((uintptr_t)j.weights & 15))
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: colors/weights must be 16-byte aligned", k);
        if (((uintptr_t)j.codes & 7) || ((uintptr_t)j.errors & 3) || ((uintptr_t)j.stats & 15))
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: misaligned codes/errors/stats", k);
    }
    if (j.block_count > (uint64_t)0xffffffffu)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: more than 2^32-1 blocks", k);
    return ICBC_OK;
}

// Per-device tuning tables (icbc_gpu_autotune / icbc_gpu_set_tuning).
std::mutex g_tuning_mutex;
std::vector<std::pair<int, icbc_tuning>> g_tuning;

bool tuning_for_current(icbc_tuning* out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    for (auto& t : g_tuning)
        if (t.first == dev) {
            *out = t.second;
            for (int l = 0; l < 10; l++)
                if (out->tile_blocks[l]) return true;
            return false;
        }
    return false;
}

void set_tuning_for_current(const icbc_tuning* t) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_tuning_mutex);
    for (auto& e : g_tuning)
        if (e.first == dev) {
            if (t) e.second = *t;
            else memset(&e.second, 0, sizeof e.second);
            return;
        }
    icbc_tuning z;
    memset(&z, 0, sizeof z);
    g_tuning.push_back({dev, t ? *t : z});
}

// Blocks per tile when the caller leaves it to the library.  With enough work
// for at least two 64-block tiles per resident workgroup, use 64 (shrinking
// only for jobs smaller than a tile).  With one or two, use 32: a workgroup
// with a single tile leaves the launch waiting for the slowest tile, while
// smaller tiles keep fewer lanes busy in the per-block stages (2048x2048 at
// level 9, 4,096 tiles of 64 for 4,096 workgroups: 32-block tiles 3.6 %
// faster, 16-block tiles 1.5 % slower).  Below that, split the blocks over
// about one tile per resident workgroup, since a tile's blocks are searched
// one after another by a single wavefront and small tiles finish sooner.
uint32_t auto_tile_blocks_impl(uint64_t job_blocks, uint64_t total_blocks, uint32_t resident) {
    uint32_t t;
    if (total_blocks >= 128ull * resident) {
        t = 64;
    } else if (total_blocks > 32ull * resident) {
        t = 32;
    } else {
        const uint64_t want = (total_blocks + resident - 1) / std::max<uint32_t>(resident, 1);
        t = 8;
        while (t < 64 && t < want) t *= 2;
        return t;
    }
    while (t > 8 && t / 2 >= job_blocks) t /= 2;
    return t;
}

// The four-cluster search schedule tuned for `level` on the current device
// (0: the default).
int32_t prune_for_level(int level) {
    if (level < 0 || level > 9) return 0;
    icbc_tuning t;
    memset(&t, 0, sizeof t);
    (void)tuning_for_current(&t);
    return t.prune_min_colors[level];
}

JobDev to_dev(const icbc_bc1_job& j, int decoder) {
    JobDev d;
    d.colors = reinterpret_cast<const float4*>(j.colors);
    d.weights = reinterpret_cast<const float4*>(j.weights);
    d.codes = reinterpret_cast<uint2*>(j.codes);
    d.errors = j.errors;
    d.stats = reinterpret_cast<int4*>(j.stats);
    d.block_count = j.block_count;
    d.level = j.level;
    d.tcm = j.three_color_mode ? 1 : 0;
    d.tcb = j.three_color_black ? 1 : 0;
    d.decoder = decoder;
    d.cw0 = j.color_weights[0];
    d.cw1 = j.color_weights[1];
    d.cw2 = j.color_weights[2];
    d.tile_blocks = j.tile_blocks ? (int32_t)j.tile_blocks : kTile;
    d.rgba8 = (j.flags & ICBC_JOB_COLORS_RGBA8) ? 1 : 0;  // the kernels read RGBA8 texels directly
    d.prune_min = prune_for_level(j.level);
    return d;
}

uint32_t tiles_of(uint64_t blocks, uint32_t tile_blocks) {
    return (uint32_t)((blocks + tile_blocks - 1) / tile_blocks);
}

// Host blob kept alive until the stream has consumed it.
void free_host_blob(void* p) { free(p); }

// Copies `bytes` from host into fresh stream-ordered device memory.
int upload_blob(const void* src, size_t bytes, hipStream_t s, void** dptr) {
    HIP_TRY(scratch_alloc((void**)dptr, bytes, s));
    void* keep = malloc(bytes);
    if (!keep) return fail(ICBC_ERR_OUT_OF_MEMORY, "host allocation of %zu bytes", bytes);
    memcpy(keep, src, bytes);
    hipError_t e = hipMemcpyAsync(*dptr, keep, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        free(keep);
        return hip_fail(e, "hipMemcpyAsync");
    }
    e = hipLaunchHostFunc(s, free_host_blob, keep);
    if (e != hipSuccess) return hip_fail(e, "hipLaunchHostFunc");
    return ICBC_OK;
}

// Host-memory jobs: device copies of the job's buffers.
struct Staged {
    void* colors = nullptr;
    void* weights = nullptr;
    void* codes = nullptr;
    void* errors = nullptr;
    void* stats = nullptr;
};

int stage_in(const icbc_bc1_job& j, hipStream_t s, Staged& st, icbc_bc1_job& dj) {
    dj = j;
    dj.flags &= ~ICBC_JOB_HOST_MEMORY;
    const uint64_t n = j.block_count;
    if (n == 0) return ICBC_OK;
    const size_t cbytes = n * ((j.flags & ICBC_JOB_COLORS_RGBA8) ? 64 : 256);
    HIP_TRY(scratch_alloc((void**)&st.colors, cbytes, s));
    HIP_TRY(hipMemcpyAsync(st.colors, j.colors, cbytes, hipMemcpyHostToDevice, s));
    dj.colors = (const float*)st.colors;
    if (j.weights) {
        HIP_TRY(scratch_alloc((void**)&st.weights, n * 64, s));
        HIP_TRY(hipMemcpyAsync(st.weights, j.weights, n * 64, hipMemcpyHostToDevice, s));
        dj.weights = (const float*)st.weights;
    }
    // codes go in too: blocks the reference leaves untouched keep their bytes
    HIP_TRY(scratch_alloc((void**)&st.codes, n * 8, s));
    HIP_TRY(hipMemcpyAsync(st.codes, j.codes, n * 8, hipMemcpyHostToDevice, s));
    dj.codes = st.codes;
    if (j.errors) {
        HIP_TRY(scratch_alloc((void**)&st.errors, n * 4, s));
        dj.errors = (float*)st.errors;
    }
    if (j.stats) {
        HIP_TRY(scratch_alloc((void**)&st.stats, n * 16, s));
        dj.stats = (int32_t*)st.stats;
    }
    return ICBC_OK;
}

int stage_out(const icbc_bc1_job& j, hipStream_t s, Staged& st) {
    const uint64_t n = j.block_count;
    if (n == 0) return ICBC_OK;
    HIP_TRY(hipMemcpyAsync(j.codes, st.codes, n * 8, hipMemcpyDeviceToHost, s));
    if (j.errors) HIP_TRY(hipMemcpyAsync(j.errors, st.errors, n * 4, hipMemcpyDeviceToHost, s));
    if (j.stats) HIP_TRY(hipMemcpyAsync(j.stats, st.stats, n * 16, hipMemcpyDeviceToHost, s));
    void* ptrs[5] = {st.colors, st.weights, st.codes, st.errors, st.stats};
    for (void* p : ptrs)
        if (p) HIP_TRY(hipFreeAsync(p, s));
    return ICBC_OK;
}

// An image job: its 4x4 blocks are gathered into the encoder's float layout
// before the encode launch (RGBA8 blocks need no pass: the kernels read them).
struct SourceConvert {
    ImageDev img{};            // image source
    float4* colors = nullptr;  // float blocks written here
    float* weights = nullptr;  // alpha as weight (ICBC_JOB_WEIGHTS_FROM_ALPHA)
    uint64_t blocks = 0;
};

bool needs_convert(const icbc_bc1_job& j) { return j.block_count && (j.flags & ICBC_JOB_IMAGE); }

// Bytes of float blocks (and weights) a converted job needs.
size_t convert_bytes(const icbc_bc1_job& j) {
    return (size_t)j.block_count * ((j.flags & ICBC_JOB_WEIGHTS_FROM_ALPHA) ? 256 + 64 : 256);
}

// Describes the conversion of job `j` into `buf` (convert_bytes(j) bytes) and
// points the device job at the converted texels.
SourceConvert describe_convert(const icbc_bc1_job& j, void* buf, JobDev& d) {
    SourceConvert c;
    c.blocks = j.block_count;
    c.colors = (float4*)buf;
    (void)image_desc(j.image, c.img, "job");  // validated before
    if (j.flags & ICBC_JOB_WEIGHTS_FROM_ALPHA) {
        c.weights = (float*)((char*)buf + (size_t)j.block_count * 256);
        d.weights = (const float4*)c.weights;
    }
    d.colors = c.colors;
    d.rgba8 = 0;
    return c;
}

hipError_t run_convert(const SourceConvert& c, hipStream_t s) {
    return launch_image_gather(c.img, c.colors, c.weights, s);
}

// Shared front half of batch/plain: validate, stage host jobs, build JobDev list.
struct Prepared {
    std::vector<icbc_bc1_job> jobs;  // device-pointer view
    std::vector<Staged> staged;      // per job (empty for device jobs)
    std::vector<JobDev> dev;
    std::vector<void*> temps;        // float blocks gathered from image inputs, freed after the launch
    bool any_host = false;
};

// Image jobs: gather their texels into float RGBA blocks in stream-ordered
// scratch and point the device job at them.
int convert_sources(Prepared& p, hipStream_t s) {
    for (size_t k = 0; k < p.jobs.size(); k++) {
        const icbc_bc1_job& j = p.jobs[k];
        if (!needs_convert(j)) continue;
        void* f = nullptr;
        HIP_TRY(scratch_alloc((void**)&f, convert_bytes(j), s));
        p.temps.push_back(f);
        HIP_TRY(run_convert(describe_convert(j, f, p.dev[k]), s));
    }
    return ICBC_OK;
}

int free_temps(Prepared& p, hipStream_t s) {
    for (void* t : p.temps) HIP_TRY(hipFreeAsync(t, s));
    p.temps.clear();
    return ICBC_OK;
}

int prepare(const icbc_bc1_job* jobs, size_t n_jobs, hipStream_t s, int decoder, Prepared& p) {
    if (n_jobs && !jobs) return fail(ICBC_ERR_INVALID_ARGUMENT, "null job array");
    for (size_t k = 0; k < n_jobs; k++) {
        int r = validate(jobs[k], k);
        if (r) return r;
    }
    p.jobs.resize(n_jobs);
    p.staged.resize(n_jobs);
    p.dev.resize(n_jobs);
    for (size_t k = 0; k < n_jobs; k++) {
        if (jobs[k].flags & ICBC_JOB_HOST_MEMORY) {
            p.any_host = true;
            int r = stage_in(jobs[k], s, p.staged[k], p.jobs[k]);
            if (r) return r;
        } else {
            p.jobs[k] = jobs[k];
        }
        p.dev[k] = to_dev(p.jobs[k], decoder);
    }
    return convert_sources(p, s);
}

int finish(const icbc_bc1_job* jobs, size_t n_jobs, hipStream_t s, Prepared& p) {
    int r = free_temps(p, s);
    if (r) return r;
    if (!p.any_host) return ICBC_OK;
    for (size_t k = 0; k < n_jobs; k++)
        if (jobs[k].flags & ICBC_JOB_HOST_MEMORY) {
            int r = stage_out(jobs[k], s, p.staged[k]);
            if (r) return r;
        }
    HIP_TRY(hipStreamSynchronize(s));
    return ICBC_OK;
}

// Fills in automatic tile sizes for a set of jobs launched together.
int choose_tiles(std::vector<JobDev>& dev, const icbc_bc1_job* jobs, bool per_job_launch) {
    uint32_t resident;
    int r = resident_workgroups(&resident);
    if (r) return r;
    uint64_t total = 0;
    for (const JobDev& d : dev) total += d.block_count;
    icbc_tuning tune;
    const bool tuned = tuning_for_current(&tune);
    for (size_t k = 0; k < dev.size(); k++)
        if (jobs[k].tile_blocks == 0) {
            const uint64_t launch_blocks = per_job_launch ? dev[k].block_count : total;
            uint32_t t = auto_tile_blocks_impl(dev[k].block_count, launch_blocks, resident);
            // a launch that fills every workgroup many times over takes the tuned
            // tile of its level, still no larger than the job needs
            const int lv = dev[k].level;
            if (tuned && lv >= 0 && lv <= 9 && tune.tile_blocks[lv] && launch_blocks >= 128ull * resident) {
                t = tune.tile_blocks[lv];
                while (t > 8 && t / 2 >= dev[k].block_count) t /= 2;
            }
            dev[k].tile_blocks = (int32_t)t;
        }
    return ICBC_OK;
}

// Relative cost of one block at each level, used to issue expensive tiles
// first (larger candidate searches and refinement at the top levels).
// Without a tuning, a fixed estimate; with one, the measured ns per block of
// every level present in the launch (when all of them are tuned).
float static_level_cost(int level) {
    static const float cost[10] = {1, 2, 40, 40, 60, 60, 60, 60, 70, 110};
    return (level >= 0 && level <= 9) ? cost[level] : 0.0f;
}

// Jobs in launch order: with `by_cost`, largest per-tile cost first (cost per
// block of the level, measured by the autotuner where it has timed every level
// present, else a static estimate, times the tile size), ties in submission
// order; otherwise submission order.
std::vector<uint32_t> job_order(const std::vector<JobDev>& dev, bool by_cost) {
    std::vector<uint32_t> order(dev.size());
    for (size_t k = 0; k < dev.size(); k++) order[k] = (uint32_t)k;
    if (!by_cost) return order;
    float cost[10];
    for (int l = 0; l < 10; l++) cost[l] = static_level_cost(l);
    icbc_tuning tune;
    if (tuning_for_current(&tune)) {
        bool all = true;
        for (const JobDev& d : dev)
            if (d.level >= 0 && d.level <= 9 && !(tune.ns_per_block[d.level] > 0.0f)) all = false;
        if (all)
            for (int l = 0; l < 10; l++) cost[l] = tune.ns_per_block[l];
    }
    auto c = [&](uint32_t k) {
        const JobDev& d = dev[k];
        return ((d.level >= 0 && d.level <= 9) ? cost[d.level] : 0.0f) * d.tile_blocks;
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return c(a) > c(b); });
    return order;
}

uint64_t tile_count(const std::vector<JobDev>& dev) {
    uint64_t total = 0;
    for (const JobDev& d : dev) total += tiles_of(d.block_count, d.tile_blocks);
    return total;
}

// Launch tails.  When the resident workgroups get only a few tiles each (one
// large image, say), a launch ends with workgroups finishing their last
// 64-block tiles one after another while the others idle.  The final blocks
// of the job launched last (the cheapest per tile) then become a job of their
// own with half-size tiles, issued after every other tile: a tile list that
// ends in smaller tiles ends more evenly.  Measured on the 4096x4096 image
// (16,384 tiles for 4,096 resident workgroups) with its last 65,536 blocks in
// 32-block tiles, interleaved runs: level 5 -1.1 %, level 8 -1.0 %, level 9
// -0.4 %; levels 0-1, whose tiles take microseconds, +3.9 %, so they are
// never split.  The split job addresses the same buffers from an offset and
// its tiles are encoded exactly as before, so results do not change.
void split_tail(std::vector<JobDev>& dev, uint32_t resident) {
    const uint64_t tiles = tile_count(dev);
    if (dev.empty() || tiles < 2ull * resident || tiles >= 16ull * resident) return;
    const uint32_t k = job_order(dev, true).back();
    JobDev& d = dev[k];
    if (d.tile_blocks < 16 || d.level < 2 || d.level > 9) return;
    const uint32_t t = (uint32_t)d.tile_blocks;
    // half as many half-size tiles as resident workgroups, a whole number of
    // the job's tiles, at most half the job
    const uint64_t tail = std::min<uint64_t>((uint64_t)(resident / 2) * (t / 2), d.block_count / 2) / t * t;
    if (tail == 0) return;
    const uint64_t keep = d.block_count - tail;
    JobDev e = d;
    d.block_count = keep;
    e.block_count = tail;
    e.tile_blocks = (int32_t)(t / 2);
    e.colors = d.colors + keep * (d.rgba8 ? 4 : 16);
    if (d.weights) e.weights = d.weights + keep * 4;
    e.codes = d.codes + keep;
    if (d.errors) e.errors = d.errors + keep;
    if (d.stats) e.stats = d.stats + keep;
    dev.push_back(e);
}

// The tile list (job, first block): every tile of a job together, jobs in
// job_order().  Tiles of one job cost the same, so this is the list sorted by
// per-tile cost with ties kept in submission order.
std::vector<uint2> make_tiles(const std::vector<JobDev>& dev, bool by_cost) {
    std::vector<uint2> tiles;
    tiles.reserve(tile_count(dev));
    for (uint32_t k : job_order(dev, by_cost))
        for (uint32_t t = 0; t < tiles_of(dev[k].block_count, dev[k].tile_blocks); t++)
            tiles.push_back(make_uint2(k, t * (uint32_t)dev[k].tile_blocks));
    return tiles;
}

// One launch over part of a job table: its kernel instance (encode_mode) and
// its range of the tile list.
struct SubLaunch {
    int mode;
    uint32_t first, count;
};

// A launch runs one kernel instance for all its jobs: levels 0-1 beside
// cluster levels run in the cluster instance's box-fit stage, 3.5x slower per
// block than a level 0-1 instance (level 1 on the 4096x4096 image: 0.36
// against 0.10 ms with the LDS-staged one), and levels 4-9
// beside levels 2-3 run the instance that carries the box-fit stage (more
// live registers, 1-2 % slower).  So jobs are grouped by the instance their
// level needs -- levels 4-9, 2-3, 0-1 (invalid levels with 0-1) -- and a
// group of at least kSplitMinWaves waves of resident workgroups gets a launch
// of its own; a smaller one joins another group's launch instead, where it
// adds no launch tail of its own.  Returns the tile list (the launches' tiles
// one after another, each in job_order) and the launches, in `out`.
constexpr uint32_t kSplitMinWaves = 2;

std::vector<uint2> make_launches(const std::vector<JobDev>& dev, bool by_cost, uint32_t resident,
                                 std::vector<SubLaunch>& out) {
    auto group_of = [](int level) { return (level >= 4 && level <= 9) ? 0 : (level >= 2 && level <= 3) ? 1 : 2; };
    std::vector<uint32_t> group(dev.size());
    uint64_t tiles[3] = {0, 0, 0};
    for (size_t k = 0; k < dev.size(); k++) {
        group[k] = (uint32_t)group_of(dev[k].level);
        tiles[group[k]] += tiles_of(dev[k].block_count, dev[k].tile_blocks);
    }
    const uint64_t big = (uint64_t)kSplitMinWaves * resident;
    auto merge = [&](int from, int into) {
        for (auto& g : group)
            if (g == (uint32_t)from) g = (uint32_t)into;
        tiles[into] += tiles[from];
        tiles[from] = 0;
    };
    const int present = (tiles[0] > 0) + (tiles[1] > 0) + (tiles[2] > 0);
    if (present > 1) {
        if (tiles[2] && tiles[2] < big) merge(2, 1);             // levels 0-1 into the box-stage launch
        if (tiles[1] && tiles[1] < big && tiles[0]) merge(1, 0);  // and that into the levels 4-9 launch
        if (tiles[0] && tiles[0] < big && tiles[1]) merge(0, 1);
    }
    std::vector<uint2> list;
    list.reserve(tile_count(dev));
    out.clear();
    const std::vector<uint32_t> order = job_order(dev, by_cost);
    for (uint32_t g = 0; g < 3; g++) {
        const uint32_t first = (uint32_t)list.size();
        std::vector<JobDev> members;
        for (uint32_t k : order) {
            if (group[k] != g) continue;
            members.push_back(dev[k]);
            for (uint32_t t = 0; t < tiles_of(dev[k].block_count, dev[k].tile_blocks); t++)
                list.push_back(make_uint2(k, t * (uint32_t)dev[k].tile_blocks));
        }
        if (list.size() > first)
            out.push_back(SubLaunch{encode_mode(members.data(), members.size()), first, (uint32_t)list.size() - first});
    }
    return list;
}

}  // namespace

// ===========================================================================
// internal interface (icbc_internal.h)

int icbc_internal_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vfail(code, fmt, ap);
    va_end(ap);
    return code;
}

int icbc_internal_ready(int* decoder) { return ensure_device_ready(decoder); }

int icbc_internal_resident(uint32_t* grid) { return resident_workgroups(grid); }

int icbc_internal_validate(const icbc_bc1_job& job) { return validate(job, 0); }

bool icbc_internal_get_tuning(icbc_tuning* out) { return tuning_for_current(out); }

void icbc_internal_set_tuning(const icbc_tuning* t) { set_tuning_for_current(t); }

namespace icbc_gpu {

JobDev job_to_device(const icbc_bc1_job& j, int decoder) { return to_dev(j, decoder); }

uint32_t auto_tile_blocks(uint64_t job_blocks, uint64_t total_blocks, uint32_t resident) {
    return auto_tile_blocks_impl(job_blocks, total_blocks, resident);
}

std::vector<uint2> tiles_for(const std::vector<JobDev>& dev, bool order_by_cost) {
    return make_tiles(dev, order_by_cost);
}

}  // namespace icbc_gpu

// ===========================================================================
// C ABI

extern "C" {

int icbc_gpu_init(int decoder) {
    if (decoder < 0 || decoder > 3) return fail(ICBC_ERR_INVALID_ARGUMENT, "decoder %d", decoder);
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (!g_tables_built) {
            build_tables(g_tables);
            g_tables_built = true;
        }
        g_decoder = decoder;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(ICBC_ERR_NO_DEVICE, "no HIP device visible");
    int d;
    return ensure_device_ready(&d);
}

int icbc_gpu_compress_bc1_batch(const icbc_bc1_job* jobs, size_t n_jobs, void* stream) {
    int decoder;
    int r = ensure_device_ready(&decoder);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    Prepared p;
    r = prepare(jobs, n_jobs, s, decoder, p);
    if (r) return r;
    r = choose_tiles(p.dev, jobs, false);
    if (r) return r;
    uint32_t resident;
    r = resident_workgroups(&resident);
    if (r) return r;
    split_tail(p.dev, resident);
    const uint64_t ntiles = tile_count(p.dev);
    if (ntiles > 0xffffffffu) return fail(ICBC_ERR_INVALID_ARGUMENT, "batch too large");
    if (ntiles) {
        const uint32_t grid = std::min<uint32_t>(resident, (uint32_t)ntiles);
        void *djobs = nullptr, *dtiles = nullptr;
        r = upload_blob(p.dev.data(), p.dev.size() * sizeof(JobDev), s, &djobs);
        if (r) return r;
        // one job: its tiles are implied (tile k = blocks [T k, T k + T)), no list
        std::vector<SubLaunch> subs{{encode_mode(p.dev.data(), p.dev.size()), 0, (uint32_t)ntiles}};
        if (p.dev.size() > 1) {
            const std::vector<uint2> tiles = make_launches(p.dev, true, resident, subs);
            r = upload_blob(tiles.data(), tiles.size() * sizeof(uint2), s, &dtiles);
            if (r) return r;
        }
        Workspace w;
        r = workspace_alloc(grid, level9_blocks(p.dev.data(), p.dev.size()), s, w);
        if (r) return r;
        for (const SubLaunch& sl : subs)
            HIP_TRY(launch_encode((const JobDev*)djobs, dtiles ? (const uint2*)dtiles + sl.first : nullptr, sl.count,
                                  w.counter(), w.slices(), std::min(grid, sl.count), sl.mode, w.refine(), s));
        HIP_TRY(hipFreeAsync(w.base, s));
        HIP_TRY(hipFreeAsync(djobs, s));
        if (dtiles) HIP_TRY(hipFreeAsync(dtiles, s));
    }
    return finish(jobs, n_jobs, s, p);
}

int icbc_gpu_compress_bc1_plain(const icbc_bc1_job* jobs, size_t n_jobs, void* stream) {
    int decoder;
    int r = ensure_device_ready(&decoder);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    Prepared p;
    r = prepare(jobs, n_jobs, s, decoder, p);
    if (r) return r;
    if (n_jobs) {
        uint32_t resident;
        r = resident_workgroups(&resident);
        if (r) return r;
        r = choose_tiles(p.dev, jobs, true);
        if (r) return r;
        uint32_t grid = 1;
        for (const JobDev& d : p.dev) grid = std::max(grid, std::min(resident, tiles_of(d.block_count, d.tile_blocks)));
        void* djobs = nullptr;
        r = upload_blob(p.dev.data(), p.dev.size() * sizeof(JobDev), s, &djobs);
        if (r) return r;
        Workspace w;
        uint64_t l9 = 0;  // one launch at a time reuses the handoff list
        for (const JobDev& d : p.dev) l9 = std::max<uint64_t>(l9, level9_blocks(&d, 1));
        r = workspace_alloc(grid, l9, s, w);
        if (r) return r;
        for (size_t k = 0; k < n_jobs; k++)
            HIP_TRY(launch_encode((const JobDev*)djobs + k, nullptr, tiles_of(p.dev[k].block_count, p.dev[k].tile_blocks),
                                  nullptr, w.slices(), grid, encode_mode(&p.dev[k], 1), w.refine(), s));
        HIP_TRY(hipFreeAsync(w.base, s));
        HIP_TRY(hipFreeAsync(djobs, s));
    }
    return finish(jobs, n_jobs, s, p);
}

// The batch call without any allocation (no device, pinned or host heap
// memory): device-memory block jobs only, within the ICBC_NOALLOC_* limits.
// Job table, tile list, queue counters, level-9 handoff and workspaces live in
// the library's per-device scratch array (bc1_encoder.hip); the job table
// reaches it as a kernel argument.  Calls that overlap on one device share
// that array and race; the caller serialises them.
static_assert(kNoAllocMaxJobs == ICBC_NOALLOC_MAX_JOBS && kNoAllocMaxTiles == ICBC_NOALLOC_MAX_TILES &&
                  kNoAllocMaxLevel9Blocks == ICBC_NOALLOC_MAX_LEVEL9_BLOCKS && kNoAllocMaxGrid == ICBC_NOALLOC_MAX_GRID,
              "no-allocation limits must match icbc_gpu.h");

int icbc_gpu_compress_bc1_batch_noalloc(const icbc_bc1_job* jobs, size_t n_jobs, void* stream) {
    if (n_jobs > kNoAllocMaxJobs)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%zu jobs: the no-allocation call takes at most %u", n_jobs,
                    kNoAllocMaxJobs);
    if (n_jobs && !jobs) return fail(ICBC_ERR_INVALID_ARGUMENT, "null jobs");
    int decoder, dev;
    int r = device_ready_no_setup(&decoder, &dev);
    if (r) return r;
    void* base = noalloc_scratch(dev);
    if (!base) return fail(ICBC_ERR_NOT_INITIALIZED, "no scratch address for device %d", dev);
    NoAllocJobs a;
    memset(&a, 0, sizeof a);
    a.n = (uint32_t)n_jobs;
    uint64_t total = 0, level9 = 0;
    for (size_t k = 0; k < n_jobs; k++) {
        const icbc_bc1_job& j = jobs[k];
        r = validate(j, k);
        if (r) return r;
        if (j.flags & (ICBC_JOB_HOST_MEMORY | ICBC_JOB_IMAGE))
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: the no-allocation call takes device-memory block "
                        "jobs only (no host memory, no image source)", k);
        a.jobs[k] = to_dev(j, decoder);
        total += j.block_count;
        if (j.level == 9) level9 += j.block_count;
    }
    if (level9 > kNoAllocMaxLevel9Blocks)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%llu level-9 blocks: the no-allocation call takes at most %llu",
                    (unsigned long long)level9, (unsigned long long)kNoAllocMaxLevel9Blocks);
    uint32_t resident;
    r = resident_workgroups(&resident);
    if (r) return r;
    const uint32_t grid_max = std::min<uint32_t>(resident, kNoAllocMaxGrid);
    // tile sizes as in the batch call (tuned for large launches), for this
    // call's residency
    icbc_tuning tune;
    const bool tuned = tuning_for_current(&tune);
    for (size_t k = 0; k < n_jobs; k++)
        if (jobs[k].tile_blocks == 0) {
            uint32_t t = auto_tile_blocks_impl(a.jobs[k].block_count, total, grid_max);
            const int lv = a.jobs[k].level;
            if (tuned && lv >= 0 && lv <= 9 && tune.tile_blocks[lv] && total >= 128ull * grid_max) {
                t = tune.tile_blocks[lv];
                while (t > 8 && t / 2 >= a.jobs[k].block_count) t /= 2;
            }
            a.jobs[k].tile_blocks = (int32_t)t;
        }
    // launch order: dearest tiles first (as the batch call), ties in submission order
    float cost[10];
    for (int l = 0; l < 10; l++) cost[l] = static_level_cost(l);
    bool all_tuned = tuned;
    for (size_t k = 0; k < n_jobs && all_tuned; k++) {
        const int lv = a.jobs[k].level;
        if (lv >= 0 && lv <= 9 && !(tune.ns_per_block[lv] > 0.0f)) all_tuned = false;
    }
    if (all_tuned)
        for (int l = 0; l < 10; l++) cost[l] = tune.ns_per_block[l];
    float job_cost[kNoAllocMaxJobs];
    for (uint32_t k = 0; k < a.n; k++) {
        const JobDev& d = a.jobs[k];
        job_cost[k] = ((d.level >= 0 && d.level <= 9) ? cost[d.level] : 0.0f) * d.tile_blocks;
        uint32_t p = k;  // stable insertion by decreasing cost
        while (p > 0 && job_cost[a.order[p - 1]] < job_cost[k]) {
            a.order[p] = a.order[p - 1];
            p--;
        }
        a.order[p] = k;
    }
    uint64_t ntiles = 0;
    for (uint32_t p = 0; p < a.n; p++) {
        a.first_tile[p] = (uint32_t)std::min<uint64_t>(ntiles, 0xffffffffu);
        const JobDev& d = a.jobs[a.order[p]];
        ntiles += tiles_of(d.block_count, d.tile_blocks);
    }
    if (ntiles > kNoAllocMaxTiles)
        return fail(ICBC_ERR_INVALID_ARGUMENT, "%llu tiles: the no-allocation call takes at most %u",
                    (unsigned long long)ntiles, kNoAllocMaxTiles);
    a.first_tile[a.n] = (uint32_t)ntiles;
    if (ntiles == 0) return ICBC_OK;
    char* b = (char*)base;
    const hipStream_t s = (hipStream_t)stream;
    HIP_TRY(launch_noalloc_setup(a, base, s));
    const JobDev* djobs = (const JobDev*)(b + kNoAllocJobsOffset);
    const uint2* dtiles = n_jobs > 1 ? (const uint2*)(b + kNoAllocTilesOffset) : nullptr;
    const uint32_t grid = std::min<uint32_t>(grid_max, (uint32_t)ntiles);
    HIP_TRY(launch_encode(djobs, dtiles, (uint32_t)ntiles, (uint32_t*)b, (float4*)(b + kNoAllocWorkspaceOffset), grid,
                          encode_mode(a.jobs, a.n), refine_list_at(b + kNoAllocRefineOffset), s));
    return ICBC_OK;
}

uint64_t icbc_gpu_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

}  // extern "C"

// ---- plans ------------------------------------------------------------------

struct icbc_bc1_plan {
    int device = 0;
    icbc_plan_options opt{};
    JobDev* djobs = nullptr;
    uint2* dtiles = nullptr;
    void* dwork = nullptr;  // tile-queue counters (kQueueBytes) + per-workgroup slices
    std::vector<SourceConvert> converts;  // image jobs, gathered again on every execute
    std::vector<void*> convert_bufs;
    uint32_t ntiles = 0;
    uint32_t grid = 0;
    std::vector<SubLaunch> launches;  // kernel instance and tile range of each launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int plan_launch(icbc_bc1_plan* pl, hipStream_t s) {
    for (const auto& c : pl->converts) HIP_TRY(run_convert(c, s));
    uint32_t* counter = pl->opt.schedule == ICBC_SCHEDULE_PERSISTENT ? (uint32_t*)pl->dwork : nullptr;
    float4* slices = (float4*)((char*)pl->dwork + kQueueBytes);
    const RefineList rl = refine_list_at((char*)pl->dwork + kQueueBytes + (size_t)pl->grid * kWorkspaceBytesPerWorkgroup);
    // the counters were zeroed when the plan was made, and every launch leaves
    // them at zero: no reset kernel in the plan's launches
    for (const SubLaunch& sl : pl->launches)
        HIP_TRY(launch_encode(pl->djobs, pl->dtiles + sl.first, sl.count, counter, slices,
                              std::min(pl->grid, sl.count), sl.mode, rl, s, true));
    return ICBC_OK;
}

}  // namespace

extern "C" {

int icbc_gpu_plan_create(const icbc_bc1_job* jobs, size_t n_jobs, const icbc_plan_options* opt,
                         void* stream, icbc_bc1_plan** out) {
    if (!out) return fail(ICBC_ERR_INVALID_ARGUMENT, "null plan pointer");
    *out = nullptr;
    int decoder;
    int r = ensure_device_ready(&decoder);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    std::vector<JobDev> dev(n_jobs);
    for (size_t k = 0; k < n_jobs; k++) {
        r = validate(jobs[k], k);
        if (r) return r;
        if (jobs[k].flags & ICBC_JOB_HOST_MEMORY)
            return fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: plans take device memory only", k);
        dev[k] = to_dev(jobs[k], decoder);
    }
    r = choose_tiles(dev, jobs, false);
    if (r) return r;
    const uint64_t ntiles = tile_count(dev);
    if (ntiles > 0xffffffffu) return fail(ICBC_ERR_INVALID_ARGUMENT, "plan too large");
    icbc_plan_options o{ICBC_SCHEDULE_PERSISTENT, 1, 0, 1};
    if (opt) {
        o = *opt;
    } else {
        // defaults: the tuned occupancy and schedule of the level that carries
        // most of the plan's measured work
        icbc_tuning tune;
        if (tuning_for_current(&tune)) {
            double share[10] = {0};
            for (const JobDev& d : dev)
                if (d.level >= 0 && d.level <= 9) share[d.level] += (double)d.block_count * tune.ns_per_block[d.level];
            int top = -1;
            for (int l = 0; l < 10; l++)
                if (tune.tile_blocks[l] && (top < 0 || share[l] > share[top])) top = l;
            uint32_t resident = 0;
            if (top >= 0 && share[top] > 0 && resident_workgroups(&resident) == ICBC_OK) {
                o.workgroups_per_cu = tune.workgroups_per_cu[top];
                if (tune.schedule[top] == ICBC_SCHEDULE_PERSISTENT ||
                    ntiles <= kGridScheduleMaxWaves * (uint64_t)resident)
                    o.schedule = tune.schedule[top];
            }
        }
    }
    icbc_bc1_plan* pl = new icbc_bc1_plan();
    pl->opt = o;
    (void)hipGetDevice(&pl->device);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, pl->device);
    if (e != hipSuccess) {
        delete pl;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    auto cleanup = [&](int code) {
        icbc_gpu_plan_destroy(pl);
        return code;
    };
    // image jobs read gathered float blocks (their job entries point there)
    for (size_t k = 0; k < n_jobs; k++) {
        if (!needs_convert(jobs[k])) continue;
        void* f = nullptr;
        if ((e = hipMalloc(&f, convert_bytes(jobs[k]))) != hipSuccess) return cleanup(hip_fail(e, "hipMalloc (plan)"));
        pl->convert_bufs.push_back(f);
        pl->converts.push_back(describe_convert(jobs[k], f, dev[k]));
    }
    const int per_cu = o.workgroups_per_cu > 0 ? std::min(o.workgroups_per_cu, encode_workgroups_per_cu())
                                               : encode_workgroups_per_cu();
    const uint32_t resident = (uint32_t)prop.multiProcessorCount * (uint32_t)per_cu;
    if (o.schedule == ICBC_SCHEDULE_PERSISTENT && o.order_by_cost) split_tail(dev, resident);
    const std::vector<uint2> tiles = make_launches(dev, o.order_by_cost != 0, resident, pl->launches);
    pl->ntiles = (uint32_t)tiles.size();
    if (o.schedule == ICBC_SCHEDULE_GRID) {
        // one workgroup per tile, no queue
        pl->grid = std::max<uint32_t>(pl->ntiles, 1u);
    } else {
        pl->grid = std::min<uint32_t>(resident, std::max<uint32_t>(pl->ntiles, 1u));
    }
    if ((e = hipMalloc(&pl->djobs, std::max<size_t>(1, dev.size()) * sizeof(JobDev))) != hipSuccess ||
        (e = hipMalloc(&pl->dtiles, std::max<size_t>(1, tiles.size()) * sizeof(uint2))) != hipSuccess ||
        (e = hipMalloc(&pl->dwork, kQueueBytes + (size_t)pl->grid * kWorkspaceBytesPerWorkgroup +
                                       refine_list_bytes(level9_blocks(dev.data(), dev.size())))) != hipSuccess)
        return cleanup(hip_fail(e, "hipMalloc (plan)"));
    // the queue and handoff-list counters start at zero; the kernels leave them so
    if ((e = hipMemset(pl->dwork, 0, kQueueBytes)) != hipSuccess ||
        (e = hipMemset((char*)pl->dwork + kQueueBytes + (size_t)pl->grid * kWorkspaceBytesPerWorkgroup, 0,
                       kRefineCounterBytes)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemset (plan counters)"));
    if (!dev.empty() &&
        (e = hipMemcpy(pl->djobs, dev.data(), dev.size() * sizeof(JobDev), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemcpy (plan jobs)"));
    if (!tiles.empty() &&
        (e = hipMemcpy(pl->dtiles, tiles.data(), tiles.size() * sizeof(uint2), hipMemcpyHostToDevice)) != hipSuccess)
        return cleanup(hip_fail(e, "hipMemcpy (plan tiles)"));
    if (o.use_graph && pl->ntiles) {
        hipStream_t cs;
        if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess)
            return cleanup(hip_fail(e, "hipStreamCreate"));
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            r = plan_launch(pl, cs);
            hipError_t e2 = hipStreamEndCapture(cs, &pl->graph);
            if (r == ICBC_OK && e2 != hipSuccess) r = hip_fail(e2, "hipStreamEndCapture");
            if (r == ICBC_OK && (e2 = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0)) != hipSuccess)
                r = hip_fail(e2, "hipGraphInstantiate");
        } else {
            r = hip_fail(e, "hipStreamBeginCapture");
        }
        (void)hipStreamDestroy(cs);
        if (r) return cleanup(r);
    }
    (void)s;
    *out = pl;
    return ICBC_OK;
}

int icbc_gpu_plan_execute(icbc_bc1_plan* pl, void* stream) {
    if (!pl) return fail(ICBC_ERR_INVALID_ARGUMENT, "null plan");
    int decoder;
    int r = ensure_device_ready(&decoder);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    if (pl->ntiles == 0) return ICBC_OK;
    if (pl->exec) {
        HIP_TRY(hipGraphLaunch(pl->exec, s));
        return ICBC_OK;
    }
    return plan_launch(pl, s);
}

int icbc_gpu_plan_destroy(icbc_bc1_plan* pl) {
    if (!pl) return ICBC_OK;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->djobs) (void)hipFree(pl->djobs);
    if (pl->dtiles) (void)hipFree(pl->dtiles);
    if (pl->dwork) (void)hipFree(pl->dwork);
    for (void* b : pl->convert_bufs) (void)hipFree(b);
    delete pl;
    return ICBC_OK;
}

uint64_t icbc_gpu_plan_tiles(const icbc_bc1_plan* pl) { return pl ? pl->ntiles : 0; }

}  // extern "C"

// ---- single block -----------------------------------------------------------

namespace {

// Buffers of a thread's single-block decode / evaluate calls: a BC1 or BC3
// block (16 B) and 64 B of RGBA8 texels in, 64 B of RGBA8 or a float out.
constexpr size_t kBlockDeviceBytes = 256;
constexpr size_t kBlockHostBytes = 256;

struct HipBlockApi {
    int stream_create(void** s) {
        const hipError_t e = hipStreamCreateWithFlags((hipStream_t*)s, hipStreamNonBlocking);
        return e == hipSuccess ? ICBC_OK : hip_fail(e, "hipStreamCreateWithFlags (single block)");
    }
    void stream_destroy(void* s) { (void)hipStreamDestroy((hipStream_t)s); }
    int device_alloc(void** p) {
        const hipError_t e = hipMalloc(p, kBlockDeviceBytes);
        return e == hipSuccess ? ICBC_OK : hip_fail(e, "hipMalloc (single block)");
    }
    void device_free(void* p) { (void)hipFree(p); }
    int host_alloc(void** p) {
        const hipError_t e = hipHostMalloc(p, kBlockHostBytes, hipHostMallocDefault);
        return e == hipSuccess ? ICBC_OK : hip_fail(e, "hipHostMalloc (single block)");
    }
    void host_free(void* p) { (void)hipHostFree(p); }
};

// decode_bc1 / evaluate_bc1_error are reentrant in the reference; each calling
// thread gets its own stream and buffers per device (never released: process
// teardown may already have destroyed the runtime when thread-local
// destructors run).
thread_local std::vector<BlockResources> t_block;

int block_context(BlockResources*& ctx) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HipBlockApi api;
    return block_resources_for(api, t_block, dev, ctx);
}

// One block: the coalescing ring (block_ring.cpp) shares launches between
// concurrent callers.
int compress_one(int level, const float* input_colors, const float* input_weights,
                 const float color_weights[3], int three_color_mode, int three_color_black,
                 void* output, float* error) {
    return icbc_internal_compress_one(level, input_colors, input_weights, color_weights, three_color_mode,
                                      three_color_black, output, error);
}

// decode of one BC1 (8 B) or BC3 (16 B) block into 64 bytes of RGBA8, on the
// calling thread's single-block buffers.
int decode_one(const void* block, unsigned char* rgba, int decoder, bool bc3) {
    if (!block || !rgba) return fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    if (decoder < 0 || decoder > 3) return fail(ICBC_ERR_INVALID_ARGUMENT, "decoder %d", decoder);
    int d0;
    int r = ensure_device_ready(&d0);
    if (r) return r;
    BlockResources* ctx;
    r = block_context(ctx);
    if (r) return r;
    const hipStream_t stream = (hipStream_t)ctx->stream;
    char* h = (char*)ctx->hbuf;
    char* d = (char*)ctx->dbuf;
    memcpy(h, block, bc3 ? 16 : 8);
    HIP_TRY(hipMemcpyAsync(d, h, 16, hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_decode((const uint2*)d, 1, (uint32_t*)(d + 64), decoder, bc3, stream));
    HIP_TRY(hipMemcpyAsync(h + 64, d + 64, 64, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    memcpy(rgba, h + 64, 64);
    return ICBC_OK;
}

// evaluate_bc1_error / evaluate_bc3_error of one block.
int evaluate_one(const unsigned char* rgba, const void* block, int decoder, bool bc3, bool alpha_blend,
                 float* error) {
    if (!block || !rgba) return fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    if (decoder < 0 || decoder > 3) return fail(ICBC_ERR_INVALID_ARGUMENT, "decoder %d", decoder);
    int d0;
    int r = ensure_device_ready(&d0);
    if (r) return r;
    BlockResources* ctx;
    r = block_context(ctx);
    if (r) return r;
    const hipStream_t stream = (hipStream_t)ctx->stream;
    char* h = (char*)ctx->hbuf;
    char* d = (char*)ctx->dbuf;
    memcpy(h, rgba, 64);
    memcpy(h + 64, block, bc3 ? 16 : 8);
    HIP_TRY(hipMemcpyAsync(d, h, 80, hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_evaluate((const uint32_t*)d, (const uint2*)(d + 64), 1, (float*)(d + 80), decoder, bc3,
                            alpha_blend, stream));
    HIP_TRY(hipMemcpyAsync(h + 80, d + 80, 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    memcpy(error, h + 80, 4);
    return ICBC_OK;
}

}  // namespace

extern "C" {

float icbc_gpu_compress_bc1(int level, const float* input_colors, const float* input_weights,
                            const float color_weights[3], int three_color_mode, int three_color_black,
                            void* output) {
    float e = NAN;
    if (compress_one(level, input_colors, input_weights, color_weights, three_color_mode,
                     three_color_black, output, &e) != ICBC_OK)
        return NAN;
    return e;
}

// ---- decode / evaluate --------------------------------------------------------

// ---- texture pre-pass -------------------------------------------------------------

uint64_t icbc_gpu_image_block_count(const icbc_image* image) {
    ImageDev im;
    if (image_desc(image, im, "image")) return 0;
    return image_block_count(im);
}

int icbc_gpu_image_to_blocks(const icbc_image* image, float* colors, float* weights, void* stream) {
    ImageDev im;
    int r = image_desc(image, im, "image");
    if (r) return r;
    if (!colors || ((uintptr_t)colors & 15) || ((uintptr_t)weights & 3))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "colors must be 16-byte aligned device memory");
    int decoder;
    r = ensure_device_ready(&decoder);
    if (r) return r;
    HIP_TRY(launch_image_gather(im, (float4*)colors, weights, (hipStream_t)stream));
    return ICBC_OK;
}

uint64_t icbc_gpu_mip_chain_bytes(uint32_t width, uint32_t height, uint32_t* levels) {
    uint64_t bytes = 0;
    uint32_t n = 0;
    uint32_t w = width, h = height;
    while (w > 4 || h > 4) {  // down to the first level that fits one 4x4 block
        downsampled_size(w, h, w, h);
        bytes += (uint64_t)w * h * 16;
        n++;
    }
    if (levels) *levels = n;
    return bytes;
}

int icbc_gpu_build_mip_chain(const icbc_image* source, float* dst, uint64_t dst_bytes, icbc_image* levels,
                             uint32_t max_levels, uint32_t* n_levels, void* stream) {
    ImageDev src;
    int r = image_desc(source, src, "source");
    if (r) return r;
    uint32_t n = 0;
    const uint64_t need = icbc_gpu_mip_chain_bytes(src.width, src.height, &n);
    if (need && (!dst || dst_bytes < need || ((uintptr_t)dst & 15)))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "mip chain needs %llu bytes of 16-byte aligned device memory",
                    (unsigned long long)need);
    if (n > max_levels || (n && !levels))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "mip chain has %u levels, room for %u", n, max_levels);
    int decoder;
    r = ensure_device_ready(&decoder);
    if (r) return r;
    ImageDev cur = src;
    char* out = (char*)dst;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t w, h;
        downsampled_size(cur.width, cur.height, w, h);
        HIP_TRY(launch_downsample(cur, (float4*)out, (hipStream_t)stream));
        icbc_image& lv = levels[k];
        lv.data = out;
        lv.format = ICBC_IMAGE_RGBA32F;
        lv.width = w;
        lv.height = h;
        lv.edge = source->edge;
        lv.pitch = (uint64_t)w * 16;
        cur = ImageDev{(const uint8_t*)out, lv.pitch, ICBC_IMAGE_RGBA32F, w, h, source->edge};
        out += (uint64_t)w * h * 16;
    }
    if (n_levels) *n_levels = n;
    return ICBC_OK;
}

namespace {

int decode_batch(const void* codes, uint64_t block_count, uint8_t* rgba, int decoder, bool bc3, void* stream) {
    if (decoder < 0 || decoder > 3) return fail(ICBC_ERR_INVALID_ARGUMENT, "decoder %d", decoder);
    if (block_count && (!codes || !rgba)) return fail(ICBC_ERR_INVALID_ARGUMENT, "null buffer");
    if (((uintptr_t)codes & (bc3 ? 15 : 7)) || ((uintptr_t)rgba & 15))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "codes must be %d-byte and rgba 16-byte aligned", bc3 ? 16 : 8);
    int d;
    int r = ensure_device_ready(&d);
    if (r) return r;
    HIP_TRY(launch_decode((const uint2*)codes, block_count, (uint32_t*)rgba, decoder, bc3, (hipStream_t)stream));
    return ICBC_OK;
}

int evaluate_batch(const uint8_t* rgba, const void* codes, uint64_t block_count, float* errors, int decoder,
                   bool bc3, bool alpha_blend, void* stream) {
    if (decoder < 0 || decoder > 3) return fail(ICBC_ERR_INVALID_ARGUMENT, "decoder %d", decoder);
    if (block_count && (!codes || !rgba || !errors)) return fail(ICBC_ERR_INVALID_ARGUMENT, "null buffer");
    if (((uintptr_t)codes & (bc3 ? 15 : 7)) || ((uintptr_t)rgba & 15) || ((uintptr_t)errors & 3))
        return fail(ICBC_ERR_INVALID_ARGUMENT, "misaligned buffer");
    int d;
    int r = ensure_device_ready(&d);
    if (r) return r;
    HIP_TRY(launch_evaluate((const uint32_t*)rgba, (const uint2*)codes, block_count, errors, decoder, bc3,
                            alpha_blend, (hipStream_t)stream));
    return ICBC_OK;
}

}  // namespace

int icbc_gpu_decode_bc1_batch(const void* codes, uint64_t block_count, uint8_t* rgba, int decoder,
                              void* stream) {
    return decode_batch(codes, block_count, rgba, decoder, false, stream);
}

int icbc_gpu_evaluate_bc1_error_batch(const uint8_t* rgba, const void* codes, uint64_t block_count,
                                      float* errors, int decoder, void* stream) {
    return evaluate_batch(rgba, codes, block_count, errors, decoder, false, false, stream);
}

int icbc_gpu_decode_bc3_batch(const void* blocks, uint64_t block_count, uint8_t* rgba, int decoder,
                              void* stream) {
    return decode_batch(blocks, block_count, rgba, decoder, true, stream);
}

int icbc_gpu_evaluate_bc3_error_batch(const uint8_t* rgba, const void* blocks, uint64_t block_count,
                                      float* errors, int alpha_blend, int decoder, void* stream) {
    return evaluate_batch(rgba, blocks, block_count, errors, decoder, true, alpha_blend != 0, stream);
}

// ---- diagnostics / plumbing ----------------------------------------------------

const char* icbc_gpu_last_error(void) { return t_last_error; }

int icbc_gpu_device_name(char* buf, size_t len) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (!buf || len == 0) return fail(ICBC_ERR_INVALID_ARGUMENT, "empty name buffer");
    // the marketing name can be empty (no product table for the device in the
    // runtime); the architecture name and CU count always identify the part
    snprintf(buf, len, "%s (%s, %d CUs)", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName,
             prop.multiProcessorCount);
    return ICBC_OK;
}

int icbc_gpu_version(void) { return 10500; }

int icbc_gpu_check_reciprocal(uint64_t* mismatches) {
    if (!mismatches) return fail(ICBC_ERR_INVALID_ARGUMENT, "mismatches is null");
    int decoder;
    int r = ensure_device_ready(&decoder);
    if (r) return r;
    unsigned long long* d = nullptr;
    HIP_TRY(hipMalloc((void**)&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = launch_check_reciprocal(d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "reciprocal check");
    *mismatches = (uint64_t)h;
    return ICBC_OK;
}

void* icbc_gpu_malloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        hip_fail(e, "hipMalloc");
        return nullptr;
    }
    return p;
}

int icbc_gpu_free(void* ptr) {
    HIP_TRY(hipFree(ptr));
    return ICBC_OK;
}

void* icbc_gpu_host_alloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        hip_fail(e, "hipHostMalloc");
        return nullptr;
    }
    return p;
}

int icbc_gpu_host_free(void* ptr) {
    HIP_TRY(hipHostFree(ptr));
    return ICBC_OK;
}

int icbc_gpu_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ICBC_OK;
}

int icbc_gpu_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ICBC_OK;
}

int icbc_gpu_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return ICBC_OK;
}

int icbc_gpu_memset(void* dst, int value, size_t bytes) {
    HIP_TRY(hipMemset(dst, value, bytes));
    return ICBC_OK;
}

void* icbc_gpu_stream_create(void) {
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        hip_fail(e, "hipStreamCreate");
        return nullptr;
    }
    return (void*)s;
}

int icbc_gpu_stream_destroy(void* stream) {
    HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return ICBC_OK;
}

int icbc_gpu_stream_synchronize(void* stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return ICBC_OK;
}

void* icbc_gpu_event_create(void) {
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        hip_fail(e, "hipEventCreate");
        return nullptr;
    }
    return (void*)ev;
}

int icbc_gpu_event_destroy(void* event) {
    HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return ICBC_OK;
}

int icbc_gpu_event_record(void* event, void* stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return ICBC_OK;
}

int icbc_gpu_event_synchronize(void* event) {
    HIP_TRY(hipEventSynchronize((hipEvent_t)event));
    return ICBC_OK;
}

float icbc_gpu_event_elapsed_ms(void* start, void* stop) {
    float ms = NAN;
    hipError_t e = hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)stop);
    if (e != hipSuccess) {
        hip_fail(e, "hipEventElapsedTime");
        return NAN;
    }
    return ms;
}

}  // extern "C"

// ===========================================================================
// Drop-in C++ API (include/icbc.h)

namespace icbc {

// The reference's API has no error channel.  On a device failure (or a call
// before a successful init) the functions return a defined value -- NaN for
// the float results, outputs untouched -- and icbc_gpu_last_error() holds the
// reason.

void init(Decoder decoder) {
    (void)icbc_gpu_init((int)decoder);  // a failure shows in the calls that follow
}

float compress_bc1(Quality level, const float* input_colors, const float* input_weights,
                   const float color_weights[3], bool three_color_mode, bool three_color_black,
                   void* output) {
    float e = 0.0f;
    if (compress_one((int)level, input_colors, input_weights, color_weights, three_color_mode ? 1 : 0,
                     three_color_black ? 1 : 0, output, &e) != ICBC_OK)
        return NAN;
    return e;
}

float compress_bc3(Quality, const float*, const float*, const float[3], bool, void*) {
    return 0.0f;  // the reference's BC3 encoder is a stub (icbc.h:3850-3853)
}

void decode_bc1(const void* block, unsigned char rgba_block[16 * 4], Decoder decoder) {
    (void)decode_one(block, rgba_block, (int)decoder, false);
}

void decode_bc3(const void* block, unsigned char rgba_block[16 * 4], Decoder decoder) {
    (void)decode_one(block, rgba_block, (int)decoder, true);
}

float evaluate_bc1_error(const unsigned char rgba_block[16 * 4], const void* block, Decoder decoder) {
    float e = 0.0f;
    if (evaluate_one(rgba_block, block, (int)decoder, false, false, &e) != ICBC_OK) return NAN;
    return e;
}

float evaluate_bc3_error(const unsigned char rgba_block[16 * 4], const void* block, bool alpha_blend,
                         Decoder decoder) {
    float e = 0.0f;
    if (evaluate_one(rgba_block, block, (int)decoder, true, alpha_blend, &e) != ICBC_OK) return NAN;
    return e;
}

}  // namespace icbc
