// Host-visible launch interface of the BC1 kernels (bc1_encoder.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icbc_gpu {

constexpr int kTile = 64;     // blocks per tile = lanes per wavefront
constexpr int kSatRows = 18;  // zero row + up to 16 prefix rows + the bound row per block
constexpr int kBoundRow = 17; // {search margin, prune allowed, 0, 0} of the four-cluster search
constexpr int kPairs = 153;    // boundary pairs c0 <= c1 <= 16 of the four-cluster partitions

// One job as the kernels see it (device memory, built by the runtime).
struct JobDev {
    const float4* colors;   // 16 float4 per block
    const float4* weights;  // 4 float4 per block, or null for all 1.0
    uint2* codes;           // col0 | col1 << 16, indices
    float* errors;          // or null
    int4* stats;            // or null
    uint64_t block_count;
    int32_t level, tcm, tcb, decoder;
    float cw0, cw1, cw2;
    int32_t tile_blocks;  // blocks per tile (8, 16, 32 or 64)
    int32_t rgba8;        // colors holds 16 RGBA8 words per block, read as c / 255.0f
    int32_t prune_min;    // four-cluster searches over at least this many colours run the
                          // pruning bound pass first (0 = kDefaultPruneMinColors)
};

// Colour count from which the four-cluster search prunes by default (the
// autotuner may pick another per level: 17 never prunes).  Measured on the
// benchmark content, pruning from 13 colours is 1-4 % faster than from 9 at
// levels 2-9: below 13 colours the bound pass costs about what it saves.
// The pruned search
// needs every lane to hold at least three candidates (C(n+3,3)-1 >= 219, so
// n >= 9): kMinPruneColors is the lowest setting.
constexpr int kDefaultPruneMinColors = 13;
constexpr int kMinPruneColors = 9;
constexpr int kNoPruneColors = 17;

// A four-cluster table entry holds the partition's boundaries c0 <= c1 <= c2
// and the index of its pair (c0, c1) among the 153 pairs, laid out so that
// the pruning bound's LDS addresses are single bit-field extracts: bits 0-11
// pair << 4 (byte offset of the pair's float4 sums), bits 12-20 c2 << 4 (byte
// offset of the third boundary's row; c2 itself at bits 16-20), bits 21-25
// c0, bits 26-30 c1.
constexpr uint32_t four_entry(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t pair) {
    return (pair << 4) | (c2 << 16) | (c0 << 21) | (c1 << 26);
}

// Tables built by the host at init (all four decoders at once).
struct HostTables {
    float mid5[32];
    float mid6[64];
    uint32_t four[976];
    uint32_t three[160];
    int32_t four_total[16];
    int32_t three_total[16];
    uint16_t match5[4][256];
    uint16_t match6[4][256];
    float dweights[4][5];
    // Per decoder, the four-cluster candidates of a block of 16 colours of
    // weight 1 whose least-squares system is nearly singular (the pruning
    // bound cannot decide them): their table indices (at most 64) and, per
    // lane l, the bit mask of rounds k with candidate l + 64 k in the set.
    uint16_t singular[4][64];
    uint16_t singular_mask[4][64];
    int32_t singular_count[4];
};

hipError_t upload_tables(const HostTables& t, hipStream_t stream);

// Workgroups of the cluster-fit encode instance that fit on one compute unit
// at once (sizes launches and workspaces; a level 0-1 instance is launched at
// its own residency, at most this grid).
int encode_workgroups_per_cu();

// Encodes `ntiles` tiles with `grid` resident workgroups (clamped to ntiles).
//   tiles != null: tile k covers blocks [tiles[k].y, +T) of job tiles[k].x
//   tiles == null: one job (jobs[0]), tile k covers blocks [T k, T k + T)
// where T is the job's tile_blocks.
//   counter != null (kQueueBytes of queue counters): workgroup g takes tile g,
//   then claims further tiles of its XCD's share from that XCD's queue (reset
//   by this call); otherwise workgroup g takes tiles g, g + grid, ...
// `workspace` holds kWorkspaceBytesPerWorkgroup bytes per workgroup of `grid`.
// Per workgroup: kLaneRows float4 rows for each of the tile's 64 blocks, in
// block order: the 4-cluster SAT (18 rows), the 3-cluster SAT (18 rows) and
// the block's reduced colours (16 rows).  A lane's rows span 832 bytes, so
// every row is its base address plus an instruction's immediate offset.
constexpr int kLaneRows = 2 * kSatRows + 16;
constexpr size_t kWorkspaceBytesPerWorkgroup = (size_t)kTile * kLaneRows * 16;

// Tile queues: one counter per XCD (workgroups are dispatched round-robin
// over the 8 XCDs, and XCD x takes tiles x, x + 8, x + 16, ... of the list),
// each in its own 128-byte line, ahead of the workgroup slices.
constexpr uint32_t kQueues = 8;
constexpr size_t kQueueStride = 128;
constexpr size_t kQueueBytes = kQueues * kQueueStride;

// Level-9 blocks handed from the encode kernel to the refinement kernel: the
// encode kernel writes the block's fitted code (or leaves the prior one) and
// appends {job, block, error after the fits}; the refinement kernel walks
// each one with a lane of its own.
struct RefineItem {
    uint32_t job;
    uint32_t block;
    float error;
    uint32_t pad;
};
struct RefineList {
    RefineItem* items;   // room for every level-9 block of the launch
    uint32_t* counters;  // [0] appended, [kQueueStride / 4] claimed
};
constexpr size_t kRefineCounterBytes = 2 * kQueueStride;
constexpr size_t refine_list_bytes(uint64_t level9_blocks) {
    return kRefineCounterBytes + (size_t)level9_blocks * sizeof(RefineItem);
}

// Kernel instances: every job of a kModeBox launch is at level 0 or 1 (or an
// invalid level); kModeCluster covers levels 0-9; kModeAll is a kModeCluster
// launch followed by the refinement kernel over its level-9 blocks.
constexpr int kModeBox = 0;
constexpr int kModeCluster = 1;
constexpr int kModeAll = 2;

// Texel formats of a launch's jobs, in bits 4-5 of a mode: the cluster
// instance is compiled once per format (all float, all RGBA8, both mixed,
// which picks per tile), so a launch of one format carries no code for the
// other.
constexpr int kTexFloat = 0;
constexpr int kTexRgba8 = 1;
constexpr int kTexMixed = 2;

// Every job of a launch has channel weights exactly (1, 1, 1): the level 0-1
// and refinement kernels' unit-weight instances skip the products by the
// weights (exact identities); the cluster kernel ignores the flag.
constexpr int kUnitWeights = 1 << 6;

inline bool unit_weights(const JobDev& j) { return j.cw0 == 1.0f && j.cw1 == 1.0f && j.cw2 == 1.0f; }

// Some job of a launch has texel weights.  A level 0-1 launch whose jobs all
// have unit channel weights and no texel weights runs the instance that reads
// each block's texels straight into registers (no LDS staging area, so more
// workgroups per compute unit); the others stage the tile's texels in LDS.
constexpr int kTexelWeights = 1 << 8;

// Some job of a launch is at level 0-3, whose box-fit stage the cluster
// instance then carries; launches at levels 4-9 only run an instance without
// it (fewer live registers across the per-block stages).
constexpr int kBoxStage = 1 << 7;

// `mode` selects the kernel instance (level mode | texel format << 4 |
// kUnitWeights | kBoxStage | kTexelWeights); encode_mode() gives the one a set
// of jobs needs.
// With kModeAll, `refine` holds room for the launch's level-9 blocks.
// The queue counters (kQueueBytes at `counter`) and the handoff list's
// counters are zeroed by a small kernel first, unless `counters_zeroed`: the
// kernels leave them at zero when they finish (the last claim on each counter
// resets it), so a resource that is reused launch after launch (a plan)
// zeroes them once, when it is made.
hipError_t launch_encode(const JobDev* jobs, const uint2* tiles, uint32_t ntiles, uint32_t* counter,
                         float4* workspace, uint32_t grid, int mode, RefineList refine, hipStream_t stream,
                         bool counters_zeroed = false);

inline int encode_mode_of_level(int level) {
    return level == 9 ? kModeAll : (level >= 2 && level <= 8) ? kModeCluster : kModeBox;
}

// Level-9 blocks of a set of (host-side) jobs.
inline uint64_t level9_blocks(const JobDev* jobs, size_t n) {
    uint64_t b = 0;
    for (size_t k = 0; k < n; k++)
        if (jobs[k].level == 9) b += jobs[k].block_count;
    return b;
}

// The kernel instance a launch over these (host-side) jobs needs.
inline int encode_mode(const JobDev* jobs, size_t n) {
    int mode = kModeBox;
    bool any_float = false, any_rgba8 = false;
    for (size_t k = 0; k < n; k++) {
        const int m = encode_mode_of_level(jobs[k].level);
        if (m > mode) mode = m;
        if (jobs[k].rgba8) any_rgba8 = true;
        else any_float = true;
    }
    const int tex = any_rgba8 ? (any_float ? kTexMixed : kTexRgba8) : kTexFloat;
    bool unit = true, box = false, weighted = false;
    for (size_t k = 0; k < n; k++) {
        if (!unit_weights(jobs[k])) unit = false;
        if (jobs[k].level >= 0 && jobs[k].level <= 3) box = true;
        if (jobs[k].weights) weighted = true;
    }
    return mode | (tex << 4) | (unit ? kUnitWeights : 0) | (box ? kBoxStage : 0) | (weighted ? kTexelWeights : 0);
}

// The one-workgroup-per-tile schedule gives every tile its own workspace
// slice; tuned defaults use it only for launches of at most this many waves
// of resident workgroups.
constexpr uint64_t kGridScheduleMaxWaves = 4;

// Counts (into *mismatches, zeroed first) the 2^32 float inputs where the
// encoder's fast reciprocal differs from IEEE-754 division on this device.
// ---------------------------------------------------------------------------
// The no-allocation batch entry point (icbc_gpu_compress_bc1_batch_noalloc):
// everything a launch needs besides the caller's inputs and outputs lives in
// one file-scope device array of bc1_encoder.hip, laid out as
//   [queue counters][job table][tile list][level-9 handoff][workgroup workspaces]
// with these maxima per call.
constexpr uint32_t kNoAllocMaxJobs = 32;
constexpr uint32_t kNoAllocMaxTiles = 65536;
constexpr uint64_t kNoAllocMaxLevel9Blocks = 262144;
constexpr uint32_t kNoAllocMaxGrid = 1024;  // resident workgroups of its launches
constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocJobsOffset = round256(kQueueBytes);
constexpr size_t kNoAllocTilesOffset = kNoAllocJobsOffset + round256(kNoAllocMaxJobs * sizeof(JobDev));
constexpr size_t kNoAllocRefineOffset = kNoAllocTilesOffset + round256(kNoAllocMaxTiles * sizeof(uint2));
constexpr size_t kNoAllocWorkspaceOffset = kNoAllocRefineOffset + round256(refine_list_bytes(kNoAllocMaxLevel9Blocks));
constexpr size_t kNoAllocScratchBytes = kNoAllocWorkspaceOffset + (size_t)kNoAllocMaxGrid * kWorkspaceBytesPerWorkgroup;

// Resolves the scratch array's address on the current device (`device`) with
// hipGetSymbolAddress; called once per device at its set-up.
hipError_t noalloc_scratch_setup(int device);
// The address resolved for `device`, or null before its set-up.
void* noalloc_scratch(int device);

// A no-allocation launch's jobs, passed to the device as a kernel argument:
// the job table, the launch order of the jobs and the first tile of each in
// that order (first_tile[n] = the tile count).
struct NoAllocJobs {
    JobDev jobs[kNoAllocMaxJobs];
    uint32_t order[kNoAllocMaxJobs];
    uint32_t first_tile[kNoAllocMaxJobs + 1];
    uint32_t n;
};
// Writes the job table and, for more than one job, the tile list into the
// scratch at `scratch`.
hipError_t launch_noalloc_setup(const NoAllocJobs& a, void* scratch, hipStream_t stream);

// Fills the per-device caches launch_encode reads (resident workgroups of the
// level 0-1 and refinement kernels), so later launches on the device take no
// first-use path.
hipError_t warm_launch_caches();

hipError_t launch_check_reciprocal(unsigned long long* mismatches, hipStream_t stream);

// decode_bc1 / evaluate_bc1_error over n blocks (8 bytes each), or with
// `bc3` decode_bc3 / evaluate_bc3_error over n BC3 blocks (16 bytes each).
hipError_t launch_decode(const uint2* codes, uint64_t n, uint32_t* rgba, int decoder, bool bc3,
                         hipStream_t stream);
hipError_t launch_evaluate(const uint32_t* rgba, const uint2* codes, uint64_t n, float* errors,
                           int decoder, bool bc3, bool alpha_blend, hipStream_t stream);

}  // namespace icbc_gpu
