// Coalesced single-block calls: icbc::compress_bc1 / icbc_gpu_compress_bc1.
//
// The reference's own caller encodes an image block by block, calling
// compress_bc1 once per block from every worker thread (icbc_test.cpp:397-409,
// ic_pfor.h:669-681).  One kernel launch, copy and stream synchronisation per
// block would leave the device idle between calls, so concurrent calls share
// launches here:
//   - each call takes the next slot of a per-device ring of page-locked host
//     buffers (colours, weights, the output's prior code), copies its block in
//     and marks the slot ready;
//   - a caller that finds fewer than kInflight launches running launches every
//     ready slot at the ring's launch frontier (a contiguous run; runs of
//     calls with equal options become one job each, so the harness's calls
//     make one job per launch) on a stream with resident launch resources
//     (level-9 list, workspace): the kernels read the blocks, the job table
//     and the tile list straight from the page-locked ring and write codes
//     and errors back into it (the memory is mapped into the device's address
//     space), so a launch is one encode launch and one event -- no copies, no
//     allocation, no host callback;
//   - every caller waits on the event of the launch holding its slot, copies
//     its code and error out and frees the slot; the first caller back from a
//     launch starts the next one for the slots that became ready meanwhile.
// No thread of its own: launching and completion ride on the callers.  Results
// are those of the batch path (blocks are independent).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <condition_variable>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/icbc_gpu.h"
#include "bc1_kernels.h"
#include "icbc_internal.h"

using namespace icbc_gpu;

namespace {

constexpr uint32_t kSlots = 8192;  // ring size (calls in flight at once)
constexpr int kInflight = 2;       // launches overlapping per device
constexpr uint32_t kMaxGrid = 512; // resident workgroups of a ring launch (its workspace)
// Small launches use small tiles: a tile's blocks are searched one after
// another by one wavefront, so fewer blocks per tile finish sooner.
uint32_t ring_tile_blocks(uint64_t launch_blocks) {
    return launch_blocks <= 64 ? 1 : launch_blocks <= 512 ? 4 : launch_blocks <= 4096 ? 8 : 64;
}
constexpr size_t kColorBytes = 256, kWeightBytes = 64, kCodeBytes = 8, kErrorBytes = 4;
constexpr size_t kSlotBytes = kColorBytes + kWeightBytes + kCodeBytes + kErrorBytes;

enum SlotState : uint8_t { kFree, kFilling, kReady, kLaunched };

struct SlotMeta {
    int32_t level = 0, tcm = 0, tcb = 0, decoder = 0;
    float cw[3] = {0, 0, 0};
    SlotState state = kFree;
    uint64_t launch = 0;  // id of the launch holding the slot
};

struct Launch {
    int stream = 0;
    hipEvent_t event = nullptr;
    uint32_t refs = 0;   // slots not collected yet
    bool done = false;   // a caller has seen its event complete
    int status = ICBC_OK;
};

// Ring of one device: structure-of-arrays (so a run of slots is a block-major
// job), host page-locked and device mirrors.
// Per stream: a job table and tile list in page-locked memory, and a level-9
// list, queue counters and workspaces in device memory, sized for a launch of
// every slot (each its own job, one block per tile).
constexpr size_t kTableJobs = kSlots;
constexpr size_t kJobsBytes = (kTableJobs * sizeof(JobDev) + 255) / 256 * 256;
constexpr size_t kTilesBytes = kSlots * sizeof(uint2);
constexpr size_t kRefineBytes = (refine_list_bytes(kSlots) + 255) / 256 * 256;
constexpr size_t kStreamBytes = kRefineBytes + kQueueBytes + kMaxGrid * kWorkspaceBytesPerWorkgroup;

struct Ring {
    int device = -1;
    uint32_t grid = 1;
    hipStream_t stream[kInflight] = {};
    bool stream_busy[kInflight] = {};
    char* sdev[kInflight] = {};   // per-stream launch resources (device)
    char* shost[kInflight] = {};  // per-stream job table + tile list (page-locked, read by the kernels)
    char* shost_dev[kInflight] = {};  // the same memory's device address
    char* h = nullptr;  // the slots (page-locked)
    char* d = nullptr;  // the slots' device address
    std::vector<SlotMeta> meta = std::vector<SlotMeta>(kSlots);
    std::vector<hipEvent_t> spare_events;
    std::map<uint64_t, Launch> launches;
    uint64_t next = 0;      // next sequence number to hand out
    uint64_t frontier = 0;  // first sequence not launched
    uint64_t freed = 0;     // every sequence below this is free
    uint64_t launch_ids = 0;
    int inflight = 0;
    uint32_t space_waiters = 0;  // callers waiting for a free slot
    std::mutex m;
    // callers waiting for their slot's launch, spread over kWaitLists
    // condition variables by slot, so that a launch wakes only its own callers
    static constexpr uint32_t kWaitLists = 64;
    std::condition_variable cv_launched[kWaitLists];
    std::condition_variable cv_space;  // slots freed

    static size_t colors_at(uint64_t i) { return i * kColorBytes; }
    static size_t weights_at(uint64_t i) { return kSlots * kColorBytes + i * kWeightBytes; }
    static size_t codes_at(uint64_t i) { return kSlots * (kColorBytes + kWeightBytes) + i * kCodeBytes; }
    static size_t errors_at(uint64_t i) { return kSlots * (kColorBytes + kWeightBytes + kCodeBytes) + i * kErrorBytes; }

    int create(int dev);
    int launch_ready();  // under m
    void release(uint64_t seq);  // under m
};

int hip_err(hipError_t e, const char* what) {
    return icbc_internal_fail(e == hipErrorOutOfMemory ? ICBC_ERR_OUT_OF_MEMORY : ICBC_ERR_DEVICE, "%s: %s", what,
                              hipGetErrorString(e));
}

// All of a ring's resources or none (a failed set-up leaves nothing behind,
// and the next call tries again).
int Ring::create(int dev) {
    device = dev;
    int r = icbc_internal_resident(&grid);
    if (r) return r;
    grid = std::min<uint32_t>(grid, kMaxGrid);
    hipError_t e = hipHostMalloc((void**)&h, kSlots * kSlotBytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&d, h, 0);
    for (int k = 0; k < kInflight && e == hipSuccess; k++) {
        e = hipStreamCreateWithFlags(&stream[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMalloc((void**)&sdev[k], kStreamBytes);
        if (e == hipSuccess) e = hipHostMalloc((void**)&shost[k], kJobsBytes + kTilesBytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&shost_dev[k], shost[k], 0);
    }
    if (e == hipSuccess) return ICBC_OK;
    for (int k = 0; k < kInflight; k++) {
        if (stream[k]) (void)hipStreamDestroy(stream[k]);
        if (sdev[k]) (void)hipFree(sdev[k]);
        if (shost[k]) (void)hipHostFree(shost[k]);
    }
    if (h) (void)hipHostFree(h);
    return hip_err(e, "single-block ring set-up");
}

void Ring::release(uint64_t seq) {
    meta[seq % kSlots].state = kFree;
    while (freed < next && meta[freed % kSlots].state == kFree && freed < frontier) freed++;
}

// Launches the run of ready slots at the frontier, if a stream is free.
int Ring::launch_ready() {
    if (inflight >= kInflight || frontier == next) return ICBC_OK;
    // the ready run at the frontier, with the decoder of its first call (the
    // decoder selected by init when the call was made; one per launch)
    uint64_t end = frontier;
    const int dec = meta[frontier % kSlots].decoder;
    while (end < next && meta[end % kSlots].state == kReady && meta[end % kSlots].decoder == dec) end++;
    if (end == frontier) return ICBC_OK;
    int k = 0;
    while (stream_busy[k]) k++;
    hipStream_t s = stream[k];
    // jobs: runs of equal options, split where the ring wraps
    const uint64_t count = end - frontier;
    const uint32_t tb = ring_tile_blocks(count);
    // job table and tile list share one upload: tiles follow the jobs
    JobDev* jobs = (JobDev*)shost[k];
    uint2* tiles = (uint2*)(shost[k] + kJobsBytes);  // moved behind the jobs below
    uint32_t njobs = 0, ntiles = 0;
    for (uint64_t a = frontier; a < end;) {
        const uint64_t ia = a % kSlots;
        const SlotMeta& ma = meta[ia];
        uint64_t b = a + 1;
        while (b < end && b % kSlots != 0) {
            const SlotMeta& mb = meta[b % kSlots];
            if (mb.level != ma.level || mb.tcm != ma.tcm || mb.tcb != ma.tcb || memcmp(mb.cw, ma.cw, sizeof ma.cw) != 0)
                break;
            b++;
        }
        JobDev& j = jobs[njobs];
        memset(&j, 0, sizeof j);
        j.colors = (const float4*)(d + colors_at(ia));
        j.weights = (const float4*)(d + weights_at(ia));
        j.codes = (uint2*)(d + codes_at(ia));
        j.errors = (float*)(d + errors_at(ia));
        j.block_count = b - a;
        j.level = ma.level;
        j.tcm = ma.tcm;
        j.tcb = ma.tcb;
        j.decoder = dec;
        j.cw0 = ma.cw[0];
        j.cw1 = ma.cw[1];
        j.cw2 = ma.cw[2];
        j.tile_blocks = (int32_t)tb;
        for (uint64_t f = 0; f < b - a; f += tb) tiles[ntiles++] = make_uint2(njobs, (uint32_t)f);
        njobs++;
        a = b;
    }
    const size_t jbytes = ((size_t)njobs * sizeof(JobDev) + 15) / 16 * 16;
    memmove(shost[k] + jbytes, tiles, ntiles * sizeof(uint2));
    const JobDev* djobs = (const JobDev*)shost_dev[k];
    const uint2* dtiles = (const uint2*)(shost_dev[k] + jbytes);
    char* res = sdev[k];
    const RefineList rl{(RefineItem*)(res + kRefineCounterBytes), (uint32_t*)res};
    uint32_t* counter = (uint32_t*)(res + kRefineBytes);
    float4* work = (float4*)(res + kRefineBytes + kQueueBytes);
    hipError_t e = launch_encode(djobs, dtiles, ntiles, counter, work, std::min<uint32_t>(grid, ntiles),
                                 encode_mode(jobs, njobs), rl, s);
    int status = ICBC_OK;
    hipEvent_t ev = nullptr;
    if (!spare_events.empty()) {
        ev = spare_events.back();
        spare_events.pop_back();
    } else if (e == hipSuccess) {
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    }
    if (e == hipSuccess && ev) e = hipEventRecord(ev, s);
    if (status == ICBC_OK && e != hipSuccess) status = hip_err(e, "single-block launch");
    const uint64_t id = ++launch_ids;
    Launch& L = launches[id];
    L.stream = k;
    L.event = ev;
    L.refs = (uint32_t)(end - frontier);
    L.status = status;
    L.done = status != ICBC_OK;  // a failed launch has nothing to wait for
    if (!L.done) {
        stream_busy[k] = true;
        inflight++;
    }
    for (uint64_t a = frontier; a < end; a++) {
        meta[a % kSlots].state = kLaunched;
        meta[a % kSlots].launch = id;
    }
    // wake the callers of the launched slots (each wait list at most once)
    uint64_t lists = 0;
    for (uint64_t a = frontier; a < end; a++) lists |= 1ull << (a % kSlots % kWaitLists);
    frontier = end;
    for (uint32_t q = 0; q < kWaitLists; q++)
        if (lists >> q & 1) cv_launched[q].notify_all();
    return status;
}

std::mutex g_rings_mutex;
std::map<int, Ring*> g_rings;  // kept until the process ends

int ring_for_current(Ring*& out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_err(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_rings_mutex);
    auto it = g_rings.find(dev);
    if (it != g_rings.end()) {
        out = it->second;
        return ICBC_OK;
    }
    Ring* r = new Ring();
    const int st = r->create(dev);
    if (st != ICBC_OK) {
        delete r;
        return st;
    }
    g_rings[dev] = r;
    out = r;
    return ICBC_OK;
}

}  // namespace

int icbc_internal_compress_one(int level, const float* colors, const float* weights, const float cw[3], int tcm,
                               int tcb, void* output, float* error) {
    if (!colors || !weights || !cw || !output)
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    int decoder;
    int r = icbc_internal_ready(&decoder);
    if (r) return r;
    Ring* ring;
    r = ring_for_current(ring);
    if (r) return r;
    std::unique_lock<std::mutex> lk(ring->m);
    if (ring->next - ring->freed >= kSlots) {
        ring->space_waiters++;
        ring->cv_space.wait(lk, [&] { return ring->next - ring->freed < kSlots; });
        ring->space_waiters--;
    }
    const uint64_t seq = ring->next++;
    const uint64_t i = seq % kSlots;
    SlotMeta& me = ring->meta[i];
    me.state = kFilling;
    me.level = level;
    me.tcm = tcm ? 1 : 0;
    me.tcb = tcb ? 1 : 0;
    me.decoder = decoder;
    memcpy(me.cw, cw, sizeof me.cw);
    lk.unlock();
    memcpy(ring->h + Ring::colors_at(i), colors, kColorBytes);
    memcpy(ring->h + Ring::weights_at(i), weights, kWeightBytes);
    memcpy(ring->h + Ring::codes_at(i), output, kCodeBytes);
    lk.lock();
    me.state = kReady;
    (void)ring->launch_ready();
    ring->cv_launched[i % Ring::kWaitLists].wait(lk, [&] {
        if (me.state == kLaunched) return true;
        (void)ring->launch_ready();  // a stream may have freed up
        return me.state == kLaunched;
    });
    const uint64_t id = me.launch;
    Launch* L = &ring->launches[id];
    if (!L->done) {
        const hipEvent_t ev = L->event;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(ev);
        lk.lock();
        L = &ring->launches[id];
        if (!L->done) {
            // first back from this launch: its stream is free for the next one
            L->done = true;
            if (e != hipSuccess) L->status = hip_err(e, "single-block launch");
            ring->stream_busy[L->stream] = false;
            ring->inflight--;
            (void)ring->launch_ready();
        }
    }
    const int status = L->status;
    if (status == ICBC_OK) {
        memcpy(output, ring->h + Ring::codes_at(i), kCodeBytes);
        memcpy(error, ring->h + Ring::errors_at(i), kErrorBytes);
    }
    ring->release(seq);
    if (--L->refs == 0) {
        if (L->event) ring->spare_events.push_back(L->event);
        ring->launches.erase(id);
    }
    if (ring->space_waiters) ring->cv_space.notify_all();
    return status;
}
