// Request server: continuous batching of BC1 encode requests on one device.
//
// A request is one texture (a host-memory icbc_bc1_job).  The server
//   - holds at most `memory_budget` bytes of device memory in all: request
//     buffers, each launch's job table, tile list and level-9 handoff list,
//     and the launch slots' workspaces are all charged against it;
//   - admits requests in arrival order while their device footprint fits the
//     budget and the admitted, not yet launched ones hold fewer than two
//     launches' worth of blocks (a request that could never fit is rejected
//     at submission);
//   - keeps request buffers resident: device memory comes from a pool of
//     size-classed blocks carved from slabs and reused; a slab goes back to
//     the driver only when all its blocks are free and another class needs
//     the room;
//   - batches every admitted, not yet launched request into one encode launch
//     (the same persistent kernel and tile queue as the batch API), with up to
//     `max_inflight` launches overlapping on separate streams, so requests
//     that arrive while others run join the next launch as soon as memory
//     and a launch slot free up;
//   - copies results back, timestamps completion and releases the budget.
// Two host threads: a scheduler (admit, upload, launch) and a reaper (wait on
// the oldest launch, complete its requests).  Results do not depend on the
// batch a request lands in: blocks are independent and each job carries its
// own options.

#include <hip/hip_runtime.h>

#include <float.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/icbc_gpu.h"
#include "bc1_kernels.h"
#include "icbc_internal.h"
#include "transfer.h"

using namespace icbc_gpu;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Fewest resident workgroups a server launches with (the workspace a small
// budget must still hold).
constexpr uint64_t kMinServerGrid = 64;

// Blocks merged into one launch at most: a million blocks fill every resident
// workgroup of the encoder many times over (1-3 ms at levels 2-9), so larger
// launches gain no throughput, only latency for their first requests and
// staging memory for pageable buffers.
constexpr uint64_t kMaxLaunchBlocks = 1u << 20;

// Requests are admitted (given device memory) only while the admitted, not
// yet launched ones hold fewer blocks than this: enough for the next two
// launches.  Under overload the backlog then waits in the host-side queue
// instead of taking device memory request by request: admitting everything
// a large budget allowed made the pool map tens of gigabytes of slabs in the
// scheduler thread (and release them again as size classes shifted), which
// halved the rate the server sustained once a backlog had formed.
constexpr uint64_t kAdmitAheadBlocks = 2 * kMaxLaunchBlocks;

// Input buffers of at least this size go by DMA (1 MiB copies run at 38 GB/s
// on MI355X, 128 KiB ones at 11 GB/s: smaller ones go through the copy
// kernel).
constexpr uint64_t kDmaMinBytes = 1u << 20;

// Device memory the runtime itself holds for each launch slot's stream (its
// queue and signal pools; 4-5 MiB measured on MI355X), charged to the budget.
constexpr size_t kStreamAllowance = 6u << 20;

// Device footprint of a request: colours (float or RGBA8, as submitted),
// weights, codes, errors, stats.
struct Layout {
    size_t colors = 0, weights = 0, codes = 0, errors = 0, stats = 0, total = 0, color_bytes = 0;
};

Layout layout_of(const icbc_bc1_job& j) {
    Layout l;
    const uint64_t n = j.block_count;
    size_t off = 0;
    l.colors = off;
    l.color_bytes = n * ((j.flags & ICBC_JOB_COLORS_RGBA8) ? 64 : 256);
    off = align_up(off + l.color_bytes, 256);
    l.weights = off;
    if (j.weights) off = align_up(off + n * 64, 256);
    l.codes = off;
    off = align_up(off + n * 8, 256);
    l.errors = off;
    if (j.errors) off = align_up(off + n * 4, 256);
    l.stats = off;
    if (j.stats) off = align_up(off + n * 16, 256);
    l.total = std::max<size_t>(off, 256);
    return l;
}

// Size-classed device blocks (powers of two from 64 KiB), reused, carved from
// slabs.  Every byte the pool commits (slabs, and raw allocations such as the
// slots' workspaces) is charged against the budget: a new slab is made only
// while the committed total stays within it, first releasing slabs whose
// blocks are all free.  Not thread-safe: the server calls it under its lock.
// The device allocator hands out memory in 2 MiB granules: every allocation
// is a whole number of them, so that the bytes charged are the bytes held.
constexpr size_t kGranule = 2u << 20;

class Pool {
public:
    explicit Pool(uint64_t budget) : budget_(budget) {}
    ~Pool() {
        for (auto& kv : slabs_) (void)hipFree(kv.second.base);
        for (auto& r : raw_) (void)hipFree(r);
    }
    static size_t class_of(size_t bytes) {
        size_t c = 64 << 10;
        while (c < bytes) c <<= 1;
        return c;
    }
    uint64_t committed() const { return committed_; }
    // Bytes held outside the pool's own allocations (the runtime's memory
    // behind a stream), charged for the pool's lifetime.
    bool charge(size_t bytes) {
        if (committed_ + bytes > budget_) return false;
        committed_ += bytes;
        return true;
    }
    // Device memory outside the classes (held until the pool is destroyed).
    hipError_t alloc_raw(size_t bytes, void** p) {
        bytes = align_up(bytes, kGranule);
        if (committed_ + bytes > budget_) return hipErrorOutOfMemory;
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return e;
        raw_.push_back(*p);
        committed_ += bytes;
        return hipSuccess;
    }
    // A block of class `cls`; hipErrorOutOfMemory when the budget has no room
    // for it now (blocks given back later may make room).
    hipError_t take(size_t cls, void** p) {
        auto& v = free_[cls];
        if (!v.empty()) {
            *p = v.back();
            v.pop_back();
            slab_of(*p).used++;
            return hipSuccess;
        }
        // slabs of up to 64 MiB (a sixteenth of a small budget), a whole
        // number of blocks and of granules (classes are powers of two)
        const size_t unit = std::max(cls, kGranule);
        size_t slab = std::max<size_t>(unit, std::min<uint64_t>(64u << 20, budget_ / 16) / unit * unit);
        if (committed_ + slab > budget_) slab = unit;
        if (committed_ + slab > budget_ && !release_empty(slab)) return hipErrorOutOfMemory;
        void* base = nullptr;
        hipError_t e = hipMalloc(&base, slab);
        if (e != hipSuccess) return e;
        committed_ += slab;
        slabs_[(uintptr_t)base] = Slab{(char*)base, slab, cls, 1};
        for (size_t off = cls; off + cls <= slab; off += cls) v.push_back((char*)base + off);
        *p = base;
        return hipSuccess;
    }
    void give(size_t cls, void* p) {
        free_[cls].push_back(p);
        slab_of(p).used--;
    }

private:
    struct Slab {
        char* base;
        size_t bytes, cls;
        uint32_t used;  // blocks taken
    };
    Slab& slab_of(void* p) {
        auto it = slabs_.upper_bound((uintptr_t)p);
        return (--it)->second;
    }
    // Frees slabs with no block in use until `need` more bytes fit the budget.
    bool release_empty(size_t need) {
        for (auto it = slabs_.begin(); it != slabs_.end() && committed_ + need > budget_;) {
            Slab& sl = it->second;
            if (sl.used != 0) {
                ++it;
                continue;
            }
            auto& v = free_[sl.cls];
            v.erase(std::remove_if(v.begin(), v.end(),
                                   [&](void* q) { return (char*)q >= sl.base && (char*)q < sl.base + sl.bytes; }),
                    v.end());
            (void)hipFree(sl.base);
            committed_ -= sl.bytes;
            it = slabs_.erase(it);
        }
        return committed_ + need <= budget_;
    }
    uint64_t budget_, committed_ = 0;
    std::map<uintptr_t, Slab> slabs_;
    std::map<size_t, std::vector<void*>> free_;
    std::vector<void*> raw_;
};

// Page-locked host staging that grows on demand, with its mapped device
// address (the copy kernels read and write it directly).
struct Staging {
    void* p = nullptr;
    void* dp = nullptr;
    size_t cap = 0;
    ~Staging() {
        if (p) (void)hipHostFree(p);
    }
    bool reserve(size_t bytes) {
        if (bytes <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = dp = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 2, 1 << 20);
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return false;
        if (hipHostGetDevicePointer(&dp, p, 0) != hipSuccess) {
            (void)hipHostFree(p);
            p = dp = nullptr;
            return false;
        }
        cap = want;
        return true;
    }
};

// The device-accessible address of `p` when it is page-locked host memory
// (mapped into the device's address space), else null: such buffers are read
// and written by the copy kernels in place, others go through staging.
void* mapped_address(const void* p) {
    if (!p) return nullptr;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return a.type == hipMemoryTypeHost ? a.devicePointer : nullptr;
}

enum State { kPending, kAdmitted, kLaunched, kDone };

struct Request {
    uint64_t id = 0;
    icbc_bc1_job job{};  // host pointers as given by the caller
    Layout lay;
    size_t cls = 0;      // pool class (also the budget charge)
    void* dev = nullptr;
    // mapped addresses of the caller's buffers (null: pageable, staged)
    const void* map_colors = nullptr;
    const void* map_weights = nullptr;
    void* map_codes = nullptr;
    void* map_errors = nullptr;
    void* map_stats = nullptr;
    State state = kPending;
    int status = ICBC_OK;
    uint32_t batch_size = 0;
    double t_submit = 0, t_admit = 0, t_launch = 0, t_done = 0;
};

// One launch slot: a stream plus resident launch resources.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t uploaded = nullptr;      // this launch's small inputs and table are on the device
    hipEvent_t uploaded_dma = nullptr;  // and its large inputs
    void* work = nullptr;  // queue counters + workspace slices (charged at creation)
    // this launch's job table + tile list and level-9 handoff list: pool
    // blocks taken with the batch, given back when it completes
    void* table = nullptr;
    size_t table_cls = 0;
    void* refine = nullptr;
    size_t refine_cls = 0;
    std::vector<uint32_t> tile_blocks;  // per request of the batch
    uint64_t ntiles = 0;
    Staging host_table;  // pinned staging for the table and the copy lists
    Staging in;   // inputs of requests whose buffers are pageable
    Staging out;  // outputs of requests whose buffers are pageable
    std::vector<size_t> out_off;  // per request: offset of its staged outputs in `out`
    std::vector<Request*> reqs;
    bool busy = false;
};

}  // namespace

struct icbc_server {
    icbc_server_options opt{};
    int decoder = 0;
    std::mutex m;
    std::condition_variable cv_sched, cv_reap, cv_done;
    std::deque<Request*> pending;
    std::vector<Request*> admitted;
    std::deque<Slot*> inflight;
    std::vector<Slot> slots;
    std::unordered_map<uint64_t, Request*> all;
    Pool pool;
    uint32_t grid = 0;  // resident workgroups of a launch (its workspace size)
    uint32_t copy_grid = 0;  // workgroups of a copy kernel
    uint64_t slot_bytes = 0;  // device bytes charged per launch slot (fixed memory and workspace)
    // Every launch's table copy and input copy kernel run in order on the
    // upload stream, its DMA copies on the other; the launch's encode waits
    // for both on its slot's stream.  So the next launch's inputs cross PCIe
    // while the current one encodes, and launches never compete with each
    // other for PCIe.
    hipStream_t upload = nullptr, upload_dma = nullptr;
    uint64_t used = 0;  // bytes of pool blocks taken (requests and launch tables)
    bool hold_admission = false;  // a formed batch waits for its launch tables' room
    uint64_t next_id = 1;
    bool stop = false;
    // the scheduler has returned: no launch can follow (the reaper stops
    // once this is set and nothing is in flight; while the scheduler enqueues
    // a launch, its batch is in no queue, so empty queues do not suffice)
    bool sched_done = false;
    int fatal = ICBC_OK;
    std::thread sched, reap;
    // counters
    uint64_t launches = 0, completed = 0, blocks_done = 0;

    explicit icbc_server(uint64_t budget) : pool(budget) {}
    void run_scheduler();
    void run_reaper();
    bool launch_resources(Slot& s, const std::vector<Request*>& reqs);
    void release_launch_resources(Slot& s);
    int launch(Slot& s, std::vector<Request*>& reqs);
};

namespace {

int set_fatal(icbc_server* sv, int code) {
    if (sv->fatal == ICBC_OK) sv->fatal = code;
    return code;
}

}  // namespace

// A launch's table block: job table, tile list, and the copy lists in (each
// request's colours, weights and prior codes) and out (codes, errors,
// counters), three segments per request each way.
constexpr size_t kSegsPerRequest = 3;
struct TableLayout {
    size_t tiles = 0, segs_in = 0, segs_out = 0, total = 0;
};
TableLayout table_layout(size_t requests, uint64_t tiles) {
    TableLayout t;
    t.tiles = align_up(requests * sizeof(JobDev), 256);
    t.segs_in = t.tiles + align_up(std::max<uint64_t>(1, tiles) * sizeof(uint2), 256);
    t.segs_out = t.segs_in + requests * kSegsPerRequest * sizeof(CopySeg);
    t.total = t.segs_out + requests * kSegsPerRequest * sizeof(CopySeg);
    return t;
}

// Device footprint of one launch besides its requests' buffers: the table
// block above and the level-9 handoff list.
struct LaunchNeeds {
    std::vector<uint32_t> tile_blocks;
    uint64_t tiles = 0;
    size_t table_bytes = 0, refine_bytes = 0;
};

LaunchNeeds launch_needs(const std::vector<Request*>& reqs, uint32_t grid) {
    LaunchNeeds n;
    uint64_t total = 0, level9 = 0;
    for (Request* r : reqs) total += r->job.block_count;
    for (Request* r : reqs) {
        const uint64_t b = r->job.block_count;
        const uint32_t t = r->job.tile_blocks ? r->job.tile_blocks : auto_tile_blocks(b, total, grid);
        n.tile_blocks.push_back(t);
        n.tiles += (b + t - 1) / t;
        if (r->job.level == 9) level9 += b;
    }
    n.table_bytes = table_layout(reqs.size(), n.tiles).total;
    n.refine_bytes = level9 ? refine_list_bytes(level9) : 0;
    return n;
}

// Takes this batch's table and handoff blocks from the pool (under the lock);
// false, with nothing taken, when the budget has no room for them now.
bool icbc_server::launch_resources(Slot& s, const std::vector<Request*>& reqs) {
    LaunchNeeds n = launch_needs(reqs, grid);
    const size_t tc = Pool::class_of(n.table_bytes);
    void* t = nullptr;
    if (pool.take(tc, &t) != hipSuccess) return false;
    void* rf = nullptr;
    size_t rc = 0;
    if (n.refine_bytes) {
        rc = Pool::class_of(n.refine_bytes);
        if (pool.take(rc, &rf) != hipSuccess) {
            pool.give(tc, t);
            return false;
        }
    }
    s.table = t;
    s.table_cls = tc;
    s.refine = rf;
    s.refine_cls = rc;
    s.tile_blocks = std::move(n.tile_blocks);
    s.ntiles = n.tiles;
    used += tc + rc;
    return true;
}

void icbc_server::release_launch_resources(Slot& s) {
    if (s.table) {
        pool.give(s.table_cls, s.table);
        used -= s.table_cls;
    }
    if (s.refine) {
        pool.give(s.refine_cls, s.refine);
        used -= s.refine_cls;
    }
    s.table = s.refine = nullptr;
    s.table_cls = s.refine_cls = 0;
}

int icbc_server::launch(Slot& s, std::vector<Request*>& reqs) {
    // pageable inputs are staged in page-locked memory first; then, on the
    // upload streams, one table copy (jobs, tiles, copy lists), a DMA copy per
    // input buffer of kDmaMinBytes or more and one copy kernel for the
    // others; on the slot's stream, once both uploads are done, the encode,
    // one copy kernel out and the completion event
    size_t in_bytes = 0, out_bytes = 0;
    for (Request* r : reqs) {
        const uint64_t n = r->job.block_count;
        if (!r->map_colors) in_bytes += align_up(r->lay.color_bytes, 256);
        if (r->job.weights && !r->map_weights) in_bytes += align_up(n * 64, 256);
        if (!r->map_codes) in_bytes += align_up(n * 8, 256);
        if (!r->map_codes) out_bytes += align_up(n * 8, 256);
        if (r->job.errors && !r->map_errors) out_bytes += align_up(n * 4, 256);
        if (r->job.stats && !r->map_stats) out_bytes += align_up(n * 16, 256);
    }
    if (!s.in.reserve(std::max<size_t>(in_bytes, 256)) || !s.out.reserve(std::max<size_t>(out_bytes, 256)))
        return ICBC_ERR_OUT_OF_MEMORY;
    std::vector<JobDev> dev;
    std::vector<CopySeg> in, out;
    size_t in_off = 0, out_off = 0;
    s.out_off.assign(reqs.size(), 0);
    // An input buffer's copy: a DMA copy from its page-locked host address
    // when large (the DMA engines keep their full rate while the encode
    // kernels hold every compute unit), else a segment of the copy kernel
    // from its device-side address (mapped in place, or staged).  A DMA'd
    // buffer keeps an empty segment, so every request has three.
    std::vector<CopySeg> dma;
    auto stage_in = [&](const void* mapped, const void* src, void* dst, size_t bytes) -> CopySeg {
        const void* host = src;
        const void* dev_src = mapped;
        if (!mapped && bytes) {
            memcpy((char*)s.in.p + in_off, src, bytes);
            host = (const char*)s.in.p + in_off;
            dev_src = (const char*)s.in.dp + in_off;
            in_off += align_up(bytes, 256);
        }
        if (bytes >= kDmaMinBytes) {
            dma.push_back({host, dst, bytes, 0});
            return {dev_src, dst, 0, 0};
        }
        return {dev_src, dst, bytes, 0};
    };
    auto stage_out = [&](void* mapped, size_t bytes) -> void* {
        if (mapped) return mapped;
        void* d = (char*)s.out.dp + out_off;
        out_off += align_up(bytes, 256);
        return d;
    };
    for (size_t k = 0; k < reqs.size(); k++) {
        Request* r = reqs[k];
        const uint64_t n = r->job.block_count;
        char* base = (char*)r->dev;
        icbc_bc1_job dj = r->job;
        dj.flags = r->job.flags & ICBC_JOB_COLORS_RGBA8;  // RGBA8 texels are read as they are
        dj.colors = (const float*)(base + r->lay.colors);
        dj.weights = r->job.weights ? (const float*)(base + r->lay.weights) : nullptr;
        dj.codes = base + r->lay.codes;
        dj.errors = r->job.errors ? (float*)(base + r->lay.errors) : nullptr;
        dj.stats = r->job.stats ? (int32_t*)(base + r->lay.stats) : nullptr;
        JobDev d = job_to_device(dj, decoder);
        d.tile_blocks = (int32_t)s.tile_blocks[k];
        dev.push_back(d);
        // in: colours, weights, prior codes (blocks the encoder leaves
        // untouched keep them); out: codes, errors, counters
        in.push_back(stage_in(r->map_colors, r->job.colors, base + r->lay.colors, n ? r->lay.color_bytes : 0));
        in.push_back(r->job.weights ? stage_in(r->map_weights, r->job.weights, base + r->lay.weights, n * 64)
                                    : CopySeg{nullptr, base + r->lay.weights, 0, 0});
        in.push_back(stage_in(r->map_codes, r->job.codes, base + r->lay.codes, n * 8));
        s.out_off[k] = out_off;
        out.push_back({base + r->lay.codes, stage_out(r->map_codes, n * 8), n * 8, 0});
        out.push_back({base + r->lay.errors, r->job.errors ? stage_out(r->map_errors, n * 4) : nullptr,
                       r->job.errors ? n * 4 : 0, 0});
        out.push_back({base + r->lay.stats, r->job.stats ? stage_out(r->map_stats, n * 16) : nullptr,
                       r->job.stats ? n * 16 : 0, 0});
    }
    const uint64_t chunks_in = number_copy_segments(in.data(), (uint32_t)in.size());
    const uint64_t chunks_out = number_copy_segments(out.data(), (uint32_t)out.size());
    std::vector<uint2> tiles = tiles_for(dev, true);
    if (tiles.size() != s.ntiles) return ICBC_ERR_DEVICE;  // the table block was sized for s.ntiles
    const TableLayout tl = table_layout(reqs.size(), tiles.size());
    if (!s.host_table.reserve(tl.total)) return ICBC_ERR_OUT_OF_MEMORY;
    char* ht = (char*)s.host_table.p;
    memcpy(ht, dev.data(), dev.size() * sizeof(JobDev));
    memcpy(ht + tl.tiles, tiles.data(), tiles.size() * sizeof(uint2));
    memcpy(ht + tl.segs_in, in.data(), in.size() * sizeof(CopySeg));
    memcpy(ht + tl.segs_out, out.data(), out.size() * sizeof(CopySeg));
    char* dt = (char*)s.table;
    for (const CopySeg& c : dma)
        if (hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, upload_dma) != hipSuccess)
            return ICBC_ERR_DEVICE;
    if (hipEventRecord(s.uploaded_dma, upload_dma) != hipSuccess ||
        hipMemcpyAsync(dt, ht, tl.total, hipMemcpyHostToDevice, upload) != hipSuccess ||
        launch_copy_segments((const CopySeg*)(dt + tl.segs_in), (uint32_t)in.size(), chunks_in, copy_grid, upload) !=
            hipSuccess ||
        hipEventRecord(s.uploaded, upload) != hipSuccess || hipStreamWaitEvent(s.stream, s.uploaded, 0) != hipSuccess ||
        hipStreamWaitEvent(s.stream, s.uploaded_dma, 0) != hipSuccess)
        return ICBC_ERR_DEVICE;
    const RefineList rl = s.refine ? RefineList{(RefineItem*)((char*)s.refine + kRefineCounterBytes), (uint32_t*)s.refine}
                                   : RefineList{nullptr, nullptr};
    if (launch_encode((const JobDev*)dt, (const uint2*)(dt + tl.tiles), (uint32_t)tiles.size(), (uint32_t*)s.work,
                      (float4*)((char*)s.work + kQueueBytes), grid, encode_mode(dev.data(), dev.size()), rl,
                      s.stream) != hipSuccess)
        return ICBC_ERR_DEVICE;
    if (launch_copy_segments((const CopySeg*)(dt + tl.segs_out), (uint32_t)out.size(), chunks_out, copy_grid,
                             s.stream) != hipSuccess ||
        hipEventRecord(s.done, s.stream) != hipSuccess)
        return ICBC_ERR_DEVICE;
    return ICBC_OK;
}

void icbc_server::run_scheduler() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        if (fatal != ICBC_OK && (!admitted.empty() || !pending.empty())) {
            // after a device failure nothing more is launched: requests not yet
            // launched end with the failure's status and give back their memory
            const double t = now_s();
            for (Request* r : admitted) {
                pool.give(r->cls, r->dev);
                r->dev = nullptr;
                used -= r->cls;
            }
            auto finish = [&](Request* r) {
                r->status = fatal;
                r->state = kDone;
                r->t_done = t;
            };
            for (Request* r : admitted) finish(r);
            for (Request* r : pending) finish(r);
            admitted.clear();
            pending.clear();
            cv_done.notify_all();
        }
        // admit in arrival order while the budget allows (the pool refuses
        // blocks the budget has no room for) and the admitted requests hold
        // fewer than kAdmitAheadBlocks blocks; not while a batch waits for the
        // room its launch tables need, which admissions would take again
        uint64_t ahead = 0;
        for (const Request* r : admitted) ahead += r->job.block_count;
        while (!pending.empty() && !hold_admission && ahead < kAdmitAheadBlocks) {
            Request* r = pending.front();
            // the pool is only touched under the lock (the reaper returns blocks)
            const hipError_t te = pool.take(r->cls, &r->dev);
            if (te != hipSuccess) {
                if (!admitted.empty() || !inflight.empty()) break;  // retry once memory frees up
                pending.pop_front();
                r->status = te == hipErrorOutOfMemory ? ICBC_ERR_BUDGET : ICBC_ERR_DEVICE;
                r->state = kDone;
                r->t_done = now_s();
                cv_done.notify_all();
                continue;
            }
            pending.pop_front();
            used += r->cls;
            r->state = kAdmitted;
            r->t_admit = now_s();
            admitted.push_back(r);
            ahead += r->job.block_count;
        }
        Slot* free_slot = nullptr;
        for (Slot& s : slots)
            if (!s.busy) {
                free_slot = &s;
                break;
            }
        if (!admitted.empty() && free_slot && fatal == ICBC_OK) {
            std::vector<Request*> batch;
            size_t take = opt.max_batch ? std::min<size_t>(opt.max_batch, admitted.size()) : admitted.size();
            // and at most kMaxLaunchBlocks blocks (a larger request runs alone)
            uint64_t blocks = admitted[0]->job.block_count;
            for (size_t k = 1; k < take; k++) {
                blocks += admitted[k]->job.block_count;
                if (blocks > kMaxLaunchBlocks) {
                    take = k;
                    break;
                }
            }
            // the launch's table and handoff list come from the budget too: a
            // batch they do not fit in is halved until they do
            bool ok = false;
            for (;;) {
                batch.assign(admitted.begin(), admitted.begin() + take);
                if (launch_resources(*free_slot, batch)) {
                    ok = true;
                    break;
                }
                if (take == 1) break;
                take = (take + 1) / 2;
            }
            if (!ok) {
                hold_admission = true;
                if (inflight.empty() && admitted.size() > 1) {
                    // the other admitted requests hold the room this launch
                    // needs: they go back to the queue (in order) and wait
                    for (size_t k = admitted.size(); k-- > 1;) {
                        Request* r = admitted[k];
                        pool.give(r->cls, r->dev);
                        r->dev = nullptr;
                        used -= r->cls;
                        r->state = kPending;
                        pending.push_front(r);
                    }
                    admitted.resize(1);
                    continue;
                }
                if (inflight.empty()) {
                    // nothing else holds memory: this request can never run
                    // within the budget
                    hold_admission = false;
                    Request* r = admitted.front();
                    admitted.erase(admitted.begin());
                    pool.give(r->cls, r->dev);
                    r->dev = nullptr;
                    used -= r->cls;
                    r->status = ICBC_ERR_BUDGET;
                    r->state = kDone;
                    r->t_done = now_s();
                    cv_done.notify_all();
                    continue;
                }
                cv_sched.wait(lk);  // a completing launch gives memory back
                continue;
            }
            admitted.erase(admitted.begin(), admitted.begin() + take);
            hold_admission = false;
            free_slot->busy = true;
            free_slot->reqs = batch;
            lk.unlock();
            const double t = now_s();
            const int r = launch(*free_slot, batch);
            lk.lock();
            for (Request* q : batch) {
                q->state = kLaunched;
                q->t_launch = t;
                q->batch_size = (uint32_t)batch.size();
            }
            if (r != ICBC_OK) {
                set_fatal(this, r);
                for (Request* q : batch) q->status = r;
            }
            launches++;
            inflight.push_back(free_slot);
            cv_reap.notify_one();
            continue;
        }
        if (stop && pending.empty() && admitted.empty()) break;
        cv_sched.wait(lk);
    }
    sched_done = true;
    cv_reap.notify_all();
}

void icbc_server::run_reaper() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        while (inflight.empty()) {
            if (sched_done) return;
            cv_reap.wait(lk);
        }
        Slot* s = inflight.front();
        lk.unlock();
        const hipError_t e = hipEventSynchronize(s->done);
        if (e == hipSuccess) {
            // outputs of pageable buffers wait in staging (mapped ones are
            // already in place)
            for (size_t k = 0; k < s->reqs.size(); k++) {
                Request* r = s->reqs[k];
                const uint64_t n = r->job.block_count;
                if (n == 0 || r->status != ICBC_OK) continue;
                const char* src = (const char*)s->out.p + s->out_off[k];
                if (!r->map_codes) {
                    memcpy(r->job.codes, src, n * 8);
                    src += align_up(n * 8, 256);
                }
                if (r->job.errors && !r->map_errors) {
                    memcpy(r->job.errors, src, n * 4);
                    src += align_up(n * 4, 256);
                }
                if (r->job.stats && !r->map_stats) memcpy(r->job.stats, src, n * 16);
            }
        }
        const double t = now_s();
        lk.lock();
        inflight.pop_front();
        for (Request* r : s->reqs) {
            if (e != hipSuccess && r->status == ICBC_OK) r->status = ICBC_ERR_DEVICE;
            r->state = kDone;
            r->t_done = t;
            pool.give(r->cls, r->dev);
            r->dev = nullptr;
            used -= r->cls;
            completed++;
            blocks_done += r->job.block_count;
        }
        if (e != hipSuccess) set_fatal(this, ICBC_ERR_DEVICE);
        release_launch_resources(*s);
        s->reqs.clear();
        s->busy = false;
        cv_done.notify_all();
        cv_sched.notify_one();
    }
}

extern "C" {

int icbc_server_create(const icbc_server_options* opt, icbc_server** out) {
    if (!out || !opt) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (opt->memory_budget < (64u << 10))
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "memory budget below 64 KiB");
    int decoder;
    int r = icbc_internal_ready(&decoder);
    if (r) return r;
    uint32_t grid;
    r = icbc_internal_resident(&grid);
    if (r) return r;
    icbc_server_options o = *opt;
    if (o.max_inflight == 0) o.max_inflight = 3;
    if (o.max_inflight > 8) o.max_inflight = 8;
    // fixed device memory of a slot besides its workspace: its stream's
    // runtime allocations
    const uint64_t slot_fixed = kStreamAllowance;
    // the launch slots (fixed memory and workspace) take at most half the
    // budget: as many resident workgroups as that allows, up to the
    // device's residency, and no fewer than kMinServerGrid (the upload
    // streams' allowances come out of the other half)
    const size_t per_wg = kWorkspaceBytesPerWorkgroup;
    const uint64_t share = o.memory_budget / 2 / o.max_inflight / kGranule * kGranule;  // per slot
    const uint64_t fit = share > kQueueBytes + slot_fixed ? (share - kQueueBytes - slot_fixed) / per_wg : 0;
    if (fit < kMinServerGrid)
        return icbc_internal_fail(ICBC_ERR_BUDGET, "memory budget %llu B is below the %llu B that %u launch slots "
                                  "need", (unsigned long long)o.memory_budget,
                                  (unsigned long long)(2ull * o.max_inflight *
                                                       (slot_fixed +
                                                        align_up(kQueueBytes + kMinServerGrid * per_wg, kGranule))),
                                  o.max_inflight);
    icbc_server* sv = new icbc_server(o.memory_budget);
    sv->opt = o;
    sv->decoder = decoder;
    // one workgroup slot per compute unit is left to the copy kernels, so that
    // a launch's copies can start beside a running encode kernel
    const uint32_t cus = std::max<uint32_t>(1, grid / (uint32_t)encode_workgroups_per_cu());
    sv->grid = (uint32_t)std::min<uint64_t>(grid > cus ? grid - cus : grid, fit);
    sv->copy_grid = cus;
    sv->slot_bytes = slot_fixed + align_up(kQueueBytes + (uint64_t)sv->grid * per_wg, kGranule);
    if (2 * kStreamAllowance + o.max_inflight * sv->slot_bytes > o.memory_budget) {
        delete sv;
        return icbc_internal_fail(ICBC_ERR_BUDGET, "memory budget %llu B is below the %llu B that the streams and "
                                  "%u launch slots hold", (unsigned long long)o.memory_budget,
                                  (unsigned long long)(2 * kStreamAllowance + o.max_inflight * sv->slot_bytes),
                                  o.max_inflight);
    }
    sv->slots.resize(sv->opt.max_inflight);
    if (hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&sv->upload_dma, hipStreamNonBlocking) != hipSuccess ||
        !sv->pool.charge(2 * kStreamAllowance)) {
        icbc_server_destroy(sv);
        return icbc_internal_fail(ICBC_ERR_OUT_OF_MEMORY, "server upload stream creation failed");
    }
    for (Slot& s : sv->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.uploaded, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.uploaded_dma, hipEventDisableTiming) != hipSuccess ||
            !sv->pool.charge(slot_fixed) ||
            sv->pool.alloc_raw(kQueueBytes + (size_t)sv->grid * per_wg, &s.work) != hipSuccess) {
            icbc_server_destroy(sv);
            return icbc_internal_fail(ICBC_ERR_OUT_OF_MEMORY, "server launch slot allocation failed");
        }
    }
    sv->sched = std::thread([sv] { sv->run_scheduler(); });
    sv->reap = std::thread([sv] { sv->run_reaper(); });
    *out = sv;
    return ICBC_OK;
}

int icbc_server_submit(icbc_server* sv, const icbc_bc1_job* job, uint64_t* ticket) {
    if (!sv || !job || !ticket) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    if (!(job->flags & ICBC_JOB_HOST_MEMORY))
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "server requests take host memory (ICBC_JOB_HOST_MEMORY)");
    int r = icbc_internal_validate(*job);
    if (r) return r;
    Request* q = new Request();
    q->job = *job;
    q->lay = layout_of(*job);
    q->cls = Pool::class_of(q->lay.total);
    q->map_colors = mapped_address(job->colors);
    q->map_weights = mapped_address(job->weights);
    q->map_codes = mapped_address(job->codes);
    q->map_errors = mapped_address(job->errors);
    q->map_stats = mapped_address(job->stats);
    q->t_submit = now_s();
    std::lock_guard<std::mutex> lk(sv->m);
    if (sv->fatal) {
        delete q;
        return icbc_internal_fail(sv->fatal, "server stopped after a device failure");
    }
    if (sv->stop) {
        delete q;
        return icbc_internal_fail(ICBC_ERR_CANCELLED, "server is shutting down");
    }
    {
        // alone in a launch: its buffers, its table and tile list, its handoff
        // list, next to the slots' workspaces
        std::vector<Request*> one{q};
        const LaunchNeeds n = launch_needs(one, sv->grid);
        const uint64_t need = q->cls + Pool::class_of(n.table_bytes) +
                              (n.refine_bytes ? Pool::class_of(n.refine_bytes) : 0);
        const uint64_t room = sv->opt.memory_budget - 2 * kStreamAllowance - sv->slots.size() * sv->slot_bytes;
        if (need > room) {
            delete q;
            return icbc_internal_fail(ICBC_ERR_BUDGET, "request needs %llu bytes, the budget leaves %llu for requests",
                                      (unsigned long long)need, (unsigned long long)room);
        }
    }
    q->id = sv->next_id++;
    sv->all[q->id] = q;
    sv->pending.push_back(q);
    *ticket = q->id;
    sv->cv_sched.notify_one();
    return ICBC_OK;
}

int icbc_server_wait(icbc_server* sv, uint64_t ticket, double timeout_s, icbc_request_info* info) {
    if (!sv) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(sv->m);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "unknown ticket %llu",
                                                        (unsigned long long)ticket);
    // the ticket is looked up again after every wake-up: another thread may
    // have collected it meanwhile
    const auto pred = [&] {
        it = sv->all.find(ticket);
        return it == sv->all.end() || it->second->state == kDone;
    };
    if (timeout_s < 0) sv->cv_done.wait(lk, pred);
    else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return ICBC_ERR_TIMEOUT;
    if (it == sv->all.end())
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "ticket %llu was collected by another wait",
                                  (unsigned long long)ticket);
    Request* q = it->second;
    if (info) {
        info->status = q->status;
        info->batch_size = q->batch_size;
        info->block_count = q->job.block_count;
        info->submit_s = q->t_submit;
        info->admit_s = q->t_admit;
        info->launch_s = q->t_launch;
        info->done_s = q->t_done;
    }
    const int status = q->status;
    sv->all.erase(it);
    delete q;
    return status;
}

int icbc_server_cancel(icbc_server* sv, uint64_t ticket) {
    if (!sv) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(sv->m);
    auto it = sv->all.find(ticket);
    if (it == sv->all.end()) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "unknown ticket");
    Request* q = it->second;
    if (q->state != kPending) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "request already admitted");
    sv->pending.erase(std::find(sv->pending.begin(), sv->pending.end(), q));
    q->state = kDone;
    q->status = ICBC_ERR_CANCELLED;
    q->t_done = now_s();
    sv->cv_done.notify_all();
    return ICBC_OK;
}

int icbc_server_stats(icbc_server* sv, icbc_server_stats_t* st) {
    if (!sv || !st) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(sv->m);
    st->launches = sv->launches;
    st->completed = sv->completed;
    st->blocks = sv->blocks_done;
    st->pending = sv->pending.size();
    st->inflight = sv->inflight.size();
    st->memory_in_use = sv->used;
    st->memory_committed = sv->pool.committed();
    return ICBC_OK;
}

int icbc_server_destroy(icbc_server* sv) {
    if (!sv) return ICBC_OK;
    {
        std::lock_guard<std::mutex> lk(sv->m);
        sv->stop = true;
        sv->cv_sched.notify_all();
        sv->cv_reap.notify_all();
    }
    if (sv->sched.joinable()) sv->sched.join();
    if (sv->reap.joinable()) sv->reap.join();
    for (Slot& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.uploaded) (void)hipEventDestroy(s.uploaded);
        if (s.uploaded_dma) (void)hipEventDestroy(s.uploaded_dma);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (hipStream_t t : {sv->upload, sv->upload_dma}) {
        if (!t) continue;
        (void)hipStreamSynchronize(t);
        (void)hipStreamDestroy(t);
    }
    for (auto& kv : sv->all) delete kv.second;
    delete sv;
    return ICBC_OK;
}

}  // extern "C"
