// Per-thread resources of the single-block entry points (icbc::compress_bc1,
// and decode / evaluate of one block): one stream, one device buffer and one
// page-locked host buffer per device the thread has used.
//
// The three are made together or not at all.  When a step fails, what was
// already made is released and nothing is recorded for the device, so the next
// call on that thread tries again instead of running on a half-made context.
//
// `Api` supplies the runtime calls (HIP in the library, a fake in the CPU
// tests): stream_create / stream_destroy, device_alloc / device_free,
// host_alloc / host_free; the create and alloc calls return 0 or an error
// status, which is passed on.  No HIP types appear here.
#pragma once

#include <vector>

struct BlockResources {
    int device = -1;
    void* stream = nullptr;
    void* dbuf = nullptr;
    void* hbuf = nullptr;
};

// Finds or makes the resources of `device` in `ctx`; `out` is valid until the
// next call with the same `ctx`.
template <typename Api>
int block_resources_for(Api& api, std::vector<BlockResources>& ctx, int device, BlockResources*& out) {
    for (BlockResources& r : ctx)
        if (r.device == device) {
            out = &r;
            return 0;
        }
    BlockResources r;
    int e = api.stream_create(&r.stream);
    if (e) return e;
    e = api.device_alloc(&r.dbuf);
    if (e) {
        api.stream_destroy(r.stream);
        return e;
    }
    e = api.host_alloc(&r.hbuf);
    if (e) {
        api.device_free(r.dbuf);
        api.stream_destroy(r.stream);
        return e;
    }
    r.device = device;
    ctx.push_back(r);
    out = &ctx.back();
    return 0;
}
