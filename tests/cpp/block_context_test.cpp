// CPU test of the single-block context setup (icbc_amd/csrc/block_context.h)
// with a fake runtime that fails at a chosen step: a failed setup leaves
// nothing recorded and nothing allocated, and the next call sets up afresh.
#include <cstdio>
#include <vector>

#include "block_context.h"

struct FakeApi {
    int fail_at = -1;  // 0 stream, 1 device buffer, 2 host buffer; -1 none
    int step = 0, live = 0, creates = 0;
    char mem[3];
    int make(void** p, int k) {
        const int s = step++;
        if (s == fail_at) return -3;
        *p = &mem[k];
        live++;
        creates++;
        return 0;
    }
    int stream_create(void** s) { return make(s, 0); }
    int device_alloc(void** p) { return make(p, 1); }
    int host_alloc(void** p) { return make(p, 2); }
    void stream_destroy(void*) { live--; }
    void device_free(void*) { live--; }
    void host_free(void*) { live--; }
};

#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            return 1;                                             \
        }                                                         \
    } while (0)

int main() {
    for (int fail_at = 0; fail_at < 3; fail_at++) {
        std::vector<BlockResources> ctx;
        FakeApi api;
        api.fail_at = fail_at;
        BlockResources* r = nullptr;
        CHECK(block_resources_for(api, ctx, 0, r) == -3);
        CHECK(ctx.empty());   // nothing recorded for the device
        CHECK(api.live == 0); // what was made is released
        // the next call tries again and succeeds
        api.fail_at = -1;
        api.step = 0;
        CHECK(block_resources_for(api, ctx, 0, r) == 0);
        CHECK(ctx.size() == 1 && r == &ctx[0] && r->device == 0);
        CHECK(r->stream && r->dbuf && r->hbuf && api.live == 3);
        // found again without new allocations
        const int before = api.creates;
        CHECK(block_resources_for(api, ctx, 0, r) == 0 && api.creates == before);
        // another device gets its own set
        CHECK(block_resources_for(api, ctx, 1, r) == 0 && ctx.size() == 2 && r->device == 1);
    }
    std::printf("ok\n");
    return 0;
}
