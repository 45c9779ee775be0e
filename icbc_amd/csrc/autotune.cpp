// Autotuner: per quality level, the launch setting (blocks per wavefront tile,
// resident workgroups per CU, schedule) and the four-cluster search schedule
// (the colour count from which the pruning bound pass runs) that encode the
// caller's sample fastest on the current device.  Each candidate is a
// graph-captured plan, timed with HIP events over a few executes after one
// warm-up; the untuned default's time is kept beside the pick.

#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/icbc_gpu.h"
#include "bc1_kernels.h"
#include "icbc_internal.h"

using icbc_gpu::encode_workgroups_per_cu;
using icbc_gpu::kGridScheduleMaxWaves;

namespace {

struct Setting {
    uint32_t tile;
    int32_t schedule;
    int32_t per_cu;  // 0 = every workgroup that fits
    int32_t prune;   // four-cluster pruning from this many colours (0 = default)
};

// Milliseconds per execute (best of `reps`) of one setting, or < 0 on failure.
// The jobs' search schedule comes from the device's tuning table (read when
// the plan is built), so `tune` is installed with the setting's for `level`.
float time_setting(std::vector<icbc_bc1_job>& jobs, const Setting& st, int level, icbc_tuning& tune, hipStream_t s,
                   hipEvent_t e0, hipEvent_t e1, int reps) {
    tune.prune_min_colors[level] = st.prune;
    icbc_internal_set_tuning(&tune);
    for (auto& j : jobs) j.tile_blocks = st.tile;
    icbc_plan_options o{st.schedule, 1, st.per_cu, 1};
    icbc_bc1_plan* plan = nullptr;
    if (icbc_gpu_plan_create(jobs.data(), jobs.size(), &o, s, &plan) != ICBC_OK) return -1.0f;
    float best = -1.0f;
    if (icbc_gpu_plan_execute(plan, s) == ICBC_OK) {
        for (int r = 0; r < reps; r++) {
            float ms = 0.0f;
            if (hipEventRecord(e0, s) != hipSuccess || icbc_gpu_plan_execute(plan, s) != ICBC_OK ||
                hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                best = -1.0f;
                break;
            }
            if (best < 0.0f || ms < best) best = ms;
        }
    }
    (void)hipStreamSynchronize(s);
    icbc_gpu_plan_destroy(plan);
    return best;
}

}  // namespace

extern "C" {

int icbc_gpu_autotune(const icbc_bc1_job* sample, size_t n_jobs, uint32_t level_mask, icbc_tuning* out) {
    if (!sample || n_jobs == 0) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "empty tuning sample");
    if (level_mask == 0 || (level_mask >> 10) != 0)
        return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "level_mask must select levels 0..9");
    uint64_t blocks = 0;
    for (size_t k = 0; k < n_jobs; k++) {
        if (sample[k].flags & ICBC_JOB_HOST_MEMORY)
            return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "job %zu: tuning samples take device memory", k);
        int r = icbc_internal_validate(sample[k]);
        if (r) return r;
        blocks += sample[k].block_count;
    }
    if (blocks == 0) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "empty tuning sample");
    int decoder;
    int r = icbc_internal_ready(&decoder);
    if (r) return r;
    const int full = encode_workgroups_per_cu();
    uint32_t resident = 0;
    r = icbc_internal_resident(&resident);
    if (r) return r;

    // One workgroup per tile needs a workspace slice per tile, so that
    // schedule is only tried (and only ever chosen) for launches of a few
    // waves of workgroups.
    std::vector<Setting> settings;
    for (uint32_t tile : {8u, 16u, 32u, 64u}) {
        settings.push_back({tile, ICBC_SCHEDULE_PERSISTENT, 0, 0});
        if (full > 1) settings.push_back({tile, ICBC_SCHEDULE_PERSISTENT, std::max(1, full / 2), 0});
        uint64_t tiles = 0;
        for (size_t k = 0; k < n_jobs; k++) tiles += (sample[k].block_count + tile - 1) / tile;
        if (tiles <= kGridScheduleMaxWaves * (uint64_t)resident) settings.push_back({tile, ICBC_SCHEDULE_GRID, 0, 0});
    }

    icbc_tuning tune;
    if (!icbc_internal_get_tuning(&tune)) memset(&tune, 0, sizeof tune);
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        if (s) (void)hipStreamDestroy(s);
        if (e0) (void)hipEventDestroy(e0);
        return icbc_internal_fail(ICBC_ERR_DEVICE, "autotune: stream/event creation failed");
    }
    std::vector<icbc_bc1_job> jobs(sample, sample + n_jobs);
    r = ICBC_OK;
    const icbc_tuning before = tune;
    for (int level = 0; level < 10 && r == ICBC_OK; level++) {
        if (!(level_mask & (1u << level))) continue;
        for (auto& j : jobs) j.level = level;
        float best = -1.0f, deflt = -1.0f;
        Setting pick{};
        auto consider = [&](const Setting& st) {
            const float ms = time_setting(jobs, st, level, tune, s, e0, e1, 3);
            if (ms < 0.0f) {
                r = icbc_internal_fail(ICBC_ERR_DEVICE, "autotune: level %d, tile %u failed", level, st.tile);
                return;
            }
            if (st.tile == 64 && st.schedule == ICBC_SCHEDULE_PERSISTENT && st.per_cu == 0 && st.prune == 0)
                deflt = ms;
            if (best < 0.0f || ms < best) {
                best = ms;
                pick = st;
            }
        };
        // stage 1: launch settings, default search schedule
        for (const Setting& st : settings) {
            consider(st);
            if (r) break;
        }
        // stage 2: the search schedule on the fastest launch setting (levels
        // 2-9 run the four-cluster search)
        if (!r && level >= 2) {
            const Setting base = pick;
            for (int32_t prune : {9, 11, 15, icbc_gpu::kNoPruneColors}) {
                consider(Setting{base.tile, base.schedule, base.per_cu, prune});
                if (r) break;
            }
        }
        if (r) break;
        tune.tile_blocks[level] = pick.tile;
        tune.schedule[level] = pick.schedule;
        tune.workgroups_per_cu[level] = pick.per_cu;
        tune.prune_min_colors[level] = pick.prune;
        tune.ns_per_block[level] = (float)(best * 1e6 / (double)blocks);
        tune.default_ns_per_block[level] = (float)(deflt * 1e6 / (double)blocks);
    }
    if (r) icbc_internal_set_tuning(&before);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (r) return r;
    icbc_internal_set_tuning(&tune);
    if (out) *out = tune;
    return ICBC_OK;
}

int icbc_gpu_get_tuning(icbc_tuning* out) {
    if (!out) return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "null tuning");
    if (!icbc_internal_get_tuning(out)) memset(out, 0, sizeof *out);
    return ICBC_OK;
}

int icbc_gpu_set_tuning(const icbc_tuning* tuning) {
    if (tuning)
        for (int l = 0; l < 10; l++) {
            const uint32_t t = tuning->tile_blocks[l];
            const int32_t pr = tuning->prune_min_colors[l];
            if ((pr != 0 && (pr < icbc_gpu::kMinPruneColors || pr > icbc_gpu::kNoPruneColors)) ||
                !(tuning->default_ns_per_block[l] >= 0.0f) ||
                (t != 0 && t != 8 && t != 16 && t != 32 && t != 64) ||
                (tuning->schedule[l] != ICBC_SCHEDULE_GRID && tuning->schedule[l] != ICBC_SCHEDULE_PERSISTENT) ||
                tuning->workgroups_per_cu[l] < 0 || !(tuning->ns_per_block[l] >= 0.0f))
                return icbc_internal_fail(ICBC_ERR_INVALID_ARGUMENT, "tuning entry for level %d is invalid", l);
        }
    icbc_internal_set_tuning(tuning);
    return ICBC_OK;
}

}  // extern "C"
