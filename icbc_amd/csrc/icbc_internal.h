// Internal interface between the runtime (icbc_runtime.cpp) and the request
// server (icbc_server.cpp).  Not part of the public ABI.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/icbc_gpu.h"
#include "bc1_kernels.h"

// Records `fmt` as this thread's last error and returns `code`.
int icbc_internal_fail(int code, const char* fmt, ...);
// Tables on the current device; the decoder selected by init.
int icbc_internal_ready(int* decoder);
// Encode workgroups resident on the current device at once.
int icbc_internal_resident(uint32_t* grid);
// Argument checks shared by every entry point taking a job.
int icbc_internal_validate(const icbc_bc1_job& job);
// One block through the coalescing ring of single-block calls (block_ring.cpp):
// the reference's compress_bc1 semantics, a status code out.
int icbc_internal_compress_one(int level, const float* colors, const float* weights, const float cw[3], int tcm,
                               int tcb, void* output, float* error);
// The current device's tuning (false when no level is tuned) / replace it.
bool icbc_internal_get_tuning(icbc_tuning* out);
void icbc_internal_set_tuning(const icbc_tuning* t);

namespace icbc_gpu {

JobDev job_to_device(const icbc_bc1_job& j, int decoder);
uint32_t auto_tile_blocks(uint64_t job_blocks, uint64_t total_blocks, uint32_t resident);
// Tile list of a set of jobs; with `order_by_cost`, expensive tiles first.
std::vector<uint2> tiles_for(const std::vector<JobDev>& dev, bool order_by_cost);

}  // namespace icbc_gpu
