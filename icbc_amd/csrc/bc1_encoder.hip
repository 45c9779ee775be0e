// BC1 encoder kernels for CDNA4 (gfx950).
//
// Work unit: a tile of 64 consecutive 4x4 blocks of one job, owned by one
// 64-lane wavefront (one workgroup).  Lane b owns block b of the tile for the
// serial, order-dependent stages (colour reduction, principal axis, sort,
// prefix sums, palette/index/error evaluation, least squares, box fit,
// refinement).  The cluster-fit searches -- the hot loop, up to 968
// candidates per block at levels 2-9 (icbc.h:2348-2743) -- run wave-
// cooperatively: for each block of the tile the 64 lanes split that block's
// candidates, read its summed-area table from LDS, and reduce (error, index)
// across the wave.  The winning candidate is recomputed by the owning lane.
//
// Bit-exactness: see bc1_device.h.  The search keeps, like the reference's
// scalar loop, the lowest-index candidate among those with the smallest error
// below FLT_MAX (NaN never wins); with no such candidate both endpoints are 0.

#include "bc1_device.h"
#include "bc1_kernels.h"

namespace icbc_gpu {

__constant__ float k_u8_to_unit[256];
__constant__ float k_mid5[32];
__constant__ float k_mid6[64];
__constant__ uint32_t k_four[976];
__constant__ uint32_t k_three[160];
__constant__ int k_four_total[16];
__constant__ int k_three_total[16];
__constant__ uint16_t k_match5[4][256];
__constant__ uint16_t k_match6[4][256];
__constant__ DecoderWeights k_dweights[4];

hipError_t upload_tables(const HostTables& t, hipStream_t stream) {
    hipError_t e;
#define UP(sym, src)                                                                          \
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(sym), &(src), sizeof(src), 0, hipMemcpyHostToDevice, \
                               stream);                                                       \
    if (e != hipSuccess) return e;
    UP(k_u8_to_unit, t.u8_to_unit);
    UP(k_mid5, t.mid5);
    UP(k_mid6, t.mid6);
    UP(k_four, t.four);
    UP(k_three, t.three);
    UP(k_four_total, t.four_total);
    UP(k_three_total, t.three_total);
    UP(k_match5, t.match5);
    UP(k_match6, t.match6);
    UP(k_dweights, t.dweights);
#undef UP
    return hipStreamSynchronize(stream);
}

// ---------------------------------------------------------------------------
// wave-wide argmin over (error, candidate index): smaller error wins, equal
// errors (including -0 == +0) resolve to the smaller index.

__device__ __forceinline__ void wave_argmin(float& best, int& idx) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ob < best || (ob == best && oi < idx)) {
            best = ob;
            idx = oi;
        }
    }
}

constexpr int kNoWinner = 0x7fffffff;

// Four-cluster search for every block of the tile whose owner passed n >= 2.
// Returns, in the owning lane, the winning candidate index or -1.
__device__ __forceinline__ int search4(const float4* lds, int my_n, f3 metric,
                                       const DecoderWeights& dw) {
    const int lane = threadIdx.x;
    int my_win = -1;
    for (int b = 0; b < kTile; b++) {
        const int nb = __builtin_amdgcn_readlane(my_n, b);
        if (nb < 2) continue;
        const float4* s = lds + b * kSatRows;
        const float4 tot = s[nb];
        const int total = k_four_total[nb - 1];
        float best = FLT_MAX;
        int besti = kNoWinner;
        for (int i = lane; i < total; i += kTile) {
            const uint32_t pk = k_four[i];
            f3 a, e;
            const float err = candidate4(s[pk & 0xff], s[(pk >> 8) & 0xff], s[pk >> 16], tot, metric,
                                         dw, a, e);
            if (err < best) {
                best = err;
                besti = i;
            }
        }
        wave_argmin(best, besti);
        if (lane == b && besti != kNoWinner) my_win = besti;
    }
    return my_win;
}

// Three-cluster search for every block whose owner passed n >= 1.
__device__ __forceinline__ int search3(const float4* lds, int my_n, f3 metric) {
    const int lane = threadIdx.x;
    int my_win = -1;
    for (int b = 0; b < kTile; b++) {
        const int nb = __builtin_amdgcn_readlane(my_n, b);
        if (nb < 1) continue;
        const float4* s = lds + b * kSatRows;
        const float4 tot = s[nb];
        const int total = k_three_total[nb - 1];
        float best = FLT_MAX;
        int besti = kNoWinner;
        for (int i = lane; i < total; i += kTile) {
            const uint32_t pk = k_three[i];
            f3 a, e;
            const float err = candidate3(s[pk & 0xff], s[(pk >> 8) & 0xff], tot, metric, a, e);
            if (err < best) {
                best = err;
                besti = i;
            }
        }
        wave_argmin(best, besti);
        if (lane == b && besti != kNoWinner) my_win = besti;
    }
    return my_win;
}

// ---------------------------------------------------------------------------
// One tile: the reference's compress_bc1 driver (icbc.h:3734-3808) per lane.

__device__ void encode_tile(const JobDev& job, uint64_t first, float4* lds) {
    const int lane = threadIdx.x;
    const uint64_t bi = first + lane;
    const bool active = bi < job.block_count;
    const Plan p = make_plan(job.level, job.tcm != 0, job.tcb != 0);
    const int dec = job.decoder;
    const f3 cw = mk(job.cw0, job.cw1, job.cw2);
    const DecoderWeights dw = k_dweights[dec];

    f3 tex[16];
    float wt[16];
    if (active) {
        const float4* src = job.colors + bi * 16;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const float4 v = src[i];
            tex[i] = mk(v.x, v.y, v.z);
        }
        if (job.weights) {
            const float4* ws = job.weights + bi * 4;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float4 v = ws[k];
                wt[4 * k + 0] = v.x;
                wt[4 * k + 1] = v.y;
                wt[4 * k + 2] = v.z;
                wt[4 * k + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) wt[i] = 1.0f;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            tex[i] = mk(0.0f, 0.0f, 0.0f);
            wt[i] = 0.0f;
        }
    }

    Code out{0u, 0u};
    bool written = false;
    float err = FLT_MAX;
    int n = 16, n3 = -1, steps = 0;
    f3 rc[16];
    float rw[16];
    bool any_black = false;
    bool done = !active;

    if (p.cluster && active) n = reduce_colors(tex, wt, p.threshold, rc, rw, any_black);

    if (!done && n == 0) {  // no weighted texel: trivial block (icbc.h:3752-3758)
        err = 0.0f;
        written = true;
        done = true;
    }
    if (!done && n == 1) {  // one colour: optimal single-colour tables (icbc.h:3761-3764)
        out = single_color(to_u8(rc[0].x), to_u8(rc[0].y), to_u8(rc[0].z), dec);
        err = code_error(tex, wt, cw, dec, out);
        written = true;
        done = true;
    }

    if (p.box && !done) {  // icbc.h:3769-3786
        f3 c0, c1;
        if (p.cluster) box_fit(rc, n, c0, c1);
        else box_fit(tex, 16, c0, c1);
        err = emit4(tex, wt, cw, dec, c0, c1, out);
        written = true;
        if (p.lsq) {
            f3 a, b;
            if (lsq_endpoints(out.idx, tex, a, b)) {
                Code o2;
                const float e2 = emit4(tex, wt, cw, dec, a, b, o2);
                if (e2 < err) {
                    err = e2;
                    out = o2;
                }
            }
        }
    }

    if (p.cluster) {  // icbc.h:3788-3801, 3537-3571
        const bool fit = !done;
        float4* my_sat = lds + lane * kSatRows;
        const f3 metric = cw * cw;
        if (fit) build_sat(rc, rw, n, my_sat);
        __syncthreads();
        const int win4 = search4(lds, fit ? n : 0, metric, dw);
        __syncthreads();

        Code o4{0u, 0u};
        float e4 = 0.0f;
        bool use3 = false;
        int m = 0;
        if (fit) {
            f3 a = mk(0.0f, 0.0f, 0.0f), b = mk(0.0f, 0.0f, 0.0f);
            if (win4 >= 0) {
                const uint32_t pk = k_four[win4];
                candidate4(my_sat[pk & 0xff], my_sat[(pk >> 8) & 0xff], my_sat[pk >> 16], my_sat[n],
                           metric, dw, a, b);
            }
            e4 = emit4(tex, wt, cw, dec, a, b, o4);
            const bool use_black = any_black && job.tcb != 0;
            use3 = p.fit3_always || (use_black && p.fit3_black_only);
            if (use3) {
                m = n;
                if (use_black) {
                    f3 c2[16];
                    float w2[16];
                    m = drop_blacks(rc, rw, n, c2, w2);
                    if (m == 0) use3 = false;
                    else build_sat(c2, w2, m, my_sat);
                }
            }
            if (use3) n3 = m;
        }
        __syncthreads();
        const int win3 = search3(lds, use3 ? m : 0, metric);
        __syncthreads();

        if (fit) {
            Code oc = o4;
            float ec = e4;
            if (use3) {
                f3 a = mk(0.0f, 0.0f, 0.0f), b = mk(0.0f, 0.0f, 0.0f);
                if (win3 >= 0) {
                    const uint32_t pk = k_three[win3];
                    candidate3(my_sat[pk & 0xff], my_sat[(pk >> 8) & 0xff], my_sat[m], metric, a, b);
                }
                Code o3;
                const float e3 = emit3(tex, wt, cw, dec, job.tcb != 0, a, b, o3);
                if (e3 < ec) {
                    ec = e3;
                    oc = o3;
                }
            }
            if (ec < err) {
                err = ec;
                out = oc;
                written = true;
            }
        }
    }

    if (p.refine && !done) {  // icbc.h:3803-3805
        if (!written) {
            // the reference refines whatever the output block holds
            const uint2 v = job.codes[bi];
            out = Code{v.x, v.y};
            written = true;
        }
        err = refine(tex, wt, cw, dec, job.tcm != 0, err, out, steps);
    }

    if (active) {
        if (written) job.codes[bi] = make_uint2(out.cols, out.idx);
        if (job.errors) job.errors[bi] = err;
        if (job.stats) job.stats[bi] = make_int4(n, n3, steps, 0);
    }
}

// ---------------------------------------------------------------------------
// kernels

__global__ __launch_bounds__(kTile) void bc1_tiles_kernel(const JobDev* __restrict__ jobs,
                                                          const uint2* __restrict__ tiles,
                                                          uint32_t ntiles) {
    __shared__ float4 lds[kTile * kSatRows];
    const uint32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const uint2 tile = tiles[t];
    encode_tile(jobs[tile.x], tile.y, lds);
}

__global__ __launch_bounds__(kTile) void bc1_job_kernel(const JobDev* __restrict__ job,
                                                        uint32_t ntiles) {
    __shared__ float4 lds[kTile * kSatRows];
    const uint32_t t = blockIdx.x;
    if (t >= ntiles) return;
    encode_tile(*job, (uint64_t)t * kTile, lds);
}

__global__ __launch_bounds__(kTile) void bc1_persistent_kernel(const JobDev* __restrict__ jobs,
                                                               const uint2* __restrict__ tiles,
                                                               uint32_t ntiles,
                                                               uint32_t* __restrict__ counter) {
    __shared__ float4 lds[kTile * kSatRows];
    for (;;) {
        uint32_t t = 0;
        if (threadIdx.x == 0) t = atomicAdd(counter, 1u);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
        if (t >= ntiles) break;  // every wave reaches this once the queue drains
        const uint2 tile = tiles[t];
        encode_tile(jobs[tile.x], tile.y, lds);
        __syncthreads();
    }
}

__global__ void reset_counter_kernel(uint32_t* counter) { *counter = 0u; }

// decode_bc1 (icbc.h:2915-2927): one thread per block, RGBA8 out.
__global__ void bc1_decode_kernel(const uint2* __restrict__ codes, uint64_t n,
                                  uint32_t* __restrict__ rgba, int decoder) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint2 c = codes[i];
    Pal8 p;
    palette8(c.x & 0xffff, c.x >> 16, decoder, p);
    const bool four = (c.x & 0xffff) > (c.x >> 16);
    uint32_t px[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t a = (k < 3 || four) ? 255u : 0u;
        px[k] = (uint32_t)p.r[k] | ((uint32_t)p.g[k] << 8) | ((uint32_t)p.b[k] << 16) | (a << 24);
    }
    uint4* dst = reinterpret_cast<uint4*>(rgba + i * 16);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t sel = (c.y >> (2 * (4 * q + k))) & 3;
            v[k] = sel == 0 ? px[0] : sel == 1 ? px[1] : sel == 2 ? px[2] : px[3];
        }
        dst[q] = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

// evaluate_bc1_error (icbc.h:3030-3046): integer squared error per texel,
// summed in float in texel order.
__global__ void bc1_evaluate_kernel(const uint32_t* __restrict__ rgba, const uint2* __restrict__ codes,
                                    uint64_t n, float* __restrict__ errors, int decoder) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint2 c = codes[i];
    Pal8 p;
    palette8(c.x & 0xffff, c.x >> 16, decoder, p);
    const uint4* src = reinterpret_cast<const uint4*>(rgba + i * 16);
    float e = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v4 = src[q];
        const uint32_t v[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t sel = (c.y >> (2 * (4 * q + k))) & 3;
            const int pr = sel == 0 ? p.r[0] : sel == 1 ? p.r[1] : sel == 2 ? p.r[2] : p.r[3];
            const int pg = sel == 0 ? p.g[0] : sel == 1 ? p.g[1] : sel == 2 ? p.g[2] : p.g[3];
            const int pb = sel == 0 ? p.b[0] : sel == 1 ? p.b[1] : sel == 2 ? p.b[2] : p.b[3];
            const int dr = pr - (int)(v[k] & 0xff);
            const int dg = pg - (int)((v[k] >> 8) & 0xff);
            const int db = pb - (int)((v[k] >> 16) & 0xff);
            e += (float)(dr * dr + dg * dg + db * db);
        }
    }
    errors[i] = e;
}

// ---------------------------------------------------------------------------
// launch glue

hipError_t launch_tiles(const JobDev* jobs, const uint2* tiles, uint32_t ntiles, hipStream_t stream) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(bc1_tiles_kernel, dim3(ntiles), dim3(kTile), 0, stream, jobs, tiles, ntiles);
    return hipGetLastError();
}

hipError_t launch_job(const JobDev* job, uint32_t ntiles, hipStream_t stream) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(bc1_job_kernel, dim3(ntiles), dim3(kTile), 0, stream, job, ntiles);
    return hipGetLastError();
}

hipError_t launch_persistent(const JobDev* jobs, const uint2* tiles, uint32_t ntiles,
                             uint32_t* counter, uint32_t grid, hipStream_t stream) {
    if (ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bc1_persistent_kernel, dim3(grid), dim3(kTile), 0, stream, jobs, tiles,
                       ntiles, counter);
    return hipGetLastError();
}

hipError_t launch_decode(const uint2* codes, uint64_t n, uint32_t* rgba, int decoder,
                         hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t threads = 256;
    hipLaunchKernelGGL(bc1_decode_kernel, dim3((uint32_t)((n + threads - 1) / threads)),
                       dim3(threads), 0, stream, codes, n, rgba, decoder);
    return hipGetLastError();
}

hipError_t launch_evaluate(const uint32_t* rgba, const uint2* codes, uint64_t n, float* errors,
                           int decoder, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t threads = 256;
    hipLaunchKernelGGL(bc1_evaluate_kernel, dim3((uint32_t)((n + threads - 1) / threads)),
                       dim3(threads), 0, stream, rgba, codes, n, errors, decoder);
    return hipGetLastError();
}

}  // namespace icbc_gpu
