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

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "bc1_device.h"
#include "bc1_kernels.h"

namespace icbc_gpu {

__constant__ float k_mid5[32];
__constant__ float k_mid6[64];
__constant__ uint32_t k_four[976];
__constant__ uint32_t k_three[160];
__constant__ int k_four_total[16];
__constant__ int k_three_total[16];
__constant__ uint16_t k_match5[4][256];
__constant__ uint16_t k_match6[4][256];
__constant__ DecoderWeights k_dweights[4];
__constant__ uint16_t k_singular[4][64];
__constant__ uint16_t k_singular_mask[4][64];
__constant__ int k_singular_count[4];

hipError_t upload_tables(const HostTables& t, hipStream_t stream) {
    hipError_t e;
#define UP(sym, src)                                                                          \
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(sym), &(src), sizeof(src), 0, hipMemcpyHostToDevice, \
                               stream);                                                       \
    if (e != hipSuccess) return e;
    UP(k_mid5, t.mid5);
    UP(k_mid6, t.mid6);
    UP(k_four, t.four);
    UP(k_three, t.three);
    UP(k_four_total, t.four_total);
    UP(k_three_total, t.three_total);
    UP(k_match5, t.match5);
    UP(k_match6, t.match6);
    UP(k_dweights, t.dweights);
    UP(k_singular, t.singular);
    UP(k_singular_mask, t.singular_mask);
    UP(k_singular_count, t.singular_count);
#undef UP
    return hipStreamSynchronize(stream);
}

constexpr int kNoWinner = 0x7fffffff;

// Fields of a four-cluster table entry (four_entry).
__device__ __forceinline__ uint32_t four_c0(uint32_t e) { return (e >> 21) & 31u; }
__device__ __forceinline__ uint32_t four_c1(uint32_t e) { return (e >> 26) & 31u; }
__device__ __forceinline__ uint32_t four_c2(uint32_t e) { return (e >> 16) & 31u; }

// Per-workgroup LDS: candidate tables and two SAT buffers (the block being
// searched and the next one).  8,928 bytes: at most 10 KiB keeps 16
// workgroups (4 waves per SIMD) resident per CU.
struct Shared {
    uint32_t four[1024];          // 976 entries, padded so 16 rounds of 64 lanes stay inside
    uint32_t three[160];
    uint16_t singular[64];        // the launch decoder's nearly singular candidates (k_singular)
    uint16_t singular_mask[64];
    float4 sat[2][kSatRows];
    float4 centred[kSatRows];     // four-cluster search: the block's rows centred for the bound pass
                                  // (one buffer: the wavefront's LDS accesses complete in order)
    union {
        struct {
            float4 c[kPairs];     // bound pass: pair sums of d (xyz) and Wa (w)
            float a[kPairs];      //             pair sums of A
        } pairs;
        uint16_t survivors[976];  // after it: the candidates left
    };
};

// The pruning bound's operands of a four-cluster table entry `e`: its pair's
// sums and the centred row of its third boundary, each addressed by one
// bit-field of the entry (a byte offset; the fields' low four bits are zero,
// so the masks are contiguous and each extract is one instruction).
__device__ __forceinline__ float4 pair_c(const Shared& sh, uint32_t e) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sh.pairs.c) + (e & 0xfffu));
}
__device__ __forceinline__ float pair_a(const Shared& sh, uint32_t e) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(sh.pairs.a) + ((e >> 2) & 0x3ffu));
}
__device__ __forceinline__ float4 row_at(const float4* rows, uint32_t e) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(rows) + ((e >> 12) & 0x1ffu));
}

// LDS of the staged level 0-1 instance: the tile's texel colours, RGB, one row of 49
// floats per block (48 + 1 pad: an odd stride keeps the 64 lanes' reads of
// texel i in distinct banks); 12,544 bytes.
constexpr int kBoxRow = 49;
struct SharedBox {
    float tex[kTile * kBoxRow];
};

// Texel source over a block's LDS row and its weights in HBM.
template <bool kWeighted>
struct TexelsLds {
    const float* c;  // 48 floats: texel i at 3 i
    const float* w;  // 16 weights (kWeighted)
    __device__ __forceinline__ f3 color(int i) const { return mk(c[3 * i], c[3 * i + 1], c[3 * i + 2]); }
    __device__ __forceinline__ float weight(int i) const { return kWeighted ? ld_global(w + i) : 1.0f; }
};

// Texel loops over LDS rows unroll by 8: the reads of 8 texels are issued
// together (a full unroll spills).
template <bool kWeighted>
struct TexelUnroll<TexelsLds<kWeighted>> {
    static constexpr int value = 8;
};


// Wave-wide minima.  Floats are reduced as order-preserving unsigned keys
// (the sign bit flipped for positive values, every bit for negative ones;
// -0 is first mapped to +0, so that the two tie): an integer min takes its
// DPP operand directly, where a float min also quiets its operands.  The
// inputs must not be NaN.
__device__ __forceinline__ uint32_t order_key(float e) {
    const uint32_t u = __float_as_uint(e + 0.0f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// One step: the DPP source's value where the lane has one, else the min's
// identity (all ones), so that the move folds into the min instruction.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t umin_step(uint32_t v) {
    return min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = umin_step<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
    v = umin_step<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
    v = umin_step<0x141, 0xf>(v);  // row_half_mirror
    v = umin_step<0x140, 0xf>(v);  // row_mirror
    v = umin_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = umin_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Minimum error over the wave (all lanes get it; of -0 and +0 either).
__device__ __forceinline__ float wave_min(float e) {
    const uint32_t k = order_key(e);
    const uint64_t at = __ballot(k == wave_min_u32(k));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e), __builtin_ctzll(at)));
}

// Wave-wide argmin over (error, candidate index): smaller error wins, equal
// errors (including -0 == +0) resolve to the smaller index.  The error keys
// are reduced alone; when one lane holds the minimum (the usual case) that
// lane's index is read out, and only ties reduce the tied lanes' indices.
// Errors here are never NaN (a lane's best only takes values below FLT_MAX; a
// lane without a winner holds FLT_MAX and kNoWinner).
__device__ __forceinline__ int wave_argmin(float e, int i) {
    const uint32_t k = order_key(e);
    const uint64_t at = __ballot(k == wave_min_u32(k));
    if (__popcll(at) == 1) return __builtin_amdgcn_readlane(i, __builtin_ctzll(at));
    return (int)wave_min_u32(k == (uint32_t)__builtin_amdgcn_readlane((int)k, __builtin_ctzll(at)) ? (uint32_t)i
                                                                                                 : (uint32_t)kNoWinner);
}

// Inclusive prefix sum of v over the wave's lanes (lane l gets the sum over
// lanes 0..l): shifted adds within each row of 16, then row broadcasts.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int scan_step(int v) {
    return v + __builtin_amdgcn_update_dpp(0, v, kCtrl, kRowMask, 0xf, true);
}

__device__ __forceinline__ int wave_inclusive_sum(int v) {
    v = scan_step<0x111, 0xf>(v);  // row_shr:1
    v = scan_step<0x112, 0xf>(v);  // row_shr:2
    v = scan_step<0x114, 0xf>(v);  // row_shr:4
    v = scan_step<0x118, 0xf>(v);  // row_shr:8
    v = scan_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = scan_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Texel source of block `bi` of a job: float colours, or RGBA8 colours when
// kRgba8 (the job's rgba8 flag), and weights.
template <bool kRgba8>
__device__ __forceinline__ TexelsOf<kRgba8> block_texels(const JobDev& job, uint64_t bi) {
    TexelsOf<kRgba8> t;
    if constexpr (kRgba8) t.c = reinterpret_cast<const uint32_t*>(job.colors) + bi * 16;
    else t.c = job.colors + bi * 16;
    t.w = job.weights ? reinterpret_cast<const float*>(job.weights + bi * 4) : nullptr;
    return t;
}

// threadIdx.x, opaque to the optimiser.  The tile loop's per-lane values
// (workspace and LDS row addresses, lane masks) derive from it; computed
// where they are used they cost a few instructions per tile, while hoisted out
// of the loop they would be held in registers across every tile's search.
__device__ __forceinline__ int tile_lane() {
    int l = threadIdx.x;
    asm volatile("" : "+v"(l));
    return l;
}

// Lane l's value of x (l wave-uniform), as a scalar.
__device__ __forceinline__ float lane_value(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// A wave-uniform pointer held in scalar registers.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(lo | (hi << 32));
}

// Texel source of one lane's block within a tile (a tile's blocks are
// consecutive blocks of one job): the tile's colours and weights as wave-
// uniform base pointers in scalar registers, plus the lane's block number, so
// every texel load is a scalar base + a 32-bit lane offset + an immediate,
// and no per-lane 64-bit address is held across the block's stages.
template <bool kRgba8>
struct TileTexels {
    static constexpr uint32_t kColorBytes = kRgba8 ? 64u : 256u;  // per block
    const char* c;  // the tile's first block
    const char* w;  // its weights (64 bytes per block), or null for all 1.0
    uint32_t blk;   // the lane's block within the tile
    __device__ __forceinline__ f3 color(int i) const {
        const char* p = c + (blk * kColorBytes + (uint32_t)i * (kColorBytes / 16));
        if constexpr (kRgba8) {
            const uint32_t v = ld_global(reinterpret_cast<const uint32_t*>(p));
            return mk(u8_to_unit(v & 0xffu), u8_to_unit((v >> 8) & 0xffu), u8_to_unit((v >> 16) & 0xffu));
        } else {
            return ld_global_rgb(reinterpret_cast<const float4*>(p));
        }
    }
    __device__ __forceinline__ float weight(int i) const {
        return w ? ld_global(reinterpret_cast<const float*>(w + (blk * 64u + (uint32_t)i * 4u))) : 1.0f;
    }
};

// Texel source of block first + blk of a job (blk < 64, first + blk inside
// the job).
template <bool kRgba8>
__device__ __forceinline__ TileTexels<kRgba8> tile_texels(const JobDev& job, uint64_t first, uint32_t blk) {
    TileTexels<kRgba8> t;
    t.c = uniform_ptr(reinterpret_cast<const char*>(job.colors) + first * TileTexels<kRgba8>::kColorBytes);
    t.w = job.weights ? uniform_ptr(reinterpret_cast<const char*>(job.weights) + first * 64u) : nullptr;
    t.blk = blk;
    return t;
}

// First block b >= from (wave-uniform) whose owner lane holds my_n >= min_n.
__device__ __forceinline__ int next_block(int my_n, int tile_blocks, int from, int min_n) {
    int b = from;
    while (b < tile_blocks && __builtin_amdgcn_readlane(my_n, b) < min_n) b++;
    return b;
}

// Wave-cooperative four-cluster search (icbc.h:2348-2743) over every block of
// the tile whose owner passed my_n >= 2: for each such block the 64 lanes
// split its C(n+3,3)-1 candidates and the DPP argmin picks the winner,
// returned in the owning lane (-1 if none).  SAT rows stream from the
// workgroup's workspace through a double-buffered LDS slot: the next block's
// rows are loaded into registers while the current block is searched.
// Block b's SAT is at ws + b * kLaneRows.
// Blocks with at least `prune_min` colours run the bound pass first (smaller
// ones have too few candidates to gain from it; the job's setting, tuned per
// level).
template <bool kUnit>
__device__ __forceinline__ int search4(Shared& sh, const float4* ws_in, int tile_blocks, int my_n, f3 cw, f3 metric,
                                       const DecoderWeights& dw, int prune_min, bool unweighted, int singular_count) {
    constexpr int kMinN = 2;
    const int lane = tile_lane();
    const float4* ws = uniform_ptr(ws_in);  // the workgroup's workspace: scalar base
    int my_win = -1;
    int b = next_block(my_n, tile_blocks, 0, kMinN);
    int nb = b < tile_blocks ? __builtin_amdgcn_readlane(my_n, b) : 0;
    float4 row = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto base = [&](int blk) { return ws + blk * kLaneRows; };
    if (b < tile_blocks && (lane <= nb || lane == kBoundRow)) row = ld_global(base(b) + lane);
    int buf = 0;
    while (b < tile_blocks) {
        if (lane <= nb || lane == kBoundRow) sh.sat[buf][lane] = row;
        if (nb >= prune_min) {
            const f3 t = mk(lane_value(row.x, nb), lane_value(row.y, nb), lane_value(row.z, nb));
            const float invW = __builtin_amdgcn_rcpf(lane_value(row.w, nb));
            if (lane <= nb) sh.centred[lane] = centred_row<kUnit>(row, invW, t, cw);
        }
        __syncthreads();
        // prefetch the next block's rows while this one is searched
        const int b2 = next_block(my_n, tile_blocks, b + 1, kMinN);
        const int nb2 = b2 < tile_blocks ? __builtin_amdgcn_readlane(my_n, b2) : 0;
        if (b2 < tile_blocks && (lane <= nb2 || lane == kBoundRow)) row = ld_global(base(b2) + lane);

        const float4* sat = sh.sat[buf];
        const float4 tot = sat[nb];
        float best = FLT_MAX;
        int besti = kNoWinner;
        const float4 brow = sat[kBoundRow];
        if (nb >= prune_min && brow.y != 0.0f) {
            const int total = (nb + 3) * (nb + 2) * (nb + 1) / 6 - 1;  // >= 219: every lane has 3+
            const int rounds = (total + kTile - 1) / kTile;
            const float margin = brow.x;
            const float W = tot.w;
            const float K0 = ((metric.x * tot.x) * tot.x + (metric.y * tot.y) * tot.y + (metric.z * tot.z) * tot.z) *
                             __builtin_amdgcn_rcpf(W);
            const BoundCoefs coefs = bound_coefs(dw);
            const float4* cen = sh.centred;
            // pair sums of every boundary pair (those with c1 > nb are not read)
            for (int p = lane; p < kPairs; p += kTile) {
                const uint32_t cc = sh.three[p] >> 16;
                pair_sums(cen[cc & 0xff], cen[cc >> 8], coefs, sh.pairs.c[p], sh.pairs.a[p]);
            }
            __syncthreads();
            // bound pass: lane l holds q of candidates l, l + 64, ... (NaN: keep);
            // its key is the one with the lowest bound (largest q)
            float q[16];
            float qkey = -INFINITY;
            int kkey = 0;  // the round of the lane's key candidate
            if (nb == 16) {
                // 968 candidates: 15 full rounds and 8 lanes of a 16th, as
                // straight-line code: no per-round branch, so the rounds form
                // one basic block without register copies between them
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    const int i = lane + kTile * k;
                    const uint32_t pk = sh.four[i];
                    const float v = candidate4_bound_q(pair_c(sh, pk), pair_a(sh, pk), row_at(cen, pk), W, coefs);
                    const bool valid = k < 15 || i < total;
                    q[k] = valid ? v : __builtin_nanf("");
                    const bool better = valid & (v > qkey);
                    qkey = better ? v : qkey;
                    kkey = better ? k : kkey;
                }
            } else {
                // software pipeline: round k computes while round k + 1's pair
                // sums and row are read, and round k + 2's table entry
                uint32_t pk = sh.four[lane];
                float4 pc = pair_c(sh, pk);
                float pa = pair_a(sh, pk);
                float4 r2 = row_at(cen, pk);
                uint32_t pk_next = sh.four[lane + kTile];
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    q[k] = __builtin_nanf("");
                    if (k < rounds) {
                        const int i = lane + kTile * k;
                        float4 pc_n = pc, r2_n = r2;
                        float pa_n = pa;
                        uint32_t pk_nn = pk_next;
                        if (k + 1 < 16) {
                            pc_n = pair_c(sh, pk_next);
                            pa_n = pair_a(sh, pk_next);
                            r2_n = row_at(cen, pk_next);
                        }
                        if (k + 2 < 16) pk_nn = sh.four[i + 2 * kTile];
                        const float v = candidate4_bound_q(pc, pa, r2, W, coefs);
                        const bool valid = i < total;
                        q[k] = valid ? v : q[k];
                        const bool better = valid & (v > qkey);
                        qkey = better ? v : qkey;
                        kkey = better ? k : kkey;
                        pk = pk_next;
                        pc = pc_n;
                        pa = pa_n;
                        r2 = r2_n;
                        pk_next = pk_nn;
                    }
                }
            }
            // threshold: the best true error among the lanes' key candidates;
            // a candidate is pruned when -K0 - W q - margin > threshold.  The
            // keys' errors also enter the argmin (the winner rule below), and
            // the keys leave the survivor list.
            float thr;
            {
                const int ikey = lane + kTile * kkey;
                const uint32_t pkey = sh.four[ikey];
                f3 a, e;
                float err = candidate4<kUnit>(sat[four_c0(pkey)], sat[four_c1(pkey)], sat[four_c2(pkey)], tot, metric,
                                       dw, a, e);
                thr = wave_min(err == err ? err : FLT_MAX);
                if (err < best) {  // the lane's first candidate: best is FLT_MAX, no winner yet
                    best = err;
                    besti = ikey;
                }
            }
            const float qthr = (-K0 - margin - thr) / W;
            // Blocks of 16 unit-weight colours: the candidates whose system is
            // nearly singular (q NaN, never pruned by q) are a fixed set, decided
            // instead by the cluster bound, one lane each
            const bool singular_set = unweighted && nb == 16;
            bool singular_keep = false;
            int singular_index = 0;
            if (singular_set && lane < singular_count) {
                singular_index = sh.singular[lane];
                const uint32_t pk = sh.four[singular_index];
                const float lb = cluster_bound<kUnit>(sat[four_c0(pk)], sat[four_c1(pk)], sat[four_c2(pk)], tot, metric);
                singular_keep = !(lb - margin > thr);  // NaN: keep (a key evaluated twice is harmless)
            }
            __syncthreads();  // the pair sums are dead: their space takes the survivor list
            // survivors to LDS, lane-major: each lane marks its kept rounds in a
            // mask, a wave prefix sum of the counts gives its first slot
            // (the lane's candidates are rounds 0 .. ceil((total - lane) / 64) - 1,
            // which also bounds them by `rounds`; total >= 219 > lane)
            uint32_t pruned = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) pruned |= (q[k] < qthr ? 1u : 0u) << k;
            const int mine_rounds = (total - lane + kTile - 1) / kTile;
            uint32_t keep = ((1u << mine_rounds) - 1u) & ~pruned & ~(1u << kkey);
            if (singular_set) keep &= ~(uint32_t)sh.singular_mask[lane];
            const int mine = __popc(keep);
            const int incl = wave_inclusive_sum(mine);
            int count = __builtin_amdgcn_readlane(incl, kTile - 1);
            int slot = incl - mine;
            while (keep) {
                const int k = __builtin_ctz(keep);
                keep &= keep - 1;
                sh.survivors[slot++] = (uint16_t)(lane + kTile * k);
            }
            if (singular_set) {  // the singular candidates the cluster bound keeps follow
                const uint64_t kept = __ballot(singular_keep);
                if (singular_keep) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(kept >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)kept, 0));
                    sh.survivors[count + rank] = (uint16_t)singular_index;
                }
                count += __popcll(kept);
            }
            __syncthreads();
            // the survivors get the exact evaluation; every candidate that could
            // reach the minimum error is among them (the list is not in candidate
            // order: equal errors keep the lower index).  A tie only counts once
            // the lane holds a winner, so, as in the reference (icbc.h:2725), an
            // error must be below FLT_MAX to win.  (Blocks that reach this pass
            // have a finite margin below 1e30, which bounds every candidate's
            // error far below FLT_MAX; the guard keeps the rule explicit.)
            for (int j = lane; j < count; j += kTile) {
                const int i = sh.survivors[j];
                const uint32_t pk = sh.four[i];
                f3 a, e;
                const float err = candidate4<kUnit>(sat[four_c0(pk)], sat[four_c1(pk)], sat[four_c2(pk)], tot,
                                                    metric, dw, a, e);
                if ((err < best) | ((err == best) & (i < besti) & (besti != kNoWinner))) {
                    best = err;
                    besti = i;
                }
            }
        } else {
            const int total = (nb + 3) * (nb + 2) * (nb + 1) / 6 - 1;  // k_four_total[nb - 1]
            // the next candidate's table entry is read while this one is evaluated
            uint32_t pk_next = sh.four[lane];
            for (int i = lane; i < total; i += kTile) {
                const uint32_t pk = pk_next;
                pk_next = sh.four[min(i + kTile, 975)];
                f3 a, e;
                const float err = candidate4<kUnit>(sat[four_c0(pk)], sat[four_c1(pk)], sat[four_c2(pk)], tot,
                                                    metric, dw, a, e);
                if (err < best) {
                    best = err;
                    besti = i;
                }
            }
        }
        const int w = wave_argmin(best, besti);
        if (lane == b && w != kNoWinner) my_win = w;
        b = b2;
        nb = nb2;
        buf ^= 1;
    }
    __syncthreads();
    return my_win;
}

// Wave-cooperative three-cluster search (icbc.h:2075-2345) over every block
// of the tile whose owner passed my_n >= 1, with consecutive blocks' candidate
// lists packed into shared rounds.  The window holds two blocks, A and B,
// whose SAT rows sit in the two LDS slots: a round evaluates the next 64
// positions of the stream "A's remaining candidates, then B's", so a block
// whose last round is partly empty shares it with the next block.  Each lane
// keeps its first minimum per block (a block's candidate indices reach a lane
// in increasing order); once A's last candidate is evaluated the DPP argmin
// picks its lowest-index minimum for the owning lane (-1 if none), B moves
// up, and the freed slot takes the rows of the next block, prefetched into
// registers one block ahead.
template <bool kUnit>
__device__ __forceinline__ int search3(Shared& sh, const float4* ws_in, int alt_offset, int my_alt, int tile_blocks,
                                       int my_n, f3 metric) {
    const int lane = tile_lane();
    const float4* ws = uniform_ptr(ws_in);
    auto base = [&](int blk) {
        return ws + blk * kLaneRows + (__builtin_amdgcn_readlane(my_alt, blk) ? alt_offset : 0);
    };
    auto total_of = [](int n) { return (n + 2) * (n + 1) / 2 - 1; };  // k_three_total[n - 1]
    auto count_of = [&](int blk) { return blk < tile_blocks ? __builtin_amdgcn_readlane(my_n, blk) : 0; };
    int my_win = -1;
    int a = next_block(my_n, tile_blocks, 0, 1);
    int b = a < tile_blocks ? next_block(my_n, tile_blocks, a + 1, 1) : tile_blocks;
    int c = b < tile_blocks ? next_block(my_n, tile_blocks, b + 1, 1) : tile_blocks;
    int na = count_of(a), nb = count_of(b), nc = count_of(c);
    int sa = 0;  // A's slot; B's is sa ^ 1
    if (a < tile_blocks && lane <= na) sh.sat[0][lane] = ld_global(base(a) + lane);
    if (b < tile_blocks && lane <= nb) sh.sat[1][lane] = ld_global(base(b) + lane);
    float4 row = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // rows of block c
    if (c < tile_blocks && lane <= nc) row = ld_global(base(c) + lane);
    __syncthreads();
    int done = 0;  // A's candidates evaluated
    float best_a = FLT_MAX, best_b = FLT_MAX;
    int bi_a = kNoWinner, bi_b = kNoWinner;
    while (a < tile_blocks) {
        const int ta = total_of(na);
        const int tb = b < tile_blocks ? total_of(nb) : 0;
        {
            const float4* sat_a = sh.sat[sa];
            const float4* sat_b = sh.sat[sa ^ 1];
            const int g = done + lane;
            const bool in_a = g < ta;
            const int ci = in_a ? g : g - ta;
            const bool in_b = !in_a && ci < tb;
            const float4* sat = in_a ? sat_a : sat_b;
            const float4 tot = in_a ? sat_a[na] : sat_b[nb];
            const uint32_t pk = sh.three[min(ci, 159)];
            f3 p0, p1;
            const float err = candidate3<kUnit>(sat[pk & 0xff], sat[(pk >> 8) & 0xff], tot, metric, p0, p1);
            if (in_a && err < best_a) {
                best_a = err;
                bi_a = ci;
            }
            if (in_b && err < best_b) {
                best_b = err;
                bi_b = ci;
            }
        }
        if (done + kTile < ta) {
            done += kTile;
            continue;
        }
        // A is complete; B has its first min(done + 64 - ta, tb) candidates
        done = min(done + kTile - ta, tb);
        for (;;) {
            const int w = wave_argmin(best_a, bi_a);
            if (lane == a && w != kNoWinner) my_win = w;
            // B moves up; the freed slot takes the prefetched rows of c
            best_a = best_b;
            bi_a = bi_b;
            best_b = FLT_MAX;
            bi_b = kNoWinner;
            const int freed = sa;
            sa ^= 1;
            a = b;
            na = nb;
            b = c;
            nb = nc;
            if (b < tile_blocks && lane <= nb) sh.sat[freed][lane] = row;
            c = b < tile_blocks ? next_block(my_n, tile_blocks, b + 1, 1) : tile_blocks;
            nc = count_of(c);
            if (c < tile_blocks && lane <= nc) row = ld_global(base(c) + lane);
            // the new A may already be complete (the round also covered all of it)
            if (a < tile_blocks && done >= total_of(na)) {
                done = 0;
                continue;
            }
            break;
        }
        __syncthreads();
    }
    __syncthreads();
    return my_win;
}

// ---------------------------------------------------------------------------
// One tile: the reference's compress_bc1 driver (icbc.h:3734-3808) per lane.
// `ws` is this workgroup's workspace: kLaneRows rows per block of the tile.

// Levels 0-1: box fit, output block, and at level 1 least-squares endpoints
// and a second output block (icbc.h:3769-3786 without reduce_colors, as at
// these levels).
template <bool kUnit, typename Src>
__device__ __forceinline__ float encode_box(const Src& t, bool lsq, f3 cw, int dec, Code& out) {
    f3 c0, c1;
    box_fit([&](int i) { return t.color(i); }, 16, c0, c1);
    float err = emit4<kUnit>(t, cw, dec, c0, c1, out);
    if (lsq) {
        f3 a, b;
        if (lsq_endpoints(out.idx, t, a, b)) {
            Code o2;
            const float e2 = emit4<kUnit>(t, cw, dec, a, b, o2);
            if (e2 < err) {
                err = e2;
                out = o2;
            }
        }
    }
    return err;
}

// Copies the colours of the tile's blocks [first, first + nblk) to LDS: the
// tile's texels are contiguous in HBM (16 float4 per block), read with
// coalesced 16-byte loads (lane l takes float4 l, l + 64, ...), all issued
// before the first is stored.
__device__ __forceinline__ void stage_tile(const JobDev& job, uint64_t first, int nblk, SharedBox& sh) {
    const int lane = threadIdx.x;
    const int nf = nblk * 16;
    if (job.rgba8) {
        // 64 B per block: lane l takes 4 texels (16 B) at l, l + 64, ...
        const uint4* src8 = reinterpret_cast<const uint4*>(job.colors) + first * 4;
        const int nq = nblk * 4;
        uint4 u[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int g = j * kTile + lane;
            u[j] = make_uint4(0u, 0u, 0u, 0u);
            if (g < nq) u[j] = ld_global(src8 + g);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = (j * kTile + lane) * 4;  // first texel of the word group
            float* d = sh.tex + (f >> 4) * kBoxRow + (f & 15) * 3;
            const uint32_t w4[4] = {u[j].x, u[j].y, u[j].z, u[j].w};
#pragma unroll
            for (int m = 0; m < 4; m++) {
                d[3 * m] = u8_to_unit(w4[m] & 0xffu);
                d[3 * m + 1] = u8_to_unit((w4[m] >> 8) & 0xffu);
                d[3 * m + 2] = u8_to_unit((w4[m] >> 16) & 0xffu);
            }
        }
        __syncthreads();
        return;
    }
    const float4* src = job.colors + first * 16;
    f3 v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int f = j * kTile + lane;
        v[j] = mk(0.0f, 0.0f, 0.0f);
        if (f < nf) v[j] = ld_global_rgb(src + f);
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int f = j * kTile + lane;
        float* d = sh.tex + (f >> 4) * kBoxRow + (f & 15) * 3;
        d[0] = v[j].x;
        d[1] = v[j].y;
        d[2] = v[j].z;
    }
    __syncthreads();
}

// A tile of a level 0-1 (or invalid-level) job.  kStaged: the tile's texels
// are staged through LDS (texel weights, or channel weights other than 1,
// which need registers the texels would take); otherwise each lane reads its
// block's 16 texels straight into registers (12-byte RGB loads, or one
// 64-byte block of RGBA8 words), and the instance holds no LDS, so more of
// its workgroups fit on a compute unit.
template <bool kUnit, bool kStaged>
__device__ __forceinline__ void encode_box_tile(const JobDev& job, uint64_t first, SharedBox& sh) {
    const int lane = threadIdx.x;
    const int nblk = (int)min<uint64_t>((uint64_t)job.tile_blocks, job.block_count - first);
    const bool active = lane < nblk;
    const uint64_t bi = first + lane;
    const Plan p = make_plan(job.level, job.tcm != 0, job.tcb != 0);
    float err = FLT_MAX;  // invalid level: FLT_MAX, output untouched (icbc.h:3766)
    if constexpr (!kStaged) {
        if (p.box && active) {
            const f3 cw = mk(job.cw0, job.cw1, job.cw2);
            Code out{0u, 0u};
            if (job.rgba8) {
                const TexelRegs<false> t(block_texels<true>(job, bi));
                err = encode_box<kUnit>(t, p.lsq, cw, job.decoder, out);
            } else {
                const TexelRegs<false> t(block_texels<false>(job, bi));
                err = encode_box<kUnit>(t, p.lsq, cw, job.decoder, out);
            }
            st_global(job.codes + bi, make_uint2(out.cols, out.idx));
        }
    } else if (p.box) {
        stage_tile(job, first, nblk, sh);
        if (active) {
            const f3 cw = mk(job.cw0, job.cw1, job.cw2);
            const float* row = sh.tex + lane * kBoxRow;
            Code out{0u, 0u};
            if (job.weights) {
                const TexelsLds<true> t{row, reinterpret_cast<const float*>(job.weights + bi * 4)};
                err = encode_box<kUnit>(t, p.lsq, cw, job.decoder, out);
            } else {
                const TexelsLds<false> t{row, nullptr};
                err = encode_box<kUnit>(t, p.lsq, cw, job.decoder, out);
            }
            st_global(job.codes + bi, make_uint2(out.cols, out.idx));
        }
        __syncthreads();  // the next tile restages the rows
    }
    if (active) {
        if (job.errors) st_global(job.errors + bi, err);
        if (job.stats) st_global(job.stats + bi, make_int4(16, -1, 0, 0));
    }
}

// Kernel instances: kModeBox serves launches whose jobs are all at levels
// 0-1 (or invalid levels), kModeCluster launches without level 9, kModeAll
// everything (the refinement walk's registers weigh only on that one).
// kRgba8: the job's colours are RGBA8 words (one instance per texel format,
// chosen per tile, so neither pays for the other's loads).  Without
// kBoxStage every job is at level 4-9 (no box-fit stage).
template <int kMode, bool kRgba8, bool kUnit, bool kBoxStage>
__device__ __forceinline__ void encode_tile(const JobDev& job, uint32_t job_index, uint64_t first, Shared& sh,
                                            float4* ws, const RefineList& rl) {
    const int lane = tile_lane();
    const int tile_blocks = job.tile_blocks;
    const bool active = lane < tile_blocks && first + lane < job.block_count;
    const Plan p = make_plan(job.level, job.tcm != 0, job.tcb != 0);
    const int dec = job.decoder;
    const f3 cw = mk(job.cw0, job.cw1, job.cw2);
    const DecoderWeights dw = k_dweights[dec];
    const TileTexels<kRgba8> tx = tile_texels<kRgba8>(job, first, active ? lane : 0);

    Code out{0u, 0u};
    bool written = false;
    float err = FLT_MAX;
    int n = 16, n3 = -1, steps = 0;
    f3 rc[16];
    float rw[16];
    bool any_black = false;
    bool done = !active;

    if (p.cluster && active) n = reduce_colors(tx, p.threshold, rc, rw, any_black);

    if (!done && n == 0) {  // no weighted texel: trivial block (icbc.h:3752-3758)
        err = 0.0f;
        written = true;
        done = true;
    }
    if (!done && n == 1) {  // one colour: optimal single-colour tables (icbc.h:3761-3764)
        out = single_color(to_u8(rc[0].x), to_u8(rc[0].y), to_u8(rc[0].z), dec);
        err = code_error<kUnit>(tx, cw, dec, out);
        written = true;
        done = true;
    }

    if (kBoxStage && p.box && !done) {  // icbc.h:3769-3786
        f3 c0, c1;
        if (p.cluster) box_fit([&](int i) { return rc[i]; }, n, c0, c1);
        else box_fit([&](int i) { return tx.color(i); }, 16, c0, c1);
        err = emit4<kUnit>(tx, cw, dec, c0, c1, out);
        written = true;
        if (p.lsq) {
            f3 a, b;
            if (lsq_endpoints(out.idx, tx, a, b)) {
                Code o2;
                const float e2 = emit4<kUnit>(tx, cw, dec, a, b, o2);
                if (e2 < err) {
                    err = e2;
                    out = o2;
                }
            }
        }
    }

    if (p.cluster) {  // icbc.h:3788-3801, 3537-3571
        // workspace slices: SATs for the 4-cluster and 3-cluster fits, and the
        // reduced colours parked in their original order for skip_blacks
        float4* sat4 = ws + lane * kLaneRows;
        float4* sat3 = sat4 + kSatRows;
        float4* park = sat4 + 2 * kSatRows;
        const bool fit = !done;
        const f3 metric = cw * cw;
        const bool use_black = fit && any_black && job.tcb != 0;
        bool use3 = fit && (p.fit3_always || (use_black && p.fit3_black_only));
        int m = n;
        const bool skip = use3 && use_black;
        if (skip) {
#pragma unroll
            for (int i = 0; i < 16; i++) park[i] = make_float4(rc[i].x, rc[i].y, rc[i].z, rw[i]);
        }
        if (fit) {
            build_sat(rc, rw, n, sat4, nullptr);
            sat4[kBoundRow] = bound_row(rc, rw, n, metric);
        }
        if (skip) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const float4 v = park[i];
                rc[i] = mk(v.x, v.y, v.z);
                rw[i] = v.w;
            }
            m = drop_blacks(rc, rw, n);
            if (m == 0) use3 = false;  // nothing but black: keep the 4-colour result
            else build_sat(rc, rw, m, sat3, nullptr);
        }
        if (use3) n3 = m;
        __syncthreads();
        // 3-cluster SATs: the skip_blacks table where one was built, else the 4-cluster one
        const int win4 = search4<kUnit>(sh, ws, tile_blocks, fit ? n : 0, cw, metric, dw,
                                        job.prune_min ? max(job.prune_min, kMinPruneColors) : kDefaultPruneMinColors,
                                        job.weights == nullptr, k_singular_count[dec & 3]);
        const int win3 = search3<kUnit>(sh, ws, kSatRows, skip ? 1 : 0, tile_blocks, use3 ? m : 0, metric);

        if (fit) {
            f3 a4 = mk(0.0f, 0.0f, 0.0f), b4 = mk(0.0f, 0.0f, 0.0f);
            if (win4 >= 0) {
                const uint32_t pk = sh.four[win4];
                candidate4<kUnit>(sat4[four_c0(pk)], sat4[four_c1(pk)], sat4[four_c2(pk)], sat4[n], metric, dw, a4,
                           b4);
            }
            f3 a3 = mk(0.0f, 0.0f, 0.0f), b3 = mk(0.0f, 0.0f, 0.0f);
            if (use3 && win3 >= 0) {
                const float4* s3 = skip ? sat3 : sat4;
                const uint32_t pk = sh.three[win3];
                candidate3<kUnit>(s3[pk & 0xff], s3[(pk >> 8) & 0xff], s3[m], metric, a3, b3);
            }
            // output_block4 and output_block3 in one pass over the texels
            Code oc, o3;
            float e3;
            float ec = emit43<kUnit>(tx, cw, dec, job.tcb != 0, use3, a4, b4, a3, b3, oc, o3, e3);
            if (use3 && e3 < ec) {
                ec = e3;
                oc = o3;
            }
            if (ec < err) {
                err = ec;
                out = oc;
                written = true;
            }
        }
        __syncthreads();  // workspace rows are reused by the next tile
    }

    // level 9 (icbc.h:3803-3805): the block goes to the refinement kernel,
    // which starts from the code written here (or, if no fit was written, the
    // one the output buffer held, as the reference refines whatever it holds)
    // the block index, recomputed from the lane: no 64-bit index is held across the tile
    const uint32_t out_lane = (uint32_t)tile_lane();
    const bool hand_off = p.refine && !done;
    const uint64_t hand_mask = __ballot(hand_off);
    if (hand_mask) {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(rl.counters, (uint32_t)__popcll(hand_mask));
        at = __builtin_amdgcn_readlane(at, 0);
        if (hand_off) {
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(hand_mask >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)hand_mask, 0));
            rl.items[at + rank] = RefineItem{job_index, (uint32_t)(first + out_lane), err, 0u};
        }
    }

    if (active) {
        if (written) st_global(uniform_ptr(job.codes + first) + out_lane, make_uint2(out.cols, out.idx));
        if (job.errors) st_global(uniform_ptr(job.errors + first) + out_lane, err);
        if (job.stats) st_global(job.stats + first + out_lane, make_int4(n, n3, steps, 0));
    }
}

// ---------------------------------------------------------------------------
// kernels

template <int kMode, typename S>
__device__ __forceinline__ void load_tables(S& sh, int decoder) {
    if constexpr (kMode != kModeBox) {
        for (int i = threadIdx.x; i < 1024; i += kTile) sh.four[i] = k_four[i < 976 ? i : 975];
        for (int i = threadIdx.x; i < 160; i += kTile) sh.three[i] = k_three[i];
        sh.singular[threadIdx.x] = k_singular[decoder][threadIdx.x];
        sh.singular_mask[threadIdx.x] = k_singular_mask[decoder][threadIdx.x];
        __syncthreads();
    }
}

// Resident workgroups encode tiles until none are left.  With `counter` the
// tiles are taken from a queue in list order (the batched path); without, a
// workgroup takes tiles blockIdx.x, blockIdx.x + gridDim.x, ...  With `tiles`
// null the launch covers one job, tile k = blocks [64 k, 64 k + 64).
// kUnit: every job of the launch has channel weights (1, 1, 1).
// kBoxStage: some job of the launch is at level 0-3 (cluster instances).
// kStaged (level 0-1 instance): see encode_box_tile.
template <int kMode, int kTex, bool kUnit, bool kBoxStage = true, bool kStaged = true>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(4, 8))) void bc1_encode_kernel(
    const JobDev* __restrict__ jobs, const uint2* __restrict__ tiles, uint32_t ntiles,
    uint32_t* __restrict__ counter, float4* __restrict__ workspace, RefineList refine) {
    __shared__ std::conditional_t<kMode == kModeBox, SharedBox, Shared> sh;
    load_tables<kMode>(sh, jobs[0].decoder & 3);  // one decoder per device (init)
    float4* ws = workspace + (size_t)blockIdx.x * (kWorkspaceBytesPerWorkgroup / sizeof(float4));
    // queue of this workgroup's XCD: its share is tiles x, x + 8, ...; the
    // first claim of each workgroup is static, later ones come from the counter
    const uint32_t x = blockIdx.x % kQueues;
    const uint32_t share = ntiles > x ? (ntiles - x + kQueues - 1) / kQueues : 0;
    const uint32_t static_claims = (gridDim.x - x + kQueues - 1) / kQueues;
    uint32_t* queue = counter ? counter + x * (kQueueStride / sizeof(uint32_t)) : nullptr;
    uint32_t j = blockIdx.x / kQueues;
    uint32_t t = blockIdx.x;
    for (;;) {
        if (queue) {
            if (j >= share) break;  // every wave reaches this once its XCD's share is taken
            t = x + j * kQueues;
        }
        if (t >= ntiles) break;
        uint32_t job = 0;
        uint64_t first = (uint64_t)t * (uint32_t)jobs[0].tile_blocks;
        if (tiles) {
            const uint2 tile = tiles[t];
            job = tile.x;
            first = tile.y;
        }
        if constexpr (kMode == kModeBox) encode_box_tile<kUnit, kStaged>(jobs[job], first, sh);  // either format
        else if constexpr (kTex == kTexMixed) {
            if (jobs[job].rgba8) encode_tile<kMode, true, kUnit, kBoxStage>(jobs[job], job, first, sh, ws, refine);
            else encode_tile<kMode, false, kUnit, kBoxStage>(jobs[job], job, first, sh, ws, refine);
        } else {
            encode_tile<kMode, kTex == kTexRgba8, kUnit, kBoxStage>(jobs[job], job, first, sh, ws, refine);
        }
        if (queue) {
            uint32_t v = 0;
            if (threadIdx.x == 0) {
                v = atomicAdd(queue, 1u);
                // A queue takes exactly `share` claims in a launch: each
                // workgroup with a static tile claims after every tile, and
                // the claims past the share end one workgroup each.  The
                // claim that returns share - 1 is the last one, so it puts the
                // counter back to zero for the next launch (a plan needs no
                // reset kernel).
                if (v == share - 1) atomicExch(queue, 0u);
            }
            j = static_claims + __builtin_amdgcn_readlane(v, 0);
        } else {
            t += gridDim.x;
        }
    }
}

// Zeroes `n` words `stride` words apart (queue counters; a kernel rather
// than a memset node, so that captured graphs replay it correctly).
__global__ void zero_words_kernel(uint32_t* p, uint32_t n, uint32_t stride) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i * stride] = 0u;
}

// refine_endpoints (icbc.h:3574-3651) for the level-9 blocks handed over by
// the encode kernel.  Each step nudges one endpoint of the current best code,
// takes the indices from the current best palette (they only change when a
// step is accepted) and keeps the step if its error is strictly lower; a walk
// ends 33 steps after its last improvement or after 256 steps.
//
// A walk runs on a group of K lanes (K a power of two, 1..64): since every
// step until the next acceptance starts from the same best code, the group
// evaluates steps i .. i+K-1 at once and keeps the first accepted one (the
// later ones are discarded, as the sequential walk would never have tried
// them from that code), or, with none accepted, moves on K steps -- never past
// the step where the walk ends.  Results are those of the sequential walk.
// K grows when there are few walks for the lanes, so short launches are not
// bound by their longest walk, and again, up to 64, once no block is left to
// start (the wavefront's last walks then share all of its lanes).  A group whose walk has ended takes the next
// block from the list (claims are gathered per wavefront).
constexpr int kRefineThreads = 256;

// Weighted squared error of fixed `indices` against the palette of thread
// `tid` in `pal` (entry-major: entry k of thread t at pal[k][t]).  Each texel
// reads its entry with one LDS load instead of selecting it from registers;
// the 64 lanes of a wavefront read 64 consecutive float4 of one row, or of
// rows 4 KiB apart, so the reads are free of bank conflicts.  Same values and
// order of operations as indexed_error.
// With kWeighted false the texels' weights are all 1.0 and the products
// w * error, which equal the errors exactly, are not formed.
template <bool kWeighted, bool kUnit, typename Src>
__device__ __forceinline__ float indexed_error_lds(const Src& t, f3 w, const float4 (*pal)[kRefineThreads], int tid,
                                                   uint32_t indices) {
    float e = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float4 p = pal[(indices >> (2 * i)) & 3][tid];
        const float te = texel_error<kUnit>(mk(p.x, p.y, p.z), t.color(i), w);
        e += kWeighted ? t.weight(i) * te : te;
    }
    return e;
}

// kUnit: every level-9 job of the launch has channel weights (1, 1, 1).
template <bool kUnit>
__global__ __launch_bounds__(kRefineThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void bc1_refine_kernel(
    const JobDev* __restrict__ jobs, RefineList rl) {
    // each thread's candidate palette, written and read back by the same lane
    __shared__ float4 s_pal[4][kRefineThreads];
    const uint32_t total = rl.counters[0];
    uint32_t* claims = rl.counters + kQueueStride / sizeof(uint32_t);
    const int lane = threadIdx.x & (kTile - 1);
    const int decoder = jobs[0].decoder;  // one decoder per device (init)
    // group size: the largest K <= 16 with total * K <= 4 * lanes of the launch
    const uint64_t lanes = (uint64_t)gridDim.x * kRefineThreads;
    int K = 1;
    while (K < 16 && (uint64_t)total * (uint64_t)(2 * K) <= 4 * lanes) K *= 2;
    int sub = lane & (K - 1);
    int leader = lane - sub;
    uint64_t group_bits = ((1ull << K) - 1) << leader;
    uint64_t leaders = 0;
    for (int l = 0; l < kTile; l += K) leaders |= 1ull << l;

    TexelRegs<true> r;
    f3 cw = mk(0.0f, 0.0f, 0.0f);
    bool tcm = false, has = false, drained = false;
    bool weighted = false;  // the walk's block has weights (else all 1.0)
    // `stale`: the walk's best code changed (or the walk just started) and its
    // indices, which every later step uses, are not recomputed yet
    bool stale = false;
    Code out{0u, 0u};
    float best = 0.0f;
    int last = 0, i = 0;
    uint32_t idx = 0, job = 0, block = 0;
    // blocks claimed by this wavefront and not started yet: [lo, hi); the
    // list counter is touched once per kTile / K blocks, not per free group
    const uint32_t chunk = (uint32_t)(kTile / K);
    const uint32_t below = total > chunk ? (total + chunk - 1) / chunk - 1 : 0;  // claims not draining the list
    const uint32_t last_claim = (below + gridDim.x * (kRefineThreads / kTile) - 1) * chunk;
    uint32_t lo = 0, hi = 0;
    // starting a walk reads its block's texels from HBM, which stalls the
    // whole wavefront: free groups start together, once at least a quarter of
    // the wavefront is free (or all of it)
    const int refill_groups = max(1, (kTile / 4) / K);
    const int all_groups = kTile / K;
    for (;;) {
        // Tail: no block is left to start (the list is drained everywhere at
        // about the same time, so every wavefront is here together), and the
        // wavefront's remaining walks would finish on a few lanes.  They spread
        // over the whole wavefront instead: walk w moves to lanes
        // [w K', (w + 1) K') for the largest K' <= 64 with walks * K' <= 64,
        // its state copied from its leader lane and its texels read again.
        if (drained && lo == hi && K < kTile) {
            const uint64_t walks = __ballot(has) & leaders;
            const int nwalks = __popcll(walks);
            int K2 = K, log2K2 = __builtin_ctz((uint32_t)K);
            while (K2 < kTile && nwalks * 2 * K2 <= kTile) {
                K2 *= 2;
                log2K2++;
            }
            if (nwalks > 0 && K2 > K) {
                int src = -1;  // the leader lane of this lane's new walk
                uint64_t left = walks;
                for (int w = 0; left; w++) {
                    const int l = __builtin_ctzll(left);
                    left &= left - 1;
                    if ((lane >> log2K2) == w) src = l;
                }
                const int from = src >= 0 ? src : lane;
                const bool moved = src >= 0;
                const uint32_t cols = (uint32_t)__shfl((int)out.cols, from, kTile);
                const uint32_t oidx = (uint32_t)__shfl((int)out.idx, from, kTile);
                best = __shfl(best, from, kTile);
                last = __shfl(last, from, kTile);
                i = __shfl(i, from, kTile);
                idx = (uint32_t)__shfl((int)idx, from, kTile);
                job = (uint32_t)__shfl((int)job, from, kTile);
                block = (uint32_t)__shfl((int)block, from, kTile);
                stale = __shfl(stale ? 1 : 0, from, kTile) != 0 && moved;
                out = Code{cols, oidx};
                has = moved;
                if (has) {
                    const JobDev& jd = jobs[job];
                    if (jd.rgba8) r.load(block_texels<true>(jd, block));
                    else r.load(block_texels<false>(jd, block));
                    cw = mk(jd.cw0, jd.cw1, jd.cw2);
                    tcm = jd.tcm != 0;
                    weighted = jd.weights != nullptr;
                }
                K = K2;
                sub = lane & (K - 1);
                leader = lane - sub;
                group_bits = (K == kTile ? ~0ull : ((1ull << K) - 1)) << leader;
                leaders = 0;
                for (int l = 0; l < kTile; l += K) leaders |= 1ull << l;
            }
        }
        // free groups (their leader lanes) start the wavefront's next blocks
        uint64_t free_mask = __ballot(!has) & leaders;
        const int nfree = __popcll(free_mask);
        if (nfree < refill_groups && nfree < all_groups) free_mask = 0;
        if (free_mask && lo == hi && !drained) {
            uint32_t base = 0;
            if (lane == 0) {
                base = atomicAdd(claims, chunk);
                // The list takes exactly (m + W) claims in a launch: m below
                // total - chunk, then one per wavefront (the claim that drains
                // it; every wavefront claims until drained, then never again).
                // The last one puts both counters back to zero for the next
                // launch's encode and refinement (a plan needs no reset
                // kernel); every wavefront read `total` before its first claim.
                if (base == last_claim) {
                    atomicExch(rl.counters, 0u);
                    atomicExch(claims, 0u);
                }
            }
            base = __builtin_amdgcn_readlane(base, 0);
            lo = min(base, total);
            hi = min(base + chunk, total);
            if (base + chunk >= total) drained = true;
        }
        if (free_mask && lo < hi) {
            const uint32_t rank = (uint32_t)__popcll(free_mask & ((1ull << leader) - 1));
            const uint32_t k = lo + rank;
            const uint32_t taken = min((uint32_t)__popcll(free_mask), hi - lo);
            lo += taken;
            if (!has && ((free_mask >> leader) & 1) && k < lo) {
                const RefineItem it = rl.items[k];
                const JobDev& jd = jobs[it.job];
                job = it.job;
                block = it.block;
                if (jd.rgba8) r.load(block_texels<true>(jd, block));
                else r.load(block_texels<false>(jd, block));
                cw = mk(jd.cw0, jd.cw1, jd.cw2);
                tcm = jd.tcm != 0;
                weighted = jd.weights != nullptr;
                const uint2 v = ld_global(jd.codes + block);
                out = Code{v.x, v.y};
                best = it.error;
                last = 0;
                i = 0;
                has = true;
                stale = true;
            }
        }
        const uint64_t active = __ballot(has);
        if (!active && lo == hi && drained) break;  // every lane idle, nothing left to start
        // index pass: the indices of a new best code are needed before the
        // walk's next step; lanes wait until a quarter of the active lanes (or
        // all of them) need it, so one pass serves many walks
        const int nstale = __popcll(__ballot(has && stale));
        if (nstale > 0 && (4 * nstale >= __popcll(active) || nstale == __popcll(active))) {
            if (has && stale) {
                f3 sel[4];
                palette_f(col0(out), col1(out), decoder, sel);
                idx = group_select_indices<kUnit>(r, cw, sel, K, sub);
                stale = false;
            }
            continue;
        }
        if (has && !stale) {
            const int limit = min(last + 33, 255);  // the walk's last step unless one is accepted
            const int step = i + sub;
            const bool valid = step <= limit;
            int dr, dg, db;
            refine_step(step & 15, dr, dg, db);
            uint32_t c0 = col0(out), c1 = col1(out);
            if ((step >> 4) & 1) c0 = nudge565(c0, dr, dg, db);
            else c1 = nudge565(c1, dr, dg, db);
            if (!tcm) {
                if (c0 == c1) c1 = nudge565(c1, 0, 1, 0);
                if (c0 < c1) {
                    const uint32_t tmp = c0;
                    c0 = c1;
                    c1 = tmp;
                }
            }
            f3 pal[4];
            palette_f(c0, c1, decoder, pal);
#pragma unroll
            for (int k = 0; k < 4; k++) s_pal[k][threadIdx.x] = make_float4(pal[k].x, pal[k].y, pal[k].z, 0.0f);
            // the unweighted form when no lane taking a step walks a weighted block
            const float e = __ballot(weighted) != 0 ? indexed_error_lds<true, kUnit>(r, cw, s_pal, threadIdx.x, idx)
                                                    : indexed_error_lds<false, kUnit>(r, cw, s_pal, threadIdx.x, idx);
            const uint64_t acc = __ballot(valid && e < best) & group_bits;
            bool finished;
            if (acc) {
                // the first accepted step of the group becomes the best code;
                // its indices are the current ones, the next steps need new ones
                const int j = __builtin_ctzll(acc >> leader);
                const uint32_t cols = __shfl(c0 | (c1 << 16), leader + j, 64);
                const float ej = __shfl(e, leader + j, 64);
                best = ej;
                out = Code{cols, idx};
                last = i + j;
                i = last + 1;
                finished = last == 255;
                stale = true;
            } else {
                finished = i + K - 1 >= limit;
                i += K;
            }
            if (finished) {
                if (sub == 0) {
                    const JobDev& jd = jobs[job];
                    const int steps = acc ? 256 : limit + 1;
                    st_global(jd.codes + block, make_uint2(out.cols, out.idx));
                    if (jd.errors) st_global(jd.errors + block, best);
                    if (jd.stats) st_global(reinterpret_cast<int32_t*>(jd.stats + block) + 2, steps);
                }
                has = false;
                stale = false;
            }
        }
    }
}

// Counts the float bit patterns where recip_rn differs from 1.0f / b (NaN
// results compare equal to each other, zeros by sign).
__global__ void bc1_check_reciprocal_kernel(unsigned long long* __restrict__ mismatches) {
    unsigned long long bad = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const float b = __uint_as_float((uint32_t)i);
        const float fast = recip_rn(b), ieee = 1.0f / b;
        const bool same = (fast != fast && ieee != ieee) || __float_as_uint(fast) == __float_as_uint(ieee);
        bad += same ? 0 : 1;
    }
    if (bad) atomicAdd(mismatches, bad);
}

// decode_bc1 (icbc.h:2915-2927), or with kBc3 decode_bc3 (icbc.h:2929-2946):
// one thread per block, RGBA8 out.  A BC3 block is 16 bytes, the alpha block
// {alpha0, alpha1, 6 index bytes} and then a BC1 block (icbc.h:260-269); the
// reference decodes its colours with the four-colour palette whatever the
// endpoint order, and gives every texel alpha0 (its alpha indices are
// unimplemented: index 0 for every texel).
template <bool kBc3>
__global__ void bc1_decode_kernel(const uint2* __restrict__ codes, uint64_t n,
                                  uint32_t* __restrict__ rgba, int decoder) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint2 c;
    uint32_t alpha0 = 255u;
    if constexpr (kBc3) {
        const uint4 b = reinterpret_cast<const uint4*>(codes)[i];
        alpha0 = b.x & 0xffu;
        c = make_uint2(b.z, b.w);
    } else {
        c = codes[i];
    }
    Pal8 p;
    palette8(c.x & 0xffff, c.x >> 16, decoder, p, kBc3);
    const bool four = kBc3 || (c.x & 0xffff) > (c.x >> 16);
    uint32_t px[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t a = kBc3 ? alpha0 : ((k < 3 || four) ? 255u : 0u);
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
// summed in float in texel order.  With kBc3, evaluate_bc3_error
// (icbc.h:3048-3071): per texel the colour error against the four-colour
// palette, scaled by alpha / 255.0f when `alpha_blend`, then the squared
// difference of alpha0 and the texel's alpha, each added in turn.
template <bool kBc3>
__global__ void bc1_evaluate_kernel(const uint32_t* __restrict__ rgba, const uint2* __restrict__ codes,
                                    uint64_t n, float* __restrict__ errors, int decoder, int alpha_blend) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint2 c;
    int alpha0 = 0;
    if constexpr (kBc3) {
        const uint4 b = reinterpret_cast<const uint4*>(codes)[i];
        alpha0 = (int)(b.x & 0xffu);
        c = make_uint2(b.z, b.w);
    } else {
        c = codes[i];
    }
    Pal8 p;
    palette8(c.x & 0xffff, c.x >> 16, decoder, p, kBc3);
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
            const float ce = (float)(dr * dr + dg * dg + db * db);
            if constexpr (kBc3) {
                const uint32_t a = v[k] >> 24;
                e += (alpha_blend ? u8_to_unit(a) : 1.0f) * ce;
                const int da = alpha0 - (int)a;
                e += (float)(da * da);
            } else {
                e += ce;
            }
        }
    }
    errors[i] = e;
}

// ---------------------------------------------------------------------------
// launch glue

int encode_workgroups_per_cu() {
    int cluster = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&cluster, bc1_encode_kernel<kModeCluster, kTexFloat, false>, kTile, 0) !=
        hipSuccess)
        cluster = 1;
    return std::max(1, cluster);
}

namespace {

// Resident workgroups of a level 0-1 instance on the current device: the
// staged one's LDS area allows fewer per CU than the cluster instance, whose
// residency sizes the launches and workspaces; the register one (no LDS) as
// many as its registers allow.
uint32_t box_grid(bool staged) {
    static std::mutex m;
    static std::vector<std::pair<int, uint32_t>> cache;  // (device * 2 + staged, grid)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    const int key = dev * 2 + (staged ? 1 : 0);
    std::lock_guard<std::mutex> lock(m);
    for (auto& c : cache)
        if (c.first == key) return c.second;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    const hipError_t e =
        staged ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc1_encode_kernel<kModeBox, kTexFloat, false>,
                                                              kTile, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     &per_cu, bc1_encode_kernel<kModeBox, kTexFloat, true, true, false>, kTile, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    const uint32_t g = (uint32_t)(cus * per_cu);
    cache.push_back({key, g});
    return g;
}

// Resident workgroups of the refinement kernel on the current device.
uint32_t refine_grid() {
    static std::mutex m;
    static std::vector<std::pair<int, uint32_t>> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lock(m);
    for (auto& c : cache)
        if (c.first == dev) return c.second;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc1_refine_kernel<false>, kRefineThreads, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const uint32_t g = (uint32_t)cus * (uint32_t)per_cu;
    cache.push_back({dev, g});
    return g;
}

}  // namespace

hipError_t launch_encode(const JobDev* jobs, const uint2* tiles, uint32_t ntiles, uint32_t* counter,
                         float4* workspace, uint32_t grid, int mode, RefineList refine, hipStream_t stream,
                         bool counters_zeroed) {
    if (ntiles == 0) return hipSuccess;
    if (grid > ntiles) grid = ntiles;
    if (counter && !counters_zeroed) {
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, counter, kQueues,
                           (uint32_t)(kQueueStride / sizeof(uint32_t)));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const int tex = (mode >> 4) & 3;
    const bool unit = (mode & kUnitWeights) != 0;
    const bool box = (mode & kBoxStage) != 0;
    const bool staged = !unit || (mode & kTexelWeights) != 0;  // level 0-1 instance
    mode &= 15;
    if (mode == kModeAll && !counters_zeroed) {
        hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, refine.counters, 2u,
                           (uint32_t)(kQueueStride / sizeof(uint32_t)));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // every schedule covers all tiles with any grid (queue claims or a
    // grid-stride walk), so a level 0-1 instance runs at its own residency
    const dim3 g(grid), bg(std::min(grid, box_grid(staged))), blk(kTile);
#define ENC(M, T, U, B) hipLaunchKernelGGL((bc1_encode_kernel<M, T, U, B>), (M) == kModeBox ? bg : g, blk, 0, stream, \
                                           jobs, tiles, ntiles, counter, workspace, refine)
    // The unit-weight instances run the level 0-1 kernels and the refinement;
    // the cluster kernel has only the general one (its unit-weight build was
    // measured slower: the compiler schedules the search worse).  Launches
    // of one texel format at levels 4-9 run cluster instances without the
    // box-fit stage.
    if (mode == kModeBox) {
        if (!staged)
            hipLaunchKernelGGL((bc1_encode_kernel<kModeBox, kTexFloat, true, true, false>), bg, blk, 0, stream, jobs,
                               tiles, ntiles, counter, workspace, refine);
        else if (unit) ENC(kModeBox, kTexFloat, true, true);
        else ENC(kModeBox, kTexFloat, false, true);
    } else if (tex == kTexRgba8) {
        if (box) ENC(kModeCluster, kTexRgba8, false, true);
        else ENC(kModeCluster, kTexRgba8, false, false);
    } else if (tex == kTexMixed) {
        ENC(kModeCluster, kTexMixed, false, true);
    } else {
        if (box) ENC(kModeCluster, kTexFloat, false, true);
        else ENC(kModeCluster, kTexFloat, false, false);
    }
#undef ENC
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || mode != kModeAll) return e;
    if (unit) hipLaunchKernelGGL(bc1_refine_kernel<true>, dim3(refine_grid()), dim3(kRefineThreads), 0, stream, jobs, refine);
    else hipLaunchKernelGGL(bc1_refine_kernel<false>, dim3(refine_grid()), dim3(kRefineThreads), 0, stream, jobs, refine);
    return hipGetLastError();
}

hipError_t warm_launch_caches() {
    (void)box_grid(true);
    (void)box_grid(false);
    (void)refine_grid();
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point.  Host code never uses the array's
// name as an address (on the host it is only a stand-in): the device address
// is resolved per device with hipGetSymbolAddress at set-up and kept below.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices];  // per device ordinal; written at set-up under the runtime's lock
}  // namespace

hipError_t noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Thread g copies job g into the table and writes tile g of the list: tile g
// belongs to the job at launch position p with first_tile[p] <= g <
// first_tile[p + 1] and covers blocks from (g - first_tile[p]) * its tile size.
__global__ void noalloc_setup_kernel(const NoAllocJobs a, JobDev* __restrict__ jobs, uint2* __restrict__ tiles,
                                     uint32_t write_tiles) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < a.n) jobs[g] = a.jobs[g];
    const uint32_t ntiles = a.first_tile[a.n];
    if (!write_tiles || g >= ntiles) return;
    uint32_t lo = 0, hi = a.n;  // first_tile[lo] <= g < first_tile[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.first_tile[mid] <= g) lo = mid;
        else hi = mid;
    }
    const uint32_t k = a.order[lo];
    tiles[g] = make_uint2(k, (g - a.first_tile[lo]) * (uint32_t)a.jobs[k].tile_blocks);
}

hipError_t launch_noalloc_setup(const NoAllocJobs& a, void* scratch_ptr, hipStream_t stream) {
    char* base = static_cast<char*>(scratch_ptr);
    const uint32_t ntiles = a.first_tile[a.n];
    const uint32_t work = std::max(a.n, a.n > 1 ? ntiles : 0u);
    if (work == 0) return hipSuccess;
    const uint32_t threads = 256;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((work + threads - 1) / threads), dim3(threads), 0, stream, a,
                       reinterpret_cast<JobDev*>(base + kNoAllocJobsOffset),
                       reinterpret_cast<uint2*>(base + kNoAllocTilesOffset), a.n > 1 ? 1u : 0u);
    return hipGetLastError();
}

hipError_t launch_check_reciprocal(unsigned long long* mismatches, hipStream_t stream) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream, (uint32_t*)mismatches, 2u, 1u);
    hipLaunchKernelGGL(bc1_check_reciprocal_kernel, dim3(4096), dim3(256), 0, stream, mismatches);
    return hipGetLastError();
}

hipError_t launch_decode(const uint2* codes, uint64_t n, uint32_t* rgba, int decoder, bool bc3,
                         hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t threads = 256;
    const dim3 grid((uint32_t)((n + threads - 1) / threads));
    if (bc3) hipLaunchKernelGGL(bc1_decode_kernel<true>, grid, dim3(threads), 0, stream, codes, n, rgba, decoder);
    else hipLaunchKernelGGL(bc1_decode_kernel<false>, grid, dim3(threads), 0, stream, codes, n, rgba, decoder);
    return hipGetLastError();
}

hipError_t launch_evaluate(const uint32_t* rgba, const uint2* codes, uint64_t n, float* errors,
                           int decoder, bool bc3, bool alpha_blend, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t threads = 256;
    const dim3 grid((uint32_t)((n + threads - 1) / threads));
    if (bc3)
        hipLaunchKernelGGL(bc1_evaluate_kernel<true>, grid, dim3(threads), 0, stream, rgba, codes, n, errors, decoder,
                           alpha_blend ? 1 : 0);
    else
        hipLaunchKernelGGL(bc1_evaluate_kernel<false>, grid, dim3(threads), 0, stream, rgba, codes, n, errors,
                           decoder, 0);
    return hipGetLastError();
}

}  // namespace icbc_gpu
