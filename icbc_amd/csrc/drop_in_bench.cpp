// Drop-in per-block benchmark: the reference harness's encode loop
// (icbc_test.cpp:389-413) against libicbc_amd.so.  Worker threads take chunks
// of 32 blocks from a shared counter (the harness's ic_pfor(b, n, 32)) and call
// icbc::compress_bc1 once per block with weights 1, channel weights {1,1,1},
// three_color_mode and three_color_black on.  Prints one JSON line: blocks per
// second over the timed passes and the per-call latency percentiles.
//
//   drop_in_bench <blocks.f32> <nblocks> <threads> <level> <repeats> <codes.out>
//
// blocks.f32 holds nblocks x 16 texels x RGBA float32 (block-major); the codes
// of the last pass go to codes.out (8 bytes per block).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "icbc.h"
#include "icbc_gpu.h"

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s blocks.f32 nblocks threads level repeats codes.out\n", argv[0]);
        return 2;
    }
    const size_t n = std::strtoull(argv[2], nullptr, 10);
    const int threads = std::atoi(argv[3]);
    const int level = std::atoi(argv[4]);
    const int repeats = std::atoi(argv[5]);
    std::vector<float> colors(n * 64);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(colors.data(), sizeof(float), colors.size(), f) != colors.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(f);
    icbc::init(icbc::Decoder_D3D10);
    std::vector<unsigned char> codes(n * 8, 0);
    std::vector<float> lat_us(n);
    std::atomic<int> bad{0};
    const float cw[3] = {1, 1, 1};

    auto pass = [&](bool record) {
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            float weights[16];
            for (float& w : weights) w = 1.0f;
            for (;;) {
                const size_t b0 = next.fetch_add(32);
                if (b0 >= n) break;
                for (size_t b = b0; b < std::min(n, b0 + 32); b++) {
                    const auto t0 = std::chrono::steady_clock::now();
                    const float e = icbc::compress_bc1((icbc::Quality)level, &colors[b * 64], weights, cw, true, true,
                                                       &codes[b * 8]);
                    const auto t1 = std::chrono::steady_clock::now();
                    if (std::isnan(e)) bad++;
                    if (record) lat_us[b] = std::chrono::duration<float, std::micro>(t1 - t0).count();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) pool.emplace_back(worker);
        for (auto& t : pool) t.join();
    };

    pass(false);  // warm-up (device set-up, first launches)
    double best = 1e30, total = 0;
    for (int r = 0; r < repeats; r++) {
        const auto t0 = std::chrono::steady_clock::now();
        pass(r == repeats - 1);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        best = std::min(best, s);
        total += s;
    }
    if (bad) {
        std::fprintf(stderr, "%d calls failed: %s\n", bad.load(), icbc_gpu_last_error());
        return 1;
    }
    FILE* o = std::fopen(argv[6], "wb");
    if (!o || std::fwrite(codes.data(), 1, codes.size(), o) != codes.size()) return 2;
    std::fclose(o);
    std::sort(lat_us.begin(), lat_us.end());
    auto pct = [&](double p) { return lat_us[std::min(n - 1, (size_t)(p * (n - 1)))]; };
    std::printf("{\"blocks\": %zu, \"threads\": %d, \"level\": %d, \"repeats\": %d, \"mblocks_per_s_avg\": %.4f, "
                "\"mblocks_per_s_best\": %.4f, \"call_latency_us_p50\": %.1f, \"call_latency_us_p99\": %.1f, "
                "\"call_latency_us_max\": %.1f}\n",
                n, threads, level, repeats, n * repeats / total / 1e6, n / best / 1e6, pct(0.5), pct(0.99),
                lat_us[n - 1]);
    return 0;
}
