// Serving benchmark client: request streams against the request server
// (icbc_server_*, include/icbc_gpu.h) driven from C++, so that the client's
// own cost per request (a few microseconds) stays far below the server's.
// bench.py writes the request pool, runs this program and checks the textures'
// alone-encoded codes against the oracle.
//
//   serve_bench <pool.bin> <alone.out> <seconds> <fractions> <budget> <inflight> <window> <sat_seconds> <seed>
//
// pool.bin: uint32 count, uint32 texel format (0 = RGBA8, 1 = float32 RGBA),
// then per texture {uint64 blocks, int32 level, int32 three_color_mode}, then
// every texture's blocks in pool order (64 or 256 bytes per block).
// alone.out (written): each texture encoded alone by the batch call -- the
// reference every served result must equal: the codes (8 bytes per block) of
// all textures in pool order, then their errors (one float per block).
// fractions: comma-separated loads relative to the measured saturation.
//
// Inputs and result buffers are page-locked (the server reads and writes them
// in place).  Saturation: `window` requests kept outstanding for
// `sat_seconds`, cycling over the pool in a seeded random order.  Then, per
// fraction, a Poisson stream at that share of the saturation rate for
// `seconds`: each request is submitted at its arrival time, its latency is
// completion minus submission (the server's timestamps).  Every served result
// is compared with the texture encoded alone.  Prints one JSON line.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "icbc_gpu.h"

namespace {

struct Texture {
    uint64_t blocks;
    int32_t level, tcm;
    const void* colors;  // page-locked
    size_t alone_off;    // first block in the alone codes
};

struct Pending {
    uint64_t ticket;
    int tex;
    uint8_t* buf;
    bool pinned;
};

struct Run {
    std::vector<icbc_request_info> infos;
    uint64_t mismatches = 0;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    return v[std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5))];
}

}  // namespace

// Seconds between progress lines of a Poisson run (SERVE_REPORT_SECONDS
// overrides, for a finer trace of a long run).
static double kReportSeconds = 10.0;

int main(int argc, char** argv) {
    if (argc != 10) {
        std::fprintf(stderr,
                     "usage: %s pool.bin alone.out seconds fractions budget inflight window sat_seconds seed\n",
                     argv[0]);
        return 2;
    }
    const double seconds = std::atof(argv[3]);
    std::vector<double> fractions;
    for (char* p = argv[4]; *p;) {
        fractions.push_back(std::strtod(p, &p));
        if (*p == ',') p++;
        else if (*p) return 2;
    }
    const uint64_t budget = std::strtoull(argv[5], nullptr, 10);
    const uint32_t inflight = (uint32_t)std::atoi(argv[6]);
    const size_t window = (size_t)std::atoi(argv[7]);
    const double sat_seconds = std::atof(argv[8]);
    std::mt19937_64 rng(std::strtoull(argv[9], nullptr, 10));

    FILE* f = std::fopen(argv[1], "rb");
    uint32_t count = 0, format = 0;
    if (!f || std::fread(&count, 4, 1, f) != 1 || std::fread(&format, 4, 1, f) != 1 || count == 0) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    const size_t bpb = format == 0 ? 64 : 256;
    std::vector<Texture> pool(count);
    uint64_t total = 0;
    for (auto& t : pool) {
        if (std::fread(&t.blocks, 8, 1, f) != 1 || std::fread(&t.level, 4, 1, f) != 1 ||
            std::fread(&t.tcm, 4, 1, f) != 1)
            return 2;
        t.alone_off = total;
        total += t.blocks;
    }
    if (icbc_gpu_init(0) != ICBC_OK) {
        std::fprintf(stderr, "init: %s\n", icbc_gpu_last_error());
        return 1;
    }
    uint8_t* colors = (uint8_t*)icbc_gpu_host_alloc(total * bpb);
    if (!colors || std::fread(colors, bpb, total, f) != total) {
        std::fprintf(stderr, "cannot load the pool\n");
        return 2;
    }
    std::fclose(f);
    for (auto& t : pool) t.colors = colors + t.alone_off * bpb;

    auto job_of = [&](const Texture& t, void* codes) {
        icbc_bc1_job j;
        std::memset(&j, 0, sizeof j);
        j.level = t.level;
        j.three_color_mode = j.three_color_black = t.tcm;
        j.flags = ICBC_JOB_HOST_MEMORY | (format == 0 ? ICBC_JOB_COLORS_RGBA8 : 0);
        j.color_weights[0] = j.color_weights[1] = j.color_weights[2] = 1.0f;
        j.block_count = t.blocks;
        j.colors = (const float*)t.colors;
        j.codes = codes;
        return j;
    };

    // every texture alone: the codes each served result must equal
    std::vector<uint8_t> alone(total * 8, 0);
    std::vector<float> alone_err(total, 0.0f);
    for (const auto& t : pool) {
        icbc_bc1_job j = job_of(t, &alone[t.alone_off * 8]);
        j.errors = &alone_err[t.alone_off];
        if (icbc_gpu_compress_bc1_batch(&j, 1, nullptr) != ICBC_OK) {
            std::fprintf(stderr, "alone: %s\n", icbc_gpu_last_error());
            return 1;
        }
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o || std::fwrite(alone.data(), 1, alone.size(), o) != alone.size() ||
        std::fwrite(alone_err.data(), sizeof(float), total, o) != total)
        return 2;
    std::fclose(o);

    // page-locked result buffers, per_texture for each texture
    const int per_texture = 16;
    uint8_t* results = (uint8_t*)icbc_gpu_host_alloc(total * 8 * per_texture);
    if (!results) return 1;
    std::vector<std::vector<uint8_t*>> free_bufs(count);
    for (uint32_t k = 0; k < count; k++)
        for (int r = 0; r < per_texture; r++)
            free_bufs[k].push_back(results + (pool[k].alone_off * per_texture + r * pool[k].blocks) * 8);
    std::mutex buf_m;
    uint64_t fallbacks = 0;

    icbc_server_options so;
    std::memset(&so, 0, sizeof so);
    so.memory_budget = budget;
    so.max_inflight = inflight;
    if (const char* e = std::getenv("SERVE_REPORT_SECONDS")) kReportSeconds = std::max(0.1, std::atof(e));
    icbc_server* srv = nullptr;
    if (icbc_server_create(&so, &srv) != ICBC_OK) {
        std::fprintf(stderr, "server: %s\n", icbc_gpu_last_error());
        return 1;
    }

    // One stream of requests: `next` returns the texture of the next request
    // once it is due (or -1 to stop); a waiter thread collects requests in
    // submission order, checks them and frees their buffers.
    std::atomic<size_t> outstanding{0};
    std::atomic<int> failed{0};
    // `expected`: about how many requests the stream will submit.  The waiter
    // records every result; its record vector is sized up front, because a
    // reallocation mid-run (about 100 MB at 2^21 records) pauses the waiter,
    // completed requests then hold their page-locked result buffers, and the
    // submitter falls back to fresh pageable ones, whose results the server
    // copies out with page faults on its completion thread.
    auto stream = [&](auto next, size_t expected) {
        Run run;
        run.infos.reserve(expected + expected / 4 + 1024);
        std::mutex qm;
        std::condition_variable qcv;
        std::deque<Pending> q;
        bool end = false;
        std::thread waiter([&] {
            for (;;) {
                Pending p;
                {
                    std::unique_lock<std::mutex> lk(qm);
                    qcv.wait(lk, [&] { return end || !q.empty(); });
                    if (q.empty()) return;
                    p = q.front();
                    q.pop_front();
                }
                icbc_request_info info;
                if (icbc_server_wait(srv, p.ticket, -1.0, &info) != ICBC_OK) failed++;
                const Texture& t = pool[p.tex];
                if (std::memcmp(p.buf, &alone[t.alone_off * 8], t.blocks * 8) != 0) run.mismatches++;
                if (p.pinned) {
                    std::lock_guard<std::mutex> lk(buf_m);
                    free_bufs[p.tex].push_back(p.buf);
                } else {
                    delete[] p.buf;
                }
                run.infos.push_back(info);
                outstanding--;
            }
        });
        for (int k; (k = next()) >= 0;) {
            const Texture& t = pool[k];
            Pending p{0, k, nullptr, true};
            {
                std::lock_guard<std::mutex> lk(buf_m);
                if (!free_bufs[k].empty()) {
                    p.buf = free_bufs[k].back();
                    free_bufs[k].pop_back();
                }
            }
            if (!p.buf) {
                p.buf = new uint8_t[t.blocks * 8];
                p.pinned = false;
                fallbacks++;
            }
            const icbc_bc1_job j = job_of(t, p.buf);
            if (icbc_server_submit(srv, &j, &p.ticket) != ICBC_OK) {
                std::fprintf(stderr, "submit: %s\n", icbc_gpu_last_error());
                failed++;
                if (!p.pinned) delete[] p.buf;
                else {
                    std::lock_guard<std::mutex> lk(buf_m);
                    free_bufs[k].push_back(p.buf);
                }
                break;
            }
            outstanding++;
            {
                std::lock_guard<std::mutex> lk(qm);
                q.push_back(p);
            }
            qcv.notify_one();
        }
        {
            std::lock_guard<std::mutex> lk(qm);
            end = true;
        }
        qcv.notify_one();
        waiter.join();
        return run;
    };
    using clk = std::chrono::steady_clock;
    auto since = [](clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); };

    // warm server: two passes over the pool grow its slabs, staging and tables
    for (int pass = 0; pass < 2; pass++) {
        uint32_t i = 0;
        stream([&] { return i < count ? (int)i++ : -1; }, count);
    }

    // saturation
    std::uniform_int_distribution<uint32_t> pick(0, count - 1);
    Run sat;
    {
        const auto t0 = clk::now();
        sat = stream(
            [&] {
                while (outstanding.load() >= window) std::this_thread::yield();
                return since(t0) < sat_seconds ? (int)pick(rng) : -1;
            },
            (size_t)(sat_seconds * 100000));
    }
    auto span_of = [](const Run& r) {
        double a = 1e300, b = -1e300;
        for (const auto& i : r.infos) {
            a = std::min(a, i.submit_s);
            b = std::max(b, i.done_s);
        }
        return b - a;
    };
    auto blocks_of = [](const Run& r) {
        uint64_t b = 0;
        for (const auto& i : r.infos) b += i.block_count;
        return b;
    };
    auto mean_batch = [](const Run& r) {
        double s = 0;
        for (const auto& i : r.infos) s += i.batch_size;
        return r.infos.empty() ? 0.0 : s / (double)r.infos.size();
    };
    const double sat_span = span_of(sat);
    const double sat_qps = (double)sat.infos.size() / sat_span;
    const double sat_mbps = (double)blocks_of(sat) / sat_span / 1e6;

    std::string runs;
    for (double frac : fractions) {
        const double rate = frac * sat_qps;
        std::exponential_distribution<double> gap(rate);
        const auto t0 = clk::now();
        double t_next = 0, t_report = kReportSeconds;
        uint64_t submitted = 0;
        const Run r = stream([&] {
            t_next += gap(rng);
            if (t_next > seconds) return -1;
            if (t_next >= t_report) {  // a progress line every 10 s of a long run
                icbc_server_stats_t st{};
                (void)icbc_server_stats(srv, &st);
                std::fprintf(stderr,
                             "serving @ %.0f req/s: %.0f s, %llu requests submitted, %llu completed, %llu queued, "
                             "%llu launches in flight, %llu launches, device memory %.1f MB in use / %.1f MB "
                             "held\n",
                             rate, t_report, (unsigned long long)submitted, (unsigned long long)st.completed,
                             (unsigned long long)st.pending, (unsigned long long)st.inflight,
                             (unsigned long long)st.launches, st.memory_in_use / 1e6, st.memory_committed / 1e6);
                std::fflush(stderr);
                t_report += kReportSeconds;
            }
            submitted++;
            const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(t_next));
            if (due - clk::now() > std::chrono::microseconds(200)) std::this_thread::sleep_until(due -
                                                                                   std::chrono::microseconds(100));
            while (clk::now() < due) {
            }
            return (int)pick(rng);
        }, (size_t)(rate * seconds));
        std::vector<double> lat;
        lat.reserve(r.infos.size());
        for (const auto& i : r.infos) lat.push_back((i.done_s - i.submit_s) * 1e3);
        char line[1024];
        std::snprintf(line, sizeof line,
                      "{\"qps_target\": %.2f, \"fraction_of_saturation\": %.3f, \"requests\": %zu, "
                      "\"mblocks_per_s\": %.3f, \"latency_ms_p50\": %.3f, \"latency_ms_p99\": %.3f, "
                      "\"latency_ms_max\": %.3f, \"mean_batch_requests\": %.2f, \"results_identical_to_alone\": %s}",
                      rate, frac, r.infos.size(), (double)blocks_of(r) / span_of(r) / 1e6,
                      pct(lat, 0.5), pct(lat, 0.99), pct(lat, 1.0), mean_batch(r), r.mismatches ? "false" : "true");
        runs += (runs.empty() ? "" : ", ") + std::string(line);
        std::fprintf(stderr, "serving @ %.0f req/s (%.0f %%): %s\n", rate, frac * 100, line);
    }
    icbc_server_destroy(srv);
    if (failed) {
        std::fprintf(stderr, "%d requests failed\n", failed.load());
        return 1;
    }
    std::printf("{\"saturation_requests_per_s\": %.2f, \"saturation_mblocks_per_s\": %.3f, "
                "\"saturation_requests\": %zu, \"saturation_window_requests\": %zu, "
                "\"saturation_mean_batch_requests\": %.2f, \"saturation_results_identical_to_alone\": %s, "
                "\"pool_textures\": %u, \"mean_blocks_per_request\": %.1f, \"memory_budget_bytes\": %llu, "
                "\"max_inflight_launches\": %u, \"seconds_per_rate\": %.1f, \"pageable_result_fallbacks\": %llu, "
                "\"runs\": [%s]}\n",
                sat_qps, sat_mbps, sat.infos.size(), window, mean_batch(sat), sat.mismatches ? "false" : "true", count,
                (double)total / count, (unsigned long long)budget, inflight, seconds,
                (unsigned long long)fallbacks, runs.c_str());
    return 0;
}
