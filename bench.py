"""Benchmark: BC1 Mblocks/s at quality level 8 (bit-exact), batched vs plain GPU.

Workload (default): the mixed-size texture batch -- 256 synthetic textures,
sides log-uniform in [64, 4096] (width and height drawn independently), each
with its full box-filtered mip chain down to 4x4 and edges padded by
replication to multiples of 4; block content 40% gradients / 30% white noise
/ 20% near-black / 10% flat; every texture encoded at quality level 8 with
three_color_mode and three_color_black on, per-texel weights 1, channel
weights (1, 1, 1) -- the reference harness's settings (icbc_test.cpp:383-408).

A step is one encode of the whole batch with inputs resident in HBM:
  batched  one persistent, graph-captured launch over all textures and mips
  plain    one launch per texture mip level (the comparison path)
`value` is the batched path's whole-batch throughput.

Also reported: the dominant kernel's roofline (algorithmic FP32 FLOPs from the
encoder's own per-block work counters over its average duration, measured
with HIP events on the launch stream), the reference encoder's own SIMD
build run multithreaded on every host core on a bounded random sample as the
CPU baseline, and, for every configuration, a parity check of a seeded sample
of the GPU's codes and errors against the oracle (any mismatch makes the run
exit non-zero) and the encoded quality as RMSE / PSNR, computed as the
reference's harness does (icbc_test.cpp:309-324, 442) against the texels
rounded to RGBA8.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
    python bench.py --workload image4096 --levels 1   # one configuration only
    python bench.py --workload weighted2048           # (profiling runs)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BC1 Mblocks/s at level 8 (bit-exact), batched vs plain GPU; p99 request latency"
PEAK_HBM_GBPS = 8000.0
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector, dense (AMD spec)
PEAK_FP32_NOFMA_TFLOPS = 78.65  # one unfused FP32 op per lane per clock (no FMA in the bit-exact build)
# algorithmic FP32 operations per unit of work (divisions count 1), from the
# reference's arithmetic: one 4-cluster candidate, one 3-cluster candidate,
# one refinement step (icbc.h:2685-2742, 2299-2344, 3614-3648)
FLOPS_FIT4 = 121
FLOPS_FIT3 = 104
FLOPS_REFINE_STEP = 888


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_flops(stats):
    n = stats[:, 0].astype(np.int64)
    n3 = stats[:, 1].astype(np.int64)
    steps = stats[:, 2].astype(np.int64)
    fit4 = np.where(n >= 2, (n + 3) * (n + 2) * (n + 1) // 6 - 1, 0)
    fit3 = np.where(n3 >= 1, (n3 + 2) * (n3 + 1) // 2 - 1, 0)
    return int((FLOPS_FIT4 * fit4 + FLOPS_FIT3 * fit3 + FLOPS_REFINE_STEP * steps).sum())


def host_threads():
    """Host cores this process may run on (the CPU baseline and the oracle use
    all of them)."""
    return len(os.sched_getaffinity(0))


class Parity:
    """Seeded-sample checks of GPU results against the oracle (test
    infrastructure: the oracle is the checker, never the thing measured)."""

    def __init__(self):
        from oracle import Oracle
        self.o = Oracle()
        self.o.init(0)
        self.threads = host_threads()
        self.checks = []

    def check(self, what, level, colors, weights, cw, tcm, tcb, codes, errors, sample=8192, seed=0,
              prior=None):
        """Compares `codes`/`errors` (host arrays of all blocks) with the oracle
        on `sample` blocks drawn with `seed` (all of them if fewer)."""
        n = len(colors)
        idx = (np.arange(n) if n <= sample
               else np.sort(np.random.default_rng(seed).choice(n, size=sample, replace=False)))
        w = None if weights is None else weights[idx]
        pr = None if prior is None else prior[idx]
        want_c, want_e, _, _ = self.o.compress_many(level, colors[idx], w, cw, tcm, tcb, nthreads=self.threads,
                                                    prior=pr)
        bad_c = int(np.any(codes[idx] != want_c, axis=1).sum())
        ge, we = errors[idx], want_e
        bad_e = int((~((ge.view(np.uint32) == we.view(np.uint32)) | (np.isnan(ge) & np.isnan(we)))).sum())
        r = {"what": what, "level": level, "blocks_checked": int(len(idx)), "code_mismatches": bad_c,
             "error_mismatches": bad_e}
        self.checks.append(r)
        log(f"parity {what} L{level}: {len(idx)} blocks, {bad_c} code / {bad_e} error mismatches")
        return r

    def summary(self):
        return {"checker": "oracle (C restatement of the reference's scalar build)",
                "blocks_checked": sum(c["blocks_checked"] for c in self.checks),
                "mismatches": sum(c["code_mismatches"] + c["error_mismatches"] for c in self.checks),
                "checks": self.checks}


def check_levels(parity, what, colors, job_levels, counts, codes, errs, tcm, tcb, per_level=2048):
    """Oracle checks of a batch whose jobs (runs of `counts` blocks) each have
    their own level: a seeded sample of `per_level` blocks per level."""
    offs = np.concatenate([[0], np.cumsum(counts)])
    out = []
    for lv in sorted(set(job_levels)):
        idx = np.concatenate([np.arange(o, o + n) for n, o, l in zip(counts, offs[:-1], job_levels) if l == lv])
        sub = np.sort(idx[np.random.default_rng(200 + lv).choice(len(idx), size=min(len(idx), per_level),
                                                                 replace=False)])
        out.append(parity.check(what, lv, colors[sub], None, (1, 1, 1), tcm, tcb, codes[sub], errs[sub],
                                sample=1 << 30))
    return out


def rgba8_of(colors, chunk=1 << 20):
    """Block-major float texels rounded to RGBA8 (the harness's 8-bit source
    image from which it encodes c / 255.0f)."""
    out = np.empty((len(colors), 64), np.uint8)
    for a in range(0, len(colors), chunk):
        c = colors[a:a + chunk].reshape(-1, 64)
        out[a:a + chunk] = np.clip(np.rint(c * np.float32(255)), 0, 255).astype(np.uint8)
    return out


def quality(g, codes_ptr, colors, stream, rgba=None):
    """RMSE / PSNR of the encoded blocks, as the reference's harness reports
    them: mse = sum of evaluate_bc1_error over blocks / (16 * blocks)
    (icbc_test.cpp:309-317), rmse = sqrt(mse), psnr = 20 log10(255 / rmse)
    (319-324), the error taken on the device (bc1_evaluate_kernel) against
    the texels rounded to RGBA8."""
    n = len(colors)
    rgba = rgba8_of(colors) if rgba is None else rgba
    drgba = g.DeviceBuffer.from_array(rgba)
    derr = g.DeviceBuffer(4 * n)
    try:
        g.evaluate_bc1_error_batch(drgba.ptr, codes_ptr, n, derr.ptr, 0, stream)
        stream.synchronize()
        errs = derr.download(np.float32, n)
    finally:
        drgba.close()
        derr.close()
    mse = float(np.sum(errs, dtype=np.float64) / (16 * n))
    rms = math.sqrt(mse)
    psnr = min(max(20 * math.log10(255.0 / rms), 0.0), 300.0) if rms else 1e10
    return {"rmse": round(rms, 4), "psnr_db": round(psnr, 3), "blocks": n,
            "against": "texels rounded to RGBA8 (icbc_test.cpp:309-324)"}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "bc1_encode_kernel_texture_batch_L8_traffic.md")


def committed_traffic(blocks, level, mixed):
    """HBM bytes per launch of the encode kernel from the committed FETCH_SIZE /
    WRITE_SIZE counter summary (copy-calibrated), when it covers this workload."""
    if mixed or level != 8 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    vals = {}
    for line in open(TRAFFIC_PROFILE):
        parts = line.strip().split(":")
        if len(parts) == 2 and parts[0].startswith("traffic_"):
            vals[parts[0]] = int(parts[1])
    if vals.get("traffic_blocks") != blocks:
        return None
    return vals.get("traffic_bytes_per_launch")


def cpu_features():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return set()
    for line in flags.splitlines():
        if line.startswith("flags"):
            return set(line.split(":", 1)[1].split())
    return set()


def cpu_baseline(blocks, level, tcm, tcb, threads, budget_cpu_s=20.0, seed=12):
    """The reference encoder's SIMD build (oracle/_ref) run multithreaded with
    the harness's chunk-of-32 scheduling on a seeded random sample of the
    workload, sized to about `budget_cpu_s` of CPU time.  Falls back to the
    oracle restatement when the reference build is absent."""
    from oracle import Oracle, Reference, reference_available
    feats = cpu_features()
    variant = None
    if {"avx512f", "avx512bw", "avx512dq", "avx512vl"} <= feats and reference_available("avx512"):
        variant = "avx512"
    elif {"avx2", "fma", "bmi2"} <= feats and reference_available("avx2"):
        variant = "avx2"
    elif reference_available("scalar"):
        variant = "scalar"
    if variant:
        enc = Reference(variant)
        enc.init(0)
        kind, label = "reference", f"icbc.h ICBC_SIMD={enc.simd_level} ({variant})"

        def run(b):
            return enc.compress_many(level, b, None, (1, 1, 1), tcm, tcb, nthreads=threads)[2]
    else:
        enc = Oracle()
        enc.init(0)
        kind, label = "port", "oracle restatement (scalar C)"

        def run(b):
            return enc.compress_many(level, b, None, (1, 1, 1), tcm, tcb, nthreads=threads)[3]
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(blocks))
    probe = blocks[np.sort(order[: min(len(blocks), 32768)])]
    t = run(probe)
    per_block_cpu = t * threads / len(probe)
    n = int(min(len(blocks), max(len(probe), budget_cpu_s / max(per_block_cpu, 1e-9))))
    sample = blocks[np.sort(order[:n])]
    secs = run(sample)
    return {"value": round(n / secs / 1e6, 4), "unit": "Mblocks/s", "cores": threads, "kind": kind,
            "sample": f"{n} blocks drawn at random (seed {seed}) from the {len(blocks)}-block workload, "
                      f"level {level}, {label}, {threads} threads (every core this process may use), "
                      f"chunks of 32 blocks; {secs:.2f} s wall"}


def write_serving_pool(path, synth, seed=5, pool_size=128, texels="rgba8"):
    """The serving workload's request pool, written for icbc_amd/serve_bench:
    textures with sides log-uniform in [16, 2048], level uniform in 1..9,
    three-colour mode on for half of them, block mix 40/30/20/10, as RGBA8 or
    float32 texels.  Returns the pool (blocks, block count, level, mode)."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(pool_size):
        w = int(round(np.exp(rng.uniform(np.log(16), np.log(2048)))))
        h = int(round(np.exp(rng.uniform(np.log(16), np.log(2048)))))
        blocks = synth.image_to_blocks(synth.make_image(seed * 10000 + k, w, h, synth.MIX_DEFAULT))
        if texels == "rgba8":
            blocks = np.clip(np.round(blocks * 255), 0, 255).astype(np.uint8)   # RGBA8 texels
        pool.append({"blocks": np.ascontiguousarray(blocks), "n": len(blocks),
                     "level": int(rng.integers(1, 10)), "tcm": bool(rng.integers(0, 2))})
    with open(path, "wb") as f:
        f.write(np.array([pool_size, 0 if texels == "rgba8" else 1], np.uint32).tobytes())
        for p in pool:
            f.write(np.array([p["n"]], np.uint64).tobytes() + np.array([p["level"], int(p["tcm"])],
                                                                       np.int32).tobytes())
        for p in pool:
            f.write(p["blocks"].tobytes())
    return pool


def serving_bench(g, synth, seconds, qps_fracs, seed=5, pool_size=128, budget=64 << 30,
                  max_inflight=3, parity=None, texels="rgba8", saturation_seconds=3.0, window=512):
    """Request streams against the request server, driven by the C++ client
    icbc_amd/serve_bench (a Python client's own cost per request would bound
    the rate it can offer): textures with sides log-uniform in [16, 2048],
    level uniform in 1..9, three-colour mode on for half of them, block mix
    40/30/20/10, stored as RGBA8 (read by the encoder as c / 255.0f) or, with
    texels="float32", as float RGBA (256 B per block).  Inputs and result
    buffers sit in page-locked host memory, which the server reads and writes
    in place, so each request's texels cross PCIe inside its latency.
    Saturation is the sustained rate with `window` requests kept outstanding
    for `saturation_seconds`; then each QPS fraction runs a Poisson stream for
    `seconds`.  Every served result is compared with the same texture encoded
    alone (by the client), and every pool texture encoded alone is checked
    block for block against the oracle (with `parity`)."""
    import subprocess
    import tempfile
    from icbc_amd import build as hip_build

    with tempfile.TemporaryDirectory() as tmp:
        src, alone_path = os.path.join(tmp, "pool.bin"), os.path.join(tmp, "alone.bin")
        pool = write_serving_pool(src, synth, seed, pool_size, texels)
        total = sum(p["n"] for p in pool)
        cmd = [hip_build.SERVE_BIN, src, alone_path, str(seconds), ",".join(str(x) for x in qps_fracs),
               str(budget), str(max_inflight), str(window), str(saturation_seconds), str(seed)]
        # the client's progress lines are passed on as they come
        proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        for line in proc.stderr:
            log(line.rstrip())
        stdout = proc.stdout.read()
        if proc.wait(timeout=600) != 0:
            raise RuntimeError(f"serve_bench failed ({proc.returncode})")
        out = json.loads(stdout.strip().splitlines()[-1])
        raw = np.fromfile(alone_path, np.uint8)
        alone, alone_err = raw[:8 * total].reshape(total, 8), raw[8 * total:].view(np.float32)
    if parity is not None:
        # the oracle reads RGBA8 texels as c / 255.0f, as the encoder does
        off = 0
        for k, p in enumerate(pool):
            f = (np.asarray(p["blocks"], np.float32) / np.float32(255) if texels == "rgba8"
                 else np.asarray(p["blocks"], np.float32))
            parity.check(f"serving pool texture {k}", p["level"], f, None, (1, 1, 1), p["tcm"], p["tcm"],
                         alone[off:off + p["n"]], alone_err[off:off + p["n"]], sample=1 << 30)
            off += p["n"]
    log(f"serving saturation: {out['saturation_requests_per_s']:.1f} requests/s, "
        f"{out['saturation_mblocks_per_s']:.1f} Mblocks/s ({out['saturation_requests']} requests, window {window}, "
        f"mean {out['saturation_mean_batch_requests']:.1f} requests per launch)")
    out["client"] = "icbc_amd/serve_bench (C++), one submitting thread and one waiting thread"
    out["input"] = (("RGBA8 texels (64 B/block)" if texels == "rgba8" else "float32 RGBA texels (256 B/block)")
                    + " and codes (8 B/block) in page-locked host memory, read and written by the server in place")
    return out


def drop_in_blocks(synth):
    """The per-block workload: the 512x512 plain-content image (16,384 blocks)
    as block-major float32 RGBA texels, the input file of icbc_amd/drop_in_bench."""
    colors, _ = synth.make_image_blocks(1, 512, 512, synth.MIX_PLAIN)
    return np.ascontiguousarray(colors, np.float32)


def drop_in_per_block(g, synth, thread_counts, level=8, repeats=3, parity=None):
    """The reference's own call pattern through the drop-in C++ API: the
    harness's loop (icbc_test.cpp:389-413) compiled against libicbc_amd.so
    (icbc_amd/drop_in_bench), worker threads taking chunks of 32 blocks and
    calling icbc::compress_bc1 once per block, over the 512x512 plain-content
    image (16,384 blocks) at level 8 with three-colour mode and transparent
    black on.  Concurrent calls share launches (the single-block ring)."""
    import subprocess
    import tempfile
    from icbc_amd import build as hip_build
    colors = drop_in_blocks(synth)
    n = len(colors)
    runs = []
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "blocks.f32")
        colors.tofile(src)
        for t in thread_counts:
            out = os.path.join(tmp, f"codes{t}.bin")
            res = subprocess.run([hip_build.BENCH_BIN, src, str(n), str(t), str(level), str(repeats), out],
                                 capture_output=True, text=True, timeout=300)
            if res.returncode != 0:
                raise RuntimeError(f"drop_in_bench failed: {res.stderr}")
            r = json.loads(res.stdout.strip().splitlines()[-1])
            codes = np.fromfile(out, np.uint8).reshape(n, 8)
            if parity is not None:
                want = g.compress_many(level, colors, None, (1, 1, 1), True, True)
                r["parity"] = parity.check(f"drop-in per-block calls, {t} threads", level, colors, None, (1, 1, 1),
                                           1, 1, codes, want[1], sample=4096, seed=400 + t)
                r["codes_identical_to_batch_path"] = bool(np.array_equal(codes, want[0]))
            runs.append(r)
            log(f"drop-in per-block, {t} threads: {r}")
    return {"workload": "512x512 plain-content image (16,384 blocks), level 8, icbc::compress_bc1 per block "
                        "from worker threads (chunks of 32), as the reference's harness calls it",
            "runs": runs}


def single_image_levels(g, synth, levels=(1, 5, 9), steps=10, seed=7, parity=None):
    """The 4096x4096 single-image configuration: one plan per level over its
    1,048,576 blocks (40/30/20/10 gradient / white noise / near-black / flat,
    three-colour mode and transparent black on), timed with HIP events; level
    1 against the HBM roofline (268 B per block: texels in, code and error
    out), the cluster levels against the FP32 one."""
    colors, _ = synth.make_image_blocks(seed, 4096, 4096, synth.MIX_DEFAULT)
    n = len(colors)
    dcol = g.DeviceBuffer.from_array(colors)
    dcode, derr, dstat = g.DeviceBuffer(8 * n), g.DeviceBuffer(4 * n), g.DeviceBuffer(16 * n)
    s = g.Stream()
    e0, e1 = g.Event(), g.Event()
    rgba = rgba8_of(colors)
    out = {}
    try:
        for lv in levels:
            job = g.make_job(lv, n, dcol, dcode, None, derr, None, (1, 1, 1), True, True)
            # one kernel per step (two at level 9) is launched directly: a graph
            # replay adds about 5 us per step on these devices (level 1: 0.094
            # against 0.089 ms), and there is no chain of kernels to capture
            plan = g.Plan([job], stream=s, use_graph=False)
            # untimed launches for 0.2 s first: the device's clocks settle over
            # the first launches (a level-1 step takes 0.112 ms on the first
            # few, 0.100 ms once settled)
            t_warm = time.perf_counter()
            while time.perf_counter() - t_warm < 0.2:
                for _ in range(10):
                    plan.execute(s)
                s.synchronize()
            e0.record(s)
            for _ in range(steps):
                plan.execute(s)
            e1.record(s)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / steps
            plan.close()
            r = {"ms": round(ms, 4), "mblocks_per_s": round(n / ms / 1e3, 1)}
            codes, errs = dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n)
            if parity is not None:
                r["parity"] = parity.check("4096x4096 image", lv, colors, None, (1, 1, 1), 1, 1, codes, errs,
                                           seed=100 + lv)
            r["quality"] = quality(g, dcode.ptr, colors, s, rgba)
            if lv <= 1:
                gbps = n * (256 + 8 + 4) / (ms * 1e-3) / 1e9
                r.update({"bound": "hbm", "achieved_gb_per_s": round(gbps, 1), "peak_gb_per_s": PEAK_HBM_GBPS,
                          "frac": round(gbps / PEAK_HBM_GBPS, 4), "algorithmic_bytes_per_block": 268})
            else:
                g.compress_batch([g.make_job(lv, n, dcol, dcode, None, derr, dstat, (1, 1, 1), True, True)],
                                 stream=s)
                s.synchronize()
                tf = algorithmic_flops(dstat.download(np.int32, (n, 4))) / (ms * 1e-3) / 1e12
                r.update({"bound": "valu-fp32", "achieved_tflops": round(tf, 3), "peak_tflops": PEAK_FP32_TFLOPS,
                          "frac": round(tf / PEAK_FP32_TFLOPS, 4)})
            out[str(lv)] = r
            log(f"4096x4096 level {lv}: {r}")
    finally:
        for b in (dcol, dcode, derr, dstat):
            b.close()
    return {"workload": "one 4096x4096 image, 1,048,576 blocks, seed 7, three_color_mode+three_color_black",
            "path": "plan per level, its kernels launched directly", "levels": out}


def mixed_level_batch(g, jobs_for, counts, texs, stream, steps=3, seed=11, parity=None, colors=None,
                      codes_of=None):
    """The batching configuration with its per-texture quality drawn from 1-9:
    the same textures and mips, an autotuned persistent graph launch against
    one launch per image."""
    rng = np.random.default_rng(seed)
    tex_level = rng.integers(1, 10, size=max(t[0] for t in texs) + 1)
    levels = [int(tex_level[t[0]]) for t in texs]
    jobs = jobs_for(levels)
    take, acc = [], 0
    for j, n in zip(jobs, counts):
        take.append(j)
        acc += int(n)
        if acc >= (1 << 20):
            break
    tuned = g.autotune(take, levels=sorted(set(levels)))
    plan = g.Plan(jobs, stream=stream)
    e = [g.Event() for _ in range(2)]

    def timed(run):
        run()
        stream.synchronize()
        e[0].record(stream)
        for _ in range(steps):
            run()
        e[1].record(stream)
        e[1].synchronize()
        return e[0].elapsed_ms(e[1]) / steps

    ms_b = timed(lambda: plan.execute(stream))
    plan.close()
    checks = None
    if parity is not None and codes_of is not None:
        codes, errs = codes_of()
        checks = check_levels(parity, "mixed-level batch", colors, levels, counts, codes, errs, True, True)
    ms_p = timed(lambda: g.compress_batch(jobs, stream=stream, plain=True))
    total = int(np.sum(counts))
    blocks_per_level = {str(lv): int(sum(int(n) for n, l in zip(counts, levels) if l == lv)) for lv in range(1, 10)}
    out = {"workload": "same 256 textures + mips, level drawn uniformly from 1-9 per texture (seed 11)",
           "blocks": total, "blocks_per_level": blocks_per_level,
           "batched_ms": round(ms_b, 3), "batched_mblocks_per_s": round(total / ms_b / 1e3, 2),
           "plain_ms": round(ms_p, 3), "plain_mblocks_per_s": round(total / ms_p / 1e3, 2),
           "batched_over_plain": round(ms_p / ms_b, 3),
           "tuning": {str(k): v for k, v in tuned.as_dict().items()},
           "parity": checks}
    log(f"mixed levels: {out}")
    return out


def weighted_level9(g, synth, steps=5, seed=4, parity=None):
    """The weighted configuration: a 2048x2048 RGBA float image whose alpha (30%
    of texels at 0) is each texel's weight, level 9, channel weights (0.3,
    0.59, 0.11), three-colour mode and transparent black on; blocks gathered
    on the device from the image (ICBC_JOB_IMAGE + weights from alpha)."""
    img = synth.make_image(seed, 2048, 2048, synth.MIX_DEFAULT)
    img[..., 3] = synth.make_weighted_alpha(seed, 2048, 2048)
    blocks = synth.image_to_blocks(img)
    di = g.DeviceImage(img)
    n = di.block_count
    cw = (0.3, 0.59, 0.11)
    dcode, derr, dstat = g.DeviceBuffer(8 * n), g.DeviceBuffer(4 * n), g.DeviceBuffer(16 * n)
    s = g.Stream()
    e0, e1 = g.Event(), g.Event()
    try:
        job = g.make_job(9, n, None, dcode, None, derr, None, cw, True, True, image=di.desc,
                         weights_from_alpha=True)
        plan = g.Plan([job], stream=s)   # device gather + encode + refinement: one graph
        t_warm = time.perf_counter()      # untimed launches for 0.2 s: the clocks settle
        while time.perf_counter() - t_warm < 0.2:
            for _ in range(5):
                plan.execute(s)
            s.synchronize()
        e0.record(s)
        for _ in range(steps):
            plan.execute(s)
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        plan.close()
        codes, errs = dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n)
        check = None
        if parity is not None:
            check = parity.check("2048x2048 weighted image", 9, blocks, np.ascontiguousarray(blocks[:, :, 3]),
                                 cw, 1, 1, codes, errs, seed=300)
        q = quality(g, dcode.ptr, blocks, s)
        g.compress_batch([g.make_job(9, n, None, dcode, None, derr, dstat, cw, True, True, image=di.desc,
                                     weights_from_alpha=True)], stream=s)
        s.synchronize()
        stats = dstat.download(np.int32, (n, 4))
    finally:
        for b in (dcode, derr, dstat, di):
            b.close()
    flops = algorithmic_flops(stats)
    tf = flops / (ms * 1e-3) / 1e12
    out = {"workload": "2048x2048 RGBA, alpha = weight (30% zero), level 9, cw (0.3, 0.59, 0.11), "
                       "gathered on the device", "blocks": n, "ms": round(ms, 4),
           "mblocks_per_s": round(n / ms / 1e3, 2),
           "roofline": {"bound": "valu-fp32", "achieved": round(tf, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tf / PEAK_FP32_TFLOPS, 4),
                        "algorithmic_gflop_per_launch": round(flops / 1e9, 3),
                        "refine_steps_per_block": round(float(stats[:, 2].mean()), 2),
                        "kernels": "image_gather_kernel + bc1_encode_kernel + bc1_refine_kernel (one plan launch)"},
           "parity": check, "quality": q}
    log(f"weighted level 9: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--textures", type=int, default=256)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--mixed-levels", action="store_true",
                    help="per-texture level drawn from 1..9 instead of the fixed level")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve", action="store_true",
                    help="run only the serving benchmark: Poisson arrivals at 50/80/95%% of saturation, 60 s each")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of the short serving run included in the default benchmark")
    ap.add_argument("--serve-profile-seconds", type=float, default=60.0,
                    help="length of each load level of the --serve profile")
    ap.add_argument("--write-drop-in-blocks", metavar="PATH",
                    help="write the per-block workload's input file for icbc_amd/drop_in_bench and exit")
    ap.add_argument("--write-serving-pool", metavar="PATH",
                    help="only write the serving workload's RGBA8 request pool to PATH (the input of "
                         "icbc_amd/serve_bench, for profiling the server on its own) and exit")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true",
                    help="skip the per-block icbc::compress_bc1 run from worker threads")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-level (1-9) batched-vs-plain run on the same textures")
    ap.add_argument("--no-levels", action="store_true",
                    help="skip the 4096x4096 single-image run at levels 1, 5 and 9")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the per-level autotuning pass before the batched plan is built")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM-counter calibration "
                         "under rocprofv3 --pmc)")
    ap.add_argument("--workload", choices=("batch", "image4096", "weighted2048"), default="batch",
                    help="batch: the full benchmark (default); image4096 / weighted2048: only that "
                         "configuration (for per-configuration profiles)")
    ap.add_argument("--levels", default="1,5,9", help="levels of the 4096x4096 image run")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the oracle checks of sampled outputs (for profiling runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the encoder shards with no exchange step, "
              "so multi-GPU figures are not produced (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import icbc_amd as g
    from icbc_amd import synth

    if args.write_drop_in_blocks:
        blocks = drop_in_blocks(synth)
        blocks.tofile(args.write_drop_in_blocks)
        log(f"wrote {len(blocks)} blocks to {args.write_drop_in_blocks}")
        return
    if args.write_serving_pool:
        pool = write_serving_pool(args.write_serving_pool, synth)
        log(f"wrote {len(pool)} textures, {sum(p['n'] for p in pool)} blocks to {args.write_serving_pool}")
        return

    g.init(g.Decoder.D3D10)
    dev = g.device_name()
    log(f"device: {dev}")
    parity = None if args.no_parity else Parity()

    def finish(line):
        """Prints the JSON line; a parity mismatch fails the run."""
        if parity is not None:
            line["parity"] = parity.summary()
        print(json.dumps(line), flush=True)
        if parity is not None and line["parity"]["mismatches"]:
            log(f"PARITY FAILURE: {line['parity']['mismatches']} mismatches against the oracle")
            sys.exit(1)

    if args.copy_calibration_mb:
        nb = args.copy_calibration_mb << 20
        src, dst = g.DeviceBuffer(nb), g.DeviceBuffer(nb)
        src.fill(1)
        g.check(g.lib().icbc_gpu_memcpy_dtod(dst.ptr, src.ptr, nb, None))
        g.check(g.lib().icbc_gpu_stream_synchronize(None))
        src.close()
        dst.close()
        log(f"calibration copy: {nb} bytes device to device")

    if args.serve:
        sec = args.serve_profile_seconds
        res = serving_bench(g, synth, sec, (0.5, 0.8, 0.95), parity=parity)
        res_f = serving_bench(g, synth, sec, (0.8,), parity=parity, texels="float32")
        top = res["runs"][-1]
        finish({"metric": "BC1 serving: Mblocks/s and p50/p99 request latency under Poisson load",
                "value": top["mblocks_per_s"], "unit": "Mblocks/s", "n_gpus": 1,
                "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                "p99_ms": top["latency_ms_p99"], "p50_ms": top["latency_ms_p50"],
                "config": {"workload": "Poisson requests, sides log-uniform 16-2048, levels 1-9, "
                                       "three_color_mode 50/50, 64 GB budget", "device": dev},
                "serving": res, "serving_float32": res_f})
        return

    if args.workload != "batch":
        # one configuration alone (the per-configuration profiles run these)
        if args.workload == "image4096":
            lv = tuple(int(x) for x in args.levels.split(","))
            res = single_image_levels(g, synth, levels=lv, steps=args.steps, parity=parity)
            first = res["levels"][str(lv[0])]
            line = {"metric": f"BC1 Mblocks/s, one 4096x4096 image, level {lv[0]}",
                    "value": first["mblocks_per_s"], "ms_per_step": first["ms"]}
        else:
            res = weighted_level9(g, synth, steps=args.steps, parity=parity)
            line = {"metric": "BC1 Mblocks/s, 2048x2048 weighted image, level 9",
                    "value": res["mblocks_per_s"], "ms_per_step": res["ms"], "roofline": res["roofline"]}
        line.update({"unit": "Mblocks/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
                     "dtype": "f32", "data": "synthetic", "config": {"workload": args.workload, "device": dev},
                     "result": res})
        finish(line)
        return

    t0 = time.time()
    levels = range(1, 10) if args.mixed_levels else [args.level]
    texs = synth.texture_set(args.seed, count=args.textures, levels=levels)
    counts = np.array([len(t[5]) for t in texs], np.int64)
    total = int(counts.sum())
    offsets = np.concatenate([[0], np.cumsum(counts)])
    colors = np.empty((total, 16, 4), np.float32)
    for (tex, mip, w, h, lv, blocks), o in zip(texs, offsets[:-1]):
        colors[o:o + len(blocks)] = blocks
    job_levels = [t[4] for t in texs]
    log(f"workload: {args.textures} textures, {len(texs)} images (with mips), {total} blocks, "
        f"generated in {time.time() - t0:.1f} s")

    tcm = tcb = True
    dcol = g.DeviceBuffer.from_array(colors)
    dcode = g.DeviceBuffer(8 * total)
    derr = g.DeviceBuffer(4 * total)
    dstat = g.DeviceBuffer(16 * total)

    def jobs(with_stats=False):
        return [g.make_job(lv, int(n), dcol.ptr + 256 * int(o), dcode.ptr + 8 * int(o), None,
                           derr.ptr + 4 * int(o), (dstat.ptr + 16 * int(o)) if with_stats else None,
                           (1, 1, 1), tcm, tcb)
                for lv, n, o in zip(job_levels, counts, offsets[:-1])]

    stream = g.Stream()
    tuning = None
    if not args.no_autotune:
        # untimed, like building the plan: tune the workload's levels on its
        # first textures (about a million blocks)
        t0 = time.time()
        all_jobs = jobs()
        take, acc = [], 0
        for j, n in zip(all_jobs, counts):
            take.append(j)
            acc += int(n)
            if acc >= (1 << 20):
                break
        tuned = g.autotune(take, levels=sorted(set(job_levels)))
        stream.synchronize()
        tuning = {"sample_blocks": acc, "seconds": round(time.time() - t0, 2),
                  "levels": {str(k): v for k, v in tuned.as_dict().items()}}
        log(f"autotune: {tuning}")
    plan = g.Plan(jobs(), use_graph=True, stream=stream) if tuning else \
        g.Plan(jobs(), schedule=g.SCHEDULE_PERSISTENT, use_graph=True, order_by_cost=True, stream=stream)
    ev = [g.Event() for _ in range(args.steps + 1)]

    def timed(run_once):
        for _ in range(args.warmup):
            run_once()
        stream.synchronize()
        ev[0].record(stream)
        for k in range(args.steps):
            run_once()
            ev[k + 1].record(stream)
        ev[-1].synchronize()
        per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(args.steps)]
        return per

    per_batched = timed(lambda: plan.execute(stream))
    ms_batched = float(np.mean(per_batched))
    codes_batched = dcode.download(np.uint8, (total, 8))
    log(f"batched: {ms_batched:.3f} ms/step ({total / ms_batched / 1e3:.2f} Mblocks/s)")
    if parity is not None:
        check_levels(parity, "texture batch (batched plan)", colors, job_levels, counts, codes_batched,
                     derr.download(np.float32, total), tcm, tcb, per_level=8192)
    batch_quality = None if args.mixed_levels else quality(g, dcode.ptr, colors, stream)

    result_plain = None
    if not args.no_plain:
        plain_jobs = jobs()
        per_plain = timed(lambda: g.compress_batch(plain_jobs, stream=stream, plain=True))
        ms_plain = float(np.mean(per_plain))
        codes_plain = dcode.download(np.uint8, (total, 8))
        same = bool(np.array_equal(codes_plain, codes_batched))
        result_plain = {"ms_per_step": round(ms_plain, 4),
                        "value": round(total / ms_plain / 1e3, 3),
                        "launches_per_step": len(texs), "codes_identical_to_batched": same}
        log(f"plain: {ms_plain:.3f} ms/step ({total / ms_plain / 1e3:.2f} Mblocks/s), identical={same}")

    # work counters for the roofline (separate, untimed pass)
    g.compress_batch(jobs(with_stats=True), stream=stream)
    stream.synchronize()
    stats = dstat.download(np.int32, (total, 4))
    flops = algorithmic_flops(stats)
    achieved = flops / (ms_batched * 1e-3) / 1e12
    roofline = {"bound": "valu-fp32", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "frac_of_unfused_peak": round(achieved / PEAK_FP32_NOFMA_TFLOPS, 4),
                "traffic": committed_traffic(total, args.level, args.mixed_levels),
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated: "
                                  "profiles/bc1_encode_kernel_texture_batch_L8_traffic.md",
                "algorithmic_bytes_per_launch": total * (256 + 8 + 4),
                "kernel": "bc1_encode_kernel",
                "algorithmic_gflop_per_launch": round(flops / 1e9, 3),
                "mean_fit4_candidates_per_block": round(float(np.mean(
                    np.where(stats[:, 0] >= 2, (stats[:, 0] + 3) * (stats[:, 0] + 2) * (stats[:, 0] + 1) // 6 - 1, 0))), 2)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        sample_lv = args.level if not args.mixed_levels else 8
        cpu = cpu_baseline(colors, sample_lv, tcm, tcb, threads)
        log(f"cpu baseline: {cpu}")

    levels_run = None
    weighted = None
    if not args.no_levels:
        levels_run = single_image_levels(g, synth, levels=tuple(int(x) for x in args.levels.split(",")),
                                         parity=parity)
        weighted = weighted_level9(g, synth, parity=parity)
    mixed = None
    if not args.no_mixed and not args.mixed_levels:
        def jobs_for(levels):
            return [g.make_job(lv, int(n), dcol.ptr + 256 * int(o), dcode.ptr + 8 * int(o), None,
                               derr.ptr + 4 * int(o), None, (1, 1, 1), tcm, tcb)
                    for lv, n, o in zip(levels, counts, offsets[:-1])]
        mixed = mixed_level_batch(g, jobs_for, counts, texs, stream, parity=parity, colors=colors,
                                  codes_of=lambda: (dcode.download(np.uint8, (total, 8)),
                                                    derr.download(np.float32, total)))

    drop_in = None
    if not args.no_drop_in:
        t = host_threads()
        drop_in = drop_in_per_block(g, synth, (t, 8 * t), parity=parity)

    serving = None
    if not args.no_serve:
        serving = {"rgba8": serving_bench(g, synth, args.serve_seconds, (0.8,), parity=parity),
                   "float32": serving_bench(g, synth, args.serve_seconds, (0.8,), parity=parity, texels="float32")}

    value = total / ms_batched / 1e3
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_batched, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("texture batch: 256 textures log-uniform 64-4096 with full mip chains, "
                                + ("levels 1-9 per texture" if args.mixed_levels else f"quality level {args.level}")
                                + ", three_color_mode+three_color_black, weights 1, cw (1,1,1)"),
                   "textures": args.textures, "images": len(texs), "blocks": total,
                   "level": "1-9" if args.mixed_levels else args.level, "seed": args.seed,
                   "path": "persistent graph-captured launch" + (", autotuned" if tuning else ""),
                   "device": dev},
        "autotune": tuning,
        "plain": result_plain,
        "batched_over_plain": (round(result_plain["ms_per_step"] / ms_batched, 4) if result_plain else None),
        "roofline": roofline,
        "quality": batch_quality,
        "cpu_baseline": cpu,
        "serving": serving,
        "drop_in_per_block": drop_in,
        "single_image_levels": levels_run,
        "weighted_level9": weighted,
        "mixed_level_batch": mixed,
    }
    plan.close()
    for b in (dcol, dcode, derr, dstat):
        b.close()
    finish(out)


if __name__ == "__main__":
    main()
