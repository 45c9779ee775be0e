"""GPU parity at the benchmark configurations' real shapes.

- The 4096x4096 single image (1,048,576 blocks, 40/30/20/10 gradient / white
  noise / near-black / flat, three-colour mode and transparent black on) at
  levels 1, 5 and 9, encoded whole on the device and checked against the
  oracle on a seeded sample of 16,384 blocks per level.
- The 2048x2048 weighted image (alpha = weight, 30% of texels at 0, channel
  weights (0.3, 0.59, 0.11), level 9) gathered on the device from the RGBA
  image: every one of its 262,144 blocks against the oracle.
- A texture set with full box-filtered mip chains and a level drawn per
  texture from 1-9, autotuned and encoded by one graph-captured persistent
  plan: every block against the oracle at its texture's level.

Bar: codes and errors bit-identical (the errors' stated tolerance, 1e-6
relative, is met with zero error).
"""
import os

import numpy as np
import pytest

from icbc_amd import synth
from oracle import prepass

pytestmark = pytest.mark.gpu

NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _say(msg):
    print(msg, flush=True)


def _same(codes, errs, want_codes, want_errs, what):
    bad = np.nonzero((codes != want_codes).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(codes)} codes differ, first {bad[:8]}"
    a, b = errs.astype(np.float64), want_errs.astype(np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    rel[both_nan | (a == b)] = 0
    assert np.all(rel <= 1e-6), f"{what}: worst relative error difference {rel.max()}"
    eq = (errs.view(np.uint32) == want_errs.view(np.uint32)) | both_nan
    assert np.all(eq), f"{what}: {int((~eq).sum())} errors not bit-identical"


@pytest.fixture(scope="module")
def image4096():
    colors, _ = synth.make_image_blocks(7, 4096, 4096, synth.MIX_DEFAULT)
    return colors


@pytest.mark.parametrize("level", [1, 5, 9])
def test_image_4096_sampled_against_oracle(gpu, oracle, image4096, level):
    colors = image4096
    n = len(colors)
    _say(f"4096x4096 L{level}: {n} blocks, plan launch")
    dcol = gpu.DeviceBuffer.from_array(colors)
    dcode, derr = gpu.DeviceBuffer(8 * n), gpu.DeviceBuffer(4 * n)
    s = gpu.Stream()
    try:
        plan = gpu.Plan([gpu.make_job(level, n, dcol, dcode, None, derr, None, (1, 1, 1), True, True)], stream=s)
        plan.execute(s)
        plan.execute(s)  # a replay gives the same output
        s.synchronize()
        plan.close()
        codes, errs = dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n)
    finally:
        for b in (dcol, dcode, derr):
            b.close()
        s.close()
    sample = np.sort(np.random.default_rng(100 + level).choice(n, size=16384, replace=False))
    want_c, want_e, _, _ = oracle.compress_many(level, colors[sample], None, (1, 1, 1), 1, 1, nthreads=NTHREADS)
    _same(codes[sample], errs[sample], want_c, want_e, f"4096x4096 L{level} sample")
    assert np.all(np.isfinite(errs)) and np.all(errs >= 0)


def test_weighted_2048_device_gather_every_block(gpu, oracle):
    img = synth.make_image(4, 2048, 2048, synth.MIX_DEFAULT)
    img[..., 3] = synth.make_weighted_alpha(4, 2048, 2048)
    blocks = prepass.gather_fast(img)
    weights = np.ascontiguousarray(blocks[:, :, 3])
    n = len(blocks)
    cw = (0.3, 0.59, 0.11)
    _say(f"2048x2048 weighted L9: {n} blocks from the device gather")
    di = gpu.DeviceImage(img)
    dcode, derr = gpu.DeviceBuffer(8 * n), gpu.DeviceBuffer(4 * n)
    s = gpu.Stream()
    try:
        assert di.block_count == n
        job = gpu.make_job(9, n, None, dcode, None, derr, None, cw, True, True, image=di.desc,
                           weights_from_alpha=True)
        plan = gpu.Plan([job], stream=s)
        plan.execute(s)
        s.synchronize()
        plan.close()
        codes, errs = dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n)
    finally:
        for b in (dcode, derr, di):
            b.close()
        s.close()
    want_c, want_e, _, _ = oracle.compress_many(9, blocks, weights, cw, 1, 1, nthreads=NTHREADS)
    _same(codes, errs, want_c, want_e, "2048x2048 weighted L9")


def test_texture_set_levels_1_to_9_autotuned_plan_every_block(gpu, oracle):
    texs = synth.texture_set(29, count=24, min_side=64, max_side=1024)
    counts = np.array([len(t[5]) for t in texs])
    offs = np.concatenate([[0], np.cumsum(counts)])
    total = int(offs[-1])
    levels = sorted({t[4] for t in texs})
    _say(f"texture set: {len(texs)} images with mips, {total} blocks, levels {levels}")
    colors = np.concatenate([t[5] for t in texs])
    dcol = gpu.DeviceBuffer.from_array(colors)
    dcode, derr = gpu.DeviceBuffer(8 * total), gpu.DeviceBuffer(4 * total)
    s = gpu.Stream()
    saved = gpu.get_tuning()
    try:
        jobs = [gpu.make_job(t[4], int(n), dcol.ptr + 256 * int(o), dcode.ptr + 8 * int(o), None,
                             derr.ptr + 4 * int(o), None, (1, 1, 1), True, True)
                for t, n, o in zip(texs, counts, offs[:-1])]
        gpu.autotune(jobs, levels=levels)
        plan = gpu.Plan(jobs, stream=s)
        for _ in range(2):
            plan.execute(s)
        s.synchronize()
        plan.close()
        codes, errs = dcode.download(np.uint8, (total, 8)), derr.download(np.float32, total)
    finally:
        gpu.set_tuning(saved)
        for b in (dcol, dcode, derr):
            b.close()
        s.close()
    for lv in levels:
        idx = np.concatenate([np.arange(o, o + n) for t, n, o in zip(texs, counts, offs[:-1]) if t[4] == lv])
        want_c, want_e, _, _ = oracle.compress_many(lv, colors[idx], None, (1, 1, 1), 1, 1, nthreads=NTHREADS)
        _same(codes[idx], errs[idx], want_c, want_e, f"texture set level {lv} ({len(idx)} blocks)")
