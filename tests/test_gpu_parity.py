"""GPU parity: the HIP encoder against the golden vectors (written by the
reference encoder) and against the oracle on seeded synthetic workloads.

Bar: 8-byte block codes bit-identical; per-block errors bit-identical as well
(the stated tolerance is 1e-6 relative, checked first so a failure reports the
size of a discrepancy; the bit-level check follows).
"""
import os

import numpy as np
import pytest

from icbc_amd import synth

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6
NTHREADS = min(16, os.cpu_count() or 1)


def _say(msg):
    print(msg, flush=True)


def assert_same(codes, errs, want_codes, want_errs, what):
    bad = np.nonzero((codes != want_codes).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(codes)} codes differ, first {bad[:8]}"
    a = errs.astype(np.float64)
    b = want_errs.astype(np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    rel[both_nan | (a == b)] = 0
    assert np.all(rel <= REL_TOL), f"{what}: error mismatch, worst rel {rel.max()} at {int(np.argmax(rel))}"
    eq = (errs.view(np.uint32) == want_errs.view(np.uint32)) | both_nan
    assert np.all(eq), f"{what}: {int((~eq).sum())} errors not bit-identical"


def test_golden_vectors(gpu, golden):
    colors, weights, prior = golden["colors"], golden["weights"], golden["prior"]
    try:
        for k, (dec, level, tcm, tcb, cwi) in enumerate(golden["cfg"]):
            gpu.init(int(dec))
            codes, errs, _ = gpu.compress_many(int(level), colors, weights, golden["cw"][cwi], tcm, tcb,
                                               prior=prior)
            assert_same(codes, errs, golden["codes"][k], golden["errors"][k],
                        f"golden config {k} {tuple(int(x) for x in golden['cfg'][k])}")
    finally:
        gpu.init(0)


def _oracle_run(oracle, level, colors, weights, cw, tcm, tcb, prior=None):
    codes, errs, stats, _ = oracle.compress_many(level, colors, weights, cw, tcm, tcb,
                                                 nthreads=NTHREADS, prior=prior, with_stats=True)
    return codes, errs, stats


@pytest.mark.parametrize("level", list(range(-1, 11)))
def test_edge_blocks_every_level(gpu, oracle, level):
    colors, weights = synth.make_edge_blocks(300 + max(level, 0), 1024)
    prior = np.random.default_rng(level + 5).integers(0, 256, size=(1024, 8), dtype=np.uint8)
    for tcm, tcb in ((0, 0), (1, 0), (1, 1)):
        for cw in ((1, 1, 1), (0.3, 0.59, 0.11)):
            want = _oracle_run(oracle, level, colors, weights, cw, tcm, tcb, prior)
            got = gpu.compress_many(level, colors, weights, cw, tcm, tcb, prior=prior, with_stats=True)
            assert_same(got[0], got[1], want[0], want[1], f"edge L{level} tcm={tcm} tcb={tcb} cw={cw}")
            assert np.array_equal(got[2][:, :3], want[2][:, :3]), "stats differ"


def _few_colour_blocks(seed, n):
    """Blocks of k distinct colours (k = 1..6, with near-black texels in some):
    short three-cluster candidate lists (2..27 entries), so that one 64-lane
    round of the packed three-cluster search completes several blocks, with
    single-colour blocks (no search) between them."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 16, 4), np.float32)
    for b in range(n):
        k = int(rng.integers(1, 7))
        pal = rng.random((k, 3), dtype=np.float32)
        if rng.random() < 0.3:
            pal[0] = rng.random(3, dtype=np.float32) * 0.02  # a black colour: skip_blacks path
        pick = rng.integers(0, k, 16)
        out[b, :, :3] = pal[pick]
        out[b, :, 3] = 1.0
    return out


@pytest.mark.parametrize("tile", [8, 64])
def test_few_colour_blocks_packed_three_cluster_rounds(gpu, oracle, tile):
    colors = _few_colour_blocks(77 + tile, 3000)
    n = len(colors)
    dc = gpu.DeviceBuffer.from_array(colors)
    dcode = gpu.DeviceBuffer(8 * n)
    derr = gpu.DeviceBuffer(4 * n)
    dst = gpu.DeviceBuffer(16 * n)
    try:
        for level, tcb in ((8, 0), (8, 1), (9, 1)):
            want = _oracle_run(oracle, level, colors, None, (1, 1, 1), 1, tcb)
            dcode.fill(0)
            job = gpu.make_job(level, n, dc, dcode, None, derr, dst, (1, 1, 1), True, bool(tcb), tile_blocks=tile)
            gpu.compress_batch([job])
            assert_same(dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n), want[0], want[1],
                        f"few-colour L{level} tcb={tcb} tile={tile}")
            assert np.array_equal(dst.download(np.int32, (n, 4))[:, :3], want[2][:, :3]), "stats differ"
    finally:
        for b in (dc, dcode, derr, dst):
            b.close()


def test_plain_image_level8(gpu, oracle):
    # 512x512, 50% gradients / 30% value noise / 20% flat, L8, three_color_mode off
    colors, _ = synth.make_image_blocks(1, 512, 512, synth.MIX_PLAIN)
    _say(f"512x512 L8: {len(colors)} blocks")
    want = _oracle_run(oracle, 8, colors, None, (1, 1, 1), 0, 0)
    got = gpu.compress_many(8, colors, None, (1, 1, 1), False, False, with_stats=True)
    assert_same(got[0], got[1], want[0], want[1], "512x512 L8")
    assert np.array_equal(got[2][:, :3], want[2][:, :3])


@pytest.mark.parametrize("level", [1, 5, 9])
def test_mixed_image_levels(gpu, oracle, level):
    colors, _ = synth.make_image_blocks(2, 256, 256, synth.MIX_DEFAULT)
    want = _oracle_run(oracle, level, colors, None, (1, 1, 1), 1, 1)
    got = gpu.compress_many(level, colors, None, (1, 1, 1), True, True, with_stats=True)
    assert_same(got[0], got[1], want[0], want[1], f"256x256 mix L{level}")
    assert np.array_equal(got[2][:, :3], want[2][:, :3])


def test_weighted_alpha_level9(gpu, oracle):
    colors, _ = synth.make_image_blocks(3, 256, 256, synth.MIX_DEFAULT)
    alpha = synth.make_weighted_alpha(3, 256, 256)
    a = alpha.reshape(64, 4, 64, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    colors[..., 3] = a
    cw = (0.3, 0.59, 0.11)
    want = _oracle_run(oracle, 9, colors, a, cw, 1, 1)
    got = gpu.compress_many(9, colors, a, cw, True, True, with_stats=True)
    assert_same(got[0], got[1], want[0], want[1], "weighted L9")


@pytest.mark.parametrize("decoder", [1, 2, 3])
def test_other_decoders(gpu, oracle, decoder):
    colors, _ = synth.make_image_blocks(4, 128, 128, synth.MIX_DEFAULT)
    ec, ew = synth.make_edge_blocks(40 + decoder, 256)
    colors = np.concatenate([colors, ec])
    weights = np.concatenate([np.ones((1024, 16), np.float32), ew])
    try:
        gpu.init(decoder)
        oracle.init(decoder)
        for level in (0, 1, 2, 4, 8, 9):
            want = _oracle_run(oracle, level, colors, weights, (1, 1, 1), 1, 1)
            got = gpu.compress_many(level, colors, weights, (1, 1, 1), True, True)
            assert_same(got[0], got[1], want[0], want[1], f"decoder {decoder} L{level}")
    finally:
        gpu.init(0)
        oracle.init(0)


def test_single_block_c_abi(gpu, oracle):
    lib = gpu.lib()
    colors, weights = synth.make_edge_blocks(77, 48)
    for level in (0, 1, 3, 8, 9, 12):
        for b in range(len(colors)):
            c = np.ascontiguousarray(colors[b].reshape(64))
            w = np.ascontiguousarray(weights[b])
            cw = np.array([1, 1, 1], np.float32)
            out = np.full(8, 0x3c, np.uint8)
            e = lib.icbc_gpu_compress_bc1(level, c.ctypes.data, w.ctypes.data, cw.ctypes.data, 1, 1,
                                          out.ctypes.data)
            want_c, want_e = oracle.compress_block(level, c, w, cw, True, True, prior=np.full(8, 0x3c, np.uint8))
            assert np.array_equal(out, want_c)
            assert np.float32(e).view(np.uint32) == np.float32(want_e).view(np.uint32) or (np.isnan(e) and np.isnan(want_e))
            # the python mirror of icbc::compress_bc1
            out2 = np.full(8, 0x3c, np.uint8)
            e2 = gpu.compress_bc1(level, c, w, cw, True, True, out2)
            assert np.array_equal(out2, want_c) and (np.float32(e2) == np.float32(want_e) or np.isnan(e2))


def test_single_block_calls_from_many_threads(gpu, oracle):
    # icbc_gpu_compress_bc1 called per block from 24 threads at once, each
    # thread with its own level and options, so that concurrent calls share
    # launches and runs of different options meet in one launch
    import threading
    colors, weights = synth.make_edge_blocks(61, 24 * 40)
    prior = np.random.default_rng(62).integers(0, 256, size=(len(colors), 8), dtype=np.uint8)
    cws = [(1, 1, 1), (0.3, 0.59, 0.11)]
    results = {}

    fn = gpu.lib().icbc_gpu_compress_bc1  # the C ABI (ctypes releases the GIL during the call)

    def worker(t):
        level, tcm, cw = t % 11 - 1, t % 3, np.array(cws[t % 2], np.float32)
        for b in range(t * 40, t * 40 + 40):
            out = prior[b].copy()
            c = np.ascontiguousarray(colors[b].reshape(64))
            w = np.ascontiguousarray(weights[b])
            e = fn(level, c.ctypes.data, w.ctypes.data, cw.ctypes.data, int(tcm > 0), int(tcm > 1), out.ctypes.data)
            results[b] = (out, e)

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(24)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    for t in range(24):
        level, tcm, cw = t % 11 - 1, t % 3, cws[t % 2]
        sl = slice(t * 40, t * 40 + 40)
        want_c, want_e, _ = _oracle_run(oracle, level, colors[sl], weights[sl], cw, tcm > 0, tcm > 1, prior[sl])
        got_c = np.array([results[b][0] for b in range(t * 40, t * 40 + 40)])
        got_e = np.array([results[b][1] for b in range(t * 40, t * 40 + 40)], np.float32)
        assert_same(got_c, got_e, want_c, want_e, f"thread {t} level {level}")


def test_drop_in_bench_program(gpu, oracle, tmp_path):
    # the C++ harness-style caller of icbc::compress_bc1 (8 threads)
    import json
    import subprocess
    from icbc_amd import build as hip_build
    colors, _ = synth.make_image_blocks(1, 128, 128, synth.MIX_PLAIN)
    src = tmp_path / "b.f32"
    np.ascontiguousarray(colors, np.float32).tofile(src)
    out = tmp_path / "codes.bin"
    res = subprocess.run([hip_build.BENCH_BIN, str(src), str(len(colors)), "8", "8", "1", str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    r = json.loads(res.stdout.strip().splitlines()[-1])
    assert r["blocks"] == len(colors) and r["mblocks_per_s_best"] > 0
    want = _oracle_run(oracle, 8, colors, None, (1, 1, 1), 1, 1)
    assert np.array_equal(np.fromfile(out, np.uint8).reshape(-1, 8), want[0])


def test_cpp_namespace_api_against_golden(gpu, golden):
    # the exported icbc:: functions themselves (include/icbc.h), called by
    # their C++ symbols as a program linked against the library would
    import ctypes as C
    lib = C.CDLL(gpu.LIB_PATH)
    init = lib._ZN4icbc4initENS_7DecoderE
    init.argtypes, init.restype = [C.c_int], None
    comp = lib._ZN4icbc12compress_bc1ENS_7QualityEPKfS2_S2_bbPv
    comp.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_bool, C.c_bool, C.c_void_p]
    comp.restype = C.c_float
    dec = lib._ZN4icbc10decode_bc1EPKvPhNS_7DecoderE
    dec.argtypes, dec.restype = [C.c_void_p, C.c_void_p, C.c_int], None
    ev = lib._ZN4icbc18evaluate_bc1_errorEPKhPKvNS_7DecoderE
    ev.argtypes, ev.restype = [C.c_void_p, C.c_void_p, C.c_int], C.c_float
    dec3 = lib._ZN4icbc10decode_bc3EPKvPhNS_7DecoderE
    dec3.argtypes, dec3.restype = [C.c_void_p, C.c_void_p, C.c_int], None
    ev3 = lib._ZN4icbc18evaluate_bc3_errorEPKhPKvbNS_7DecoderE
    ev3.argtypes, ev3.restype = [C.c_void_p, C.c_void_p, C.c_bool, C.c_int], C.c_float
    colors, weights, prior = golden["colors"], golden["weights"], golden["prior"]
    try:
        for k, (decoder, level, tcm, tcb, cwi) in enumerate(golden["cfg"]):
            if k % 3:
                continue   # every third configuration, block by block
            init(int(decoder))
            cw = np.ascontiguousarray(golden["cw"][cwi], np.float32)
            for b in range(0, len(colors), 8):
                c = np.ascontiguousarray(colors[b].reshape(64))
                w = np.ascontiguousarray(weights[b])
                out = np.array(prior[b], np.uint8, copy=True)
                e = comp(int(level), c.ctypes.data, w.ctypes.data, cw.ctypes.data, bool(tcm), bool(tcb),
                         out.ctypes.data)
                assert np.array_equal(out, golden["codes"][k][b]), f"config {k} block {b}"
                want = golden["errors"][k][b]
                assert np.float32(e).view(np.uint32) == want.view(np.uint32) or (np.isnan(e) and np.isnan(want))
        for d in range(4):
            for i in range(0, len(golden["dec_codes"]), 16):
                code = np.ascontiguousarray(golden["dec_codes"][i])
                rgba = np.zeros(64, np.uint8)
                dec(code.ctypes.data, rgba.ctypes.data, d)
                assert np.array_equal(rgba, golden["dec_rgba"][d][i])
                src = np.ascontiguousarray(golden["eval_rgba"][i])
                e = ev(src.ctypes.data, code.ctypes.data, d)
                assert np.float32(e) == golden["eval_err"][d][i]
                b3 = np.ascontiguousarray(golden["bc3_blocks"][i])
                dec3(b3.ctypes.data, rgba.ctypes.data, d)
                assert np.array_equal(rgba, golden["bc3_rgba"][d][i])
                src3 = np.ascontiguousarray(golden["bc3_eval_rgba"][i])
                for blend in (False, True):
                    e = ev3(src3.ctypes.data, b3.ctypes.data, blend, d)
                    assert np.float32(e) == golden["bc3_eval_err"][int(blend)][d][i]
    finally:
        gpu.init(0)


def test_batch_plain_and_plans_agree(gpu, oracle):
    # several jobs with different levels, sizes and options, on device memory
    specs = [(8, 700, (1, 1, 1), 1, 1), (1, 64, (1, 1, 1), 0, 0), (9, 333, (0.3, 0.59, 0.11), 1, 1),
             (5, 1000, (1, 1, 1), 1, 1), (0, 1, (1, 1, 1), 0, 0), (3, 129, (1, 2, 1), 1, 0),
             (12, 50, (1, 1, 1), 1, 1), (2, 0, (1, 1, 1), 1, 1)]
    host = []
    for k, (level, n, cw, tcm, tcb) in enumerate(specs):
        c, _ = synth.make_image_blocks(500 + k, 4 * max(n, 1), 4, synth.MIX_DEFAULT)
        c = c[:n]
        w = np.random.default_rng(k).random((n, 16)).astype(np.float32) if k % 2 else None
        host.append((c, w))
    bufs = []

    def device_jobs():
        jobs = []
        for (level, n, cw, tcm, tcb), (c, w) in zip(specs, host):
            dc = gpu.DeviceBuffer.from_array(c)
            dw = gpu.DeviceBuffer.from_array(w) if w is not None else None
            dcode = gpu.DeviceBuffer(max(8 * n, 8))
            dcode.fill(0x77)
            derr = gpu.DeviceBuffer(max(4 * n, 4))
            bufs.append((dc, dw, dcode, derr))
            jobs.append(gpu.make_job(level, n, dc, dcode, dw, derr, None, cw, tcm, tcb))
        return jobs

    def collect(first):
        out = []
        for (level, n, *_), (dc, dw, dcode, derr) in zip(specs, bufs[first:first + len(specs)]):
            out.append((dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n)))
        return out

    results = {}
    jobs = device_jobs()
    gpu.compress_batch(jobs)
    results["batch"] = collect(0)
    jobs = device_jobs()
    gpu.compress_batch(jobs, plain=True)
    results["plain"] = collect(len(specs))
    for name, kw in (("persistent_graph", dict(schedule=1, use_graph=True)),
                     ("persistent", dict(schedule=1, use_graph=False)),
                     ("grid_graph", dict(schedule=0, use_graph=True, order_by_cost=False))):
        first = len(bufs)
        jobs = device_jobs()
        plan = gpu.Plan(jobs, **kw)
        stream = gpu.Stream()
        for _ in range(3):   # replays must give the same bytes every time
            plan.execute(stream)
            stream.synchronize()
            results[name] = collect(first)
        plan.close()
        stream.close()
    for (level, n, cw, tcm, tcb), (c, w), k in zip(specs, host, range(len(specs))):
        prior = np.full((n, 8), 0x77, np.uint8)
        want = _oracle_run(oracle, level, c, w, cw, tcm, tcb, prior)
        for name, res in results.items():
            assert_same(res[k][0], res[k][1], want[0], want[1], f"{name} job {k} L{level}")


@pytest.mark.parametrize("tile_blocks", [8, 16, 32, 64, 0])
def test_tile_sizes_agree(gpu, oracle, tile_blocks):
    colors, _ = synth.make_image_blocks(9, 96, 80, synth.MIX_DEFAULT)   # 480 blocks: ragged tiles
    ec, ew = synth.make_edge_blocks(91, 100)
    colors = np.concatenate([colors, ec])
    weights = np.concatenate([np.ones((480, 16), np.float32), ew])
    n = len(colors)
    for level in (1, 5, 8, 9):
        want = _oracle_run(oracle, level, colors, weights, (1, 1, 1), 1, 1)
        dc = gpu.DeviceBuffer.from_array(colors)
        dw = gpu.DeviceBuffer.from_array(weights)
        dcode = gpu.DeviceBuffer(8 * n)
        dcode.fill(0)
        derr = gpu.DeviceBuffer(4 * n)
        job = gpu.make_job(level, n, dc, dcode, dw, derr, None, (1, 1, 1), True, True, tile_blocks=tile_blocks)
        for plain in (False, True):
            gpu.compress_batch([job], plain=plain)
            assert_same(dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n), want[0], want[1],
                        f"tile_blocks={tile_blocks} L{level} plain={plain}")


def test_rgba8_input_matches_float(gpu, oracle):
    # 8-bit texels are read as c / 255.0f, exactly as the reference's harness
    # converts them before calling compress_bc1 (icbc_test.cpp:401-405)
    colors, _ = synth.make_image_blocks(12, 128, 96, synth.MIX_DEFAULT)
    c8 = np.clip(np.round(colors * 255), 0, 255).astype(np.uint8)
    c8[..., 3] = 255
    cf = c8.astype(np.float32) / np.float32(255)
    for level in (0, 1, 4, 8, 9):
        want = _oracle_run(oracle, level, cf, None, (1, 1, 1), 1, 1)
        got8 = gpu.compress_many(level, c8, None, (1, 1, 1), True, True)
        gotf = gpu.compress_many(level, cf, None, (1, 1, 1), True, True)
        assert_same(got8[0], got8[1], want[0], want[1], f"rgba8 L{level}")
        assert_same(gotf[0], gotf[1], want[0], want[1], f"float L{level}")


@pytest.mark.parametrize("formats", ["mixed", "rgba8"])
def test_rgba8_and_float_jobs_in_one_batch(gpu, oracle, formats):
    # one launch holding 8-bit and float jobs reads each job in its own
    # format; an all-8-bit launch takes the 8-bit instance of the kernels
    specs = [(1, 300, True, False), (1, 200, False, True), (8, 500, True, True), (9, 257, False, False),
             (9, 190, True, True), (0, 64, True, False), (5, 130, False, False), (8, 77, True, False)]
    if formats == "rgba8":
        specs = [(lv, n, True, w) for lv, n, _, w in specs]
    host = []
    for k, (level, n, is8, weighted) in enumerate(specs):
        c, _ = synth.make_image_blocks(700 + k, 4 * n, 4, synth.MIX_DEFAULT)
        c8 = np.clip(np.round(c * 255), 0, 255).astype(np.uint8)
        cf = c8.astype(np.float32) / np.float32(255)
        w = np.random.default_rng(k).random((n, 16)).astype(np.float32) if weighted else None
        host.append((c8 if is8 else cf, cf, w))
    bufs = []

    def device_jobs():
        jobs = []
        for (level, n, is8, _), (c, _, w) in zip(specs, host):
            dc = gpu.DeviceBuffer.from_array(c)
            dw = gpu.DeviceBuffer.from_array(w) if w is not None else None
            dcode = gpu.DeviceBuffer(8 * n)
            dcode.fill(0x77)
            derr = gpu.DeviceBuffer(4 * n)
            bufs.append((dc, dw, dcode, derr))
            jobs.append(gpu.make_job(level, n, dc, dcode, dw, derr, None, (1, 1, 1), True, True, rgba8=is8))
        return jobs

    def collect(first):
        return [(dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n))
                for (_, n, *_), (_, _, dcode, derr) in zip(specs, bufs[first:first + len(specs)])]

    results = {}
    for name in ("batch", "plain", "plan"):
        first = len(bufs)
        jobs = device_jobs()
        _say(f"{formats} {name}")
        if name == "plan":
            plan = gpu.Plan(jobs)
            stream = gpu.Stream()
            for _ in range(2):
                plan.execute(stream)
                stream.synchronize()
                results[name] = collect(first)
            plan.close()
            stream.close()
        else:
            gpu.compress_batch(jobs, plain=name == "plain")
            results[name] = collect(first)
    for k, ((level, n, is8, _), (_, cf, w)) in enumerate(zip(specs, host)):
        want = _oracle_run(oracle, level, cf, w, (1, 1, 1), 1, 1, np.full((n, 8), 0x77, np.uint8))
        for name, res in results.items():
            assert_same(res[k][0], res[k][1], want[0], want[1],
                        f"{formats} {name} job {k} L{level} {'rgba8' if is8 else 'float'}")


def test_decode_and_evaluate(gpu, golden, oracle):
    codes = np.ascontiguousarray(golden["dec_codes"])
    n = len(codes)
    rgba = np.ascontiguousarray(golden["eval_rgba"])
    dcodes = gpu.DeviceBuffer.from_array(codes)
    drgba = gpu.DeviceBuffer(64 * n)
    dsrc = gpu.DeviceBuffer.from_array(rgba)
    derr = gpu.DeviceBuffer(4 * n)
    for dec in range(4):
        gpu.decode_bc1_batch(dcodes.ptr, n, drgba.ptr, dec)
        assert np.array_equal(drgba.download(np.uint8, (n, 64)), golden["dec_rgba"][dec])
        gpu.evaluate_bc1_error_batch(dsrc.ptr, dcodes.ptr, n, derr.ptr, dec)
        assert np.array_equal(derr.download(np.float32, n).view(np.uint32), golden["eval_err"][dec].view(np.uint32))
    # single-block mirrors
    assert np.array_equal(gpu.decode_bc1(codes[3].tobytes(), 2), golden["dec_rgba"][2, 3])
    assert gpu.evaluate_bc1_error(rgba[5], codes[5].tobytes(), 1) == golden["eval_err"][1, 5]
    for b in (dcodes, drgba, dsrc, derr):
        b.close()


def test_bc3_decode_and_evaluate(gpu, golden):
    # decode_bc3 / evaluate_bc3_error (icbc.h:2929-2946, 3048-3071) against
    # the reference-written vectors: 128 random BC3 blocks, every decoder,
    # with and without alpha blending
    blocks = np.ascontiguousarray(golden["bc3_blocks"])
    n = len(blocks)
    dblk = gpu.DeviceBuffer.from_array(blocks)
    drgba = gpu.DeviceBuffer(64 * n)
    dsrc = gpu.DeviceBuffer.from_array(np.ascontiguousarray(golden["bc3_eval_rgba"]))
    derr = gpu.DeviceBuffer(4 * n)
    try:
        for dec in range(4):
            gpu.decode_bc3_batch(dblk.ptr, n, drgba.ptr, dec)
            assert np.array_equal(drgba.download(np.uint8, (n, 64)), golden["bc3_rgba"][dec])
            for blend in range(2):
                gpu.evaluate_bc3_error_batch(dsrc.ptr, dblk.ptr, n, derr.ptr, blend, dec)
                assert np.array_equal(derr.download(np.float32, n).view(np.uint32),
                                      golden["bc3_eval_err"][blend, dec].view(np.uint32)), (dec, blend)
        assert np.array_equal(gpu.decode_bc3(blocks[7].tobytes(), 3), golden["bc3_rgba"][3, 7])
        e = gpu.evaluate_bc3_error(golden["bc3_eval_rgba"][9], blocks[9].tobytes(), True, 1)
        assert np.float32(e) == golden["bc3_eval_err"][1, 1, 9]
    finally:
        for b in (dblk, drgba, dsrc, derr):
            b.close()


def test_full_size_image_level8_properties(gpu, oracle):
    # 4096x4096 (1,048,576 blocks) at L8: a seeded sample against the oracle,
    # determinism across runs, and size-independent invariants
    colors, cat = synth.make_image_blocks(7, 4096, 4096, synth.MIX_DEFAULT)
    n = len(colors)
    _say(f"4096x4096 L8: {n} blocks")
    dcol = gpu.DeviceBuffer.from_array(colors)
    runs = []
    for _ in range(2):
        dcode = gpu.DeviceBuffer(8 * n)
        derr = gpu.DeviceBuffer(4 * n)
        dst = gpu.DeviceBuffer(16 * n)
        gpu.compress_batch([gpu.make_job(8, n, dcol, dcode, None, derr, dst, (1, 1, 1), True, True)])
        runs.append((dcode.download(np.uint8, (n, 8)), derr.download(np.float32, n), dst.download(np.int32, (n, 4))))
        for b in (dcode, derr, dst):
            b.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1].view(np.uint32), runs[1][1].view(np.uint32))
    codes, errs, stats = runs[0]
    sample = np.random.default_rng(8).choice(n, size=8192, replace=False)
    want = _oracle_run(oracle, 8, colors[sample], None, (1, 1, 1), 1, 1)
    assert_same(codes[sample], errs[sample], want[0], want[1], "4096x4096 L8 sample")
    assert np.all(np.isfinite(errs)) and np.all(errs >= 0)
    flat = cat == synth.CATEGORIES.index("flat")
    assert np.all(stats[flat, 0] == 1)          # flat blocks reduce to one colour
    assert np.all((stats[:, 0] >= 1) & (stats[:, 0] <= 16))


def test_fast_reciprocal_matches_ieee_division_for_every_float(gpu):
    # the cluster fits take 1/det from the hardware estimate plus one Newton
    # step; all 2^32 inputs must give the IEEE-754 quotient bit for bit
    assert gpu.check_reciprocal() == 0


def _adversarial_blocks(seed, n):
    """Blocks built to stress the four-cluster pruning: many distinct colours
    (n >= 9 after reduction) on near-degenerate lines, weights spread over
    seven decades, colours outside [0, 1], tiny perturbations."""
    rng = np.random.default_rng(seed)
    out = np.empty((n, 16, 4), np.float32)
    wts = np.empty((n, 16), np.float32)
    for b in range(n):
        kind = b % 5
        if kind == 0:    # points on a line with tiny off-line jitter
            a, d = rng.random(3), rng.random(3) - 0.5
            tt = np.sort(rng.random(16))
            c = a + tt[:, None] * d + rng.normal(0, 1e-4, (16, 3))
        elif kind == 1:  # two tight clusters plus a few stragglers
            p, q = rng.random(3), rng.random(3)
            c = np.where(rng.random((16, 1)) < 0.5, p, q) + rng.normal(0, 3e-3, (16, 3))
        elif kind == 2:  # out-of-range colours
            c = rng.uniform(-2, 3, (16, 3))
        elif kind == 3:  # a gradient with one texel far away
            a, d = rng.random(3), rng.random(3) - 0.5
            c = a + np.linspace(0, 1, 16)[:, None] * d
            c[rng.integers(16)] = rng.random(3) * 4 - 1
        else:            # white noise
            c = rng.random((16, 3))
        out[b, :, :3] = c
        out[b, :, 3] = 1
        w = 10.0 ** rng.uniform(-7, 0, 16) if b % 3 == 0 else rng.random(16) + 0.01
        wts[b] = w
    return out, wts


def test_nonfinite_blocks_every_level(gpu, oracle):
    # NaN / infinite colours and weights, huge and denormal colours (the
    # oracle follows the reference build on these, see test_oracle.py)
    colors, weights = synth.make_nonfinite_blocks(41, 2000)
    prior = np.random.default_rng(42).integers(0, 256, size=(len(colors), 8), dtype=np.uint8)
    with np.errstate(all="ignore"):
        for level in range(10):
            for tcm in (False, True):
                _say(f"non-finite L{level} tcm={tcm}")
                want = _oracle_run(oracle, level, colors, weights, (1, 1, 1), tcm, tcm, prior)
                got = gpu.compress_many(level, colors, weights, (1, 1, 1), tcm, tcm, prior=prior)
                assert_same(got[0], got[1], want[0], want[1], f"non-finite L{level} tcm={tcm}")


def test_levels_0_1_register_instance_edge_blocks(gpu, oracle):
    # level 0-1 launches with unit channel weights and no texel weights run the
    # instance that reads texels into registers: edge-case and non-finite
    # colours, float and RGBA8, with random prior output bytes, against the
    # oracle (the weighted tests above reach the LDS-staged instance)
    ec, _ = synth.make_edge_blocks(61, 1500)
    nc, _ = synth.make_nonfinite_blocks(62, 500)
    colors = np.concatenate([ec, nc])
    prior = np.random.default_rng(63).integers(0, 256, size=(len(colors), 8), dtype=np.uint8)
    c8 = np.clip(np.nan_to_num(colors[:len(ec)], nan=0.0, posinf=1.0, neginf=0.0) * 255, 0, 255).round().astype(np.uint8)
    cf8 = c8.astype(np.float32) / np.float32(255)
    with np.errstate(all="ignore"):
        for level in (0, 1):
            for tcm in (False, True):
                _say(f"register instance L{level} tcm={tcm}")
                want = _oracle_run(oracle, level, colors, None, (1, 1, 1), tcm, tcm, prior)
                got = gpu.compress_many(level, colors, None, (1, 1, 1), tcm, tcm, prior=prior)
                assert_same(got[0], got[1], want[0], want[1], f"float L{level} tcm={tcm}")
                want8 = _oracle_run(oracle, level, cf8, None, (1, 1, 1), tcm, tcm, prior[:len(ec)])
                got8 = gpu.compress_many(level, c8, None, (1, 1, 1), tcm, tcm, prior=prior[:len(ec)])
                assert_same(got8[0], got8[1], want8[0], want8[1], f"rgba8 L{level} tcm={tcm}")


@pytest.mark.parametrize("prune", [0, 9])
@pytest.mark.parametrize("cw", [(1, 1, 1), (0.3, 0.59, 0.11), (0, 1, 0), (5, 0.001, 1)])
def test_pruned_search_adversarial_blocks(gpu, oracle, cw, prune):
    # prune 0: the default schedule (blocks of 13+ colours pruned); 9: every
    # block of 9+ colours through the bound pass
    blocks, weights = _adversarial_blocks(31, 6000)
    t = gpu.Tuning()
    for lv in range(10):
        t.tile_blocks[lv], t.schedule[lv], t.ns_per_block[lv], t.prune_min_colors[lv] = 64, 1, 1.0, prune
    try:
        if prune:
            gpu.set_tuning(t)
        for level, tcm in ((5, False), (8, True), (8, False)):
            want = oracle.compress_many(level, blocks, weights, cw, tcm, tcm, nthreads=8)
            got = gpu.compress_many(level, blocks, weights, cw, tcm, tcm)
            bad = np.nonzero(np.any(got[0] != want[0], axis=1))[0]
            assert len(bad) == 0, f"level {level} cw {cw}: {len(bad)} blocks differ, first {bad[:5]}"
            assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
    finally:
        gpu.set_tuning(None)


def test_refinement_group_size_does_not_change_results(gpu, oracle):
    # a 4096x2048 level-9 image hands ~0.5 M walks to the refinement kernel,
    # which then walks one block per lane; a few of its blocks alone are
    # walked by groups of lanes evaluating steps speculatively.  Both must give
    # the sequential walk's codes and errors.
    blocks, _ = synth.make_image_blocks(17, 4096, 2048, synth.MIX_DEFAULT)
    big = gpu.compress_many(9, blocks, None, (1, 1, 1), True, True)
    rng = np.random.default_rng(3)
    pick = np.sort(rng.choice(len(blocks), 2048, replace=False))
    small = gpu.compress_many(9, blocks[pick], None, (1, 1, 1), True, True)
    assert np.array_equal(big[0][pick], small[0])
    assert np.array_equal(big[1][pick].view(np.uint32), small[1].view(np.uint32))
    want = oracle.compress_many(9, blocks[pick], None, (1, 1, 1), True, True, nthreads=8)
    assert np.array_equal(small[0], want[0])
    assert np.array_equal(small[1].view(np.uint32), want[1].view(np.uint32))


def test_sweep_decoders_levels_options(gpu, oracle):
    # every decoder x level x three-colour setting, random channel weights,
    # on one mixed set: image content (some blocks weighted), edge cases and
    # non-finite inputs, with random prior output bytes
    rng = np.random.default_rng(101)
    c1, _ = synth.make_image_blocks(101, 256, 256, synth.MIX_DEFAULT)
    ec, ew = synth.make_edge_blocks(103, 1500)
    nc, nw = synth.make_nonfinite_blocks(104, 500)
    colors = np.concatenate([c1, ec, nc])
    w1 = np.where(rng.random((len(c1), 1)) < 0.3, rng.random((len(c1), 16)), 1.0).astype(np.float32)
    weights = np.concatenate([w1, ew, nw])
    prior = rng.integers(0, 256, size=(len(colors), 8), dtype=np.uint8)
    try:
        for dec in range(4):
            gpu.init(dec)
            oracle.init(dec)
            _say(f"sweep decoder {dec}")
            for level in range(10):
                for tcm, tcb in ((0, 0), (1, 0), (1, 1)):
                    cw = tuple(float(x) for x in rng.random(3) * 2) if rng.random() < 0.5 else (1.0, 1.0, 1.0)
                    with np.errstate(all="ignore"):
                        want = _oracle_run(oracle, level, colors, weights, cw, tcm, tcb, prior)
                        got = gpu.compress_many(level, colors, weights, cw, tcm, tcb, prior=prior)
                        assert_same(got[0], got[1], want[0], want[1],
                                    f"decoder {dec} L{level} tcm={tcm} tcb={tcb} cw={cw}")
    finally:
        gpu.init(0)
        oracle.init(0)
