"""GPU tests of the device texture pre-pass: block gather (both edge modes,
RGBA8 and float, padded pitch, alpha weights) and box-filtered mip chains
against the CPU restatement (oracle/prepass.py), bit for bit; and image jobs
encoding exactly like the same texels handed over as float blocks, through the
batch, plain and plan (graph) paths."""
import numpy as np
import pytest

from icbc_amd import synth
from oracle import prepass

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 7), (6, 1), (4, 4), (5, 6), (13, 9), (61, 35), (256, 130)]


def _image(seed, w, h, u8):
    rng = np.random.default_rng(seed)
    if u8:
        return rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    return rng.random((h, w, 4), dtype=np.float32)


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("u8", [False, True])
@pytest.mark.parametrize("edge", [0, 1])
def test_gather_matches_restatement(gpu, w, h, u8, edge):
    img = _image(w * 31 + h, w, h, u8)
    want = prepass.gather_fast(img, edge)
    di = gpu.DeviceImage(img, edge=edge)
    try:
        assert di.block_count == len(want)
        if len(want) == 0:
            return
        got, alpha = gpu.image_to_blocks(di.desc, alpha_weights=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.array_equal(alpha, want[:, :, 3])
    finally:
        di.close()


def test_gather_with_padded_pitch(gpu):
    img = _image(3, 23, 11, False)
    di = gpu.DeviceImage(img, pitch=23 * 16 + 48)
    try:
        got = gpu.image_to_blocks(di.desc)
    finally:
        di.close()
    assert np.array_equal(got, prepass.gather_fast(img))


@pytest.mark.parametrize("w,h", [(13, 9), (1, 37), (64, 1), (517, 300), (1024, 768)])
@pytest.mark.parametrize("u8", [False, True])
def test_mip_chain_matches_restatement(gpu, w, h, u8):
    img = _image(w + 7 * h, w, h, u8)
    want = prepass.mip_chain(img)
    di = gpu.DeviceImage(img)
    try:
        chain = gpu.MipChain(di.desc)
        try:
            assert len(chain.levels) == len(want)
            for k, lv in enumerate(want):
                assert (chain.levels[k].width, chain.levels[k].height) == (lv.shape[1], lv.shape[0])
                assert np.array_equal(chain.download(k).view(np.uint32), lv.view(np.uint32)), f"level {k + 1}"
        finally:
            chain.close()
    finally:
        di.close()


def _encode_blocks(gpu, level, blocks, weights, cw=(1, 1, 1), tcm=True, tcb=True):
    codes, errs, _ = gpu.compress_many(level, blocks, weights, cw, tcm, tcb)
    return codes, errs


def _encode_image_job(gpu, level, desc, n, cw=(1, 1, 1), tcm=True, tcb=True, alpha_weights=False,
                      path="batch"):
    dcode = gpu.DeviceBuffer(8 * n)
    derr = gpu.DeviceBuffer(4 * n)
    try:
        job = gpu.make_job(level, n, None, dcode, None, derr, None, cw, tcm, tcb, image=desc,
                           weights_from_alpha=alpha_weights)
        s = gpu.Stream()
        if path == "plan":
            plan = gpu.Plan([job], schedule=gpu.SCHEDULE_PERSISTENT, use_graph=True, stream=s)
            outs = []
            for _ in range(3):
                dcode.fill(0)
                plan.execute(s)
                s.synchronize()
                outs.append((dcode.download(np.uint8, (n, 8)), derr.download(np.float32, (n,))))
            plan.close()
            for o in outs[1:]:
                assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])
            return outs[0]
        gpu.compress_batch([job], stream=s, plain=(path == "plain"))
        s.synchronize()
        return dcode.download(np.uint8, (n, 8)), derr.download(np.float32, (n,))
    finally:
        dcode.close()
        derr.close()


@pytest.mark.parametrize("path", ["batch", "plain", "plan"])
@pytest.mark.parametrize("u8,edge", [(False, 0), (True, 0), (True, 1)])
def test_image_job_encodes_like_float_blocks(gpu, path, u8, edge):
    w, h = 203, 141
    if u8:
        img = (synth.make_image(11, w, h, synth.MIX_DEFAULT) * 255 + 0.5).astype(np.uint8)
    else:
        img = synth.make_image(11, w, h, synth.MIX_DEFAULT)
    blocks = prepass.gather_fast(img, edge)
    want = _encode_blocks(gpu, 8, blocks, None)
    di = gpu.DeviceImage(img, edge=edge)
    try:
        got = _encode_image_job(gpu, 8, di.desc, len(blocks), path=path)
    finally:
        di.close()
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


def test_alpha_weighted_image_job(gpu, oracle):
    # weighted mode: 30% of texels at alpha 0, the rest weight = alpha, L9, cw (0.3, 0.59, 0.11)
    w, h = 96, 80
    rgb = synth.make_image(4, w, h, synth.MIX_DEFAULT)
    img = rgb.copy()
    img[..., 3] = synth.make_weighted_alpha(4, w, h)
    blocks = prepass.gather_fast(img)
    weights = np.ascontiguousarray(blocks[:, :, 3])
    cw = (0.3, 0.59, 0.11)
    want = oracle.compress_many(9, blocks, weights, cw, True, True, nthreads=8)
    di = gpu.DeviceImage(img)
    try:
        got = _encode_image_job(gpu, 9, di.desc, len(blocks), cw=cw, alpha_weights=True)
    finally:
        di.close()
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


def test_mip_levels_encode_like_host_levels(gpu):
    img = synth.make_image(21, 300, 170, synth.MIX_DEFAULT)
    levels = prepass.mip_chain(img)
    di = gpu.DeviceImage(img)
    chain = gpu.MipChain(di.desc)
    try:
        for k, lv in enumerate(levels):
            blocks = prepass.gather_fast(lv)
            want = _encode_blocks(gpu, 5, blocks, None)
            got = _encode_image_job(gpu, 5, chain.levels[k], len(blocks))
            assert np.array_equal(got[0], want[0]), f"level {k + 1}"
    finally:
        chain.close()
        di.close()


def test_image_job_argument_checks(gpu):
    img = _image(1, 16, 16, True)
    di = gpu.DeviceImage(img)
    dcode = gpu.DeviceBuffer(8 * 16)
    try:
        bad_count = gpu.make_job(8, 15, None, dcode, image=di.desc)
        with pytest.raises(gpu.IcbcError) as e:
            gpu.compress_batch([bad_count])
        assert e.value.code == gpu.ERR_INVALID_ARGUMENT and "blocks" in str(e.value)
        alpha_without_image = gpu.make_job(8, 16, dcode, dcode, weights_from_alpha=True)
        with pytest.raises(gpu.IcbcError):
            gpu.compress_batch([alpha_without_image])
        bad = gpu.Image(di.desc.data, 7, 16, 16, 0, 64)  # unknown format
        with pytest.raises(gpu.IcbcError):
            gpu.compress_batch([gpu.make_job(8, 16, None, dcode, image=bad)])
    finally:
        dcode.close()
        di.close()


def test_golden_dds_regression(gpu, tmp_path):
    # the reference harness's -cmp mode (icbc_test.cpp:213-231): encode the
    # 512x512 plain-content image at level 8 and compare the block data with
    # the reference-encoded DDS fixture; the whole written file must match too
    import os
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_512x512_L8.dds")
    img = synth.blocks_to_image(synth.make_image_blocks(1, 512, 512, synth.MIX_PLAIN)[0], 512, 512)
    di = gpu.DeviceImage(img)
    try:
        codes, _ = _encode_image_job(gpu, 8, di.desc, di.block_count, tcm=False, tcb=False)
    finally:
        di.close()
    assert gpu.compare_dds(golden, 512, 512, codes)
    out = tmp_path / "c1.dds"
    gpu.write_container(str(out), gpu.CONTAINER_DDS, 512, 512, codes)
    assert out.read_bytes() == open(golden, "rb").read()


def test_plan_with_image_rgba8_and_float_jobs(gpu, oracle):
    # one plan: an 8-bit device image gathered on every execute, 8-bit blocks
    # read by the kernels directly, and float blocks, at levels 8 and 9
    w, h = 150, 90
    img = (synth.make_image(31, w, h, synth.MIX_DEFAULT) * 255 + 0.5).astype(np.uint8)
    img_blocks = prepass.gather_fast(img)
    c8 = np.clip(np.round(synth.make_image_blocks(32, 256, 64, synth.MIX_DEFAULT)[0] * 255), 0, 255).astype(np.uint8)
    cf = synth.make_image_blocks(33, 192, 64, synth.MIX_DEFAULT)[0]
    di = gpu.DeviceImage(img)
    specs = [(9, len(img_blocks)), (8, len(c8)), (9, len(cf))]
    bufs = [gpu.DeviceBuffer.from_array(c8), gpu.DeviceBuffer.from_array(cf)]
    outs = [(gpu.DeviceBuffer(8 * n), gpu.DeviceBuffer(4 * n)) for _, n in specs]
    try:
        for dcode, _ in outs:
            dcode.fill(0)
        jobs = [gpu.make_job(9, len(img_blocks), None, outs[0][0], None, outs[0][1], None, (1, 1, 1), True, True,
                             image=di.desc),
                gpu.make_job(8, len(c8), bufs[0], outs[1][0], None, outs[1][1], None, (1, 1, 1), True, True,
                             rgba8=True),
                gpu.make_job(9, len(cf), bufs[1], outs[2][0], None, outs[2][1], None, (1, 1, 1), True, True)]
        plan = gpu.Plan(jobs)
        stream = gpu.Stream()
        for _ in range(2):
            plan.execute(stream)
            stream.synchronize()
        plan.close()
        stream.close()
        sources = [img_blocks, c8.astype(np.float32) / np.float32(255), cf]
        for k, ((level, n), src) in enumerate(zip(specs, sources)):
            want = oracle.compress_many(level, src, None, (1, 1, 1), True, True, nthreads=8,
                                        prior=np.zeros((n, 8), np.uint8))
            codes = outs[k][0].download(np.uint8, (n, 8))
            errs = outs[k][1].download(np.float32, n)
            assert np.array_equal(codes, want[0]), f"job {k}"
            assert np.array_equal(errs.view(np.uint32), want[1].view(np.uint32)), f"job {k}"
    finally:
        di.close()
        for b in bufs:
            b.close()
        for a, b in outs:
            a.close()
            b.close()
