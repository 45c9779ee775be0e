"""CPU pins of the pre-pass restatement (oracle/prepass.py) and the host-side
helpers of the library's pre-pass API (no device needed)."""
import ctypes as C

import numpy as np
import pytest

import icbc_amd
from oracle import prepass

SIZES = [(1, 1), (1, 7), (6, 1), (4, 4), (5, 6), (13, 9), (17, 32)]


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("edge", [prepass.EDGE_REPLICATE, prepass.EDGE_ROUND_DOWN])
def test_gather_loops_match_vectorised(w, h, edge):
    rng = np.random.default_rng(w * 100 + h)
    img = rng.random((h, w, 4), dtype=np.float32)
    assert np.array_equal(prepass.gather(img, edge), prepass.gather_fast(img, edge))


def test_gather_texel_order_follows_harness():
    # texel j = 4*yy + xx of block (bx, by), blocks row-major (icbc_test.cpp:361-372);
    # every texel holds its own coordinates
    w, h = 10, 7
    yy, xx = np.mgrid[0:h, 0:w]
    img = np.stack([xx, yy, xx * 0, xx * 0 + 1], -1).astype(np.float32)
    rd = prepass.gather(img, prepass.EDGE_ROUND_DOWN)
    assert rd.shape == ((w // 4) * (h // 4), 16, 4)
    b = 1  # block (bx=1, by=0); texel yy=2, xx=3 is image (x=7, y=2)
    assert rd[b, 4 * 2 + 3, 0] == 7 and rd[b, 4 * 2 + 3, 1] == 2
    rep = prepass.gather(img, prepass.EDGE_REPLICATE)
    assert rep.shape == (3 * 2, 16, 4)
    last = rep[-1]  # block (2, 1): columns 8, 9, 9, 9 and rows 4, 5, 6, 6
    assert list(last[:4, 0]) == [8, 9, 9, 9] and list(last[::4, 1]) == [4, 5, 6, 6]


def test_rgba8_conversion_is_correctly_rounded_quotient():
    v = np.arange(256, dtype=np.uint8).reshape(1, 64, 4)
    got = prepass.to_float(v).reshape(-1)
    from fractions import Fraction
    for c in range(256):
        exact = Fraction(c, 255)
        f = np.float32(c / 255.0)
        # nearest float32: no representable float32 is strictly closer
        lo, hi = np.nextafter(f, np.float32(0)), np.nextafter(f, np.float32(2))
        assert abs(Fraction(float(f)) - exact) <= min(abs(Fraction(float(lo)) - exact), abs(Fraction(float(hi)) - exact))
        assert got[c] == f


@pytest.mark.parametrize("w,h", SIZES + [(33, 2), (2, 33)])
def test_downsample_loops_match_vectorised(w, h):
    rng = np.random.default_rng(w * 7 + h)
    img = rng.random((h, w, 4), dtype=np.float32)
    assert np.array_equal(prepass.downsample(img), prepass.downsample_fast(img))


def test_box_filter_is_the_exact_mean_on_integer_texels():
    # integer texels: the float sums and the quarter are exact, so each output
    # is the true mean of its 2x2 footprint (a dropped or repeated term fails)
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (9, 11, 4)).astype(np.float32)
    out = prepass.downsample(img)
    ref = img[:8, :10].reshape(4, 2, 5, 2, 4).astype(np.float64).mean(axis=(1, 3))
    assert np.array_equal(out, ref.astype(np.float32))
    row = img[:1]
    assert np.array_equal(prepass.downsample(row)[0], (row[0, 0:10:2] + row[0, 1:10:2]) / 2)


def test_constant_image_keeps_its_value_down_the_chain():
    img = np.full((37, 53, 4), 0.3, np.float32)
    for lv in prepass.mip_chain(img):
        assert np.all(lv == np.float32(0.3))


@pytest.mark.parametrize("w,h", [(13, 9), (4096, 64), (4, 4), (1, 100), (517, 300)])
def test_library_chain_size_matches_restatement(w, h):
    levels = prepass.mip_chain(np.zeros((h, w, 4), np.float32))
    nbytes, n = icbc_amd.mip_chain_bytes(w, h)
    assert n == len(levels)
    assert nbytes == sum(l.shape[0] * l.shape[1] * 16 for l in levels)


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("edge", [0, 1])
def test_library_block_count_matches_restatement(w, h, edge):
    desc = icbc_amd.Image(16, icbc_amd.IMAGE_RGBA32F, w, h, edge, w * 16)  # not dereferenced
    assert icbc_amd.image_block_count(desc) == len(prepass.gather_fast(np.zeros((h, w, 4), np.float32), edge))
    bad = icbc_amd.Image(16, icbc_amd.IMAGE_RGBA32F, w, h, edge, w * 16 - 16)  # pitch too small
    assert icbc_amd.image_block_count(bad) == 0
