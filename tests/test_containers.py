"""CPU tests of the DDS / KTX writers (host code of the library): headers
against the reference's layouts (icbc_test.cpp:155-211, 233-279), mip-level
fields, the -cmp style comparison (icbc_test.cpp:213-231), and the committed
reference-encoded DDS fixture against the oracle."""
import os
import struct

import numpy as np
import pytest

import icbc_amd as g
from icbc_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DDS = os.path.join(HERE, "golden", "c1_512x512_L8.dds")


def _dims(w, h, k):
    return max(1, w >> k), max(1, h >> k)


def _blocks(w, h):
    return ((w + 3) // 4) * ((h + 3) // 4)


def _levels(w, h, n, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, (_blocks(*_dims(w, h, k)), 8), dtype=np.uint8) for k in range(n)]


def _parse_dds(b):
    f = struct.unpack("<31I", b[:124])
    return {"magic": b[:4], "size": f[1], "flags": f[2], "height": f[3], "width": f[4], "linear": f[5],
            "depth": f[6], "mips": f[7], "reserved": f[8:19], "pf_size": f[19], "pf_flags": f[20],
            "pf_fourcc": b[84:88], "pf_rest": f[22:27], "caps": f[27:31],
            "notused": struct.unpack("<I", b[124:128])[0]}


def test_dds_single_level_matches_reference_layout():
    w, h = 37, 21
    codes = _levels(w, h, 1)
    b = g.container_bytes(g.CONTAINER_DDS, w, h, codes)
    d = _parse_dds(b)
    assert d["magic"] == b"DDS " and d["size"] == 124
    assert d["flags"] == 0x1 | 0x1000 | 0x4 | 0x2 | 0x80000
    assert (d["width"], d["height"], d["depth"], d["mips"]) == (w, h, 0, 1)
    assert d["linear"] == 8 * ((w + 3) // 4) * ((h + 3) // 4)
    assert d["pf_size"] == 32 and d["pf_flags"] == 0x4 and d["pf_fourcc"] == b"DXT1" and d["pf_rest"] == (0,) * 5
    assert d["caps"] == (0x1000, 0, 0, 0) and d["notused"] == 0 and d["reserved"] == (0,) * 11
    assert b[128:] == codes[0].tobytes()


def test_dds_mip_levels():
    w, h, n = 300, 70, 7
    codes = _levels(w, h, n, 1)
    b = g.container_bytes(g.CONTAINER_DDS, w, h, codes)
    d = _parse_dds(b)
    assert d["mips"] == n and d["flags"] & 0x20000
    assert d["caps"][0] == 0x1000 | 0x8 | 0x400000
    assert b[128:] == b"".join(c.tobytes() for c in codes)
    assert len(b) == g.lib().icbc_container_size(g.CONTAINER_DDS, w, h, n)


def test_ktx_levels_and_image_sizes():
    w, h, n = 64, 20, 5
    codes = _levels(w, h, n, 2)
    b = g.container_bytes(g.CONTAINER_KTX, w, h, codes)
    assert b[:12] == bytes([0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31, 0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A])
    f = struct.unpack("<13I", b[12:64])
    assert f[0] == 0x04030201 and f[1] == 0 and f[2] == 1 and f[3] == 0
    assert f[4] == 0x83F0 and f[5] == 0x1908 and (f[6], f[7], f[8]) == (w, h, 0)
    assert (f[9], f[10], f[11], f[12]) == (0, 1, n, 0)
    at = 64
    for k, c in enumerate(codes):
        size = struct.unpack("<I", b[at:at + 4])[0]
        lw, lh = _dims(w, h, k)
        assert size == 8 * _blocks(lw, lh) == c.nbytes  # the KTX 1.1 imageSize of level k
        assert b[at + 4:at + 4 + size] == c.tobytes()
        at += 4 + size
    assert at == len(b)


def test_wrong_block_count_is_rejected():
    with pytest.raises(ValueError):
        g.container_bytes(g.CONTAINER_DDS, 16, 16, [np.zeros((15, 8), np.uint8)])
    assert g.lib().icbc_container_size(7, 16, 16, 1) == 0
    assert g.lib().icbc_container_encode(0, 0, 16, 1, None, None, 0) == g.ERR_INVALID_ARGUMENT


def test_write_and_compare_dds(tmp_path):
    w, h = 48, 40
    codes = _levels(w, h, 3, 3)
    p = tmp_path / "t.dds"
    g.write_container(str(p), g.CONTAINER_DDS, w, h, codes)
    assert p.read_bytes() == g.container_bytes(g.CONTAINER_DDS, w, h, codes)
    assert g.compare_dds(str(p), w, h, codes[0])
    other = codes[0].copy()
    other[5, 3] ^= 1
    assert not g.compare_dds(str(p), w, h, other)
    k = tmp_path / "t.ktx"
    g.write_container(str(k), g.CONTAINER_KTX, w, h, codes)
    assert k.read_bytes() == g.container_bytes(g.CONTAINER_KTX, w, h, codes)


def test_golden_dds_fixture_matches_oracle_and_writer(oracle):
    data = open(GOLDEN_DDS, "rb").read()
    colors, _ = synth.make_image_blocks(1, 512, 512, synth.MIX_PLAIN)
    codes, _, _, _ = oracle.compress_many(8, colors, None, (1, 1, 1), False, False, nthreads=8)
    assert data == g.container_bytes(g.CONTAINER_DDS, 512, 512, codes)
    assert g.compare_dds(GOLDEN_DDS, 512, 512, codes)
