"""Writes tests/golden/bc1_golden.npz from the reference encoder.

The reference (castano/icbc v1.05) has no tests or fixtures of its own, so the
vectors are produced here by the reference header itself, compiled at
ICBC_SIMD=0 by oracle/Makefile into oracle/_ref/libicbc_ref_scalar.so, and
called through its public API (icbc.h:19, 38, 43, 48).  Run in a checkout
where /root/reference exists:

    make -C oracle && python tests/golden/gen_golden.py

Contents (all inputs are seeded synthetic blocks, block-major 16 x RGBA):
  colors, weights           the input blocks (edge cases, image content, weighted)
  prior                     the output buffer before each call (to check that
                            invalid levels leave it untouched)
  cfg                       (decoder, level, three_color_mode, three_color_black, cw_index)
  cw                        channel-weight sets indexed by cw_index
  codes, errors             reference outputs per config
  tables_*, midpoints*      the reference's init tables (decoder 0)
  dec_codes, dec_rgba       decode_bc1 outputs for every decoder
  eval_rgba, eval_err       evaluate_bc1_error outputs for every decoder
  bc3_blocks, bc3_rgba      decode_bc3 outputs for every decoder (random 16-byte blocks)
  bc3_eval_rgba, bc3_eval_err
                            evaluate_bc3_error outputs, [alpha_blend][decoder]

The input generator is frozen in this script (it does not use the package's
synthetic-content module, which may change), so running the script again
writes the same inputs; tests/test_oracle.py checks that it does, and that the
committed outputs equal the reference build's on those inputs.

and tests/golden/c1_512x512_L8.dds: the 512x512 plain-content image (seed 1,
50/30/20 gradient / value noise / flat) encoded by the reference at level 8
with three_color_mode off, weights 1, channel weights 1, in the DDS layout of
the reference's output_dxt_dds (icbc_test.cpp:155-211; its uninitialised
reserved words written as zeros) -- the file its -cmp regression mode compares
against (icbc_test.cpp:213-231).
"""
import os
import struct
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bc1_golden.npz")
DDS_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_512x512_L8.dds")


# ---------------------------------------------------------------------------
# frozen input generator (block-major 16 x RGBA float32, texel = 4*y + x)


def _value_noise(rng, h, w, cell=8):
    gh, gw = h // cell + 2, w // cell + 2
    lattice = rng.random((gh, gw, 3), dtype=np.float32)
    ys = (np.arange(h, dtype=np.float32) + 0.5) / cell
    xs = (np.arange(w, dtype=np.float32) + 0.5) / cell
    y0 = np.floor(ys).astype(np.int64)
    x0 = np.floor(xs).astype(np.int64)
    fy, fx = ys - y0, xs - x0
    sy = (fy * fy * (3 - 2 * fy))[:, None, None]
    sx = (fx * fx * (3 - 2 * fx))[None, :, None]
    a, b = lattice[y0][:, x0], lattice[y0][:, x0 + 1]
    c, d = lattice[y0 + 1][:, x0], lattice[y0 + 1][:, x0 + 1]
    top = a + (b - a) * sx
    bot = c + (d - c) * sx
    return (top + (bot - top) * sy).astype(np.float32)


def _gradients(rng, n):
    e0 = rng.random((n, 3), dtype=np.float32)
    e1 = rng.random((n, 3), dtype=np.float32)
    ang = rng.random(n, dtype=np.float32) * np.float32(2 * np.pi)
    yy, xx = np.mgrid[0:4, 0:4].astype(np.float32)
    proj = (np.cos(ang)[:, None] * (xx.reshape(1, 16) - 1.5) + np.sin(ang)[:, None] * (yy.reshape(1, 16) - 1.5))
    lo, hi = proj.min(axis=1, keepdims=True), proj.max(axis=1, keepdims=True)
    t = (proj - lo) / np.maximum(hi - lo, 1e-6)
    return (e0[:, None, :] + (e1 - e0)[:, None, :] * t[:, :, None]).astype(np.float32)


def image_blocks(seed, width, height, mix):
    """Blocks of a width x height image; mix = {category: probability} over
    gradient, value_noise, white_noise, near_black (1-6 texels below 0.02) and
    flat."""
    rng = np.random.default_rng(seed)
    nb = (width // 4) * (height // 4)
    names = list(mix)
    p = np.array([mix[k] for k in names], np.float64)
    cat = np.array(names)[rng.choice(len(names), size=nb, p=p / p.sum())]
    rgb = np.empty((nb, 16, 3), np.float32)
    sel = cat == "gradient"
    rgb[sel] = _gradients(rng, int(sel.sum()))
    sel = cat == "value_noise"
    if sel.any():
        f = _value_noise(rng, height, width)
        vb = f.reshape(height // 4, 4, width // 4, 4, 3).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 3)[sel]
        rgb[sel] = np.clip(vb + rng.uniform(-0.05, 0.05, size=vb.shape).astype(np.float32), 0.0, 1.0)
    sel = cat == "white_noise"
    rgb[sel] = rng.random((int(sel.sum()), 16, 3), dtype=np.float32)
    sel = cat == "flat"
    rgb[sel] = np.repeat(rng.random((int(sel.sum()), 1, 3), dtype=np.float32), 16, axis=1)
    sel = np.nonzero(cat == "near_black")[0]
    if len(sel):
        base = rng.random((len(sel), 16, 3), dtype=np.float32)
        k = rng.integers(1, 7, size=len(sel))
        dark = np.argsort(np.argsort(rng.random((len(sel), 16)), axis=1), axis=1) < k[:, None]
        base[dark] = rng.random((int(dark.sum()), 3), dtype=np.float32) * np.float32(0.02)
        rgb[sel] = base
    colors = np.empty((nb, 16, 4), np.float32)
    colors[..., :3] = rgb
    colors[..., 3] = 1.0
    return colors


def edge_blocks(seed, n):
    """Corner cases: zero / negative weights, two colours, colours within the
    merge threshold, saturated and out-of-range values, near-black and
    all-dark blocks, 8-bit ties, grey ramps, weighted and single-texel blocks,
    three tight clusters."""
    rng = np.random.default_rng(seed)
    colors = np.empty((n, 16, 4), np.float32)
    weights = np.ones((n, 16), np.float32)
    for b, k in enumerate(rng.integers(0, 14, size=n)):
        c = rng.random((16, 3), dtype=np.float32)
        w = np.ones(16, np.float32)
        if k == 0:
            w[:] = 0
        elif k == 1:
            w = rng.random(16, dtype=np.float32)
            w[rng.random(16) < 0.4] = 0
            w[rng.random(16) < 0.1] = -1
        elif k == 2:
            a, d = rng.random((2, 3), dtype=np.float32)
            m = rng.random(16) < 0.5
            c[m], c[~m] = a, d
        elif k == 3:
            c = rng.random(3, dtype=np.float32) + rng.uniform(-1 / 200, 1 / 200, size=(16, 3)).astype(np.float32)
        elif k == 4:
            c = (rng.random((16, 3)) < 0.5).astype(np.float32)
        elif k == 5:
            c[rng.random(16) < 0.4] *= np.float32(0.1)
        elif k == 6:
            c *= np.float32(0.11)
        elif k == 7:
            c = rng.uniform(-0.25, 1.25, size=(16, 3)).astype(np.float32)
        elif k == 8:
            c = np.round(c * 255) / np.float32(255)
        elif k == 9:
            c = (rng.integers(0, 4, size=(16, 3)) * 85).astype(np.float32) / 255
        elif k == 10:
            c = np.repeat(np.linspace(0, 1, 16, dtype=np.float32)[:, None], 3, axis=1)
        elif k == 11:
            w = rng.random(16, dtype=np.float32) * 4
        elif k == 12:
            w[:] = 0
            w[rng.integers(0, 16)] = 1
        else:
            cen = rng.random((3, 3), dtype=np.float32)
            c = cen[rng.integers(0, 3, size=16)] + rng.normal(0, 0.01, size=(16, 3)).astype(np.float32)
        colors[b, :, :3] = c
        colors[b, :, 3] = 1.0
        weights[b] = w
    return colors.astype(np.float32), weights.astype(np.float32)


def alpha_field(seed, width, height, zero_fraction=0.3):
    """Thresholded value noise: zero_fraction of texels at 0, the rest in (0, 1]."""
    rng = np.random.default_rng(seed + 7919)
    f = _value_noise(rng, height, width)[..., 0]
    cut = np.quantile(f, zero_fraction)
    a = np.clip(np.where(f <= cut, 0.0, (f - cut) / max(1e-6, float(f.max() - cut))), 0.0, 1.0).astype(np.float32)
    a[(a > 0) & (a < 1e-3)] = 1e-3
    return a


def bc3_inputs(seed, n):
    """Random 16-byte BC3 blocks (endpoint orders c0 > c1, c0 < c1 and equal
    all present) and RGBA8 texels with random alpha."""
    rng = np.random.default_rng(seed)
    blocks = rng.integers(0, 256, size=(n, 16), dtype=np.uint8)
    cols = blocks[:, 8:12].copy().view(np.uint16)
    cols[n // 4: n // 2, 1] = cols[n // 4: n // 2, 0]     # equal endpoints
    blocks[:, 8:12] = cols.view(np.uint8)
    rgba = rng.integers(0, 256, size=(n, 64), dtype=np.uint8)
    return blocks, rgba


PLAIN = {"gradient": 0.5, "value_noise": 0.3, "flat": 0.2}
DEFAULT = {"gradient": 0.4, "white_noise": 0.3, "near_black": 0.2, "flat": 0.1}


def dds_header(width, height):
    """DDS_HEADER of icbc_test.cpp:157-196 for one DXT1 level."""
    fourcc = lambda s: struct.unpack("<I", s)[0]  # noqa: E731
    flags = 0x1 | 0x1000 | 0x4 | 0x2 | 0x80000  # CAPS | PIXELFORMAT | WIDTH | HEIGHT | LINEARSIZE
    linear = 8 * ((width + 3) // 4) * ((height + 3) // 4)
    head = struct.pack("<7I", fourcc(b"DDS "), 124, flags, height, width, linear, 0)
    head += struct.pack("<I", 1) + b"\0" * 44                     # mipmapcount, reserved[11]
    head += struct.pack("<8I", 32, 0x4, fourcc(b"DXT1"), 0, 0, 0, 0, 0)
    head += struct.pack("<5I", 0x1000, 0, 0, 0, 0)                  # caps1..4, notused
    assert len(head) == 128
    return head


def write_c1_dds(ref):
    colors = image_blocks(1, 512, 512, PLAIN)
    ref.init(0)
    codes, _, _ = ref.compress_many(8, colors, None, np.ones(3, np.float32), 0, 0)
    with open(DDS_OUT, "wb") as f:
        f.write(dds_header(512, 512) + np.ascontiguousarray(codes, np.uint8).tobytes())


def inputs():
    parts_c, parts_w = [], []
    c = image_blocks(101, 64, 48, DEFAULT)                   # 192 blocks
    parts_c.append(c)
    parts_w.append(np.ones((len(c), 16), np.float32))
    c = image_blocks(102, 32, 32, PLAIN)                     # 64 blocks
    parts_c.append(c)
    parts_w.append(np.ones((len(c), 16), np.float32))
    c, w = edge_blocks(103, 192)                             # 192 blocks
    parts_c.append(c)
    parts_w.append(w)
    c = image_blocks(104, 32, 32, DEFAULT)                   # 64 weighted blocks
    a = alpha_field(104, 32, 32).reshape(8, 4, 8, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    c[..., 3] = a
    parts_c.append(c)
    parts_w.append(a.astype(np.float32))
    return np.concatenate(parts_c), np.concatenate(parts_w)


def main():
    from oracle import Reference
    ref = Reference("scalar")
    colors, weights = inputs()
    n = len(colors)
    cws = np.array([[1, 1, 1], [0.3, 0.59, 0.11]], np.float32)
    configs = []
    for level in range(10):
        configs.append((0, level, 1, 1, 0))
        configs.append((0, level, 0, 0, 0))
    for level in (2, 5, 8, 9):
        configs.append((0, level, 1, 0, 0))
    for level in (1, 5, 8, 9):
        configs.append((0, level, 1, 1, 1))
    configs.append((0, 10, 1, 1, 0))   # invalid level: FLT_MAX, output untouched
    configs.append((0, -1, 0, 0, 0))
    for dec in (1, 2, 3):
        for level in (1, 5, 8, 9):
            configs.append((dec, level, 1, 1, 0))

    rng = np.random.default_rng(105)
    prior = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
    codes = np.zeros((len(configs), n, 8), np.uint8)
    errors = np.zeros((len(configs), n), np.float32)
    for k, (dec, level, tcm, tcb, cwi) in enumerate(configs):
        ref.init(dec)
        codes[k], errors[k], _ = ref.compress_many(level, colors, weights, cws[cwi], tcm, tcb,
                                                   prior=prior)

    ref.init(0)
    t = ref.tables()
    m5, m6 = ref.midpoints()

    # decode / evaluate on a slice of the level-8 codes, for every decoder
    k8 = configs.index((0, 8, 1, 1, 0))
    dec_codes = codes[k8, :128]
    eval_rgba = np.clip(np.round(colors[:128] * 255), 0, 255).astype(np.uint8).reshape(128, 64)
    dec_rgba = np.zeros((4, 128, 64), np.uint8)
    eval_err = np.zeros((4, 128), np.float32)
    for dec in range(4):
        for i in range(128):
            dec_rgba[dec, i] = ref.decode(dec_codes[i], dec)
            eval_err[dec, i] = ref.evaluate_error(eval_rgba[i], dec_codes[i], dec)

    bc3_blocks, bc3_eval_rgba = bc3_inputs(106, 128)
    bc3_rgba = np.zeros((4, 128, 64), np.uint8)
    bc3_eval_err = np.zeros((2, 4, 128), np.float32)
    for dec in range(4):
        for i in range(128):
            bc3_rgba[dec, i] = ref.decode_bc3(bc3_blocks[i], dec)
            for blend in range(2):
                bc3_eval_err[blend, dec, i] = ref.evaluate_bc3_error(bc3_eval_rgba[i], bc3_blocks[i], blend, dec)

    match5 = np.zeros((4, 256, 2), np.uint8)
    match6 = np.zeros((4, 256, 2), np.uint8)
    dweights = np.zeros((4, 5), np.float32)
    for dec in range(4):
        ref.init(dec)
        td = ref.tables()
        match5[dec], match6[dec], dweights[dec] = td["match5"], td["match6"], td["weights"]

    np.savez_compressed(
        OUT, colors=colors, weights=weights, prior=prior, cfg=np.array(configs, np.int32),
        cw=cws, codes=codes, errors=errors, tables_four=t["four"], tables_three=t["three"],
        tables_four_total=t["four_total"], tables_three_total=t["three_total"],
        tables_match5=match5, tables_match6=match6, tables_weights=dweights,
        midpoints5=m5, midpoints6=m6, dec_codes=dec_codes, dec_rgba=dec_rgba,
        eval_rgba=eval_rgba, eval_err=eval_err, bc3_blocks=bc3_blocks, bc3_rgba=bc3_rgba,
        bc3_eval_rgba=bc3_eval_rgba, bc3_eval_err=bc3_eval_err)
    print(f"wrote {OUT}: {n} blocks x {len(configs)} configs, {os.path.getsize(OUT)} bytes")
    write_c1_dds(ref)
    print(f"wrote {DDS_OUT}: {os.path.getsize(DDS_OUT)} bytes")


if __name__ == "__main__":
    main()
