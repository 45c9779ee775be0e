"""Seeded synthetic texture content for tests and benchmarks.

The reference's harness encodes the Kodak photographs (icbc_test.cpp:449-474),
which are not available offline; every workload here is generated from a seed
instead.  Images are float32 RGBA in [0, 1] (alpha = 1 unless stated) and are
handed to the encoder either as an image or as block-major arrays: 16 texels x
RGBA per 4x4 block, texel j = 4*y + x (the layout of icbc.h's compress_bc1).

Block categories:
  gradient    two random endpoints, texels interpolated along a random direction
  value_noise bilinear (smoothstep) value noise on an 8-pixel lattice, +-0.05 jitter
  white_noise independent uniform texels
  near_black  a gradient or noise block with 1-6 texels set below 0.02
  flat        one uniform colour
"""
from __future__ import annotations

import numpy as np

CATEGORIES = ("gradient", "value_noise", "white_noise", "near_black", "flat")

# content mixes of the benchmark workloads
MIX_PLAIN = {"gradient": 0.5, "value_noise": 0.3, "flat": 0.2}
MIX_DEFAULT = {"gradient": 0.4, "white_noise": 0.3, "near_black": 0.2, "flat": 0.1}


def _value_noise(rng, h, w, cell=8):
    gh, gw = h // cell + 2, w // cell + 2
    lattice = rng.random((gh, gw, 3), dtype=np.float32)
    ys = (np.arange(h, dtype=np.float32) + 0.5) / cell
    xs = (np.arange(w, dtype=np.float32) + 0.5) / cell
    y0 = np.floor(ys).astype(np.int64)
    x0 = np.floor(xs).astype(np.int64)
    fy = ys - y0
    fx = xs - x0
    sy = (fy * fy * (3 - 2 * fy))[:, None, None]
    sx = (fx * fx * (3 - 2 * fx))[None, :, None]
    a = lattice[y0][:, x0]
    b = lattice[y0][:, x0 + 1]
    c = lattice[y0 + 1][:, x0]
    d = lattice[y0 + 1][:, x0 + 1]
    top = a + (b - a) * sx
    bot = c + (d - c) * sx
    return (top + (bot - top) * sy).astype(np.float32)


def _gradient_blocks(rng, n):
    e0 = rng.random((n, 3), dtype=np.float32)
    e1 = rng.random((n, 3), dtype=np.float32)
    ang = rng.random(n, dtype=np.float32) * np.float32(2 * np.pi)
    yy, xx = np.mgrid[0:4, 0:4].astype(np.float32)
    proj = (np.cos(ang)[:, None] * (xx.reshape(1, 16) - 1.5)
            + np.sin(ang)[:, None] * (yy.reshape(1, 16) - 1.5))
    lo = proj.min(axis=1, keepdims=True)
    hi = proj.max(axis=1, keepdims=True)
    t = (proj - lo) / np.maximum(hi - lo, 1e-6)
    return (e0[:, None, :] + (e1 - e0)[:, None, :] * t[:, :, None]).astype(np.float32)


def _blocks_from_image(img):
    h, w, _ = img.shape
    return (img.reshape(h // 4, 4, w // 4, 4, 3).transpose(0, 2, 1, 3, 4)
            .reshape(-1, 16, 3))


def make_image_blocks(seed, width, height, mix):
    """Texel colours of a width x height image as block-major (nblocks, 16, 4)
    float32 plus the per-block category index, drawn from ``mix``."""
    assert width % 4 == 0 and height % 4 == 0
    rng = np.random.default_rng(seed)
    nb = (width // 4) * (height // 4)
    names = list(mix.keys())
    probs = np.array([mix[k] for k in names], dtype=np.float64)
    probs /= probs.sum()
    cat_local = rng.choice(len(names), size=nb, p=probs)
    cat = np.array([CATEGORIES.index(names[i]) for i in cat_local], dtype=np.int8)
    rgb = np.empty((nb, 16, 3), dtype=np.float32)

    sel = cat == CATEGORIES.index("gradient")
    rgb[sel] = _gradient_blocks(rng, int(sel.sum()))

    sel = cat == CATEGORIES.index("value_noise")
    if sel.any():
        field = _value_noise(rng, height, width)
        vb = _blocks_from_image(field)[sel]
        vb = vb + rng.uniform(-0.05, 0.05, size=vb.shape).astype(np.float32)
        rgb[sel] = np.clip(vb, 0.0, 1.0)

    sel = cat == CATEGORIES.index("white_noise")
    rgb[sel] = rng.random((int(sel.sum()), 16, 3), dtype=np.float32)

    sel = cat == CATEGORIES.index("flat")
    rgb[sel] = np.repeat(rng.random((int(sel.sum()), 1, 3), dtype=np.float32), 16, axis=1)

    sel = np.nonzero(cat == CATEGORIES.index("near_black"))[0]
    if len(sel):
        base = _gradient_blocks(rng, len(sel))
        noisy = rng.random(len(sel)) < 0.5
        base[noisy] = rng.random((int(noisy.sum()), 16, 3), dtype=np.float32)
        k = rng.integers(1, 7, size=len(sel))
        # k distinct texels per block: the k smallest of 16 random keys
        ranks = np.argsort(np.argsort(rng.random((len(sel), 16)), axis=1), axis=1)
        dark = ranks < k[:, None]
        base[dark] = rng.random((int(dark.sum()), 3), dtype=np.float32) * np.float32(0.02)
        rgb[sel] = base

    colors = np.empty((nb, 16, 4), dtype=np.float32)
    colors[..., :3] = rgb
    colors[..., 3] = 1.0
    return colors, cat


def blocks_to_image(colors, width, height):
    """Inverse of the block gather: (nblocks, 16, 4) -> (height, width, 4)."""
    bw, bh = width // 4, height // 4
    return (colors.reshape(bh, bw, 4, 4, 4).transpose(0, 2, 1, 3, 4)
            .reshape(height, width, 4))


def make_weighted_alpha(seed, width, height, zero_fraction=0.3):
    """Alpha field for the weighted workload: thresholded value noise with
    ``zero_fraction`` of texels at alpha 0 (weight 0) and the rest in (0, 1]."""
    rng = np.random.default_rng(seed + 7919)
    field = _value_noise(rng, height, width)[..., 0]
    cut = np.quantile(field, zero_fraction)
    alpha = np.where(field <= cut, 0.0, (field - cut) / max(1e-6, float(field.max() - cut)))
    alpha = np.clip(alpha, 0.0, 1.0).astype(np.float32)
    alpha[(alpha > 0) & (alpha < 1e-3)] = 1e-3
    return alpha


def make_edge_blocks(seed, n):
    """Blocks aimed at the encoder's corner cases: zero / negative / NaN
    weights, duplicates, two-colour and near-threshold blocks, saturated and
    out-of-range values, near-black texels and exact ties."""
    rng = np.random.default_rng(seed)
    colors = np.empty((n, 16, 4), dtype=np.float32)
    weights = np.ones((n, 16), dtype=np.float32)
    kinds = rng.integers(0, 14, size=n)
    for b in range(n):
        k = kinds[b]
        c = rng.random((16, 3), dtype=np.float32)
        w = np.ones(16, dtype=np.float32)
        if k == 0:      # all weights zero
            w[:] = 0
        elif k == 1:    # some zero / negative weights, random positive rest
            w = rng.random(16, dtype=np.float32)
            w[rng.random(16) < 0.4] = 0
            w[rng.random(16) < 0.1] = -1
        elif k == 2:    # two colours
            a, d = rng.random((2, 3), dtype=np.float32)
            m = rng.random(16) < 0.5
            c[m] = a
            c[~m] = d
        elif k == 3:    # one colour plus tiny perturbations (merge threshold)
            a = rng.random(3, dtype=np.float32)
            c = a + rng.uniform(-1.0 / 200, 1.0 / 200, size=(16, 3)).astype(np.float32)
        elif k == 4:    # saturated corners
            c = (rng.random((16, 3)) < 0.5).astype(np.float32)
        elif k == 5:    # near black texels
            c[rng.random(16) < 0.4] *= np.float32(0.1)
        elif k == 6:    # all black-ish
            c *= np.float32(0.11)
        elif k == 7:    # out of range values
            c = rng.uniform(-0.25, 1.25, size=(16, 3)).astype(np.float32)
        elif k == 8:    # quantised to 8-bit like real images
            c = np.round(c * 255) / np.float32(255)
        elif k == 9:    # few distinct 8-bit levels, lots of ties
            lv = (rng.integers(0, 4, size=(16, 3)) * 85).astype(np.float32) / 255
            c = lv
        elif k == 10:   # grey ramp
            g = np.linspace(0, 1, 16, dtype=np.float32)
            c = np.repeat(g[:, None], 3, axis=1)
        elif k == 11:   # weighted random
            w = rng.random(16, dtype=np.float32) * 4
        elif k == 12:   # single weighted texel
            w[:] = 0
            w[rng.integers(0, 16)] = 1
        else:           # 3 clusters
            cen = rng.random((3, 3), dtype=np.float32)
            c = cen[rng.integers(0, 3, size=16)] + rng.normal(0, 0.01, size=(16, 3)).astype(np.float32)
        colors[b, :, :3] = c
        colors[b, :, 3] = 1.0
        weights[b] = w
    return colors.astype(np.float32), weights.astype(np.float32)



def make_nonfinite_blocks(seed, n):
    """Blocks with non-finite and extreme inputs: NaN or infinite colour
    channels and weights, colours of magnitude 1e30, denormal colours, and
    NaN texels whose weight is zero (dropped before any arithmetic)."""
    rng = np.random.default_rng(seed)
    colors = rng.random((n, 16, 4), dtype=np.float32)
    colors[..., 3] = 1.0
    weights = np.ones((n, 16), dtype=np.float32)
    for b in range(n):
        k = b % 10
        t = rng.integers(0, 16)
        ch = rng.integers(0, 3)
        if k == 0:
            colors[b, t, ch] = np.nan
        elif k == 1:
            colors[b, t, ch] = np.inf
        elif k == 2:
            colors[b, t, ch] = -np.inf
        elif k == 3:
            colors[b, :, :3] = np.nan
        elif k == 4:
            weights[b, t] = np.nan
        elif k == 5:
            weights[b, t] = np.inf
        elif k == 6:
            colors[b, rng.random(16) < 0.5, :3] *= np.float32(1e30)
        elif k == 7:
            colors[b, :, :3] *= np.float32(1e-40)
        elif k == 8:
            colors[b, t, :3] = np.nan
            weights[b, t] = 0.0
        else:
            colors[b, t, ch] = np.nan
            colors[b, (t + 5) % 16, (ch + 1) % 3] = np.inf
    return colors, weights

# ---------------------------------------------------------------------------
# textures with mip chains (the batched workload)


def image_to_blocks(img):
    """(h, w, 4) float32 image -> block-major (nblocks, 16, 4), padding the
    right and bottom edges by replication to multiples of 4."""
    h, w, _ = img.shape
    ph, pw = (h + 3) // 4 * 4, (w + 3) // 4 * 4
    if ph != h or pw != w:
        img = np.pad(img, ((0, ph - h), (0, pw - w), (0, 0)), mode="edge")
    return np.ascontiguousarray(
        img.reshape(ph // 4, 4, pw // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4))


def make_image(seed, width, height, mix):
    """A width x height RGBA image (any size) with the block content ``mix``."""
    pw, ph = (width + 3) // 4 * 4, (height + 3) // 4 * 4
    blocks, _ = make_image_blocks(seed, pw, ph, mix)
    return blocks_to_image(blocks, pw, ph)[:height, :width]


def box_downsample(img):
    """2x2 box filter; an odd trailing row/column is dropped (sizes floor to
    at least 1)."""
    h, w, c = img.shape
    if h == 1 and w == 1:
        return img
    if h == 1:
        return (img[:, 0:w - 1:2] + img[:, 1:w:2]) * np.float32(0.5)
    if w == 1:
        return (img[0:h - 1:2] + img[1:h:2]) * np.float32(0.5)
    a = img[0:h - 1:2, 0:w - 1:2]
    b = img[0:h - 1:2, 1:w:2]
    c_ = img[1:h:2, 0:w - 1:2]
    d = img[1:h:2, 1:w:2]
    return ((a + b) + (c_ + d)) * np.float32(0.25)


def mip_chain(img):
    """Full box-filtered chain from ``img`` down to the first level that fits
    in one 4x4 block."""
    levels = [img]
    while levels[-1].shape[0] > 4 or levels[-1].shape[1] > 4:
        levels.append(box_downsample(levels[-1]))
    return levels


def texture_set(seed, count=256, min_side=64, max_side=4096, mix=None, levels=range(1, 10)):
    """The batched workload: ``count`` textures with sides drawn log-uniform in
    [min_side, max_side] (width and height independent), each with its full
    mip chain, and a quality level drawn uniformly from ``levels`` per texture.
    Returns a list of (texture, mip, width, height, level, blocks)."""
    mix = mix or MIX_DEFAULT
    rng = np.random.default_rng(seed)
    lo, hi = np.log(min_side), np.log(max_side)
    sides = np.exp(rng.uniform(lo, hi, size=(count, 2))).round().astype(int)
    qual = rng.choice(list(levels), size=count)
    out = []
    for t in range(count):
        w, h = int(sides[t, 0]), int(sides[t, 1])
        img = make_image(seed * 1000 + t, w, h, mix)
        for m, lv in enumerate(mip_chain(img)):
            out.append((t, m, lv.shape[1], lv.shape[0], int(qual[t]), image_to_blocks(lv)))
    return out
