"""CPU restatement of the texture pre-pass: image -> 4x4 blocks and
box-filtered mip levels.

TEST INFRASTRUCTURE ONLY (the checker for the device pre-pass kernels);
the product never imports it.

- Block gather follows the reference harness, icbc_test.cpp:353-380: blocks
  in row-major order over the image, texel j = 4*yy + xx inside a block, each
  RGBA8 channel read as c / 255.0f (icbc_test.cpp:401-405).  The harness
  rounds the block grid down (icbc_test.cpp:354-358); the batching workload
  instead pads the right and bottom edges by replicating the last column and
  row, which is the other edge mode here.
- The mip levels are not in the reference: the batching workload defines them
  as a 2x2 box filter in float32, ((a + b) + (c + d)) * 0.25 per channel, an
  odd last row or column dropped, a dimension of 1 filtered along the other
  only ((a + b) * 0.5), down to the first level that fits one 4x4 block.

Written as explicit per-texel loops over small arrays (clarity over speed);
``gather_fast`` / ``downsample_fast`` are numpy forms pinned against them.
"""
import numpy as np

EDGE_REPLICATE = 0
EDGE_ROUND_DOWN = 1


def to_float(img):
    """RGBA8 -> float32 c / 255.0f (correctly rounded float32 division)."""
    if img.dtype == np.uint8:
        return img.astype(np.float32) / np.float32(255.0)
    return img.astype(np.float32, copy=False)


def block_grid(width, height, edge):
    if edge == EDGE_ROUND_DOWN:
        return width // 4, height // 4
    return (width + 3) // 4, (height + 3) // 4


def gather(img, edge=EDGE_REPLICATE):
    """(h, w, 4) image -> (n, 16, 4) float32 blocks, by explicit loops
    (icbc_test.cpp:361-380 with the edge mode above)."""
    f = to_float(img)
    h, w, _ = f.shape
    bx, by = block_grid(w, h, edge)
    out = np.empty((bx * by, 16, 4), np.float32)
    b = 0
    for y in range(by):
        for x in range(bx):
            for yy in range(4):
                for xx in range(4):
                    sx = min(4 * x + xx, w - 1)
                    sy = min(4 * y + yy, h - 1)
                    out[b, 4 * yy + xx] = f[sy, sx]
            b += 1
    return out


def gather_fast(img, edge=EDGE_REPLICATE):
    f = to_float(img)
    h, w, _ = f.shape
    bx, by = block_grid(w, h, edge)
    ys = np.minimum(np.arange(by * 4), h - 1)
    xs = np.minimum(np.arange(bx * 4), w - 1)
    g = f[ys][:, xs]
    return np.ascontiguousarray(g.reshape(by, 4, bx, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4))


def downsampled_size(w, h):
    if w == 1 and h == 1:
        return 1, 1
    if h == 1:
        return w // 2, 1
    if w == 1:
        return 1, h // 2
    return w // 2, h // 2


def downsample(img):
    """One box-filtered level, by explicit loops."""
    f = to_float(img)
    h, w, _ = f.shape
    dw, dh = downsampled_size(w, h)
    out = np.empty((dh, dw, 4), np.float32)
    half, quarter = np.float32(0.5), np.float32(0.25)
    for y in range(dh):
        for x in range(dw):
            if w == 1 and h == 1:
                out[y, x] = f[0, 0]
            elif h == 1:
                out[y, x] = (f[0, 2 * x] + f[0, 2 * x + 1]) * half
            elif w == 1:
                out[y, x] = (f[2 * y, 0] + f[2 * y + 1, 0]) * half
            else:
                a, b = f[2 * y, 2 * x], f[2 * y, 2 * x + 1]
                c, d = f[2 * y + 1, 2 * x], f[2 * y + 1, 2 * x + 1]
                out[y, x] = ((a + b) + (c + d)) * quarter
    return out


def downsample_fast(img):
    f = to_float(img)
    h, w, _ = f.shape
    if w == 1 and h == 1:
        return f.copy()
    if h == 1:
        return (f[:, 0:w - 1:2] + f[:, 1:w:2]) * np.float32(0.5)
    if w == 1:
        return (f[0:h - 1:2] + f[1:h:2]) * np.float32(0.5)
    a, b = f[0:h - 1:2, 0:w - 1:2], f[0:h - 1:2, 1:w:2]
    c, d = f[1:h:2, 0:w - 1:2], f[1:h:2, 1:w:2]
    return ((a + b) + (c + d)) * np.float32(0.25)


def mip_chain(img, fast=True):
    """Every level below ``img`` down to the first that fits one 4x4 block."""
    step = downsample_fast if fast else downsample
    levels = []
    cur = to_float(img)
    while cur.shape[0] > 4 or cur.shape[1] > 4:
        cur = step(cur)
        levels.append(cur)
    return levels
