"""icbc_amd -- MI355X (gfx950) BC1 encoder, a drop-in for castano/icbc v1.05.

Python mirror of the reference's public API (icbc.h:10-52) over the C ABI of
``libicbc_amd.so`` (include/icbc_gpu.h):

    import icbc_amd as icbc
    icbc.init(icbc.Decoder.D3D10)
    err = icbc.compress_bc1(icbc.Quality.Level8, colors64, weights16, (1, 1, 1),
                            True, True, out8)

plus batched entry points over numpy arrays or device buffers.  Every call
runs on the GPU; there is no CPU encoder in this package.  If the shared
library is missing or no gfx950 device is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libicbc_amd.so")


class Decoder(enum.IntEnum):  # icbc.h:12-17
    D3D10 = 0
    NVIDIA = 1
    AMD = 2
    Intel = 3


class Quality(enum.IntEnum):  # icbc.h:21-36
    Level0 = 0
    Level1 = 1
    Level2 = 2
    Level3 = 3
    Level4 = 4
    Level5 = 5
    Level6 = 6
    Level7 = 7
    Level8 = 8
    Level9 = 9
    Fast = 1
    Default = 8
    Max = 9


OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_NOT_INITIALIZED = -2
ERR_OUT_OF_MEMORY = -3
ERR_DEVICE = -4
ERR_NO_DEVICE = -5
ERR_BUDGET = -6
ERR_CANCELLED = -7
ERR_TIMEOUT = -8

JOB_HOST_MEMORY = 1
JOB_COLORS_RGBA8 = 2
JOB_IMAGE = 4
JOB_WEIGHTS_FROM_ALPHA = 8
IMAGE_RGBA8 = 0
IMAGE_RGBA32F = 1
EDGE_REPLICATE = 0
EDGE_ROUND_DOWN = 1
SCHEDULE_GRID = 0
SCHEDULE_PERSISTENT = 1


class IcbcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"icbc_amd error {code}: {msg}")
        self.code = code


class Image(C.Structure):
    """icbc_image (include/icbc_gpu.h): an RGBA8 or RGBA32F image in device memory."""
    _fields_ = [("data", C.c_void_p), ("format", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("edge", C.c_uint32), ("pitch", C.c_uint64)]


class Job(C.Structure):
    """icbc_bc1_job (include/icbc_gpu.h)."""
    _fields_ = [("level", C.c_int32), ("three_color_mode", C.c_int32),
                ("three_color_black", C.c_int32), ("flags", C.c_uint32),
                ("color_weights", C.c_float * 3), ("tile_blocks", C.c_uint32),
                ("block_count", C.c_uint64), ("colors", C.c_void_p), ("weights", C.c_void_p),
                ("codes", C.c_void_p), ("errors", C.c_void_p), ("stats", C.c_void_p),
                ("image", C.POINTER(Image))]


class ServerOptions(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("max_batch", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("batch_size", C.c_uint32), ("block_count", C.c_uint64),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("launch_s", C.c_double),
                ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("completed", C.c_uint64), ("blocks", C.c_uint64),
                ("pending", C.c_uint64), ("inflight", C.c_uint64), ("memory_in_use", C.c_uint64),
                ("memory_committed", C.c_uint64)]


class Tuning(C.Structure):
    """icbc_tuning (include/icbc_gpu.h): per-level launch settings."""
    _fields_ = [("tile_blocks", C.c_uint32 * 10), ("workgroups_per_cu", C.c_int32 * 10),
                ("schedule", C.c_int32 * 10), ("ns_per_block", C.c_float * 10),
                ("prune_min_colors", C.c_int32 * 10), ("default_ns_per_block", C.c_float * 10)]

    def as_dict(self):
        out = {}
        for lv in range(10):
            if not self.tile_blocks[lv]:
                continue
            d = {"tile_blocks": self.tile_blocks[lv], "workgroups_per_cu": self.workgroups_per_cu[lv],
                 "schedule": "persistent" if self.schedule[lv] == 1 else "grid",
                 "prune_min_colors": self.prune_min_colors[lv] or 13,
                 "ns_per_block": round(float(self.ns_per_block[lv]), 4),
                 "default_ns_per_block": round(float(self.default_ns_per_block[lv]), 4)}
            if self.default_ns_per_block[lv] > 0:
                d["speedup_over_default"] = round(float(self.default_ns_per_block[lv] / self.ns_per_block[lv]), 4)
            out[lv] = d
        return out


class PlanOptions(C.Structure):
    _fields_ = [("schedule", C.c_int32), ("use_graph", C.c_int32),
                ("workgroups_per_cu", C.c_int32), ("order_by_cost", C.c_int32)]


# exported symbols of the C ABI (include/icbc_gpu.h), with ctypes signatures
_SIGNATURES = {
    "icbc_gpu_init": (C.c_int, [C.c_int]),
    "icbc_gpu_compress_bc1_batch": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p]),
    "icbc_gpu_compress_bc1_plain": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p]),
    "icbc_gpu_compress_bc1_batch_noalloc": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p]),
    "icbc_gpu_noalloc_scratch_bytes": (C.c_uint64, []),
    "icbc_gpu_compress_bc1": (C.c_float, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_int, C.c_void_p]),
    "icbc_gpu_plan_create": (C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_void_p)]),
    "icbc_gpu_plan_execute": (C.c_int, [C.c_void_p, C.c_void_p]),
    "icbc_gpu_plan_destroy": (C.c_int, [C.c_void_p]),
    "icbc_gpu_plan_tiles": (C.c_uint64, [C.c_void_p]),
    "icbc_gpu_decode_bc1_batch": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]),
    "icbc_gpu_evaluate_bc1_error_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                                    C.c_int, C.c_void_p]),
    "icbc_gpu_decode_bc3_batch": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]),
    "icbc_gpu_evaluate_bc3_error_batch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                                    C.c_int, C.c_int, C.c_void_p]),
    "icbc_gpu_last_error": (C.c_char_p, []),
    "icbc_gpu_device_name": (C.c_int, [C.c_char_p, C.c_size_t]),
    "icbc_gpu_version": (C.c_int, []),
    "icbc_gpu_check_reciprocal": (C.c_int, [C.POINTER(C.c_uint64)]),
    "icbc_gpu_autotune": (C.c_int, [C.c_void_p, C.c_size_t, C.c_uint32, C.POINTER(Tuning)]),
    "icbc_gpu_get_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "icbc_gpu_set_tuning": (C.c_int, [C.POINTER(Tuning)]),
    "icbc_container_size": (C.c_uint64, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]),
    "icbc_container_data_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "icbc_container_encode": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_uint64]),
    "icbc_container_write": (C.c_int, [C.c_int, C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "icbc_container_compare_dds": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "icbc_gpu_image_block_count": (C.c_uint64, [C.POINTER(Image)]),
    "icbc_gpu_image_to_blocks": (C.c_int, [C.POINTER(Image), C.c_void_p, C.c_void_p, C.c_void_p]),
    "icbc_gpu_mip_chain_bytes": (C.c_uint64, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "icbc_gpu_build_mip_chain": (C.c_int, [C.POINTER(Image), C.c_void_p, C.c_uint64, C.POINTER(Image),
                                           C.c_uint32, C.POINTER(C.c_uint32), C.c_void_p]),
    "icbc_gpu_malloc": (C.c_void_p, [C.c_size_t]),
    "icbc_gpu_free": (C.c_int, [C.c_void_p]),
    "icbc_gpu_host_alloc": (C.c_void_p, [C.c_size_t]),
    "icbc_gpu_host_free": (C.c_int, [C.c_void_p]),
    "icbc_gpu_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "icbc_gpu_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "icbc_gpu_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "icbc_gpu_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "icbc_gpu_stream_create": (C.c_void_p, []),
    "icbc_gpu_stream_destroy": (C.c_int, [C.c_void_p]),
    "icbc_gpu_stream_synchronize": (C.c_int, [C.c_void_p]),
    "icbc_gpu_event_create": (C.c_void_p, []),
    "icbc_gpu_event_destroy": (C.c_int, [C.c_void_p]),
    "icbc_gpu_event_record": (C.c_int, [C.c_void_p, C.c_void_p]),
    "icbc_gpu_event_synchronize": (C.c_int, [C.c_void_p]),
    "icbc_gpu_event_elapsed_ms": (C.c_float, [C.c_void_p, C.c_void_p]),
    "icbc_server_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "icbc_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "icbc_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.c_void_p]),
    "icbc_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
    "icbc_server_stats": (C.c_int, [C.c_void_p, C.c_void_p]),
    "icbc_server_destroy": (C.c_int, [C.c_void_p]),
}

_lib = None


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IcbcError(ERR_NO_DEVICE, f"{LIB_PATH} is not built; run python -m icbc_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().icbc_gpu_last_error().decode()


def check(code):
    if code != OK:
        raise IcbcError(code, last_error())
    return code


# ---------------------------------------------------------------------------
# the reference's API


def init(decoder=Decoder.D3D10):
    """icbc::init (icbc.h:19, 3813)."""
    check(lib().icbc_gpu_init(int(decoder)))


def _ptr(a):
    return a.ctypes.data


def compress_bc1(level, input_colors, input_weights, color_weights, three_color_mode,
                 three_color_black, output):
    """icbc::compress_bc1 (icbc.h:48, 3846) for one block.

    ``input_colors``: 64 floats (16 texels x RGBA); ``input_weights``: 16
    floats; ``color_weights``: 3 floats; ``output``: a writable 8-byte buffer
    (numpy uint8 array / bytearray).  Returns the block error; FLT_MAX with
    ``output`` untouched for an invalid level.
    """
    c = np.ascontiguousarray(input_colors, dtype=np.float32).reshape(1, 16, 4)
    w = np.ascontiguousarray(input_weights, dtype=np.float32).reshape(1, 16)
    out = np.frombuffer(output, dtype=np.uint8) if not isinstance(output, np.ndarray) else output
    if out.nbytes < 8 or not out.flags.c_contiguous or not out.flags.writeable:
        raise ValueError("output must be a writable contiguous 8-byte buffer")
    codes, errors, _ = compress_many(level, c, w, color_weights, three_color_mode, three_color_black,
                                     prior=out.view(np.uint8)[:8].reshape(1, 8))
    out.view(np.uint8)[:8] = codes[0]
    return float(errors[0])


def compress_bc3(level, input_colors, input_weights, color_weights, six_alpha_mode, output):
    """icbc::compress_bc3 (icbc.h:49, 3850): a stub upstream that encodes nothing."""
    return 0.0


# ---------------------------------------------------------------------------
# device buffers and streams (plumbing for callers without a HIP runtime)


class DeviceBuffer:
    """Device allocation owned by Python; freed on close()/GC."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = lib().icbc_gpu_malloc(max(1, self.nbytes))
        if not p:
            raise IcbcError(ERR_OUT_OF_MEMORY, last_error())
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(lib().icbc_gpu_memcpy_htod(self.ptr + offset, _ptr(a), a.nbytes))

    def download(self, dtype, shape, offset=0):
        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        check(lib().icbc_gpu_memcpy_dtoh(_ptr(out), self.ptr + offset, out.nbytes))
        return out

    def fill(self, byte):
        check(lib().icbc_gpu_memset(self.ptr, int(byte), self.nbytes))

    def close(self):
        if self.ptr:
            lib().icbc_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A numpy array in page-locked host memory (icbc_gpu_host_alloc)."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        p = lib().icbc_gpu_host_alloc(max(1, nbytes))
        if not p:
            raise IcbcError(ERR_OUT_OF_MEMORY, last_error())
        self.ptr = p
        buf = (C.c_uint8 * max(1, nbytes)).from_address(p)
        self.array = np.frombuffer(buf, dtype=np.uint8, count=nbytes).view(dtype).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            lib().icbc_gpu_host_free(self.ptr)
            self.ptr = None


def pinned_copy(a):
    """Copy of `a` in page-locked host memory (keep the returned holder alive)."""
    h = PinnedArray(a.shape, a.dtype)
    h.array[...] = a
    return h


class Stream:
    def __init__(self):
        self.handle = lib().icbc_gpu_stream_create()
        if not self.handle:
            raise IcbcError(ERR_DEVICE, last_error())

    def synchronize(self):
        check(lib().icbc_gpu_stream_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib().icbc_gpu_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = lib().icbc_gpu_event_create()
        if not self.handle:
            raise IcbcError(ERR_DEVICE, last_error())

    def record(self, stream=None):
        check(lib().icbc_gpu_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self):
        check(lib().icbc_gpu_event_synchronize(self.handle))

    def elapsed_ms(self, end):
        return float(lib().icbc_gpu_event_elapsed_ms(self.handle, end.handle))

    def close(self):
        if self.handle:
            lib().icbc_gpu_event_destroy(self.handle)
            self.handle = None


def device_name():
    buf = C.create_string_buffer(256)
    check(lib().icbc_gpu_device_name(buf, 256))
    return buf.value.decode()


def check_reciprocal():
    """Count of float32 inputs where the encoder's fast reciprocal differs from
    IEEE-754 division on the current device (0 is required for bit-exact codes)."""
    n = C.c_uint64(0)
    check(lib().icbc_gpu_check_reciprocal(C.byref(n)))
    return n.value


# ---------------------------------------------------------------------------
# batches


def make_job(level, block_count, colors, codes, weights=None, errors=None, stats=None,
             color_weights=(1.0, 1.0, 1.0), three_color_mode=False, three_color_black=False,
             host=False, tile_blocks=0, rgba8=False, image=None, weights_from_alpha=False):
    """Builds an icbc_bc1_job; pointer arguments are ints (device) or numpy arrays (host).
    With ``image`` (an Image), texels are gathered from it and ``colors`` must be None."""
    def addr(x):
        if x is None:
            return None
        if isinstance(x, np.ndarray):
            return x.ctypes.data
        if isinstance(x, DeviceBuffer):
            return x.ptr
        return int(x)
    j = Job()
    j.level = int(level)
    j.three_color_mode = int(bool(three_color_mode))
    j.three_color_black = int(bool(three_color_black))
    j.flags = ((JOB_HOST_MEMORY if host else 0) | (JOB_COLORS_RGBA8 if rgba8 else 0)
               | (JOB_IMAGE if image is not None else 0) | (JOB_WEIGHTS_FROM_ALPHA if weights_from_alpha else 0))
    for k in range(3):
        j.color_weights[k] = float(color_weights[k])
    j.tile_blocks = int(tile_blocks)
    j.block_count = int(block_count)
    j.colors = addr(colors)
    j.weights = addr(weights)
    j.codes = addr(codes)
    j.errors = addr(errors)
    j.stats = addr(stats)
    if image is not None:
        j.image = C.pointer(image)
        j._image_ref = image  # the job points into it
    return j


# ---------------------------------------------------------------------------
# texture pre-pass

class DeviceImage:
    """A (h, w, 4) uint8 or float32 image uploaded to device memory, with its
    icbc_image descriptor in ``.desc``."""

    def __init__(self, img, edge=EDGE_REPLICATE, pitch=None):
        img = np.ascontiguousarray(img)
        if img.ndim != 3 or img.shape[2] != 4 or img.dtype not in (np.uint8, np.float32):
            raise ValueError("image must be (h, w, 4) uint8 or float32")
        h, w, _ = img.shape
        tb = img.dtype.itemsize * 4
        self.pitch = int(pitch) if pitch else w * tb
        self.buf = DeviceBuffer(self.pitch * h)
        if self.pitch == w * tb:
            self.buf.upload(img)
        else:
            for y in range(h):
                self.buf.upload(img[y], offset=y * self.pitch)
        self.desc = Image(self.buf.ptr, IMAGE_RGBA8 if img.dtype == np.uint8 else IMAGE_RGBA32F,
                          w, h, int(edge), self.pitch)

    @property
    def block_count(self):
        return image_block_count(self.desc)

    def close(self):
        self.buf.close()


def image_block_count(desc):
    return int(lib().icbc_gpu_image_block_count(C.byref(desc)))


def image_to_blocks(desc, alpha_weights=False, stream=None):
    """Device gather of an Image into (n, 16, 4) float32 blocks (and (n, 16)
    alpha weights); returns numpy arrays."""
    n = image_block_count(desc)
    colors = DeviceBuffer(n * 256)
    weights = DeviceBuffer(n * 64) if alpha_weights else None
    try:
        check(lib().icbc_gpu_image_to_blocks(C.byref(desc), colors.ptr, weights.ptr if weights else None,
                                             stream.handle if stream else None))
        if stream:
            stream.synchronize()
        out = colors.download(np.float32, (n, 16, 4))
        w = weights.download(np.float32, (n, 16)) if weights else None
    finally:
        colors.close()
        if weights:
            weights.close()
    return (out, w) if alpha_weights else out


def mip_chain_bytes(width, height):
    """(bytes, levels) of the float RGBA chain below a width x height image."""
    n = C.c_uint32(0)
    b = lib().icbc_gpu_mip_chain_bytes(int(width), int(height), C.byref(n))
    return int(b), n.value


class MipChain:
    """Every box-filtered level below ``source`` (an Image) built on the device:
    ``.levels`` holds their Image descriptors (RGBA32F, tightly packed)."""

    def __init__(self, source, stream=None):
        nbytes, n = mip_chain_bytes(source.width, source.height)
        self.buf = DeviceBuffer(max(nbytes, 16))
        descs = (Image * max(n, 1))()
        got = C.c_uint32(0)
        check(lib().icbc_gpu_build_mip_chain(C.byref(source), self.buf.ptr, nbytes, descs, n, C.byref(got),
                                             stream.handle if stream else None))
        if stream:
            stream.synchronize()
        self.levels = [descs[k] for k in range(got.value)]

    def download(self, k):
        d = self.levels[k]
        return self.buf.download(np.float32, (d.height, d.width, 4), offset=d.data - self.buf.ptr)

    def close(self):
        self.buf.close()


# ---------------------------------------------------------------------------
# autotuning

def autotune(sample_jobs, levels=range(10)):
    """Tunes the listed levels on device-memory sample jobs; returns the
    current device's Tuning."""
    arr = _job_array(sample_jobs)
    mask = 0
    for lv in levels:
        mask |= 1 << int(lv)
    t = Tuning()
    check(lib().icbc_gpu_autotune(C.addressof(arr), len(sample_jobs), mask, C.byref(t)))
    return t


def get_tuning():
    t = Tuning()
    check(lib().icbc_gpu_get_tuning(C.byref(t)))
    return t


def set_tuning(t):
    check(lib().icbc_gpu_set_tuning(C.byref(t) if t is not None else None))


# ---------------------------------------------------------------------------
# containers

CONTAINER_DDS = 0
CONTAINER_KTX = 1


def _level_codes(level_codes):
    if isinstance(level_codes, np.ndarray):
        level_codes = [level_codes]
    return np.ascontiguousarray(np.concatenate([np.asarray(c, np.uint8).reshape(-1) for c in level_codes])), len(level_codes)


def container_bytes(kind, width, height, level_codes):
    """The DDS / KTX file of a BC1 texture (``level_codes``: one (n, 8) uint8
    array per mip level, level 0 first) as bytes."""
    data, levels = _level_codes(level_codes)
    need = int(lib().icbc_container_data_bytes(width, height, levels))
    if need != data.nbytes:
        raise ValueError(f"{levels} levels of {width}x{height} hold {need} bytes of blocks, got {data.nbytes}")
    size = int(lib().icbc_container_size(int(kind), width, height, levels))
    out = np.empty(size, np.uint8)
    check(lib().icbc_container_encode(int(kind), width, height, levels, data.ctypes.data, out.ctypes.data, size))
    return out.tobytes()


def write_container(path, kind, width, height, level_codes):
    data, levels = _level_codes(level_codes)
    need = int(lib().icbc_container_data_bytes(width, height, levels))
    if need != data.nbytes:
        raise ValueError(f"{levels} levels of {width}x{height} hold {need} bytes of blocks, got {data.nbytes}")
    check(lib().icbc_container_write(int(kind), os.fsencode(path), width, height, levels, data.ctypes.data))


def compare_dds(path, width, height, codes):
    """True if the top level of the DDS file equals ``codes`` (the reference's -cmp check)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    r = lib().icbc_container_compare_dds(os.fsencode(path), width, height, codes.ctypes.data)
    if r < 0:
        check(r)
    return r == 1


def _job_array(jobs):
    arr = (Job * len(jobs))(*jobs)
    return arr


def compress_batch(jobs, stream=None, plain=False):
    """icbc_gpu_compress_bc1_batch / _plain over a list of Job."""
    arr = _job_array(jobs)
    fn = lib().icbc_gpu_compress_bc1_plain if plain else lib().icbc_gpu_compress_bc1_batch
    check(fn(C.addressof(arr), len(jobs), stream.handle if stream else None))


def compress_many(level, colors, weights=None, color_weights=(1.0, 1.0, 1.0),
                  three_color_mode=False, three_color_black=False, prior=None, with_stats=False,
                  plain=False):
    """Encodes block-major host arrays: colors (n, 16, 4) float32 -- or uint8,
    read as c / 255.0f -- and weights (n, 16) or None.  Returns
    (codes uint8[n, 8], errors float32[n], stats int32[n, 4] | None)."""
    rgba8 = np.asarray(colors).dtype == np.uint8
    colors = np.ascontiguousarray(colors, dtype=np.uint8 if rgba8 else np.float32)
    n = colors.shape[0]
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    codes = np.zeros((n, 8), np.uint8) if prior is None else np.array(prior, np.uint8, copy=True)
    errors = np.zeros(n, np.float32)
    stats = np.zeros((n, 4), np.int32) if with_stats else None
    job = make_job(level, n, colors, codes, w, errors, stats, color_weights, three_color_mode,
                   three_color_black, host=True, rgba8=rgba8)
    compress_batch([job], plain=plain)
    return codes, errors, stats


class Plan:
    """Persistent batched plan over device-memory jobs (icbc_gpu_plan_*)."""

    def __init__(self, jobs, schedule=None, use_graph=True, workgroups_per_cu=0,
                 order_by_cost=True, stream=None):
        """With every option at its default the library picks them (the
        device's tuning when one is installed, else persistent + graph)."""
        self._jobs = _job_array(jobs)
        if schedule is None and use_graph and not workgroups_per_cu and order_by_cost:
            opt = None
        else:
            opt = PlanOptions(SCHEDULE_PERSISTENT if schedule is None else schedule, int(bool(use_graph)),
                              workgroups_per_cu, int(bool(order_by_cost)))
        h = C.c_void_p()
        check(lib().icbc_gpu_plan_create(C.addressof(self._jobs), len(jobs), C.byref(opt) if opt else None,
                                         stream.handle if stream else None, C.byref(h)))
        self.handle = h

    @property
    def tiles(self):
        return int(lib().icbc_gpu_plan_tiles(self.handle))

    def execute(self, stream=None):
        check(lib().icbc_gpu_plan_execute(self.handle, stream.handle if stream else None))

    def close(self):
        if self.handle:
            lib().icbc_gpu_plan_destroy(self.handle)
            self.handle = None


def decode_bc1_batch(codes_dev, block_count, rgba_dev, decoder=Decoder.D3D10, stream=None):
    check(lib().icbc_gpu_decode_bc1_batch(int(codes_dev), int(block_count), int(rgba_dev), int(decoder),
                                          stream.handle if stream else None))


def evaluate_bc1_error_batch(rgba_dev, codes_dev, block_count, errors_dev, decoder=Decoder.D3D10,
                             stream=None):
    check(lib().icbc_gpu_evaluate_bc1_error_batch(int(rgba_dev), int(codes_dev), int(block_count),
                                                  int(errors_dev), int(decoder),
                                                  stream.handle if stream else None))


def decode_bc3_batch(blocks_dev, block_count, rgba_dev, decoder=Decoder.D3D10, stream=None):
    """icbc::decode_bc3 (icbc.h:2929-2946) over 16-byte BC3 blocks in device memory."""
    check(lib().icbc_gpu_decode_bc3_batch(int(blocks_dev), int(block_count), int(rgba_dev), int(decoder),
                                          stream.handle if stream else None))


def evaluate_bc3_error_batch(rgba_dev, blocks_dev, block_count, errors_dev, alpha_blend,
                             decoder=Decoder.D3D10, stream=None):
    """icbc::evaluate_bc3_error (icbc.h:3048-3071) over 16-byte BC3 blocks in device memory."""
    check(lib().icbc_gpu_evaluate_bc3_error_batch(int(rgba_dev), int(blocks_dev), int(block_count),
                                                  int(errors_dev), int(bool(alpha_blend)), int(decoder),
                                                  stream.handle if stream else None))


def _one_block(block, size):
    b = np.frombuffer(bytes(block), np.uint8)
    if len(b) < size:
        raise ValueError(f"a block is {size} bytes")
    return np.ascontiguousarray(b[:size])


def decode_bc1(block, decoder=Decoder.D3D10):
    """icbc::decode_bc1 (icbc.h:38, 3820) for one block -> 64 RGBA8 bytes."""
    d = DeviceBuffer(128)
    d.upload(_one_block(block, 8))
    decode_bc1_batch(d.ptr, 1, d.ptr + 64, decoder)
    out = d.download(np.uint8, 64, offset=64)
    d.close()
    return out


def decode_bc3(block, decoder=Decoder.D3D10):
    """icbc::decode_bc3 (icbc.h:39, 3824) for one 16-byte block -> 64 RGBA8 bytes."""
    d = DeviceBuffer(128)
    d.upload(_one_block(block, 16))
    decode_bc3_batch(d.ptr, 1, d.ptr + 64, decoder)
    out = d.download(np.uint8, 64, offset=64)
    d.close()
    return out


def evaluate_bc1_error(rgba_block, block, decoder=Decoder.D3D10):
    """icbc::evaluate_bc1_error (icbc.h:43, 3837) for one block."""
    d = DeviceBuffer(128)
    d.upload(np.ascontiguousarray(rgba_block, np.uint8).reshape(64))
    d.upload(_one_block(block, 8), offset=64)
    evaluate_bc1_error_batch(d.ptr, d.ptr + 64, 1, d.ptr + 80, decoder)
    e = float(d.download(np.float32, 1, offset=80)[0])
    d.close()
    return e


def evaluate_bc3_error(rgba_block, block, alpha_blend, decoder=Decoder.D3D10):
    """icbc::evaluate_bc3_error (icbc.h:44, 3841) for one 16-byte block."""
    d = DeviceBuffer(128)
    d.upload(np.ascontiguousarray(rgba_block, np.uint8).reshape(64))
    d.upload(_one_block(block, 16), offset=64)
    evaluate_bc3_error_batch(d.ptr, d.ptr + 64, 1, d.ptr + 96, alpha_blend, decoder)
    e = float(d.download(np.float32, 1, offset=96)[0])
    d.close()
    return e


# ---------------------------------------------------------------------------
# request server (continuous batching under a device-memory budget)


class Server:
    """icbc_server_*: textures submitted as requests are admitted by memory
    budget, batched with whatever else is waiting, and run while earlier
    launches are still on the device."""

    def __init__(self, memory_budget, max_inflight=3, max_batch=0):
        opt = ServerOptions(int(memory_budget), int(max_inflight), int(max_batch))
        h = C.c_void_p()
        check(lib().icbc_server_create(C.byref(opt), C.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, level, colors, weights=None, color_weights=(1.0, 1.0, 1.0),
               three_color_mode=False, three_color_black=False, codes=None, errors=None,
               tile_blocks=0, with_errors=True):
        """Queues one texture (block-major host arrays, float32 or uint8 RGBA).
        Output arrays in page-locked memory (PinnedArray) are written in place
        by the device; others are staged.  Returns a ticket."""
        rgba8 = colors.dtype == np.uint8
        colors = np.ascontiguousarray(colors, dtype=np.uint8 if rgba8 else np.float32)
        n = colors.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        codes = np.zeros((n, 8), np.uint8) if codes is None else codes
        if errors is None and with_errors:
            errors = np.zeros(n, np.float32)
        job = make_job(level, n, colors, codes, w, errors, None, color_weights, three_color_mode,
                       three_color_black, host=True, tile_blocks=tile_blocks, rgba8=rgba8)
        t = C.c_uint64()
        check(lib().icbc_server_submit(self.handle, C.byref(job), C.byref(t)))
        self._keep[t.value] = (colors, w, codes, errors)
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """Waits for a request; returns (codes, errors, info) or raises."""
        info = RequestInfo()
        r = lib().icbc_server_wait(self.handle, ticket, float(timeout), C.byref(info))
        if r == ERR_TIMEOUT:
            raise TimeoutError(f"request {ticket} still running")
        if r == ERR_INVALID_ARGUMENT:   # unknown or already collected ticket
            raise IcbcError(r, last_error())
        colors, w, codes, errors = self._keep.pop(ticket)
        if r != OK:
            raise IcbcError(r, last_error() if r != info.status else f"request failed ({r})")
        return codes, errors, info

    def cancel(self, ticket):
        check(lib().icbc_server_cancel(self.handle, ticket))

    def stats(self):
        st = ServerStats()
        check(lib().icbc_server_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().icbc_server_destroy(self.handle)
            self.handle = None
