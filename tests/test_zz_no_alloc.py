"""GPU test of the no-allocation batch entry point
(icbc_gpu_compress_bc1_batch_noalloc): called through the C ABI, its codes
and errors equal those of icbc_gpu_compress_bc1_batch bit for bit, on two
calls in a row (the second smaller, so stale scratch contents would show),
it leaves the device's free memory unchanged, and calls beyond its limits are
rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from icbc_amd import synth

pytestmark = pytest.mark.gpu


def _hip_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


class _Case:
    """Jobs over device inputs with two sets of outputs (same prior bytes)."""

    def __init__(self, gpu, specs, seed):
        self.gpu = gpu
        self.bufs = []
        self.jobs = {"noalloc": [], "batch": []}
        self.outs = {"noalloc": [], "batch": []}
        rng = np.random.default_rng(seed)
        for k, (level, n, weighted, rgba8, tcm) in enumerate(specs):
            colors, _ = synth.make_image_blocks(seed * 100 + k, 4 * n, 4, synth.MIX_DEFAULT)
            if rgba8:
                src = np.clip(np.rint(colors * 255), 0, 255).astype(np.uint8).reshape(n, 64)
            else:
                src = colors
            dcol = self._buf(gpu.DeviceBuffer.from_array(np.ascontiguousarray(src)))
            dw = None
            if weighted:
                w = rng.random((n, 16), dtype=np.float32)
                w[rng.random((n, 16)) < 0.3] = 0.0
                dw = self._buf(gpu.DeviceBuffer.from_array(w))
            prior = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
            cw = (0.3, 0.59, 0.11) if weighted else (1, 1, 1)
            for path in ("noalloc", "batch"):
                dcode = self._buf(gpu.DeviceBuffer.from_array(prior))
                derr = self._buf(gpu.DeviceBuffer(4 * n))
                self.jobs[path].append(gpu.make_job(level, n, dcol, dcode, dw, derr, None, cw, tcm, tcm,
                                                    rgba8=rgba8))
                self.outs[path].append((dcode, derr, n))

    def _buf(self, b):
        self.bufs.append(b)
        return b

    def results(self, path):
        return [(d.download(np.uint8, (n, 8)), e.download(np.float32, n)) for d, e, n in self.outs[path]]

    def close(self):
        for b in self.bufs:
            b.close()


def _noalloc(gpu, jobs, stream):
    arr = gpu._job_array(jobs)
    return gpu.lib().icbc_gpu_compress_bc1_batch_noalloc(C.addressof(arr), len(jobs), stream.handle)


def _assert_same(got, want, what):
    for k, ((gc, ge), (wc, we)) in enumerate(zip(got, want)):
        assert np.array_equal(gc, wc), f"{what}: job {k} codes differ"
        same = (ge.view(np.uint32) == we.view(np.uint32)) | (np.isnan(ge) & np.isnan(we))
        assert np.all(same), f"{what}: job {k} errors differ"


def test_noalloc_matches_batch_two_calls(gpu):
    assert gpu.lib().icbc_gpu_noalloc_scratch_bytes() < (64 << 20)
    s = gpu.Stream()
    big = [(8, 3000, False, False, True), (9, 1200, True, False, True), (1, 900, False, True, False),
           (5, 400, False, False, True), (10, 50, False, False, True), (0, 77, True, False, False),
           (2, 333, False, True, True), (9, 65, False, False, False)]
    small = [(9, 150, False, False, True), (8, 40, True, True, True)]
    single = [(7, 2500, False, False, True)]
    cases = [_Case(gpu, big, 1), _Case(gpu, small, 2), _Case(gpu, single, 3)]
    try:
        # warm-up on a throwaway case: first launches of a process may set up
        # runtime state (code objects, kernel argument pools)
        warm = _Case(gpu, small, 4)
        assert _noalloc(gpu, warm.jobs["noalloc"], s) == 0, gpu.last_error()
        s.synchronize()
        warm.close()
        free0 = _hip_free_bytes()
        for c in cases:
            print(f"no-allocation call: {len(c.jobs['noalloc'])} jobs", flush=True)
            assert _noalloc(gpu, c.jobs["noalloc"], s) == 0, gpu.last_error()
        s.synchronize()
        free1 = _hip_free_bytes()
        for c in cases:
            gpu.compress_batch(c.jobs["batch"], stream=s)
        s.synchronize()
        for i, c in enumerate(cases):
            _assert_same(c.results("noalloc"), c.results("batch"), f"case {i}")
        assert free1 == free0, f"device free memory changed by {free0 - free1} bytes over no-allocation calls"
    finally:
        for c in cases:
            c.close()
        s.close()


def test_noalloc_rejects_calls_beyond_its_limits(gpu):
    lib = gpu.lib()
    s = gpu.Stream()
    d = gpu.DeviceBuffer(1 << 16)
    try:
        job = gpu.make_job(8, 16, d.ptr, d.ptr + 8192, None, None, None, (1, 1, 1), False, False)
        assert _noalloc(gpu, [job] * 33, s) == gpu.ERR_INVALID_ARGUMENT
        assert "at most 32" in gpu.last_error()
        big9 = gpu.make_job(9, 262145, d.ptr, d.ptr + 8192, None, None, None, (1, 1, 1), False, False)
        assert _noalloc(gpu, [big9], s) == gpu.ERR_INVALID_ARGUMENT
        host = np.zeros((16, 16, 4), np.float32)
        hcodes = np.zeros((16, 8), np.uint8)
        hjob = gpu.make_job(8, 16, host, hcodes, None, None, None, (1, 1, 1), False, False, host=True)
        assert _noalloc(gpu, [hjob], s) == gpu.ERR_INVALID_ARGUMENT
        many = gpu.make_job(1, 8 * 65537, d.ptr, d.ptr + 8192, None, None, None, (1, 1, 1), False, False)
        many.tile_blocks = 8
        assert _noalloc(gpu, [many], s) == gpu.ERR_INVALID_ARGUMENT
        assert "tiles" in gpu.last_error()
        assert lib.icbc_gpu_compress_bc1_batch_noalloc(None, 0, s.handle) == 0
    finally:
        d.close()
        s.close()
