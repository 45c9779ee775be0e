"""GPU tests of the request server (continuous batching under a memory
budget): every request's result is the same alone and in the mix (and equal
to the oracle), the device memory the server holds stays within its budget
(measured with hipMemGetInfo), admission is FIFO, oversize and cancelled
requests get defined statuses, and launches overlap."""
import ctypes
import threading
import time

import numpy as np
import pytest

from icbc_amd import synth

pytestmark = pytest.mark.gpu


def _requests(seed, count, max_side=256):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        w = int(np.exp(rng.uniform(np.log(16), np.log(max_side))))
        h = int(np.exp(rng.uniform(np.log(16), np.log(max_side))))
        img = synth.make_image(seed * 100 + k, w, h, synth.MIX_DEFAULT)
        blocks = synth.image_to_blocks(img)
        level = int(rng.integers(0, 10))
        tcm = bool(rng.integers(0, 2))
        weights = rng.random((len(blocks), 16)).astype(np.float32) if k % 5 == 0 else None
        reqs.append((level, blocks, weights, tcm))
    return reqs


def _oracle(oracle, level, blocks, weights, tcm):
    codes, errs, _, _ = oracle.compress_many(level, blocks, weights, (1, 1, 1), tcm, tcm, nthreads=8)
    return codes, errs


def test_results_same_alone_and_in_mix(gpu, oracle):
    reqs = _requests(1, 40)
    srv = gpu.Server(memory_budget=64 << 20, max_inflight=3)
    try:
        tickets = [srv.submit(lv, b, w, (1, 1, 1), tcm, tcm) for lv, b, w, tcm in reqs]
        mixed = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    batch_sizes = [info.batch_size for _, _, info in mixed]
    assert max(batch_sizes) > 1, "requests were never batched together"
    for (lv, b, w, tcm), (codes, errs, info) in zip(reqs, mixed):
        assert info.status == 0
        alone = gpu.compress_many(lv, b, w, (1, 1, 1), tcm, tcm)
        want = _oracle(oracle, lv, b, w, tcm)
        assert np.array_equal(codes, alone[0]) and np.array_equal(errs.view(np.uint32), alone[1].view(np.uint32))
        assert np.array_equal(codes, want[0]) and np.array_equal(errs.view(np.uint32), want[1].view(np.uint32))
    assert st["completed"] == len(reqs)


def _device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


# the device's free-memory count moves in 2 MiB granules
SLACK = 2 << 20


def _warm_runtime(gpu):
    """Process-wide device memory the runtime sets up on first use -- the
    library's code object and each hardware queue's scratch for the encoder's
    register spills -- is taken before measuring: every kernel instance runs
    once on each of several streams (streams share the device's few
    hardware queues)."""
    colors, _ = synth.make_image_blocks(11, 256, 64, synth.MIX_DEFAULT)
    w = np.random.default_rng(11).random((len(colors), 16)).astype(np.float32)
    streams = [gpu.Stream() for _ in range(8)]
    try:
        for s in streams:
            # one launch per set of levels: the level 0-1 instance, the cluster
            # instances without and with the box-fit stage (a launch mixing
            # levels 0-3 with 4-9 runs the latter, which spills most), and the
            # refinement kernel
            for lvs, wt in (((1,), None), ((1,), w), ((8,), None), ((9,), w), ((2, 8), None), ((3, 9), w)):
                outs = [np.zeros((len(colors), 8), np.uint8) for _ in lvs]
                jobs = [gpu.make_job(lv, len(colors), colors, out, wt, None, None, (1, 1, 1), True, True, host=True)
                        for lv, out in zip(lvs, outs)]
                gpu.compress_batch(jobs, stream=s)
            s.synchronize()
    finally:
        for s in streams:
            s.close()


@pytest.mark.parametrize("budget,inflight", [(48 << 20, 1), (64 << 20, 2), (96 << 20, 3)])
def test_device_memory_within_budget(gpu, budget, inflight):
    # the budget bounds every device byte the server holds -- request buffers,
    # launch tables, level-9 lists and slot workspaces -- checked against the
    # device's own free-memory count while requests (more than the budget
    # leaves for them, levels 0-9, some weighted) flow through
    reqs = _requests(2, 300, max_side=512)
    need = sum(len(b) * (256 + 8 + 4) for _, b, _, _ in reqs)
    assert need > budget // 2   # more than the room the launch slots leave for requests
    _warm_runtime(gpu)
    free0 = _device_free_bytes()
    srv = gpu.Server(memory_budget=budget, max_inflight=inflight)
    peak = {"committed": 0, "in_use": 0, "device": 0}
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            st = srv.stats()
            peak["committed"] = max(peak["committed"], st["memory_committed"])
            peak["in_use"] = max(peak["in_use"], st["memory_in_use"])
            peak["device"] = max(peak["device"], free0 - _device_free_bytes())
            time.sleep(0.0005)

    th = threading.Thread(target=watch)
    th.start()
    try:
        tickets = [srv.submit(lv, b, w, (1, 1, 1), tcm, tcm) for lv, b, w, tcm in reqs]
        results = [srv.wait(t) for t in tickets]
        peak["device"] = max(peak["device"], free0 - _device_free_bytes())
    finally:
        stop.set()
        th.join()
        srv.close()
    print(f"budget {budget}: peak {peak}", flush=True)
    assert 0 < peak["in_use"] <= peak["committed"] <= budget, peak
    assert peak["device"] <= budget + SLACK, peak
    infos = [r[2] for r in results]
    # FIFO admission: admit times follow submit order
    admits = [i.admit_s for i in infos]
    assert all(a <= b for a, b in zip(admits, admits[1:]))
    assert all(i.status == 0 for i in infos)
    for (lv, b, w, tcm), (codes, errs, _) in zip(reqs[:12], results[:12]):
        alone = gpu.compress_many(lv, b, w, (1, 1, 1), tcm, tcm)
        assert np.array_equal(codes, alone[0])


def test_launches_overlap(gpu):
    # a long level-9 texture, then small requests: they are admitted and
    # launched on the second slot while the long one is still on the device.
    # The first round warms both launch slots (staging buffers, pool slabs).
    big, _ = synth.make_image_blocks(6, 2048, 2048, synth.MIX_DEFAULT)
    small = _requests(7, 16, max_side=64)
    srv = gpu.Server(memory_budget=1 << 30, max_inflight=2)
    try:
        for _ in range(2):
            tb = srv.submit(9, big, None, (1, 1, 1), True, True)
            ts = [srv.submit(lv, b, w, (1, 1, 1), tcm, tcm) for lv, b, w, tcm in small]
            big_info = srv.wait(tb)[2]
            infos = [srv.wait(t)[2] for t in ts]
    finally:
        srv.close()
    later = [i for i in infos if i.launch_s > big_info.launch_s]
    detail = (f"big launch {big_info.launch_s:.6f} done {big_info.done_s:.6f}; small "
              + ", ".join(f"({i.launch_s:.6f},{i.done_s:.6f},b{i.batch_size})" for i in infos))
    assert later, "small requests all joined the long request's launch: " + detail
    assert any(i.launch_s < big_info.done_s for i in later), "no launch overlapped the long one: " + detail


def test_destroy_drains_outstanding_requests(gpu):
    # requests still queued, admitted or on the device when the server is
    # destroyed complete first: their (staged) outputs are written before
    # destroy returns, as each would be encoded alone
    reqs = _requests(12, 24, max_side=128)
    srv = gpu.Server(memory_budget=64 << 20, max_inflight=2)
    try:
        tickets = [srv.submit(lv, b, w, (1, 1, 1), tcm, tcm) for lv, b, w, tcm in reqs]
        kept = [srv._keep[t] for t in tickets]
    finally:
        srv.close()
    for (lv, b, w, tcm), (_, _, codes, errs) in zip(reqs, kept):
        alone = gpu.compress_many(lv, b, w, (1, 1, 1), tcm, tcm)
        assert np.array_equal(codes, alone[0]), f"level {lv}, {len(b)} blocks: codes not written"
        assert np.array_equal(errs.view(np.uint32), alone[1].view(np.uint32))


def test_oversize_request_rejected(gpu):
    # a budget below the launch slots' minimum workspace is refused at creation
    with pytest.raises(gpu.IcbcError) as e:
        gpu.Server(memory_budget=1 << 20)
    assert e.value.code == gpu.ERR_BUDGET
    srv = gpu.Server(memory_budget=32 << 20, max_inflight=1)
    try:
        big = np.zeros((131072, 16, 4), np.float32)   # 32 MiB of texels
        with pytest.raises(gpu.IcbcError) as e:
            srv.submit(8, big)
        assert e.value.code == gpu.ERR_BUDGET
    finally:
        srv.close()


def test_cancel_pending_request(gpu):
    # a budget that holds one request at a time keeps the later ones pending
    blocks, _ = synth.make_image_blocks(3, 1024, 1024, synth.MIX_DEFAULT)   # 65536 blocks, ~17 MiB
    srv = gpu.Server(memory_budget=112 << 20, max_inflight=1)
    try:
        t1 = srv.submit(9, blocks, None, (1, 1, 1), True, True)
        t2 = srv.submit(9, blocks, None, (1, 1, 1), True, True)
        t3 = srv.submit(8, blocks[:64])
        srv.cancel(t3)
        with pytest.raises(gpu.IcbcError) as e:
            srv.wait(t3)
        assert e.value.code == gpu.ERR_CANCELLED
        c1, e1, _ = srv.wait(t1)
        c2, e2, _ = srv.wait(t2)
        assert np.array_equal(c1, c2)
    finally:
        srv.close()


def test_invalid_level_request(gpu):
    blocks, _ = synth.make_image_blocks(4, 32, 32, synth.MIX_DEFAULT)
    srv = gpu.Server(memory_budget=32 << 20, max_inflight=1)
    try:
        prior = np.full((len(blocks), 8), 0xab, np.uint8)
        t = srv.submit(11, blocks, codes=prior)
        codes, errs, info = srv.wait(t)
        assert info.status == 0
        assert np.all(codes == 0xab)
        assert np.all(errs == np.finfo(np.float32).max)
    finally:
        srv.close()


def test_concurrent_clients(gpu, oracle):
    reqs = _requests(5, 24)
    srv = gpu.Server(memory_budget=64 << 20, max_inflight=2)
    results = {}

    def client(k):
        lv, b, w, tcm = reqs[k]
        t = srv.submit(lv, b, w, (1, 1, 1), tcm, tcm)
        results[k] = srv.wait(t)

    try:
        ths = [threading.Thread(target=client, args=(k,)) for k in range(len(reqs))]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    finally:
        srv.close()
    for k, (lv, b, w, tcm) in enumerate(reqs):
        want = _oracle(oracle, lv, b, w, tcm)
        assert np.array_equal(results[k][0], want[0])


def test_identity_under_load(gpu):
    # many overlapping requests over a small pool of textures, inputs in
    # page-locked memory: every served result equals the texture encoded alone
    reqs = _requests(8, 20, max_side=192)
    pinned = [gpu.pinned_copy(b) for _, b, _, _ in reqs]
    alone = [gpu.compress_many(lv, p.array, w, (1, 1, 1), tcm, tcm)[0]
             for (lv, _, w, tcm), p in zip(reqs, pinned)]
    srv = gpu.Server(memory_budget=96 << 20, max_inflight=3)
    rng = np.random.default_rng(9)
    try:
        order = rng.integers(0, len(reqs), size=400)
        tickets = []
        for k in order:
            lv, _, w, tcm = reqs[k]
            codes = np.empty((len(pinned[k].array), 8), np.uint8)
            tickets.append((srv.submit(lv, pinned[k].array, w, (1, 1, 1), tcm, tcm, codes=codes), k))
        for t, k in tickets:
            codes, _, info = srv.wait(t)
            assert np.array_equal(codes, alone[k]), f"request {t} (texture {k}, batch of {info.batch_size})"
        assert srv.stats()["launches"] > 3
    finally:
        srv.close()
        for p in pinned:
            p.close()


def test_mapped_and_staged_buffers(gpu, oracle):
    # one batch mixing page-locked buffers (read and written in place by the
    # copy kernels through their device mapping) with pageable ones (staged),
    # at 16-byte aligned and unaligned addresses, with odd block counts
    reqs = _requests(12, 24, max_side=160)
    # and buffers of 1 MiB or more (copied by DMA), page-locked and pageable
    for k in range(3):
        big, _ = synth.make_image_blocks(40 + k, 320, 256, synth.MIX_DEFAULT)   # 5,120 blocks, 1.25 MiB
        reqs.append((8 + k % 2, big, None, bool(k % 2)))
    rng = np.random.default_rng(12)
    holders = []

    def place(a, pinned, shift):
        if not pinned:
            return a.copy()
        h = gpu.PinnedArray((a.nbytes + 16,), np.uint8)
        holders.append(h)
        v = h.array[shift:shift + a.nbytes].view(a.dtype).reshape(a.shape)
        v[...] = a
        return v

    srv = gpu.Server(memory_budget=96 << 20, max_inflight=2)
    try:
        tickets = []
        for k, (lv, b, w, tcm) in enumerate(reqs):
            if k % 3 == 0 and len(b) % 2 == 0:  # an odd block count
                b = b[: len(b) - 1]
                w = None if w is None else w[: len(b)]
            pin = rng.integers(0, 2, size=4).astype(bool)
            shift = int(rng.choice([0, 4, 8]))
            n = len(b)
            colors = place(np.ascontiguousarray(b, np.float32), pin[0], shift)
            wts = None if w is None else place(np.ascontiguousarray(w, np.float32), pin[1], shift)
            codes = place(np.zeros((n, 8), np.uint8), pin[2], shift)
            errs = place(np.zeros(n, np.float32), pin[3], shift)
            tickets.append((srv.submit(lv, colors, wts, (1, 1, 1), tcm, tcm, codes=codes, errors=errs),
                            lv, np.array(colors), None if wts is None else np.array(wts), tcm))
        for t, lv, b, w, tcm in tickets:
            codes, errs, info = srv.wait(t)
            assert info.status == 0
            want = _oracle(oracle, lv, b, w, tcm)
            assert np.array_equal(codes, want[0]), f"request {t} codes"
            assert np.array_equal(errs.view(np.uint32), want[1].view(np.uint32)), f"request {t} errors"
    finally:
        srv.close()
        for h in holders:
            h.close()


def test_two_waits_on_one_ticket(gpu):
    # one wait collects the request; the other, blocked on the same ticket,
    # gets an argument error instead of touching the released request
    blocks, _ = synth.make_image_blocks(6, 512, 512, synth.MIX_DEFAULT)
    srv = gpu.Server(memory_budget=64 << 20)
    outcome = []

    def waiter(t):
        try:
            outcome.append(("ok", srv.wait(t)[2].status))
        except gpu.IcbcError as e:
            outcome.append(("error", e.code))

    try:
        t = srv.submit(9, blocks, None, (1, 1, 1), True, True)
        ths = [threading.Thread(target=waiter, args=(t,)) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
    finally:
        srv.close()
    assert sorted(outcome) == [("error", gpu.ERR_INVALID_ARGUMENT), ("ok", 0)], outcome


def test_serve_bench_program(gpu, oracle, tmp_path):
    # the C++ serving client (icbc_amd/serve_bench) on a small RGBA8 pool:
    # every served result equals the texture encoded alone, which equals the oracle
    import json
    import subprocess
    from icbc_amd import build as hip_build
    reqs = _requests(14, 12, max_side=160)
    pool = [(lv, np.clip(np.round(b * 255), 0, 255).astype(np.uint8), tcm) for lv, b, _, tcm in reqs]
    src, alone = tmp_path / "pool.bin", tmp_path / "alone.bin"
    with open(src, "wb") as f:
        f.write(np.array([len(pool), 0], np.uint32).tobytes())
        for lv, b, tcm in pool:
            f.write(np.array([len(b)], np.uint64).tobytes() + np.array([lv, int(tcm)], np.int32).tobytes())
        for _, b, _ in pool:
            f.write(np.ascontiguousarray(b).tobytes())
    res = subprocess.run([hip_build.SERVE_BIN, str(src), str(alone), "0.3", "0.5", str(256 << 20), "2", "32", "0.3",
                          "3"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    r = json.loads(res.stdout.strip().splitlines()[-1])
    assert r["saturation_results_identical_to_alone"] and r["saturation_requests"] > 0
    assert all(run["results_identical_to_alone"] for run in r["runs"])
    total = sum(len(b) for _, b, _ in pool)
    raw = np.fromfile(alone, np.uint8)
    codes, errs = raw[:8 * total].reshape(total, 8), raw[8 * total:].view(np.float32)
    off = 0
    for lv, b, tcm in pool:
        want = _oracle(oracle, lv, b.astype(np.float32) / np.float32(255), None, tcm)
        assert np.array_equal(codes[off:off + len(b)], want[0])
        assert np.array_equal(errs[off:off + len(b)].view(np.uint32), want[1].view(np.uint32))
        off += len(b)


def test_backlog_waits_in_the_host_queue(gpu):
    # a burst far above what the device holds at once, with a budget that
    # could hold all of it: requests are admitted only while fewer than two
    # launches' worth of blocks (2^21) wait, so the memory in use stays near
    # that bound and the backlog waits in the host-side queue
    blocks, _ = synth.make_image_blocks(21, 512, 512, synth.MIX_DEFAULT)   # 16,384 blocks
    pinned = gpu.pinned_copy(np.clip(np.round(blocks * 255), 0, 255).astype(np.uint8))
    srv = gpu.Server(memory_budget=16 << 30, max_inflight=3)
    peak = {"in_use": 0, "pending": 0}
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            st = srv.stats()
            peak["in_use"] = max(peak["in_use"], st["memory_in_use"])
            peak["pending"] = max(peak["pending"], st["pending"])
            time.sleep(0.0002)

    th = threading.Thread(target=watch)
    th.start()
    try:
        outs = [np.empty((len(blocks), 8), np.uint8) for _ in range(600)]   # 9.8 M blocks in all
        tickets = [srv.submit(9, pinned.array, None, (1, 1, 1), True, True, codes=o, with_errors=False)
                   for o in outs]
        for t in tickets:
            srv.wait(t)
    finally:
        stop.set()
        th.join()
        srv.close()
        pinned.close()
    alone = gpu.compress_many(9, np.clip(np.round(blocks * 255), 0, 255).astype(np.uint8), None, (1, 1, 1), True,
                              True)[0]
    assert all(np.array_equal(o, alone) for o in outs)
    # in use: admitted requests (admission stops once 2^21 blocks wait, so at
    # most one request past that), 3 launches of at most 2^20 blocks each,
    # every request's 64 B texels and 8 B codes in a power-of-two class, each
    # launch's level-9 list (16 B per block, a 32 MiB class) and small tables:
    # about 0.75 GiB, against the 1.2 GiB all 600 requests would take at once
    per_request = 2 << 20   # 16,384 x 72 B = 1.1 MiB -> its 2 MiB class
    waiting = (1 << 21) // len(blocks) + 1
    launched = 3 * ((1 << 20) // len(blocks))
    bound = (waiting + launched) * per_request + 3 * (32 << 20) + (16 << 20)
    assert bound < len(outs) * per_request
    assert 0 < peak["in_use"] <= bound, (peak, bound)
    assert peak["pending"] > 50, peak
