"""GPU tests of the autotuner: it returns a valid per-level table, installs
it for the device, and changes only speed -- codes and errors from tuned
launches equal the untuned ones."""
import numpy as np
import pytest

from icbc_amd import synth

pytestmark = pytest.mark.gpu


def _sample(gpu, n_tex=6):
    texs = synth.texture_set(17, count=n_tex, min_side=64, max_side=512)
    counts = [len(t[5]) for t in texs]
    total = sum(counts)
    colors = np.concatenate([t[5] for t in texs])
    dcol = gpu.DeviceBuffer.from_array(colors)
    dcode = gpu.DeviceBuffer(8 * total)
    derr = gpu.DeviceBuffer(4 * total)
    jobs, off = [], 0
    for t, n in zip(texs, counts):
        jobs.append(gpu.make_job(t[4], n, dcol.ptr + 256 * off, dcode.ptr + 8 * off, None, derr.ptr + 4 * off,
                                 None, (1, 1, 1), True, True))
        off += n
    return jobs, total, (dcol, dcode, derr)


def test_autotune_table_and_identical_results(gpu):
    jobs, total, (dcol, dcode, derr) = _sample(gpu)
    s = gpu.Stream()
    try:
        gpu.set_tuning(None)
        plan = gpu.Plan(jobs, stream=s)
        plan.execute(s)
        s.synchronize()
        base = dcode.download(np.uint8, (total, 8)), derr.download(np.float32, (total,))
        plan.close()

        t = gpu.autotune(jobs, levels=[1, 8, 9])
        table = t.as_dict()
        assert sorted(table) == [1, 8, 9]
        for lv, e in table.items():
            assert e["tile_blocks"] in (8, 16, 32, 64) and e["ns_per_block"] > 0
        assert gpu.get_tuning().as_dict() == table
        for lv, e in table.items():
            # the untuned default is among the timed settings: the pick is never slower
            assert 0 < e["ns_per_block"] <= e["default_ns_per_block"]
            assert e["prune_min_colors"] in (9, 11, 13, 15, 17)
        # level 9 costs more per block than level 1
        assert table[9]["ns_per_block"] > table[1]["ns_per_block"]

        # tuned defaults (tile sizes, plan occupancy/schedule, cost order): same results
        dcode.fill(0)
        plan = gpu.Plan(jobs, stream=s)
        for _ in range(2):
            plan.execute(s)
            s.synchronize()
            assert np.array_equal(dcode.download(np.uint8, (total, 8)), base[0])
            assert np.array_equal(derr.download(np.float32, (total,)).view(np.uint32), base[1].view(np.uint32))
        plan.close()
        dcode.fill(0)
        gpu.compress_batch(jobs, stream=s)
        s.synchronize()
        assert np.array_equal(dcode.download(np.uint8, (total, 8)), base[0])
    finally:
        gpu.set_tuning(None)
        for b in (dcol, dcode, derr):
            b.close()
    assert gpu.get_tuning().as_dict() == {}


@pytest.mark.parametrize("prune", [9, 11, 15, 17])
def test_search_schedule_does_not_change_results(gpu, oracle, prune):
    # the four-cluster search prunes from another colour count (17: never):
    # codes and errors equal the oracle's at every cluster level
    colors, _ = synth.make_image_blocks(23, 128, 128, synth.MIX_DEFAULT)
    edge, weights = synth.make_edge_blocks(24, 1024)
    t = gpu.Tuning()
    for lv in range(10):
        t.tile_blocks[lv], t.schedule[lv], t.ns_per_block[lv], t.prune_min_colors[lv] = 64, 1, 1.0, prune
    try:
        gpu.set_tuning(t)
        for lv in (2, 5, 8, 9):
            for c, w in ((colors, None), (edge, weights)):
                got = gpu.compress_many(lv, c, w, (1, 1, 1), True, True)
                want = oracle.compress_many(lv, c, w, (1, 1, 1), True, True, nthreads=8)
                assert np.array_equal(got[0], want[0]), (prune, lv)
                assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32)), (prune, lv)
    finally:
        gpu.set_tuning(None)


def test_set_tuning_validates(gpu):
    t = gpu.Tuning()
    t.tile_blocks[3] = 12
    with pytest.raises(gpu.IcbcError):
        gpu.set_tuning(t)
    t.tile_blocks[3] = 16
    t.prune_min_colors[3] = 5   # below the pruned search's minimum of 9 colours
    with pytest.raises(gpu.IcbcError):
        gpu.set_tuning(t)
    t.prune_min_colors[3] = 0
    t.tile_blocks[3] = 16
    t.schedule[3] = 1
    t.ns_per_block[3] = 1.0
    try:
        gpu.set_tuning(t)
        assert gpu.get_tuning().as_dict()[3]["tile_blocks"] == 16
    finally:
        gpu.set_tuning(None)
