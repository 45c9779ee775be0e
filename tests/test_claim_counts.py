"""The tile queues and the level-9 handoff list put their own claim counters
back to zero: the claim that returns the last value a launch can hand out
resets the counter (icbc_amd/csrc/bc1_encoder.hip), so plans launch no reset
kernel.  These tests replay the kernels' claim rules on the CPU over random
interleavings of workgroups and wavefronts and check that the number of
claims, and so the last claim's value, is the one the kernels compare with."""
import random

import pytest


def test_queue_takes_exactly_its_share_of_claims():
    # bc1_encode_kernel: workgroup g is on queue x = g % 8; its first tile is
    # static (index g // 8 of the queue's share), after each tile it claims
    # v = counter++ and goes on with index static_claims + v while that is
    # below the share.  The kernels reset the counter on the claim v == share-1.
    rng = random.Random(1)
    for _ in range(3000):
        grid, ntiles = rng.randint(1, 300), rng.randint(0, 3000)
        for x in range(min(8, grid)):
            share = (ntiles - x + 7) // 8 if ntiles > x else 0
            static = (grid - x + 7) // 8
            index = {g: g // 8 for g in range(grid) if g % 8 == x}
            running = [g for g, j in index.items() if j < share]
            counter, claims = 0, []
            while running:
                g = rng.choice(running)
                claims.append(counter)
                index[g] = static + counter
                counter += 1
                if index[g] >= share:
                    running.remove(g)
            assert len(claims) == share
            if share:
                assert claims[-1] == share - 1 == max(claims)


@pytest.mark.parametrize("chunk", [64, 32, 16, 8, 4])
def test_handoff_list_last_claim(chunk):
    # bc1_refine_kernel: each wavefront claims `chunk` blocks at a time
    # (base = counter; counter += chunk) until a claim drains the list
    # (base + chunk >= total), then never again; the kernels reset both list
    # counters on the claim whose base is (below + wavefronts - 1) * chunk.
    rng = random.Random(chunk)
    for _ in range(3000):
        waves, total = rng.randint(1, 64), rng.randint(0, 5000)
        below = (total + chunk - 1) // chunk - 1 if total > chunk else 0
        last = (below + waves - 1) * chunk
        counter, claims, running = 0, [], list(range(waves))
        while running:
            w = rng.choice(running)
            claims.append(counter)
            counter += chunk
            if claims[-1] + chunk >= total:
                running.remove(w)
        assert claims[-1] == last and claims.count(last) == 1
