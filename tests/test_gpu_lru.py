"""LRUCache on the GPU vs the oracle: every slot's age, availability and item,
and every item's slot, bit-exact after each update (js/lru_cache.js)."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from gpu_common import dev, say  # noqa: F401

pytestmark = pytest.mark.gpu


def _needed(total, rng, frac):
    f = np.zeros(total, np.uint32)
    lo = int(rng.integers(0, total))
    f[lo:lo + int(rng.integers(1, max(2, int(total * frac))))] = 1
    f[rng.uniform(0, 1, total) < frac / 4] = 1
    return f


@pytest.mark.parametrize("total,initial,steps", [(100, 32, 10), (1000, 64, 15), (5000, 200, 12),
                                                 (20000, 20000, 6), (70000, 3500, 8)])
def test_lru_sequence_matches_oracle(dev, total, initial, steps):
    say("lru", total, initial, steps)
    rng = np.random.default_rng(total + initial)
    gpu = bcmc_amd.LRUCache(dev, None, None, initial, 256, total)
    cpu = oracle.LRU(initial, 256, total)
    for step in range(steps):
        needed = _needed(total, rng, 0.3)
        n_g, ids_g = gpu.update(dev.buffer_from(needed))
        n_c, ids_c = cpu.update(needed)
        assert n_g == n_c, step
        if n_c:
            assert np.array_equal(ids_g.download(np.uint32, n_g), ids_c)
        for a, b, what in zip(gpu.state(), cpu.state(), ("age", "avail", "item", "item_slot")):
            assert np.array_equal(a, b), (step, what)
        if step == steps // 2:
            gpu.reset()
            cpu.reset()


def test_lru_no_new_items_is_age_only(dev):
    gpu = bcmc_amd.LRUCache(dev, None, None, 32, 256, 64)
    cpu = oracle.LRU(32, 256, 64)
    f = np.zeros(64, np.uint32)
    f[:5] = 1
    for _ in range(3):
        assert gpu.update(dev.buffer_from(f))[0] == cpu.update(f)[0]
    for a, b in zip(gpu.state(), cpu.state()):
        assert np.array_equal(a, b)


def test_lru_long_run_after_reset_matches_oracle(dev):
    """150 updates after a reset, with growth and many equal ages: the
    eviction order sorts compressed ages after a reset (one radix pass for the
    first 127 updates, two from then on) and must give the oracle's slots at
    every step, before the reset (constructor ages, raw keys) as well."""
    total, initial = 3000, 96
    rng = np.random.default_rng(7)
    gpu = bcmc_amd.LRUCache(dev, None, None, initial, 256, total)
    cpu = oracle.LRU(initial, 256, total)

    def step(tag):
        needed = _needed(total, rng, 0.02 + 0.2 * rng.uniform())
        n_g, ids_g = gpu.update(dev.buffer_from(needed))
        n_c, ids_c = cpu.update(needed)
        assert n_g == n_c, tag
        if n_c:
            assert np.array_equal(ids_g.download(np.uint32, n_g), ids_c), tag

    def same_state(tag):
        for a, b, what in zip(gpu.state(), cpu.state(), ("age", "avail", "item", "item_slot")):
            assert np.array_equal(a, b), (tag, what)

    for k in range(4):
        step(("before reset", k))
    same_state("before reset")
    gpu.reset()
    cpu.reset()
    for k in range(150):
        step(k)
        if k % 10 == 0 or 125 <= k <= 130:
            same_state(k)
    same_state("end")


def test_lru_perf_tracker_has_the_reference_keys(dev):
    """LRUCache.update(needed, perfTracker): one entry per update under each
    of the reference's 11 lru* keys; counts equal the oracle's; with no new
    items every stage after the count is 0 (lru_cache.js:468-480)."""
    total, initial = 5000, 100
    gpu = bcmc_amd.LRUCache(dev, None, None, initial, 256, total)
    cpu = oracle.LRU(initial, 256, total)
    size0 = len(cpu.state()[0])
    needed = np.zeros(total, np.uint32)
    needed[1000:2500] = 1
    perf = {}
    for _ in range(2):
        n_g, _ = gpu.update(dev.buffer_from(needed), perf)
        assert n_g == cpu.update(needed)[0]
    assert list(perf) == list(bcmc_amd.LRU_PERF_KEYS)
    assert all(len(v) == 2 for v in perf.values())
    assert perf["lruNumNewItems"] == [1500, 0]
    assert perf["lruNumSlotsAvailable"][0] == size0
    assert perf["lruGrowCache"][0] > 0 and perf["lruTotalSortTime"][0] > 0 and perf["lruWriteNewItems"][0] > 0
    assert perf["lruMarkNewItems"][1] > 0
    assert all(perf[k][1] == 0 for k in bcmc_amd.LRU_PERF_KEYS[3:])
