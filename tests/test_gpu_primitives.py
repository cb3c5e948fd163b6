"""Scan, compaction and sort on the GPU vs the oracle (bit-exact)."""
import numpy as np
import pytest

import bcmc_amd
import oracle
from gpu_common import dev, say  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 5, 511, 512, 4095, 4096, 4097, 100000, 262145, 1000003])
def test_exclusive_scan(dev, n):
    say("scan n", n)
    rng = np.random.default_rng(n)
    data = rng.integers(0, 1000, n).astype(np.uint32)
    pipe = bcmc_amd.ExclusiveScanPipeline(dev)
    scanner = pipe.prepareInput(data)
    total = scanner.scan(n)
    got = scanner.inputBuf.download(np.uint32, pipe.getAlignedSize(n))
    want, wtotal = oracle.exclusive_scan(data)
    assert total == wtotal
    assert np.array_equal(got[:n], want)
    if n < pipe.getAlignedSize(n):
        assert got[n] == wtotal  # the reference leaves the total at data[n]


def test_exclusive_scan_wraps(dev):
    data = np.full(9000, 0x40000000, np.uint32)
    pipe = bcmc_amd.ExclusiveScanPipeline(dev)
    sc = pipe.prepareInput(data)
    total = sc.scan(9000)
    want, wt = oracle.exclusive_scan(data)
    assert total == wt and np.array_equal(sc.inputBuf.download(np.uint32, 9000), want)


def test_scan_repeated_launches_reuse_state(dev):
    # the look-back state resets itself between launches (epochs)
    pipe = bcmc_amd.ExclusiveScanPipeline(dev)
    rng = np.random.default_rng(3)
    for it in range(40):
        n = int(rng.integers(1, 70000))
        data = rng.integers(0, 3, n).astype(np.uint32)
        sc = pipe.prepareInput(data)
        assert sc.scan(n) == int(data.sum())


@pytest.mark.parametrize("n", [1, 100, 4096, 70001])
def test_compact(dev, n):
    say("compact n", n)
    rng = np.random.default_rng(n + 1)
    flags = (rng.uniform(0, 1, n) < 0.4).astype(np.uint32)
    offs, total = oracle.exclusive_scan(flags)
    fb, ob = dev.buffer_from(flags), dev.buffer_from(offs)
    out = dev.create_buffer(max(total, 1) * 4)
    bcmc_amd.StreamCompact(dev).compactActiveIDs(n, fb, ob, out)
    assert np.array_equal(out.download(np.uint32, total), np.flatnonzero(flags))
    out2 = dev.create_buffer(n * 4)
    assert bcmc_amd.StreamCompactor(dev).compactFlags(n, fb, out2) == total
    assert np.array_equal(out2.download(np.uint32, total), np.flatnonzero(flags))


@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 1000, 5000, 70000])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("keyset", ["distinct", "few", "ages", "with_max"])
def test_sort(dev, n, reverse, keyset):
    say("sort", n, reverse, keyset)
    rng = np.random.default_rng(n * 3 + reverse)
    if keyset == "distinct":
        keys = rng.permutation(4 * n)[:n].astype(np.uint32)
    elif keyset == "few":
        keys = rng.integers(0, 5, n).astype(np.uint32)
    elif keyset == "ages":
        keys = (100000 + rng.integers(0, 3, n)).astype(np.uint32)
    else:
        keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        keys[rng.uniform(0, 1, n) < 0.2] = 0xFFFFFFFF
    vals = rng.permutation(n).astype(np.uint32)
    sorter = bcmc_amd.RadixSorter(dev)
    a = sorter.getAlignedSize(n)
    kpad = np.zeros(a, np.uint32)
    vpad = np.zeros(a, np.uint32)
    kpad[:n], vpad[:n] = keys, vals
    kb, vb = dev.buffer_from(kpad), dev.buffer_from(vpad)
    sorter.sort(kb, vb, n, reverse)
    want, _ = oracle.sort_by_key(keys, vals, reverse)
    assert np.array_equal(vb.download(np.uint32, n), want)


def test_scan_after_scratch_growth(dev):
    """Each larger scan reallocates the context's look-back state; the first
    scan on the new state must not see the previous state's status words."""
    d = bcmc_amd.Device(0)
    rng = np.random.default_rng(12)
    for n in (3000, 70000, 300000, 1200000, 2500000):
        data = rng.integers(0, 1000, n).astype(np.uint32)
        pipe = bcmc_amd.ExclusiveScanPipeline(d)
        sc = pipe.prepareInput(data)
        total = sc.scan(n)
        want, wt = oracle.exclusive_scan(data)
        assert total == wt, n
        assert np.array_equal(sc.inputBuf.download(np.uint32, n), want), n
