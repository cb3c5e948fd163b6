"""The scan entry point that allocates nothing (bcmc_scan_exclusive_u32_noalloc),
called through the C ABI and checked against bcmc_scan_exclusive_u32 and the
oracle on two calls in a row, the second with a smaller input."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import bcmc_amd
import oracle
from gpu_common import ROOT, dev, say  # noqa: F401

pytestmark = pytest.mark.gpu


def _scan_both(dev, data, n):
    """Scans `data` (aligned length) with both entry points; returns their
    results and totals."""
    L = bcmc_amd.lib()
    aligned = len(data)
    out = []
    for fn in (L.bcmc_scan_exclusive_u32, L.bcmc_scan_exclusive_u32_noalloc):
        buf = dev.buffer_from(data)
        total = ctypes.c_uint32()
        rc = fn(dev.h, buf.h, aligned, n, ctypes.byref(total))
        assert rc == 0, L.bcmc_last_error()
        out.append((buf.download(np.uint32, aligned), total.value))
        buf.destroy()
    return out


@pytest.mark.parametrize("sizes", [(3_000_011, 700_001), (1 << 22, 8192), (100_000, 1)])
def test_noalloc_scan_matches_allocating_scan(dev, sizes):
    rng = np.random.default_rng(sizes[0])
    for n in sizes:
        say(f"noalloc scan n={n}")
        aligned = int(bcmc_amd.lib().bcmc_scan_aligned_size(n))
        data = np.zeros(aligned, np.uint32)
        data[:n] = rng.integers(0, 1 << 20, n, dtype=np.uint32)
        (a, ta), (b, tb) = _scan_both(dev, data, n)
        assert np.array_equal(a, b) and ta == tb
        ref, tref = oracle.exclusive_scan(data[:n])
        assert np.array_equal(b[:n], ref) and tb == tref
        if n < aligned:
            assert b[n] == tref


def test_noalloc_scan_rejects_sizes_past_its_maximum(dev):
    L = bcmc_amd.lib()
    buf = dev.buffer_from(np.zeros(512, np.uint32))
    n = (1 << 30) + 512
    total = ctypes.c_uint32()
    rc = L.bcmc_scan_exclusive_u32_noalloc(dev.h, buf.h, n, n, ctypes.byref(total))
    assert rc == -1 and b"maximum" in L.bcmc_last_error()
    buf.destroy()


_CHILD = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import bcmc_amd
L = bcmc_amd.lib()
dev = bcmc_amd.Device(0)
data = np.arange(1 << 20, dtype=np.uint32) % 7
buf = dev.buffer_from(data)   # the only device allocation the injection lets through
t = ctypes.c_uint32()
for n in (1 << 20, 5000):
    assert L.bcmc_scan_exclusive_u32_noalloc(dev.h, buf.h, 1 << 20, n, ctypes.byref(t)) == 0, \
        L.bcmc_last_error()
    buf.upload(data)
want = int(data[:5000].sum())
assert t.value == want, (t.value, want)
# the allocating scan needs its context scratch: the injected failure hits it
rc = L.bcmc_scan_exclusive_u32(dev.h, buf.h, 1 << 20, 5000, ctypes.byref(t))
assert rc == -3, rc
print("child ok", flush=True)
"""


def test_noalloc_scan_makes_no_device_allocation():
    """In a fresh process whose second device allocation is made to fail
    (BCMC_FAIL_ALLOC_AFTER=2), the no-allocation scan runs twice while the
    allocating scan fails on its scratch."""
    env = dict(os.environ, BCMC_FAIL_ALLOC_AFTER="2")
    say("noalloc scan: child process with injected allocation failure")
    r = subprocess.run([sys.executable, "-c", _CHILD, os.path.join(ROOT, "webgpu-bcmc_amd", "python")],
                       env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "child ok" in r.stdout, r.stdout + r.stderr
