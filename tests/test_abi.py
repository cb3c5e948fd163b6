"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point include/bcmc.h declares, and reports a missing device cleanly."""
import ctypes
import os
import subprocess

import pytest

import bcmc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    syms = bcmc_amd.exported_symbols()
    assert len(syms) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", bcmc_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = bcmc_amd.lib()
    for s in syms:
        assert getattr(lib, s) is not None


def test_version_and_pure_helpers():
    lib = bcmc_amd.lib()
    assert b"gfx950" in lib.bcmc_version()
    # getAlignedSize of ExclusiveScanPipeline (512) and RadixSorter (pow2 chunks of 64)
    assert [lib.bcmc_scan_aligned_size(n) for n in (0, 1, 512, 513)] == [0, 512, 512, 1024]
    assert [lib.bcmc_sort_aligned_size(n) for n in (1, 64, 65, 129, 1000)] == [64, 64, 128, 256, 1024]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_an_error_not_a_fallback():
    with pytest.raises(bcmc_amd.BcmcError) as e:
        bcmc_amd.Device(0)
    assert e.value.code == -6


def test_invalid_arguments_return_codes():
    lib = bcmc_amd.lib()
    h = ctypes.c_void_p()
    assert lib.bcmc_ctx_create(0, None) == -1
    assert lib.bcmc_buf_alloc(None, 16, ctypes.byref(h)) == -1
    assert b"invalid argument" in lib.bcmc_last_error()
    assert lib.bcmc_mc_compute_surface(None, 0.5, None, None) == -1


def test_tri_table_packed_in_library_source_matches_oracle_header():
    # Both copies of the table are generated from the same data; the hash is
    # checked in test_oracle.py.  Here: the kernel header is what mc.hip uses.
    src = open(os.path.join(ROOT, "webgpu-bcmc_amd", "csrc", "mc_device.h")).read()
    assert "BCMC_TRI_EDGES_INIT" in src and "BCMC_TRI_COUNT_INIT" in src
