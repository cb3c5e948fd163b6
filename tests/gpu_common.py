"""Shared helpers for the GPU parity tests (HIP path vs the CPU oracle)."""
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = np.load(os.path.join(ROOT, "tests", "golden", "zfp_fixtures.npz"))
META = json.loads(bytes(FIX["meta"]).decode())

_dev = None


def device():
    global _dev
    import bcmc_amd
    if _dev is None:
        _dev = bcmc_amd.Device(0)
    return _dev


@pytest.fixture(scope="module")
def dev():
    return device()


@pytest.fixture
def replay_dev():
    """The device with graph replay of the query chains on for one test."""
    d = device()
    d.set_graph_replay(True)
    yield d
    d.set_graph_replay(False)


def say(*a):
    print(*a, flush=True)
