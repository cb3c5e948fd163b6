"""The isovalue benchmarks of the JS facade (js/run_benchmark.js:1-47, driven
as js/bcmc.js:319-370 drives them) on a stand-in pipeline: reset first, 100
isovalues in the reference's order, one computeSurface per changed value,
the perf arrays and the saved file name."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_benchmark_protocol_on_a_stand_in(tmp_path):
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "benchmark_protocol.js"), str(tmp_path),
                        "mock"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads((tmp_path / "out.json").read_text())
    # sweeps: the value moves by (hi - lo) / 100 per step, accumulated in doubles
    for kind, start, sign in (("sweepUp", 0.1, 1), ("sweepDown", 0.9, -1)):
        want, cur = [], start
        for _ in range(100):
            cur += sign * ((0.9 - 0.1) / 100)
            want.append(cur)
        o = out[kind]
        assert o["calls"] == ["reset"] + want
        assert o["fileName"] == f"bcmc-explicit-skull-{kind}.json"
        assert o["perf"]["isovalue"] == want
        assert len(o["perf"]["totalTime"]) == 100 and all(t >= 0 for t in o["perf"]["totalTime"])
        assert o["perf"]["memUse"] == [{"mc": {"total": 1}}] * 100
        assert o["perf"]["numVertices"] == [7] * 100
    # random: U[lo, hi) from the given source; without the cache the LRU is
    # reset before every step
    seed, want = 1, []
    for _ in range(100):
        seed = (seed * 48271) % 2147483647
        want.append(seed / 2147483647 * 0.5 + 0.25)
    o = out["random"]
    assert o["calls"] == ["reset"] + [c for w in want for c in ("reset", w)]
    assert o["fileName"] == "bcmc-explicit-volume-random.json"
    assert out["unknownThrows"] and out["alias"]


def test_python_mirror_benchmark_values():
    import bcmc_amd
    up = bcmc_amd.isovalue_benchmark_values("sweepUp", (0.1, 0.9))
    down = bcmc_amd.isovalue_benchmark_values("sweepDown", (0.1, 0.9))
    assert len(up) == len(down) == 100
    assert abs(up[-1] - 0.9) < 1e-12 and abs(down[-1] - 0.1) < 1e-12
    assert all(b > a for a, b in zip(up, up[1:])) and all(b < a for a, b in zip(down, down[1:]))
    vals = iter([0.0, 0.5, 0.999])
    r = bcmc_amd.isovalue_benchmark_values("random", (0.2, 0.6), random=lambda: next(vals, 0.25))
    assert r[:3] == [0.2, 0.5 * (0.6 - 0.2) + 0.2, 0.999 * (0.6 - 0.2) + 0.2] and len(r) == 100
    with pytest.raises(ValueError):
        bcmc_amd.isovalue_benchmark_values("zigzag", (0, 1))
