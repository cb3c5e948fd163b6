"""zfp volume files named the reference's way (js/volumes.js:7-64 names,
js/volumes.js:147-152 dims rule): dims and rate from the name, headerless
stream bytes from the file."""
import numpy as np
import pytest

import bcmc_amd
from bcmc_amd import synth


@pytest.mark.parametrize("name,dims,rate", [
    ("plane_x_32x32x32_float32.gen.crate2", (32, 32, 32), 2),
    ("fuel_64x64x64_uint8.raw.crate4", (64, 64, 64), 4),
    ("duct_193x194x1000_float32.raw.crate2", (193, 194, 1000), 2),
    ("richtmyer_meshkov_2048x2048x1920_uint8.raw.crate1", (2048, 2048, 1920), 1),
    ("/data/volumes/jicf_q_1408x1080x1100_float32.raw.crate2", (1408, 1080, 1100), 2),
    ("skull_256x256x256_uint8.raw", (256, 256, 256), None),
])
def test_crate_info(name, dims, rate):
    assert bcmc_amd.crate_info(name) == (dims, rate)


def test_crate_info_rejects_names_without_dims():
    with pytest.raises(ValueError):
        bcmc_amd.crate_info("volume.crate2")


def test_load_crate_round_trip(tmp_path):
    f = synth.field(synth.GAUSS, 4, (13, 10, 7), 3, (2, 4, 0.5, 1))
    s = synth.encode(f, 4 * 64)
    p = tmp_path / "blob_13x10x7_float32.raw.crate4"
    p.write_bytes(s.tobytes())
    data, rate, dims = bcmc_amd.load_crate(str(p))
    assert (rate, dims) == (4, (13, 10, 7))
    assert np.array_equal(data, s)
    assert bcmc_amd.load_crate(str(p), rate=2)[1] == 2
    with pytest.raises(ValueError):
        bcmc_amd.load_crate(str(p), rate=8)  # too short for rate 8
    q = tmp_path / "blob_13x10x7_float32.raw"
    q.write_bytes(s.tobytes())
    with pytest.raises(ValueError):
        bcmc_amd.load_crate(str(q))  # no rate in the name and none given


def test_stream_bytes():
    assert bcmc_amd.stream_bytes((16, 16, 16), 64) == 512
    assert bcmc_amd.stream_bytes((12, 8, 8), 96) == 144
    assert bcmc_amd.stream_bytes((5, 4, 3), 9) == 8
