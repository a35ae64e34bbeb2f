"""MRC2014 I/O and the reference's RELION5 loaders (host code)."""
import struct

import numpy as np
import pytest

from matcha_amd import mrc

RNG = np.random.default_rng(5)


@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.int8, np.uint16, np.float16])
def test_write_read_round_trip(tmp_path, dtype):
    a = (RNG.standard_normal((5, 6, 7)) * 20).astype(dtype)
    p = tmp_path / "v.mrc"
    mrc.write_mrc(p, a, voxel_size=2.5)
    data, h = mrc.read_mrc(p)
    assert h.shape == (5, 6, 7) and data.shape == (5, 6, 7)
    np.testing.assert_array_equal(np.asarray(data), a)
    np.testing.assert_allclose(h.voxel_size, (2.5, 2.5, 2.5), rtol=1e-6)
    # the header as the format defines it: NX first, x fastest in the data
    raw = open(p, "rb").read(1024)
    assert struct.unpack("<3i", raw[:12]) == (7, 6, 5)
    assert raw[208:212] == b"MAP "


def test_big_endian_and_extended_header(tmp_path):
    a = RNG.standard_normal((3, 4, 5)).astype(">f4")
    hdr = bytearray(1024)
    struct.pack_into(">4i", hdr, 0, 5, 4, 3, 2)
    struct.pack_into(">3i", hdr, 28, 5, 4, 3)
    struct.pack_into(">3f", hdr, 40, 5.0, 4.0, 3.0)
    struct.pack_into(">i", hdr, 92, 64)
    hdr[208:212] = b"MAP "
    hdr[212:214] = bytes([0x11, 0x11])
    p = tmp_path / "be.mrc"
    p.write_bytes(bytes(hdr) + b"\0" * 64 + a.tobytes())
    data, h = mrc.read_mrc(p, mmap=False)
    assert h.endian == ">" and h.nsymbt == 64
    np.testing.assert_array_equal(data, a)


def test_truncated_and_bad_files_raise(tmp_path):
    p = tmp_path / "t.mrc"
    mrc.write_mrc(p, np.zeros((4, 4, 4), np.float32))
    p.write_bytes(p.read_bytes()[:1100])
    with pytest.raises(ValueError):
        mrc.read_mrc(p)
    q = tmp_path / "bad.mrc"
    q.write_bytes(b"\0" * 2000)
    with pytest.raises(ValueError):
        mrc.read_mrc(q)


def test_relion5_loaders(tmp_path):
    v = RNG.standard_normal((8, 8, 8)).astype(np.float32)
    p = tmp_path / "TS_01_000001_data.mrc"
    mrc.write_mrc(p, v)
    np.testing.assert_array_equal(mrc.load_subtomogram(p), v)
    w = RNG.random((16, 16, 16)).astype(np.float32)
    wp = mrc.weights_path(str(p))
    assert wp.endswith("TS_01_000001_weights.mrc")
    mrc.write_mrc(wp, w)
    np.testing.assert_array_equal(mrc.load_ctf_relion5(wp), w[:8, :, :9])
    np.testing.assert_array_equal(mrc.load_ctf_relion5(wp, transpose=True), np.transpose(w, (2, 1, 0))[:8, :, :9])
