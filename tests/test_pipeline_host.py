"""Threaded particle reader (host side; the device staging is tested on the GPU)."""
import numpy as np
import pytest
import torch

from matcha_amd import mrc
from matcha_amd.pipeline import ParticleBatches, particle_bytes

RNG = np.random.default_rng(9)


def make_files(tmp_path, n, N, bad=()):
    paths, vols, ctfs = [], [], []
    for i in range(n):
        p = tmp_path / f"TS_01_{i:06d}_data.mrc"
        v = RNG.standard_normal((N, N, N)).astype(np.float32)
        # RELION5 weights: the loader keeps [:n0 // 2, :, :n0 // 2 + 1] of a (2N, N, N//2+1) file
        w = RNG.random((2 * N, N, N // 2 + 1)).astype(np.float32)
        if i in bad:
            p.write_bytes(b"not an mrc file" * 100)
        else:
            mrc.write_mrc(p, v)
        mrc.write_mrc(mrc.weights_path(str(p)), w)
        paths.append(str(p))
        vols.append(v)
        ctfs.append(w[:N])
    return paths, vols, ctfs


@pytest.mark.parametrize("workers,depth", [(1, 2), (4, 3)])
def test_batches_in_order_with_weights(tmp_path, workers, depth):
    paths, vols, ctfs = make_files(tmp_path, 11, 8)
    pb = ParticleBatches(paths, batch=4, device="cpu", workers=workers, depth=depth)
    seen = []
    for b in pb:
        assert b.vols.shape[0] == len(b.paths) <= 4
        for j, p in enumerate(b.paths):
            i = paths.index(p)
            np.testing.assert_array_equal(b.vols[j].numpy(), vols[i])
            np.testing.assert_array_equal(b.ctfs[j].numpy(), ctfs[i])
        seen += b.paths
    assert seen == paths and not pb.errors


def test_bad_files_are_reported_and_skipped(tmp_path):
    paths, vols, _ = make_files(tmp_path, 9, 8, bad=(2, 5))
    pb = ParticleBatches(paths, batch=3, device="cpu", workers=3)
    seen = []
    for b in pb:
        for j, p in enumerate(b.paths):
            np.testing.assert_array_equal(b.vols[j].numpy(), vols[paths.index(p)])
        seen += b.paths
    assert seen == [p for i, p in enumerate(paths) if i not in (2, 5)]
    assert sorted(paths.index(p) for p, _ in pb.errors) == [2, 5]


def test_without_weights_and_early_stop(tmp_path):
    paths, vols, _ = make_files(tmp_path, 10, 8)
    pb = ParticleBatches(paths, batch=2, device="cpu", ctf=False, workers=2)
    it = iter(pb)
    b = next(it)
    assert b.ctfs is None and b.paths == paths[:2]
    it.close()   # the reader thread stops without the rest being consumed


def test_particle_bytes_scale():
    small = particle_bytes(64, 1006, 19, 16)
    big = particle_bytes(96, 8252, 42, 32)
    assert 0 < small < big
    # the projection operand and the volume dominate at box 96: tens of MB per particle
    assert 5e6 < big < 1e8
