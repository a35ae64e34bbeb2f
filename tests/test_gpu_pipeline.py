"""Device staging of the particle reader: contents, stream ordering, batch probe."""
import numpy as np
import pytest
import torch

from matcha_amd import mrc

pytestmark = pytest.mark.gpu


def _files(tmp_path, n, N, seed=0):
    rng = np.random.default_rng(seed)
    paths, vols, ctfs = [], [], []
    for i in range(n):
        p = tmp_path / f"TS_02_{i:06d}_data.mrc"
        v = rng.standard_normal((N, N, N)).astype(np.float32)
        w = rng.random((2 * N, N, N // 2 + 1)).astype(np.float32)
        mrc.write_mrc(p, v)
        mrc.write_mrc(mrc.weights_path(str(p)), w)
        paths.append(str(p))
        vols.append(v)
        ctfs.append(w[:N])
    return paths, vols, ctfs


def test_device_batches_match_files_and_respect_compute_order(gpu_device, tmp_path):
    from matcha_amd.pipeline import ParticleBatches
    paths, vols, ctfs = _files(tmp_path, 21, 32)
    pb = ParticleBatches(paths, batch=4, device=gpu_device, workers=4, depth=2)
    big = torch.randn(2048, 2048, device=gpu_device)
    sums, names = [], []
    for b in pb:
        assert b.vols.is_cuda and b.vols.shape[1:] == (32, 32, 32) and b.ctfs.shape[1:] == (32, 32, 17)
        # a long op ahead of the reads of this batch on the compute stream: a copy of a
        # later batch into this device slot must not overtake it
        for _ in range(4):
            big = torch.tanh(big @ big) * 0.5
        sums.append(torch.stack([b.vols.double().sum(dim=(1, 2, 3)), b.ctfs.double().sum(dim=(1, 2, 3))], 1))
        names += b.paths
    torch.cuda.synchronize()
    got = torch.cat(sums).cpu().numpy()
    assert names == paths
    want = np.array([[v.astype(np.float64).sum(), c.astype(np.float64).sum()] for v, c in zip(vols, ctfs)])
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-6)


def test_probe_batch_size(gpu_device):
    from matcha_amd.pipeline import probe_batch_size
    b = probe_batch_size(96, 8252, 42, 32, gpu_device)
    assert b >= 64 and (b & (b - 1)) == 0


def test_align_paths_end_to_end(gpu_device, tmp_path):
    # particles written as MRC files, read by the pipeline, aligned, recorded for the STAR join
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from matcha_amd.pipeline import align_paths
    from matcha_amd.synthetic import Workload, euler_zyz_to_matrix, geodesic_deg
    N, B = 48, 4
    wl = Workload(N, 20, (1.5, 3.0), 5, gpu_device, mask_radius=21)
    vols, R, t = wl.particles(6, 2, None)
    paths = []
    for i in range(6):
        p = tmp_path / f"TS_03_{i:06d}_data.mrc"
        mrc.write_mrc(p, vols[i].cpu().numpy())
        mrc.write_mrc(mrc.weights_path(str(p)), np.ones((2 * N, N, N // 2 + 1), np.float32))
        paths.append(str(p))
    ccm = CrossCorrelationMatcher(N, gpu_device, batchsize=B, bandlimit=12, radius=21)
    ccm.set_template(wl.template)
    m = Matcha(B, gpu_device, 12, {"Li": [8], "oversampling_factor_K": 2})
    sm = ShiftMatcher({"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1,
                       "shift_search_radius": 2}, gpu_device)
    sm.set_reference(wl.template.cpu().numpy())
    rec = align_paths(paths, ccm, m, sm, batch=B, num_alternations=2, half="half1", mode="global")
    df = rec.frame
    assert len(df) == 6 * 2 and sorted(set(df["path"])) == sorted(paths)
    last = df[df["alternation_index"] == 1].set_index("path").loc[paths]
    from matcha_amd import relion
    tr = np.stack(last["rotation"].to_numpy())
    P = np.array([[0.0, 0, 1], [0, 1, 0], [1, 0, 0]])
    est = P @ relion.quat_to_matrix(relion.quat_from_euler_zyz(tr)) @ P
    assert np.median(geodesic_deg(R, est)) < 1.5
    sh = np.stack(last["translation"].to_numpy())
    assert np.array_equal(sh.astype(int), t[:, ::-1])
