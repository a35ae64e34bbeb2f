"""GPU parity of the alternation resampler and the device rotation tracker
(HIP C-ABI) against the FP64 oracle and the host tracker rule.

Bars: resampled particles within 1e-4 absolute of the oracle (unit-variance
data, FP32 trilinear weights); trackers within 1e-6 of the FP64 host rule.
"""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from oracle import resample as RS

pytestmark = pytest.mark.gpu

SHIFTS = [(0.0, 0.0, 0.0), (2.0, -1.0, 3.0), (0.5, -1.25, 2.75), (-3.3, 0.6, -0.1)]


def _volumes(rng, B, Nv):
    # smooth-ish random volumes: trilinear sampling of white noise is a harsh test already
    return (rng.standard_normal((B, Nv, Nv, Nv)) * 2.0 + 0.7).astype(np.float32)


@pytest.mark.parametrize("N,Nv,radius", [(32, 32, 14.0), (24, 30, 11.5)])
def test_extract_rotated_matches_oracle(gpu_device, N, Nv, radius):
    from matcha_amd.align import extract_rotated_batch
    rng = np.random.default_rng(N + Nv)
    B = len(SHIFTS)
    vols = _volumes(rng, B, Nv)
    shifts = np.array(SHIFTS, dtype=np.float32)
    rots = Rotation.random(B, random_state=N).as_matrix()
    rots[0] = np.eye(3)
    got = extract_rotated_batch(torch.tensor(vols, device=gpu_device), torch.tensor(shifts, device=gpu_device),
                                torch.tensor(rots, dtype=torch.float32, device=gpu_device), radius, N=N)
    got = got.cpu().numpy()
    for b in range(B):
        want = RS.extract_rotated(vols[b].astype(np.float64), shifts[b].astype(np.float64), rots[b], radius) \
            if Nv == N else RS.rotate_volume(RS.normalise_masked(RS.extract_patch(vols[b].astype(np.float64),
                                                                                  shifts[b], N=N),
                                                                 RS.spherical_mask(N, radius)), rots[b])
        err = np.max(np.abs(got[b] - want))
        assert err < 1e-4, (b, err)


def test_identity_at_zero_shift_is_masked_normalisation(gpu_device):
    from matcha_amd.align import extract_rotated_batch
    rng = np.random.default_rng(3)
    N, B, r = 32, 2, 12.0
    vols = _volumes(rng, B, N)
    eye = torch.eye(3, device=gpu_device).expand(B, 3, 3)
    got = extract_rotated_batch(torch.tensor(vols, device=gpu_device), None, eye, r).cpu().numpy()
    mask = RS.spherical_mask(N, r)
    for b in range(B):
        np.testing.assert_allclose(got[b], RS.normalise_masked(vols[b].astype(np.float64), mask), atol=2e-5)
        m = mask > 0
        assert abs(got[b][m].mean()) < 1e-5 and abs(got[b][m].std() - 1) < 1e-5


def test_tracker_update_matches_host_rule(gpu_device):
    from matcha_amd import relion
    from matcha_amd.align import RotationTracker
    rng = np.random.default_rng(5)
    B = 37
    prior = relion.quat_normalized(rng.standard_normal((B, 4)))
    tr = RotationTracker(B, gpu_device, prior)
    np.testing.assert_allclose(tr.rot_particle.cpu().numpy(), relion.quat_to_matrix(prior), atol=1e-6)
    host = prior
    for _ in range(3):
        ang = np.stack([rng.uniform(-np.pi, np.pi, B), rng.uniform(0, np.pi, B), rng.uniform(-np.pi, np.pi, B)])
        t = [torch.tensor(a, dtype=torch.float32, device=gpu_device) for a in ang]
        tr.update(*t)
        host, _ = relion.update_rotation_estimate(*[a.astype(np.float32).astype(np.float64) for a in ang], host)
        q = tr.q.cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(q, host, atol=2e-6)
        Rh = relion.quat_to_matrix(host)
        np.testing.assert_allclose(tr.rot_particle.cpu().numpy(), Rh, atol=2e-6)
        np.testing.assert_allclose(tr.rot_reference.cpu().numpy(), np.swapaxes(Rh, 1, 2), atol=2e-6)


def test_resampler_rejects_bad_arguments(gpu_device):
    from matcha_amd import _lib
    v = torch.zeros((1, 16, 16, 16), device=gpu_device)
    rot = torch.eye(3, device=gpu_device).reshape(1, 9).contiguous()
    sh = torch.zeros((1, 3), device=gpu_device)
    part = torch.empty((1, 16, 4), device=gpu_device)
    out = torch.empty_like(v)
    lib = _lib.lib()
    s = _lib.stream_handle(gpu_device)
    # a shift without a work buffer
    assert lib.mch_extract_rotated(v.data_ptr(), v.stride(0), 16, 1, 16, sh.data_ptr(), rot.data_ptr(), 6.0,
                                   None, part.data_ptr(), out.data_ptr(), out.stride(0), s) != 0
    # patch larger than the particle
    assert lib.mch_extract_rotated(v.data_ptr(), v.stride(0), 16, 1, 20, None, rot.data_ptr(), 6.0,
                                   None, part.data_ptr(), out.data_ptr(), 8000, s) != 0
    assert b"bad arguments" in lib.mch_last_error()
    assert lib.mch_tracker_update(None, None, None, 0, None, None, None, s) != 0
