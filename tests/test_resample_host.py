"""Oracle of the alternation resampler (extraction, masked normalisation,
rotation), pinned by a library routine and closed forms."""
import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial.transform import Rotation

from oracle import resample as RS

RNG = np.random.default_rng(11)


@pytest.mark.parametrize("shift", [(0, 0, 0), (1, -2, 3), (0.25, -1.5, 2.75), (-3.3, 0.6, -0.1)])
def test_extract_patch_matches_scipy_trilinear(shift):
    N = 12
    vol = RNG.standard_normal((N, N, N))
    got = RS.extract_patch(vol, shift)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    coords = np.stack([z + shift[2], y + shift[1], x + shift[0]]).astype(np.float64)
    want = ndimage.map_coordinates(vol, coords, order=1, mode="grid-constant", cval=0.0)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_extract_patch_centres_on_larger_volume():
    vol = RNG.standard_normal((15, 16, 17))
    got = RS.extract_patch(vol, (0, 0, 0), N=8)
    np.testing.assert_array_equal(got, vol[7 - 4:7 + 4, 8 - 4:8 + 4, 8 - 4:8 + 4])


@pytest.mark.parametrize("N", [10, 11])
def test_rotation_by_90_degrees_about_z_is_a_voxel_permutation(N):
    # x' = c - (y - c), y' = x - c + c, z' = z about c = N // 2: out[z, y, x] = vol[z, x, 2c - y]
    vol = RNG.standard_normal((N, N, N))
    M = Rotation.from_euler("z", 90, degrees=True).as_matrix()
    got = RS.rotate_volume(vol, M)
    c = N // 2
    want = np.zeros_like(vol)
    for y in range(N):
        src = 2 * c - y
        if 0 <= src < N:
            want[:, y, :] = vol[:, :, src]
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_rotation_is_exact_on_affine_volumes():
    N = 16
    c = N // 2
    a = np.array([0.3, -0.7, 1.1])      # coefficients of (x, y, z)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    vol = a[0] * x + a[1] * y + a[2] * z + 2.0
    for M in Rotation.random(4, random_state=3).as_matrix():
        got = RS.rotate_volume(vol, M)
        d = np.stack([x - c, y - c, z - c]).reshape(3, -1)
        p = M @ d + c
        inside = np.all((p >= 0) & (p <= N - 1), axis=0).reshape(N, N, N)
        want = (a @ p).reshape(N, N, N) + 2.0
        assert inside.sum() > N ** 3 // 4
        np.testing.assert_allclose(got[inside], want[inside], atol=1e-10)
    # identity leaves the volume unchanged
    np.testing.assert_allclose(RS.rotate_volume(vol, np.eye(3)), vol, atol=1e-12)


def test_masked_normalisation_moments_and_invariance():
    N, r = 14, 5.5
    mask = RS.spherical_mask(N, r)
    # binary sphere about N // 2, strict inequality
    assert mask[N // 2, N // 2, N // 2] == 1 and mask[N // 2, N // 2, N // 2 + 5] == 1
    assert mask[N // 2, N // 2, N // 2 + 6] == 0
    p = RNG.standard_normal((N, N, N)) * 3 + 1
    out = RS.normalise_masked(p, mask)
    m = mask > 0
    assert abs(out[m].mean()) < 1e-12
    assert abs(np.sqrt((out[m] ** 2).mean()) - 1) < 1e-12
    assert np.all(out[~m] == 0)
    np.testing.assert_allclose(RS.normalise_masked(2.5 * p - 4, mask), out, atol=1e-12)
    np.testing.assert_array_equal(RS.normalise_masked(p, np.zeros_like(mask)), 0 * p)


def test_extract_rotated_composes_the_three_steps():
    N, r = 12, 5.0
    vol = RNG.standard_normal((N, N, N))
    M = Rotation.random(random_state=5).as_matrix()
    want = RS.rotate_volume(RS.normalise_masked(RS.extract_patch(vol, (1, 0, -2)), RS.spherical_mask(N, r)), M)
    np.testing.assert_array_equal(RS.extract_rotated(vol, (1, 0, -2), M, r), want)
    # identity rotation at zero shift is the normalised, masked particle
    np.testing.assert_allclose(RS.extract_rotated(vol, (0, 0, 0), np.eye(3), r),
                               RS.normalise_masked(vol, RS.spherical_mask(N, r)), atol=1e-12)
