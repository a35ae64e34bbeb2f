"""Host-side shift preparation checked on the CPU."""
import numpy as np
import pytest

from oracle import fle, shift as S


def test_host_reference_matches_oracle():
    # the product's FP64 host preparation of the Fourier reference
    from matcha_amd.core.ShiftMatcher import fourier_reference
    rng = np.random.default_rng(1)
    tmpl = fle.blob_volume(rng, 32, 10, sigma_range=(1.5, 3.0))
    a = fourier_reference(tmpl, 32, 1.3)
    b = S.load_reference_full(tmpl, 32, 1.3)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))


def test_shift_oracle_closed_form():
    # a circular integer shift against the unshifted reference returns that shift
    N = 32
    rng = np.random.default_rng(0)
    tmpl = fle.blob_volume(rng, N, 10, sigma_range=(1.5, 3.0))
    refF = S.load_reference_full(tmpl, N, 1.0)
    t = np.array([2, -3, 1])
    sub = np.roll(tmpl, tuple(t), axis=(0, 1, 2))
    sh, sc = S.estimate_shift(refF, sub[None], np.zeros((1, 1, 3)), np.eye(3)[None], N=N)
    assert np.array_equal(sh[0], t[::-1].astype(float))
    sh2, _ = S.estimate_shift(refF, sub[None], t[::-1].reshape(1, 1, 3).astype(float), np.eye(3)[None], N=N)
    assert np.array_equal(sh2[0], np.zeros(3))


def test_shift_frame_of_search_orientations():
    # the shift search with rb = P R^T P (R = particle rotation in the expansion's
    # frame, P = axis reversal) recovers the particle's shift
    from matcha_amd.synthetic import blob_params, haar_rotations
    N = 32
    rng = np.random.default_rng(3)
    c, s, a = blob_params(rng, N, 12, (1.5, 3.0))
    ax = np.arange(N) - N // 2
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")

    def render(cen):
        v = np.zeros((N, N, N))
        for j in range(len(s)):
            v += a[j] * np.exp(-((X - cen[j, 0]) ** 2 + (Y - cen[j, 1]) ** 2 + (Z - cen[j, 2]) ** 2) / (2 * s[j] ** 2))
        return v

    refF = S.load_reference_full(render(c), N, 1.0)
    P = np.eye(3)[::-1]
    for _ in range(2):
        R = haar_rotations(rng, 1)[0]
        t = rng.integers(-3, 4, 3)
        sh, sc = S.estimate_shift(refF, render(c @ R.T + t)[None], np.zeros((1, 1, 3)), (P @ R.T @ P)[None], N=N)
        assert np.array_equal(sh[0], t[::-1].astype(float)) and sc[0] > 0.95


@pytest.mark.parametrize("up,s0,s1,s2", [(8, 0.375, -1.125, 0.0), (4, -0.5, 0.75, 0.25), (2, 1.0, 0.0, -0.5)])
def test_oracle_upsampled_refinement_closed_form(up, s0, s1, s2):
    # ref = 1 and sub = e^{-2 pi i (k0' s0 + k1' s1) / N} e^{-2 pi i k2' s2 / H}, with signed
    # frequencies (numpy fftfreq order) and, along the half axis, the reference's own
    # convention fftfreq(N/2 + 1): the local matrix DFT is then a product of Dirichlet kernels
    # whose real part peaks exactly at u = floor(R/2) + up * s, so the refinement returns
    # (s0, s1, s2) with score 1.  A wrong sign, offset or frequency convention moves the peak.
    from oracle import shift as S
    N = 16
    H = N // 2 + 1
    kf = np.fft.fftfreq(N) * N
    ramp = np.exp(-2j * np.pi * (kf[:, None] * s0 + kf[None, :] * s1) / N)
    ramp2 = np.exp(-2j * np.pi * np.fft.fftfreq(H) * s2)
    sub = ramp[:, :, None] * ramp2[None, None, :]
    ref = np.ones((N, N, H), dtype=np.complex128)
    shift, score = S.compute_shift_local(ref, sub, upsample=up, window=2 * 2 + 1)
    assert np.array_equal(shift, np.array([s0, s1, s2]))
    assert score == pytest.approx(1.0, rel=1e-12)


# --- reference preparation (ShiftMatcher.py:377-453), pinned by single modes ---

def _mode(N, k):
    ax = np.arange(N)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    return np.cos(2 * np.pi * (k[0] * Z + k[1] * Y + k[2] * X) / N)


@pytest.mark.parametrize("k,factor", [
    ((3, 0, 0), 1.0),                  # inside the pass band (radius 7)
    ((0, 7, 0), 1.0),                  # on its edge: 0.5 - 0.5 cos(pi) = 1
    ((0, 0, 8), 0.5),                  # mid cosine band: 0.5 - 0.5 cos(pi / 2)
    ((2, -2, -8), None),               # in the band, |k| = sqrt(72): 0.5 - 0.5 cos(pi (9 - |k|) / 2)
    ((9, 1, 0), 0.0),                  # beyond radius_p = 9
])
def test_soft_mask_scales_a_single_fourier_mode(k, factor):
    # box 32 at 1 A/px low-passed at 4 A with a 2 px edge: radius 32 / 4 - 1 = 7,
    # radius_p = 9 (in Fourier pixels); a real mode cos(2 pi k.x / N) has its two
    # components at +-k, both at |k|, so it comes back scaled by the mask there
    N = 32
    if factor is None:
        factor = 0.5 - 0.5 * np.cos(np.pi * (9 - np.sqrt(72.0)) / 2)
    v = _mode(N, k)
    out = S.apply_soft_mask(v, N, 1.0, 4.0, 2.0)
    assert np.max(np.abs(out - factor * v)) < 1e-12


def test_soft_mask_off_for_non_positive_resolution():
    v = _mode(16, (1, 2, 3))
    assert S.apply_soft_mask(v, 16, 1.0, 0.0) is v


@pytest.mark.parametrize("pos,value", [
    ((16, 16, 16), 1.0),               # the centre: sinc(0) = 1
    ((16, 16, 24), np.pi ** 2 / 8),    # r = 8 / 32 = 1/4: sinc = 2 sqrt(2) / pi
    ((8, 16, 16), np.pi ** 2 / 8),     # the same along z, on the negative side
    ((0, 16, 16), np.pi ** 2 / 4),     # r = 1/2: sinc = 2 / pi
])
def test_gridding_correction_divides_by_sinc_squared(pos, value):
    # a delta at array index pos (centre N/2) comes back as 1 / sinc^2(|pos - N/2| / N)
    N = 32
    v = np.zeros((N, N, N))
    v[pos] = 1.0
    out = S.gridding_correct(v, N, 1.0)
    assert out[pos] == pytest.approx(value, rel=1e-12)
    out[pos] = 0
    assert not out.any()


def test_full_mask_is_the_ball_of_radius_half_box():
    # mask_full over unshifted FFT indices: frequencies (k mod N, wrapped to
    # [-N/2, N/2)) with |k|^2 <= (N/2)^2, counted by brute force
    import itertools
    N = 8
    m = S.get_mask_full((N, N, N))
    want = np.zeros((N, N, N), dtype=bool)
    for i, j, l in itertools.product(range(N), repeat=3):
        f = [x if x < N // 2 else x - N for x in (i, j, l)]
        want[i, j, l] = sum(x * x for x in f) <= (N // 2) ** 2
    assert np.array_equal(m, want)
