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
