"""Pins of the FP64 SO(3) search restatement (oracle/so3.py)."""
import numpy as np
import pytest

from oracle import so3, wigner as W
from tests.so3_helpers import geodesic_deg, peaked_sigma, random_sigma


def test_bank_derivatives_match_finite_differences():
    rng = np.random.default_rng(3)
    L = 7
    sig = random_sigma(rng, 2, L)
    a = rng.uniform(0, 2 * np.pi, (2, 3))
    b = rng.uniform(0.2, 3.0, (2, 3))
    g = rng.uniform(0, 2 * np.pi, (2, 3))
    D = W.wigner_D(a.ravel(), b.ravel(), g.ravel(), L).reshape(2, 3, -1)
    grad, H = so3.derivatives(D, so3.coeff_bank(sig, L), a)

    def f(bi, x):
        return so3.score(np.array([x[0]]), np.array([x[1]]), np.array([x[2]]), sig[bi], L)[0]

    for bi in range(2):
        for c in range(3):
            x0 = np.array([a[bi, c], b[bi, c], g[bi, c]])
            h = 1e-5
            fd = np.array([(f(bi, x0 + h * e) - f(bi, x0 - h * e)) / (2 * h) for e in np.eye(3)])
            assert np.max(np.abs(fd - grad[bi, c])) < 1e-7 * max(1.0, np.abs(fd).max())
            h2 = 1e-4
            fh = np.zeros((3, 3))
            for i, ei in enumerate(np.eye(3)):
                for k, ek in enumerate(np.eye(3)):
                    fh[i, k] = (f(bi, x0 + h2 * ei + h2 * ek) - f(bi, x0 + h2 * ei - h2 * ek)
                                - f(bi, x0 - h2 * ei + h2 * ek) + f(bi, x0 - h2 * ei - h2 * ek)) / (4 * h2 * h2)
            assert np.max(np.abs(fh - H[bi, c])) < 1e-5 * max(1.0, np.abs(fh).max())


def test_sofft_grid_equals_direct_sum():
    rng = np.random.default_rng(4)
    L0, K = 4, 2
    sig = random_sigma(rng, 1, L0)
    f = so3.sofft_grid(sig, L0, K)
    n = 2 * L0 * K
    for (j, ai, gi) in [(0, 0, 0), (3, 5, 7), (n - 1, 2, 9), (7, 15, 1)]:
        beta = np.pi * j / (n - 1)
        direct = so3.score(np.array([2 * np.pi * ai / n]), np.array([beta]),
                           np.array([2 * np.pi * gi / n]), sig[0], L0 - 1)[0] / n ** 2
        assert f[0, j, ai, gi] == pytest.approx(direct, abs=1e-12)


def test_march_recovers_peak_and_is_monotone():
    rng = np.random.default_rng(5)
    a0, b0, g0 = 1.234, 0.987, 4.321
    sig = peaked_sigma(rng, 12, a0, b0, g0, noise=0.02)[None, :]
    f = so3.sofft_grid(sig, 4, 2)
    ids, vals = so3.topk_grid(f, 4)
    a, b, g = so3.ids_to_angles(ids, f.shape[1])
    s_prev = so3.score(a.ravel(), b.ravel(), g.ravel(), sig[0], 8)
    ra, rb, rg, rs = so3.optimize(a, b, g, sig, 8, 5)
    assert np.all(rs.ravel() >= s_prev - 1e-12)  # best score never decreases
    A, B_, G, S = so3.search_orientations(sig, 12, [4, 8], candidates=4, num_steps=5, K=2)
    assert geodesic_deg(A, B_, G, a0, b0, g0)[0] < 1.0


def test_fixed_point_at_exact_peak():
    a0, b0, g0 = 0.5, 1.2, 2.0
    sig = peaked_sigma(None, 10, a0, b0, g0)[None, :]
    ra, rb, rg, rs = so3.optimize(np.array([[a0]]), np.array([[b0]]), np.array([[g0]]), sig, 10, 3)
    assert geodesic_deg(ra[0], rb[0], rg[0], a0, b0, g0)[0] < 1e-6


def test_march_schedule_skip_rule():
    assert so3.march_schedule([8, 16, 24], 32) == [8, 16, 24, 32]
    assert so3.march_schedule([8, 16, 24, 32], 32) == [8, 16, 24, 32]
    assert so3.march_schedule([30, 40, 65], 40) == [30, 40]
