"""The torch-CPU FP32 restatement (oracle/torch_cpu.py, timed as the CPU
baseline) against the FP64 oracle and the reference's own grid sizes."""
import numpy as np
import pytest
import torch

from oracle import fle, so3, shift as SH, torch_cpu as TC, wigner as W
from tests.so3_helpers import geodesic_deg, peaked_sigma, random_sigma


@pytest.mark.parametrize("N,bl,want", [(64, 16, (64, 106, 108)), (96, 32, (91, 180, 182)),
                                       (128, 48, (119, 252, 254))])
def test_grid_sizes_are_the_references(N, bl, want):
    ls, ks, ms, mds, lmds, cs, ne = fle.lap_eig_ball(N, bl)
    assert TC.grid_sizes(N, 1e-4, float(lmds[-1]), int(ls.max())) == want


def test_clenshaw_curtis_is_exact_on_polynomials():
    theta, w = TC.clenshaw_curtis(11)
    x = np.cos(theta)
    for k in range(11):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-13


@pytest.mark.parametrize("N,bl,radius", [(24, 8, 11.0), (20, 6, 9.5)])
def test_nufft_projection_matches_brute_force(N, bl, radius):
    rng = np.random.default_rng(N)
    vols = np.stack([fle.blob_volume(rng, N, 8, sigma_range=(1.0, 2.5)) + 0.3 * rng.standard_normal((N, N, N))
                     for _ in range(2)])
    tf = TC.TorchFLE(N, bl, radius)
    got = tf.evaluate_t(torch.tensor(vols, dtype=torch.float32)).numpy()
    ref = fle.evaluate_t(vols, N, bl, radius)
    # the reference's eps = 1e-4 expansion (NUFFT at 1e-4, quadrature, interpolation)
    assert np.max(np.abs(got - ref)) < 5e-4 * np.max(np.abs(ref))


def test_wigner_d_recurrence_matches_oracle():
    betas = np.array([0.0, 0.3, 1.7, np.pi, -0.4, 3.6])
    got = TC.wigner_d_torch(torch.tensor(betas), 14).numpy()
    ref = W.wigner_d(betas, 14)
    assert np.max(np.abs(got - ref)) < 2e-5


def test_sigma_matches_oracle():
    N, bl, radius = 20, 6, 9.5
    tf = TC.TorchFLE(N, bl, radius)
    rng = np.random.default_rng(2)
    aT = (rng.standard_normal(tf.ne) + 1j * rng.standard_normal(tf.ne)).astype(np.complex64)
    aV = (rng.standard_normal((3, tf.ne)) + 1j * rng.standard_normal((3, tf.ne))).astype(np.complex64)
    got = TC.TorchSigma(tf, tf.lmax)(torch.tensor(aT), torch.tensor(aV)).numpy()
    ref = fle.compute_sigma(aT.astype(complex), aV.astype(complex), tf.ls, tf.mds, tf.lmax)
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref))


def test_search_matches_oracle():
    rng = np.random.default_rng(7)
    truth = [(1.0, 0.8, 2.0), (4.0, 2.2, 5.5)]
    sig = np.stack([peaked_sigma(rng, 16, *t, noise=0.03) for t in truth])
    ts = TC.TorchSearch(16, [8, 12], 2)
    A, B, G, S = (x.numpy().astype(np.float64) for x in ts.search(torch.tensor(sig, dtype=torch.complex64)))
    ra, rb, rg, rs = so3.search_orientations(sig, 16, [8, 12], candidates=10, num_steps=5, K=2)
    assert geodesic_deg(A, B, G, ra, rb, rg).max() < 0.1
    assert np.max(np.abs(S - rs) / np.abs(rs)) < 1e-4


def test_grid_and_bank_derivatives_match_oracle():
    rng = np.random.default_rng(8)
    sig = random_sigma(rng, 2, 8)
    ts = TC.TorchSearch(8, [4], 2)
    f = so3.sofft_grid(sig, 4, 2)
    ids, vals = so3.topk_grid(f, 10)
    ra, rb, rg = so3.ids_to_angles(ids, f.shape[1])
    a, b, g = ts.sofft_topk(torch.tensor(sig, dtype=torch.complex64))
    assert np.allclose(a.numpy(), ra, atol=1e-6) and np.allclose(b.numpy(), rb, atol=1e-6)
    D = ts._D(a, b, g, 8)
    grad, H = ts._derivatives(torch.bmm(D, ts._bank(torch.tensor(sig, dtype=torch.complex64), 8).transpose(1, 2)), a)
    Dr = W.wigner_D(ra.ravel(), rb.ravel(), rg.ravel(), 8).reshape(2, 10, -1)
    gr, Hr = so3.derivatives(Dr, so3.coeff_bank(sig, 8), ra)
    scale = np.abs(gr).max()
    assert np.max(np.abs(grad.numpy() - gr)) < 1e-5 * scale
    assert np.max(np.abs(H.numpy() - Hr)) < 1e-5 * np.abs(Hr).max()


def test_shift_search_matches_oracle():
    from matcha_amd.synthetic import haar_rotations
    N = 32
    rng = np.random.default_rng(3)
    tmpl = fle.blob_volume(rng, N, 10, sigma_range=(1.5, 3.0))
    refF = SH.load_reference_full(tmpl, N, 1.0)
    subs = np.stack([np.roll(tmpl, (2, -1, 3), axis=(0, 1, 2)) + 0.2 * rng.standard_normal((N, N, N)),
                     rng.standard_normal((N, N, N))])
    rots = np.stack([np.eye(3), haar_rotations(rng, 1)[0]])
    prior = np.array([[0.0, 0.0, 0.0], [1.0, -2.0, 0.0]])
    want, wsc = SH.estimate_shift(refF, subs, prior.reshape(2, 1, 3), rots, N=N)
    ss = TC.ShiftSearch(tmpl, N)
    got, gsc = ss.search(torch.tensor(subs, dtype=torch.float32), torch.tensor(rots), torch.tensor(prior))
    assert np.array_equal(got.numpy(), want)
    assert np.allclose(gsc.numpy(), wsc, rtol=1e-4)
