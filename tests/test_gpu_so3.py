"""GPU parity of the SO(3) search (HIP C-ABI) against the FP64 oracle.

Tolerances (FP32 path): scores 1e-4 relative, angles 0.1 degree geodesic
(the targets stated for this path); grid values 1e-5 of the grid's range.
"""
import numpy as np
import pytest
import torch

from oracle import so3, wigner as W
from tests.so3_helpers import geodesic_deg, peaked_sigma, random_sigma

pytestmark = pytest.mark.gpu


def _sig_t(sig, dev):
    return torch.tensor(np.asarray(sig), dtype=torch.complex64, device=dev)


def _matcha(dev, B, L_max, Li, C=10, K=2, steps=5):
    from matcha_amd.core.Matcha import Matcha
    return Matcha(B, dev, L_max, {"Li": Li, "candidates": C, "num_steps": steps,
                                  "oversampling_factor_K": K})


@pytest.mark.parametrize("L", [3, 8, 16, 32])
def test_evaluate_score_grad_hessian(gpu_device, L):
    rng = np.random.default_rng(10 + L)
    B, C = 2, 3
    sig = random_sigma(rng, B, 32)
    a = rng.uniform(0, 2 * np.pi, (B, C))
    b = rng.uniform(-0.3, np.pi + 0.3, (B, C))
    b[0, 0] = 0.0
    b[1, 1] = np.pi
    g = rng.uniform(0, 2 * np.pi, (B, C))
    m = _matcha(gpu_device, B, 32, [8], C=C)
    from matcha_amd.core.Matcha import so3_evaluate
    out = so3_evaluate(m, _sig_t(sig, gpu_device), torch.tensor(a, device=gpu_device),
                       torch.tensor(b, device=gpu_device), torch.tensor(g, device=gpu_device), L)
    out = out.cpu().numpy().astype(np.float64)
    # oracle uses the FP32-rounded angles the kernel saw
    a32, b32, g32 = (x.astype(np.float32).astype(np.float64) for x in (a, b, g))
    D = W.wigner_D(a32.ravel(), b32.ravel(), g32.ravel(), L).reshape(B, C, -1)
    s_ref = np.einsum('bcn,bn->bc', D, sig[:, :W.dsize(L)]).real
    grad, H = so3.derivatives(D, so3.coeff_bank(sig, L), a32)
    ref = np.stack([s_ref, grad[..., 0], grad[..., 1], grad[..., 2], H[..., 0, 0], H[..., 0, 1],
                    H[..., 0, 2], H[..., 1, 1], H[..., 1, 2], H[..., 2, 2]], axis=-1)
    scale = np.sqrt((np.abs(sig[:, :W.dsize(L)]) ** 2).sum(axis=1))[:, None, None]
    # derivative columns scale with l and l^2
    colscale = np.array([1, L, L, L, L * L, L * L, L * L, L * L, L * L, L * L], dtype=np.float64)
    err = np.abs(out - ref) / (scale * np.maximum(colscale, 1))
    assert err.max() < 2e-5, (err.max(axis=(0, 1)))


@pytest.mark.parametrize("L0,K", [(4, 2), (8, 4)])
def test_sofft_grid_matches_oracle(gpu_device, L0, K):
    from matcha_amd.core.SOFFT import SOFFT
    rng = np.random.default_rng(L0 * 7 + K)
    B = 2
    sig = random_sigma(rng, B, L0)
    sof = SOFFT(L0, gpu_device, B, oversampling_factor=K)
    f = sof.eval(_sig_t(sig, gpu_device)).cpu().numpy()
    ref = so3.sofft_grid(sig, L0, K)
    rng_ = ref.max() - ref.min()
    assert np.max(np.abs(f - ref)) < 1e-5 * rng_


def test_sofft_topk_values_and_angles(gpu_device):
    from matcha_amd.core.SOFFT import SOFFT
    rng = np.random.default_rng(21)
    B, L0, K, C = 3, 8, 4, 10
    sig = random_sigma(rng, B, L0)
    sof = SOFFT(L0, gpu_device, B, oversampling_factor=K, candidates=C)
    a, b, g, s = (x.cpu().numpy() for x in sof.topk(_sig_t(sig, gpu_device), C))
    f = so3.sofft_grid(sig, L0, K)
    ids, vals = so3.topk_grid(f, C)
    ra, rb, rg = so3.ids_to_angles(ids, f.shape[1])
    assert np.max(np.abs(s - vals)) < 1e-5 * (f.max() - f.min())
    # where consecutive oracle values are well separated the picks must agree exactly
    n = f.shape[1]
    for bi in range(B):
        for c in range(C):
            gap = min(abs(vals[bi, c] - vals[bi, c - 1]) if c else np.inf,
                      abs(vals[bi, c] - vals[bi, c + 1]) if c + 1 < C else np.inf)
            if gap > 1e-4 * (f.max() - f.min()):
                assert abs(b[bi, c] - rb[bi, c]) < 1e-6 and abs(a[bi, c] - ra[bi, c]) < 1e-6
                assert abs(g[bi, c] - rg[bi, c]) < 1e-6


def test_newton_march_matches_oracle(gpu_device):
    rng = np.random.default_rng(31)
    B, C, L = 2, 4, 16
    sig = np.stack([peaked_sigma(rng, 16, 2.1, 1.0, 0.4, noise=0.05),
                    peaked_sigma(rng, 16, 5.0, 2.5, 3.3, noise=0.05)])
    # starting points a few degrees off the peaks
    a0 = np.array([[2.1, 2.15, 2.0, 2.2], [5.0, 5.05, 4.95, 5.1]])
    b0 = np.array([[1.0, 1.05, 0.95, 1.1], [2.5, 2.45, 2.55, 2.4]])
    g0 = np.array([[0.4, 0.45, 0.35, 0.5], [3.3, 3.35, 3.25, 3.2]])
    a0, b0, g0 = (x.astype(np.float32).astype(np.float64) for x in (a0, b0, g0))
    m = _matcha(gpu_device, B, 16, [8], C=C)
    t = lambda x: torch.tensor(x, dtype=torch.float32, device=gpu_device)
    A, Bt, G, S = (x.cpu().numpy().astype(np.float64)
                   for x in m.march(_sig_t(sig, gpu_device), t(a0), t(b0), t(g0), levels=[L]))
    ra, rb, rg, rs = so3.optimize(a0, b0, g0, sig, L, 5)
    geo = geodesic_deg(A.ravel(), Bt.ravel(), G.ravel(), ra.ravel(), rb.ravel(), rg.ravel())
    assert geo.max() < 0.1, geo
    assert np.max(np.abs(S - rs) / np.abs(rs)) < 1e-4


@pytest.mark.parametrize("L_max,Li,K", [(16, [8, 12], 2), (32, [8, 16, 24], 4)])
def test_search_orientations_matches_oracle(gpu_device, L_max, Li, K):
    rng = np.random.default_rng(L_max)
    B = 3
    truth = [(1.0, 0.8, 2.0), (4.0, 2.2, 5.5), (0.3, 1.6, 3.9)]
    sig = np.stack([peaked_sigma(rng, L_max, *t, noise=0.03) for t in truth])
    m = _matcha(gpu_device, B, L_max, Li, C=10, K=K)
    A, Bt, G, S = (x.cpu().numpy().astype(np.float64)
                   for x in m.search_orientations(_sig_t(sig, gpu_device)))
    ra, rb, rg, rs = so3.search_orientations(sig, L_max, Li, candidates=10, num_steps=5, K=K)
    geo_ref = geodesic_deg(A, Bt, G, ra, rb, rg)
    assert geo_ref.max() < 0.1, geo_ref
    assert np.max(np.abs(S - rs) / np.abs(rs)) < 1e-4
    geo_truth = geodesic_deg(A, Bt, G, *[np.array(x) for x in zip(*truth)])
    assert geo_truth.max() < 1.0


def test_bad_inputs_raise(gpu_device):
    m = _matcha(gpu_device, 2, 16, [8])
    with pytest.raises(ValueError):
        m.search_orientations(torch.zeros((2, 10), dtype=torch.complex64, device=gpu_device))
    with pytest.raises(TypeError):
        m.search_orientations(torch.zeros((2, W.dsize(16)), dtype=torch.complex128, device=gpu_device))
    with pytest.raises(ValueError):
        m.search_orientations(torch.zeros((3, W.dsize(16)), dtype=torch.complex64, device=gpu_device))


def test_zero_sigma_singular_hessian_is_defined(gpu_device):
    # sigma = 0 -> H = 0 (singular): the reference skips the step; scores stay 0
    m = _matcha(gpu_device, 2, 16, [8])
    A, Bt, G, S = m.search_orientations(torch.zeros((2, W.dsize(16)), dtype=torch.complex64,
                                                    device=gpu_device))
    assert torch.all(S == 0) and torch.all(torch.isfinite(A)) and torch.all(torch.isfinite(Bt))


def test_host_wigner_table_matches_oracle(gpu_device):
    import ctypes
    from matcha_amd import _lib
    L = 12
    for beta in (0.0, 0.4, 2.0, np.pi):
        buf = (ctypes.c_double * W.dsize(L))()
        _lib.check(_lib.lib().mch_wigner_d_host(beta, L, buf))
        ref = W.wigner_d(np.array([beta]), L)[0]
        assert np.max(np.abs(np.array(buf[:]) - ref)) < 1e-12
