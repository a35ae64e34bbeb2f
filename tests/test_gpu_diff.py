"""GPU differentiable score (HIP C-ABI through torch.autograd) against the FP64
oracle and against central finite differences.

Bars: score within 1e-4 relative of the FP64 oracle; gradient in
(omega, t) within 1e-3 relative (of its norm) of the FP64 reference gradient
(central differences of the oracle) and of FP32 central differences with
steps 1e-3 rad / 1e-2 px; template gradient: the exact directional change of
the (linear) score within 1e-4 of |s| + |change|.
"""
import numpy as np
import pytest
import torch

from oracle import diff as DF
from oracle import fle

pytestmark = pytest.mark.gpu

N, BL, RADIUS, L = 24, 8, 11.0, 7


@pytest.fixture(scope="module")
def setup(gpu_device):
    from matcha_amd.diff import ScoreEngine
    rng = np.random.default_rng(42)
    tmpl = fle.blob_volume(rng, N, 8, sigma_range=(1.0, 2.5))
    parts = np.stack([fle.blob_volume(rng, N, 8, sigma_range=(1.0, 2.5)) for _ in range(3)]).astype(np.float32)
    eng = ScoreEngine(N, BL, RADIUS, L, 4, gpu_device)
    aT = eng.template_coefficients(torch.tensor(tmpl, dtype=torch.float32, device=gpu_device))
    omega = np.array([[0.4, -1.1, 0.7], [2.0, 0.3, -0.9], [1e-5, -2e-5, 1e-5]])  # the last: beta ~ 0
    t = np.array([[0.3, -0.45, 0.2], [-1.35, 0.6, 0.15], [0.25, 0.25, -0.7]])
    return eng, aT, parts, omega, t


def _gpu(eng, omega, t, aT, parts, dev, template_gradient=False):
    out, G = eng.score_and_gradients(torch.tensor(omega, dtype=torch.float32, device=dev),
                                     torch.tensor(t, dtype=torch.float32, device=dev), aT,
                                     torch.tensor(parts, device=dev), template_gradient)
    return out.cpu().numpy().astype(np.float64), G


def test_score_and_gradient_match_fp64_reference(gpu_device, setup):
    eng, aT, parts, omega, t = setup
    out, _ = _gpu(eng, omega, t, aT, parts, gpu_device)
    aT64 = aT.cpu().numpy().astype(np.complex128)
    for b in range(len(parts)):
        pr = DF.Problem(parts[b], aT64, N, BL, RADIUS, L)
        s = pr.score(omega[b], t[b])
        assert abs(out[b, 0] - s) < 1e-4 * abs(s), (b, out[b, 0], s)
        ref = pr.fd_gradient(omega[b], t[b])
        err = np.abs(out[b, 1:] - ref).max() / np.abs(ref).max()
        assert err < 1e-3, (b, out[b, 1:], ref)


def test_gradient_matches_fp32_central_differences(gpu_device, setup):
    eng, aT, parts, omega, t = setup
    out, _ = _gpu(eng, omega, t, aT, parts, gpu_device)
    fd = np.zeros((len(parts), 6))
    for i in range(3):
        for j, (h, arr) in enumerate([(1e-3, omega), (1e-2, t)]):
            e = np.zeros(3)
            e[i] = h
            p = _gpu(eng, omega + (e if j == 0 else 0), t + (e if j == 1 else 0), aT, parts, gpu_device)[0][:, 0]
            m = _gpu(eng, omega - (e if j == 0 else 0), t - (e if j == 1 else 0), aT, parts, gpu_device)[0][:, 0]
            fd[:, 3 * j + i] = (p - m) / (2 * h)
    for b in range(len(parts)):
        err = np.abs(out[b, 1:] - fd[b]).max() / np.abs(fd[b]).max()
        assert err < 1e-3, (b, out[b, 1:], fd[b])


def test_template_gradient_is_the_exact_directional_change(gpu_device, setup):
    eng, aT, parts, omega, t = setup
    out, G = _gpu(eng, omega, t, aT, parts, gpu_device, template_gradient=True)
    rng = np.random.default_rng(1)
    delta = torch.tensor(rng.standard_normal(eng.ne) + 1j * rng.standard_normal(eng.ne),
                         dtype=torch.complex64, device=gpu_device) * aT.abs().mean()
    out2, _ = _gpu(eng, omega, t, aT + delta, parts, gpu_device)
    pred = (torch.conj(G) * delta[None]).real.sum(dim=1).double().cpu().numpy()
    change = out2[:, 0] - out[:, 0]
    for b in range(len(parts)):
        assert abs(pred[b] - change[b]) < 1e-4 * (abs(out[b, 0]) + abs(change[b])), (b, pred[b], change[b])
    # against the FP64 oracle's directional change
    aT64 = aT.cpu().numpy().astype(np.complex128)
    d64 = delta.cpu().numpy().astype(np.complex128)
    for b in range(len(parts)):
        ref = DF.Problem(parts[b], aT64, N, BL, RADIUS, L).template_directional(omega[b], t[b], d64)
        assert abs(pred[b] - ref) < 1e-4 * (abs(out[b, 0]) + abs(ref)), (b, pred[b], ref)


def test_autograd_function(gpu_device, setup):
    eng, aT, parts, omega, t = setup
    om = torch.tensor(omega, dtype=torch.float32, device=gpu_device, requires_grad=True)
    tt = torch.tensor(t, dtype=torch.float32, device=gpu_device, requires_grad=True)
    a = aT.clone().requires_grad_(True)
    w = torch.tensor([0.5, -2.0, 1.25], device=gpu_device)
    s = eng.score(om, tt, a, torch.tensor(parts, device=gpu_device))
    (s * w).sum().backward()
    out, G = _gpu(eng, omega, t, aT, parts, gpu_device, template_gradient=True)
    wn = w.cpu().numpy()
    np.testing.assert_allclose(s.detach().cpu().numpy(), out[:, 0], rtol=1e-6)
    np.testing.assert_allclose(om.grad.cpu().numpy(), wn[:, None] * out[:, 1:4], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(tt.grad.cpu().numpy(), wn[:, None] * out[:, 4:7], rtol=1e-5, atol=1e-6)
    want = (w[:, None].to(torch.complex64) * G).sum(dim=0)
    np.testing.assert_allclose(a.grad.cpu().numpy(), want.cpu().numpy(), rtol=1e-4, atol=1e-6 * float(want.abs().max()))
    # no template gradient requested: G is not formed
    om2 = torch.tensor(omega, dtype=torch.float32, device=gpu_device, requires_grad=True)
    eng.score(om2, tt.detach(), aT, torch.tensor(parts, device=gpu_device)).sum().backward()
    np.testing.assert_allclose(om2.grad.cpu().numpy(), out[:, 1:4], rtol=1e-5, atol=1e-6)


def test_score_is_maximal_at_the_planted_pose(gpu_device):
    # a particle that is the template rotated by R and moved by an integer shift:
    # the score's gradient vanishes (to the interpolation error) at the true pose
    from matcha_amd.diff import ScoreEngine
    from matcha_amd.synthetic import Workload
    n, bl, r, ll = 48, 12, 21.0, 12
    wl = Workload(n, 20, (1.5, 3.0), 7, gpu_device, mask_radius=r)
    vols, R, t = wl.particles(2, 3, None)
    eng = ScoreEngine(n, bl, r, ll, 2, gpu_device)
    aT = eng.template_coefficients(wl.template)
    from scipy.spatial.transform import Rotation
    omega = torch.tensor(Rotation.from_matrix(R).as_rotvec(), dtype=torch.float32, device=gpu_device)
    shift = torch.tensor(t[:, ::-1].copy(), dtype=torch.float32, device=gpu_device)
    out, _ = eng.score_and_gradients(omega, shift, aT, vols)
    out = out.cpu().numpy()
    # relative size of the gradient at the optimum vs at a pose 3 degrees away
    omega_off = omega + torch.tensor([0.05, 0.0, 0.0], device=gpu_device)
    off, _ = eng.score_and_gradients(omega_off, shift, aT, vols)
    off = off.cpu().numpy()
    for b in range(2):
        assert np.linalg.norm(out[b, 1:4]) < 0.1 * np.linalg.norm(off[b, 1:4]), (out[b], off[b])
        assert out[b, 0] > off[b, 0]


def test_gradient_at_box128_L48_matches_fp32_central_differences(gpu_device):
    # the differentiation workload's shape: box 128, FLE bandlimit 48 (lmax 67),
    # score to L = 48, FP32 coefficients; gradients in (omega, t) against FP32
    # central differences at 1e-3 rad / 1e-2 px, a few degrees and a fraction of
    # a voxel away from the planted pose
    from matcha_amd.diff import ScoreEngine
    from matcha_amd.synthetic import Workload
    from scipy.spatial.transform import Rotation
    n, bl, r, ll, B = 128, 48, 60.0, 48, 2
    wl = Workload(n, 200, (1.0, 5.0), 5, gpu_device, mask_radius=r)
    vols, R, t = wl.particles(B, 6, -6.0)
    eng = ScoreEngine(n, bl, r, ll, B, gpu_device)
    assert eng.lmax == 67
    aT = eng.template_coefficients(wl.template)
    rng = np.random.default_rng(9)
    omega = np.stack([(Rotation.from_rotvec(rng.normal(0, 0.03, 3)) * Rotation.from_matrix(Ri)).as_rotvec()
                      for Ri in R])
    shift = t[:, ::-1] + rng.uniform(-0.4, 0.4, (B, 3))

    def run(om, sh):
        out, _ = eng.score_and_gradients(torch.tensor(om, dtype=torch.float32, device=gpu_device),
                                         torch.tensor(sh, dtype=torch.float32, device=gpu_device), aT, vols)
        return out.cpu().numpy().astype(np.float64)

    out = run(omega, shift)
    fd = np.zeros((B, 6))
    for i in range(3):
        for j, h in enumerate((1e-3, 1e-2)):
            e = np.zeros(3)
            e[i] = h
            p = run(omega + (e if j == 0 else 0), shift + (e if j == 1 else 0))[:, 0]
            m = run(omega - (e if j == 0 else 0), shift - (e if j == 1 else 0))[:, 0]
            fd[:, 3 * j + i] = (p - m) / (2 * h)
    for b in range(B):
        err = np.abs(out[b, 1:] - fd[b]).max() / np.abs(fd[b]).max()
        assert err < 1e-3, (b, out[b, 1:], fd[b])


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_split_precision_projection_keeps_the_bars(gpu_device, precision):
    # the projection of the patch and its tangent volumes in the 16-bit hi/lo
    # split GEMM: score and (omega, t) gradient against the FP64 oracle to the
    # FP32 path's bars (measured here: score 2e-7, gradient 2e-6 relative)
    from matcha_amd.diff import ScoreEngine
    rng = np.random.default_rng(42)
    tmpl = fle.blob_volume(rng, N, 8, sigma_range=(1.0, 2.5))
    parts = np.stack([fle.blob_volume(rng, N, 8, sigma_range=(1.0, 2.5)) for _ in range(3)]).astype(np.float32)
    omega = np.array([[0.4, -1.1, 0.7], [2.0, 0.3, -0.9], [1e-5, -2e-5, 1e-5]])
    t = np.array([[0.3, -0.45, 0.2], [-1.35, 0.6, 0.15], [0.25, 0.25, -0.7]])
    eng = ScoreEngine(N, BL, RADIUS, L, 4, gpu_device, compute_precision=precision)
    aT = eng.template_coefficients(torch.tensor(tmpl, dtype=torch.float32, device=gpu_device))
    out, _ = _gpu(eng, omega, t, aT, parts, gpu_device)
    aT64 = aT.cpu().numpy().astype(np.complex128)
    for b in range(len(parts)):
        pr = DF.Problem(parts[b], aT64, N, BL, RADIUS, L)
        s = pr.score(omega[b], t[b])
        assert abs(out[b, 0] - s) < 1e-4 * abs(s), (b, out[b, 0], s)
        ref = pr.fd_gradient(omega[b], t[b])
        assert np.abs(out[b, 1:] - ref).max() < 1e-3 * np.abs(ref).max(), (b, out[b, 1:], ref)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n, bl, radius, nv, bp", [(24, 8, 11.0, 24, 3), (24, 8, 13.0, 28, 70),
                                                    (32, 10, 16.0, 36, 5)])
def test_fused_tangent_projection_matches_volumes_through_hbm(gpu_device, precision, n, bl, radius, nv, bp):
    # the patch and tangent volumes formed inside the operand build against
    # the same four volumes written by mch_shift_tangent and projected: the
    # operand values are the same trilinear formulas, so the coefficients
    # agree to FP32 rounding of the class sums (70 particles: 280 volumes, so
    # the real and imaginary operand rows span three 256-row tiles; radius 13
    # in box 24: images beyond the box; nv > n: patches cut from larger volumes;
    # shifts on cell faces and negative)
    from matcha_amd.diff import ScoreEngine
    rng = np.random.default_rng(7)
    eng = ScoreEngine(n, bl, radius, 4, 4 * bp, gpu_device, compute_precision=precision)
    vols = torch.tensor(rng.standard_normal((bp, nv, nv, nv)), dtype=torch.float32, device=gpu_device)
    t = rng.uniform(-2.5, 2.5, (bp, 3))
    t[0] = (1.0, -2.0, 0.0)
    t = torch.tensor(t, dtype=torch.float32, device=gpu_device)
    fused = eng.tangent_coefficients(vols, t, fused=True)
    ref = eng.tangent_coefficients(vols, t, fused=False)
    scale = ref.abs().amax(dim=1, keepdim=True).clamp_min(1e-30)
    err = ((fused - ref).abs() / scale).max().item()
    assert err < 2e-6, err
