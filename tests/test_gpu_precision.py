"""Precision of the search against the FP64 oracle, per projection precision.

The GPU path (projection in FP32, BF16x3 or FP16x3 -- 16-bit hi/lo split
operands, three MFMA products, FP32 accumulation -- then sigma and the SO(3)
search in FP32 with FP64 Wigner-d tables) against the oracle's whole chain in
FP64 (brute-force projection, sigma, SOFFT grid, Newton march), on the same
float32 particles.

Bars, for every precision: orientations within 0.1 degree geodesic of the FP64
result and scores within 1e-4 relative (the targets stated for the FP32
path); coefficients at box 128 / bandlimit 48 within 2e-5 (FP32, FP16x3) and
8e-5 (BF16x3) of the coefficient vector's norm against FP64 brute force.
"""
import numpy as np
import pytest
import torch

from oracle import fle, so3
from tests.so3_helpers import geodesic_deg

pytestmark = pytest.mark.gpu

MODES = ["fp32", "bf16x3", "fp16x3"]
COEFF_TOL = {"fp32": 2e-5, "fp16x3": 2e-5, "bf16x3": 8e-5}


@pytest.fixture(scope="module")
def small_problem(gpu_device):
    """Box 40, bandlimit 16 (lmax 19), search L = 16 (march 8, 12 -> 16):
    4 rotated blob particles, no shift, 6 dB, and the FP64 oracle's result."""
    from matcha_amd.synthetic import Workload
    N, bl, radius = 40, 16, 18
    wl = Workload(N, 24, (1.5, 3.0), 41, gpu_device, mask_radius=radius)
    vols, R, _ = wl.particles(4, 0, 6.0)
    tmpl = wl.template.cpu().numpy().astype(np.float64)
    v64 = vols.cpu().numpy().astype(np.float64)
    ls, ks, ms, mds, lmds, cs, ne = fle.lap_eig_ball(N, bl)
    lmax = int(ls.max())
    aT = fle.evaluate_t(tmpl[None], N, bl, radius)[0]
    aV = fle.evaluate_t(v64, N, bl, radius)
    sigma = fle.compute_sigma(aT, aV, ls, mds, lmax)
    ref = so3.search_orientations(sigma, 16, [8, 12], candidates=10, num_steps=5, K=2)
    return dict(N=N, bl=bl, radius=radius, wl=wl, vols=vols, R=R, ref=ref)


@pytest.mark.parametrize("mode", MODES)
def test_search_end_to_end_vs_fp64_oracle(gpu_device, small_problem, mode):
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    p = small_problem
    ccm = CrossCorrelationMatcher(p["N"], gpu_device, batchsize=4, bandlimit=p["bl"], radius=p["radius"],
                                  compute_precision=mode)
    ccm.set_template(p["wl"].template)
    m = Matcha(4, gpu_device, 16, {"Li": [8, 12], "candidates": 10, "num_steps": 5, "oversampling_factor_K": 2})
    sig, _ = ccm.get_sigma(p["vols"], max_degree=16)
    A, Bt, G, S = (x.cpu().numpy().astype(np.float64) for x in m.search_orientations(sig[:4]))
    ra, rb, rg, rs = p["ref"]
    geo = geodesic_deg(A, Bt, G, ra, rb, rg)
    assert geo.max() < 0.1, geo
    assert np.max(np.abs(S - rs) / np.abs(rs)) < 1e-4
    # and the oracle itself found the planted rotations
    from matcha_amd.synthetic import euler_zyz_to_matrix
    from matcha_amd.synthetic import geodesic_deg as geo_R
    assert np.max(geo_R(p["R"], euler_zyz_to_matrix(ra, rb, rg))) < 1.0


@pytest.mark.parametrize("bl", [48, 65])
def test_projection_box128_vs_fp64_subset(gpu_device, bl):
    # bandlimit 48 (the precision workload) and 65 (the reference's production
    # correlator bandlimit, max(Li) of its config; lmax 92, a 17.9 GB table)
    from matcha_amd.core.FLEBasis3D import FLEBasis3D
    N, radius = 128, 60
    rng = np.random.default_rng(128)
    vol = fle.blob_volume(rng, N, 60, sigma_range=(1.0, 5.0)) + 0.5 * rng.standard_normal((N, N, N))
    vols = vol[None].astype(np.float32)
    f = FLEBasis3D(N, bl, 1e-4, batchsize=2, radius=radius, device=gpu_device)
    pick = sorted(set([0, 1, f.ne - 1, int(np.argmax(f.ls)), int(np.argmax(np.abs(f.ms)))]
                      + list(rng.choice(f.ne, 12, replace=False))))
    ref = fle.evaluate_t_subset(vols.astype(np.float64), N, bl, radius, pick)[0]
    got = {}
    for mode in MODES:
        f.set_compute_precision(mode)
        got[mode] = f.evaluate_t(torch.tensor(vols, device=gpu_device)).cpu().numpy()[0]
    scale = np.linalg.norm(got["fp32"])
    for mode in MODES:
        err = np.max(np.abs(got[mode][pick] - ref)) / scale
        assert err < COEFF_TOL[mode], (mode, err)
        if mode != "fp32":
            assert np.max(np.abs(got[mode] - got["fp32"])) < (COEFF_TOL[mode] + 2e-5) * scale
