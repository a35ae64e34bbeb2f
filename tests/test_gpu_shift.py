"""GPU parity of the FFT shift match (HIP C-ABI) against the FP64 oracle.

Bar: integer shifts bit-exact; scores within 1e-4 relative (FP32 path).
"""
import numpy as np
import pandas as pd
import pytest
import torch

from oracle import fle, shift as S

pytestmark = pytest.mark.gpu


def _matcher(dev, N, B, **kw):
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    cfg = {"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1,
           "shift_search_radius": 2, "voxel_size": 1.0, "do_ctf_correction": True}
    cfg.update(kw)
    return ShiftMatcher(cfg, dev, torch.complex64)


def _rand_quats(rng, B):
    q = rng.standard_normal((B, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


@pytest.mark.parametrize("N", [32, 48])
def test_circular_shift_closed_form(gpu_device, N):
    rng = np.random.default_rng(N)
    tmpl = fle.blob_volume(rng, N, 12, sigma_range=(1.5, 3.0))
    shifts = [(2, -3, 1), (0, 0, 0), (-5, 4, 3)]
    vols = np.stack([np.roll(tmpl, t, axis=(0, 1, 2)) for t in shifts]).astype(np.float32)
    sm = _matcher(gpu_device, N, len(shifts))
    sm.set_reference(tmpl)
    q = np.tile([1.0, 0, 0, 0], (len(shifts), 1))
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None,
                              shift_zyx=torch.zeros((len(shifts), 1, 3), device=gpu_device),
                              rotation_tracker=q)
    got = sh.cpu().numpy()
    for g, t in zip(got, shifts):
        assert tuple(g.astype(int)) == tuple(t[::-1])
        assert np.all(g == np.round(g))
    assert np.all(sc.cpu().numpy() > 0.9)


@pytest.mark.parametrize("N,B", [(32, 4), (64, 2), (80, 2), (112, 2)])
def test_matches_oracle_random_rotations(gpu_device, N, B):
    rng = np.random.default_rng(100 + N)
    tmpl = fle.blob_volume(rng, N, 15, sigma_range=(1.5, 3.5))
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    q = _rand_quats(rng, B)
    rot = tracker_to_matrices(q)
    vols = []
    for b in range(B):
        t = rng.integers(-4, 5, 3)
        v = np.roll(tmpl, tuple(t), axis=(0, 1, 2)) + 0.5 * rng.standard_normal(tmpl.shape)
        vols.append(v)
    vols = np.stack(vols).astype(np.float32)
    prior = np.zeros((B, 1, 3))
    prior[0, 0] = (1.0, -2.0, 0.0)
    sm = _matcher(gpu_device, N, B)
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None,
                              shift_zyx=torch.tensor(prior, device=gpu_device), rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), prior, rot, N=N)
    assert np.array_equal(sh.cpu().numpy(), rsh.astype(np.float32))
    assert np.max(np.abs(sc.cpu().numpy() - rsc) / np.abs(rsc)) < 1e-4


def test_fractional_prior_and_ctf(gpu_device):
    N, B = 32, 2
    rng = np.random.default_rng(5)
    tmpl = fle.blob_volume(rng, N, 12, sigma_range=(1.5, 3.0))
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    q = _rand_quats(rng, B)
    rot = tracker_to_matrices(q)
    vols = np.stack([np.roll(tmpl, (1, 2, -1), axis=(0, 1, 2)),
                     np.roll(tmpl, (-2, 0, 3), axis=(0, 1, 2))]).astype(np.float32)
    prior = np.array([[[0.5, -1.25, 0.0]], [[0.0, 0.0, 2.0]]])
    ctf = rng.uniform(0.2, 1.0, (B, N, N, N // 2 + 1)).astype(np.float32)
    sm = _matcher(gpu_device, N, B)
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device),
                              ctfs=torch.tensor(ctf, device=gpu_device),
                              shift_zyx=torch.tensor(prior, device=gpu_device), rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), prior, rot, N=N, ctfs=ctf.astype(np.float64))
    assert np.array_equal(sh.cpu().numpy(), rsh.astype(np.float32))
    assert np.max(np.abs(sc.cpu().numpy() - rsc) / np.abs(rsc)) < 1e-4


def test_full_size_box96(gpu_device):
    N, B = 96, 3
    rng = np.random.default_rng(96)
    tmpl = fle.blob_volume(rng, N, 60, sigma_range=(1.0, 5.0))
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    q = _rand_quats(rng, B)
    q[0] = (1.0, 0.0, 0.0, 0.0)
    rot = tracker_to_matrices(q)
    vols = np.stack([np.roll(tmpl, tuple(rng.integers(-6, 7, 3)), axis=(0, 1, 2))
                     + 0.3 * rng.standard_normal(tmpl.shape) for _ in range(B)]).astype(np.float32)
    sm = _matcher(gpu_device, N, B)
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None,
                              shift_zyx=None, rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), np.zeros((B, 1, 3)), rot, N=N)
    assert np.array_equal(sh.cpu().numpy(), rsh.astype(np.float32))
    assert np.max(np.abs(sc.cpu().numpy() - rsc) / np.abs(rsc)) < 1e-4


@pytest.mark.parametrize("N", [144, 160, 192])
def test_large_boxes_match_oracle(gpu_device, N):
    # boxes beyond 128 (line lengths 12 x 12, 10 x 16, 12 x 16 in the LDS FFTs):
    # a planted integer shift under a rotation near the identity (a clear
    # peak) and one under a random rotation, bit-exact against the oracle
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    rng = np.random.default_rng(N)
    tmpl = fle.blob_volume(rng, N, 16, sigma_range=(2.0, 5.0))
    q = np.array([[np.cos(0.01), np.sin(0.01), 0.0, 0.0], _rand_quats(rng, 1)[0]])
    rot = tracker_to_matrices(q)
    t = (5, -7, 3)
    vols = np.stack([np.roll(tmpl, t, axis=(0, 1, 2)) + 0.2 * rng.standard_normal(tmpl.shape),
                     np.roll(tmpl, (-2, 1, 4), axis=(0, 1, 2))]).astype(np.float32)
    sm = _matcher(gpu_device, N, 2)
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None, shift_zyx=None,
                              rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), np.zeros((2, 1, 3)), rot, N=N)
    assert np.array_equal(sh.cpu().numpy(), rsh.astype(np.float32))
    assert tuple(sh.cpu().numpy()[0].astype(int)) == t[::-1]
    assert np.max(np.abs(sc.cpu().numpy() - rsc) / np.abs(rsc)) < 1e-4


def test_subvoxel_refinement_beyond_box128_is_rejected(gpu_device):
    sm = _matcher(gpu_device, 144, 1, shift_upsample_factor=2)
    with pytest.raises(RuntimeError, match="up to 128"):
        sm.set_reference(np.zeros((144, 144, 144), dtype=np.float32))


def test_extract_patches_matches_oracle(gpu_device):
    from matcha_amd.core.ShiftMatcher import extract_shifted_patches
    N = 32
    rng = np.random.default_rng(9)
    vols = rng.standard_normal((3, N, N, N)).astype(np.float32)
    shifts = np.array([[2.0, -1.0, 3.0], [0.5, 0.0, -1.25], [0.0, 0.0, 0.0]], dtype=np.float32)
    got = extract_shifted_patches(torch.tensor(vols, device=gpu_device),
                                  torch.tensor(shifts, device=gpu_device)).cpu().numpy()
    for b in range(3):
        ref = S.extract_shifted_patch(vols[b].astype(np.float64), shifts[b][::-1].astype(np.float64))
        assert np.max(np.abs(got[b] - ref)) < 1e-5


@pytest.mark.parametrize("N", [48, 96])
def test_extract_patches_integer_shifts_exact(gpu_device, N):
    # integral shifts are a shifted copy with zeros outside: exact, including
    # shifts that leave part or all of the volume (per axis and both signs)
    from matcha_amd.core.ShiftMatcher import extract_shifted_patches
    rng = np.random.default_rng(N)
    B = 8
    vols = rng.standard_normal((B, N, N, N)).astype(np.float32)
    shifts = rng.integers(-N // 3, N // 3 + 1, (B, 3)).astype(np.float32)
    shifts[0] = (0, 0, 0)
    shifts[1] = (N - 1, -(N - 1), 1)
    shifts[2] = (-N, 0, 0)
    shifts[3] = (3, 0, -N - 2)
    got = extract_shifted_patches(torch.tensor(vols, device=gpu_device),
                                  torch.tensor(shifts, device=gpu_device)).cpu().numpy()
    for b in range(B):
        ref = S.extract_shifted_patch(vols[b].astype(np.float64), shifts[b][::-1].astype(np.float64))
        assert np.array_equal(got[b], ref.astype(np.float32)), b


@pytest.mark.parametrize("mode", ["reference", "global"])
def test_alternation_recovers_rotation_and_shift(gpu_device, tmp_path, mode):
    from scipy.spatial.transform import Rotation
    from matcha_amd import relion
    from matcha_amd.align import Alternation
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from matcha_amd.synthetic import Workload, euler_zyz_to_matrix, geodesic_deg
    N, B = 64, 8
    wl = Workload(N, 40, (1.5, 4.0), 3, gpu_device, mask_radius=28)
    vols, R, t = wl.particles(B, 4, 0.0)
    ccm = CrossCorrelationMatcher(N, gpu_device, batchsize=B, bandlimit=16, radius=28)
    ccm.set_template(wl.template)
    m = Matcha(B, gpu_device, 16, {"Li": [8, 12], "oversampling_factor_K": 2})
    sm = ShiftMatcher({"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1},
                      gpu_device)
    sm.set_reference(wl.template.cpu().numpy())
    # three alternations, the reference default (num_alternations: 3): at L = 16 the first
    # orientation is too coarse for every shift, the second fixes the shifts
    pipe = Alternation(ccm, m, sm, num_alternations=3, mode=mode)
    a, b, g, ns, shift, ss = pipe.run(vols)
    P = np.array([[0.0, 0, 1], [0, 1, 0], [1, 0, 0]])
    if mode == "global":
        est = euler_zyz_to_matrix(*(x.cpu().numpy().astype(np.float64) for x in (a, b, g)))
    else:
        # the composed tracker: R(tracker) = P (R_1 R_2 ...) P on the particles' axis-0-first frame
        est = P @ relion.quat_to_matrix(pipe.tracker.q.cpu().numpy().astype(np.float64)) @ P
        assert len(pipe.history[-1]) == 7
    err = geodesic_deg(R, est)
    if mode == "global":
        assert np.max(err) < 1.0, err
    else:
        # the reference's loop searches the residual rotation of pre-rotated particles,
        # i.e. near beta = 0 where the Euler Newton step is singular and skipped: its
        # accuracy at L = 16 is grid-limited for some particles
        assert np.max(err) < 3.0 and np.median(err) < 1.5, err
    # particles were moved by +t along array axes (0,1,2); the shift comes back flipped
    assert np.array_equal(shift.cpu().numpy().astype(int), t[:, ::-1])
    # RELION STAR output of the same run: Rz(psi) Ry(tilt) Rz(rot) = Rx(pi) P R P Rx(pi)
    # with P the x<->z exchange, origins = -shift * pixel size
    names = [f"TS_01/{i}" for i in range(B)]
    star_in = {"optics": pd.DataFrame({"rlnImagePixelSize": [1.5]}),
               "particles": pd.DataFrame({"rlnTomoParticleName": names})}
    rec = relion.records_from_history(names, pipe.history)
    assert relion.join_to_star(star_in, rec, tmp_path / "out.star", num_alternations=3)
    df = relion.read_star(tmp_path / "out.star")["particles"]
    RX = np.diag([1.0, -1.0, -1.0])
    M = Rotation.from_euler("ZYZ", df[["rlnAnglePsi", "rlnAngleTilt", "rlnAngleRot"]].to_numpy(),
                            degrees=True).as_matrix()
    # the STAR angles carry exactly the estimate (FP32 tracker -> FP64 host conversion)
    assert np.max(geodesic_deg(M, RX @ P @ est @ P @ RX)) < 1e-3
    org = df[["rlnOriginXAngst", "rlnOriginYAngst", "rlnOriginZAngst"]].to_numpy()
    np.testing.assert_allclose(org, -1.5 * t[:, ::-1], atol=1e-6)


def test_reference_mode_first_alternation_matches_global(gpu_device):
    # identity tracker, zero shift: the resampled particle is the particle normalised over
    # the mask, which rescales sigma and changes only l = 0, so the search agrees
    from matcha_amd.align import Alternation
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from matcha_amd.synthetic import Workload, euler_zyz_to_matrix, geodesic_deg
    N, B = 64, 8
    wl = Workload(N, 40, (1.5, 4.0), 4, gpu_device, mask_radius=28)
    vols, R, t = wl.particles(B, 0, 0.0)
    ccm = CrossCorrelationMatcher(N, gpu_device, batchsize=B, bandlimit=16, radius=28)
    ccm.set_template(wl.template)
    m = Matcha(B, gpu_device, 16, {"Li": [8, 12], "oversampling_factor_K": 2})
    sm = ShiftMatcher({"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1}, gpu_device)
    sm.set_reference(wl.template.cpu().numpy())
    outs = {}
    for mode in ("reference", "global"):
        a, b, g, ns, shift, ss = Alternation(ccm, m, sm, num_alternations=1, mode=mode).run(vols)
        outs[mode] = euler_zyz_to_matrix(*(x.cpu().numpy().astype(np.float64) for x in (a, b, g)))
        assert np.array_equal(shift.cpu().numpy().astype(int), np.zeros((B, 3), int))
    assert np.max(geodesic_deg(outs["reference"], outs["global"])) < 0.25


@pytest.mark.parametrize("N,B,up,radius", [(64, 3, 8, 2), (32, 2, 4, 2), (48, 2, 2, 3)])
def test_upsampled_refinement_matches_oracle(gpu_device, N, B, up, radius):
    # sub-voxel refinement (the reference's local matrix DFT, its production setting is
    # up 8, radius 2): shifts in multiples of 1/up must match the FP64 oracle exactly
    rng = np.random.default_rng(200 + N + up)
    tmpl = fle.blob_volume(rng, N, 15, sigma_range=(1.5, 3.5))
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    q = _rand_quats(rng, B)
    rot = tracker_to_matrices(q)
    vols = np.stack([np.roll(tmpl, tuple(rng.integers(-radius, radius + 1, 3)), axis=(0, 1, 2))
                     + 0.4 * rng.standard_normal(tmpl.shape) for _ in range(B)]).astype(np.float32)
    prior = np.zeros((B, 1, 3))
    prior[0, 0] = (0.5, -1.0, 0.0)
    sm = _matcher(gpu_device, N, B, shift_upsample_factor=up, shift_search_radius=radius)
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None,
                              shift_zyx=torch.tensor(prior, device=gpu_device), rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), prior, rot, N=N, upsample=up, radius=radius)
    got = sh.cpu().numpy()
    assert np.array_equal(got, rsh.astype(np.float32)), (got, rsh)
    assert np.all(got * up == np.round(got * up))
    assert np.max(np.abs(sc.cpu().numpy() - rsc) / np.abs(rsc)) < 1e-4


def test_upsampled_refinement_unsupported_grid_raises(gpu_device):
    # the refinement grid R = ceil(up (2 r + 1)) must fit the box (R <= N)
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    sm = ShiftMatcher({"N": 32, "num_subtomograms_per_batch": 1, "shift_upsample_factor": 16,
                       "shift_search_radius": 2}, gpu_device)
    with pytest.raises(RuntimeError):
        sm.set_reference(np.zeros((32, 32, 32), dtype=np.float32))


def _reduced_case(rng, N, B, ctf, rotate):
    from matcha_amd.core.ShiftMatcher import tracker_to_matrices
    tmpl = fle.blob_volume(rng, N, max(12, N // 2), sigma_range=(1.0, 4.0))
    q = _rand_quats(rng, B) if rotate else np.tile([1.0, 0.0, 0.0, 0.0], (B, 1))
    rot = tracker_to_matrices(q)
    vols = np.stack([np.roll(tmpl, tuple(rng.integers(-6, 7, 3)), axis=(0, 1, 2))
                     + 0.5 * rng.standard_normal(tmpl.shape) for _ in range(B)]).astype(np.float32)
    prior = np.zeros((B, 1, 3))
    prior[0, 0] = (0.5, -1.25, 0.0)
    ctfs = rng.uniform(0.2, 1.0, (B, N, N, N // 2 + 1)).astype(np.float32) if ctf else None
    return tmpl, q, rot, vols, prior, ctfs


def _search_and_oracle(dev, N, B, tmpl, q, rot, vols, prior, ctfs, **cfg):
    sm = _matcher(dev, N, B, **cfg)
    sm.set_reference(tmpl)
    kw = dict(subtomos=torch.tensor(vols, device=dev),
              ctfs=None if ctfs is None else torch.tensor(ctfs, device=dev),
              shift_zyx=torch.tensor(prior, device=dev), rotation_tracker=q)
    sh, sc = sm.search_shifts(**kw)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), prior, rot, N=N,
                                ctfs=None if ctfs is None else ctfs.astype(np.float64))
    return sm, kw, sh, sc, rsh, rsc


@pytest.mark.parametrize("N,B,ctf", [(32, 4, True), (48, 8, False), (48, 4, True), (96, 6, False),
                                     (128, 2, False)])
def test_reduced_precision_clear_peaks_match_oracle(gpu_device, N, B, ctf):
    # reduced precision (FP16 reference table, BF16 work spectrum, FP32
    # arithmetic) where the correlation has a clear peak (the reference at the
    # particles' own orientation): each stored value carries at most 2^-9
    # relative rounding, so the normalised scores stay within 2e-3 of the FP64
    # oracle and the lags, whose peak stands far above that, are exact
    rng = np.random.default_rng(300 + N + B)
    case = _reduced_case(rng, N, B, ctf, rotate=False)
    sm, kw, sh, sc, rsh, rsc = _search_and_oracle(gpu_device, N, B, *case, shift_precision="reduced")
    assert np.array_equal(sh.cpu().numpy(), rsh.astype(np.float32))
    assert np.max(np.abs(sc.cpu().numpy() - rsc)) < 2e-3
    # back on FP32: the FP32 path's own results, bit for bit
    sm.set_precision("fp32")
    sh32, sc32 = sm.search_shifts(**kw)
    fresh = _matcher(gpu_device, N, B)
    fresh.set_reference(case[0])
    shf, scf = fresh.search_shifts(**kw)
    assert torch.equal(sh32, shf) and torch.equal(sc32, scf)
    assert np.array_equal(sh32.cpu().numpy(), rsh.astype(np.float32))


@pytest.mark.parametrize("N,B", [(48, 8), (96, 4)])
def test_reduced_precision_flat_correlations(gpu_device, N, B):
    # at orientations unrelated to the particles the correlation surface is
    # flat and near-tied lags are common: reduced precision then returns a lag
    # whose correlation is within 2e-3 of the oracle's maximum, here a
    # neighbour of the oracle's lag
    rng = np.random.default_rng(500 + N + B)
    case = _reduced_case(rng, N, B, False, rotate=True)
    _, _, sh, sc, rsh, rsc = _search_and_oracle(gpu_device, N, B, *case, shift_precision="reduced")
    assert np.max(np.abs(sc.cpu().numpy() - rsc)) < 2e-3
    assert np.max(np.abs(sh.cpu().numpy() - rsh)) <= 1.0


def test_reduced_precision_upsampled(gpu_device):
    # sub-voxel refinement in reduced precision (FP16 reference table; this
    # path keeps its work spectrum in FP32): within one refinement step (1/up)
    # of the FP64 oracle on every axis, scores within 1e-3
    N, B, up, radius = 64, 4, 8, 2
    rng = np.random.default_rng(404)
    tmpl, q, rot, vols, prior, _ = _reduced_case(rng, N, B, False, rotate=True)
    vols = np.stack([np.roll(tmpl, tuple(rng.integers(-radius, radius + 1, 3)), axis=(0, 1, 2))
                     + 0.4 * rng.standard_normal(tmpl.shape) for _ in range(B)]).astype(np.float32)
    sm = _matcher(gpu_device, N, B, shift_upsample_factor=up, shift_search_radius=radius,
                  shift_precision="reduced")
    sm.set_reference(tmpl)
    sh, sc = sm.search_shifts(subtomos=torch.tensor(vols, device=gpu_device), ctfs=None,
                              shift_zyx=torch.tensor(prior, device=gpu_device), rotation_tracker=q)
    refF = S.load_reference_full(tmpl, N, 1.0).astype(np.complex64).astype(np.complex128)
    rsh, rsc = S.estimate_shift(refF, vols.astype(np.float64), prior, rot, N=N, upsample=up, radius=radius)
    got = sh.cpu().numpy()
    assert np.all(got * up == np.round(got * up))
    assert np.max(np.abs(got - rsh)) <= 1.0 / up + 1e-6, (got, rsh)
    assert np.max(np.abs(sc.cpu().numpy() - rsc)) < 1e-3


def test_shift_precision_rejects_unknown_mode(gpu_device):
    with pytest.raises(ValueError):
        _matcher(gpu_device, 32, 1, shift_precision="fp16")


def test_reference_mode_alternations_match_oracle_chain(gpu_device):
    # the reference's alternation loop (align_subtomograms.py:316-361) against its
    # FP64 restatement built from the pinned oracle pieces: per alternation the
    # rotated, normalised patches (oracle.resample), the FLE projection and sigma
    # (oracle.fle), the search (oracle.so3), the tracker update (the host rule of
    # rotation_ops.py:23-34) and the shift search on the original particles at the
    # accumulated shift (oracle.shift); every alternation's angles within 0.1
    # degree, shifts equal, trackers within 1e-4
    from scipy.spatial.transform import Rotation
    from matcha_amd import relion
    from matcha_amd.align import Alternation
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from oracle import resample as RS, so3
    from tests.so3_helpers import geodesic_deg
    N, B, bl, radius, L, Li, K, C, n_alt = 32, 4, 12, 14, 10, [4, 6, 8], 2, 10, 2
    rng = np.random.default_rng(77)
    tmpl = fle.blob_volume(rng, N, 10, sigma_range=(1.5, 3.0))
    vols = []
    for b in range(B):
        M = Rotation.random(random_state=int(rng.integers(1 << 30))).as_matrix()
        v = np.roll(RS.rotate_volume(tmpl, M), tuple(rng.integers(-2, 3, 3)), axis=(0, 1, 2))
        vols.append(v + 0.3 * rng.standard_normal(v.shape))
    vols = np.stack(vols).astype(np.float32)
    ccm = CrossCorrelationMatcher(N, gpu_device, batchsize=B, bandlimit=bl, radius=radius)
    ccm.set_template(tmpl.astype(np.float32))
    m = Matcha(B, gpu_device, L, {"Li": Li, "candidates": C, "num_steps": 5, "oversampling_factor_K": K})
    sm = ShiftMatcher({"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1,
                       "shift_search_radius": 2, "voxel_size": 1.0}, gpu_device)
    sm.set_reference(tmpl.astype(np.float32))
    pipe = Alternation(ccm, m, sm, num_alternations=n_alt, mode="reference")
    pipe.run(torch.tensor(vols, device=gpu_device))
    # the FP64 chain on the same particles
    aT = fle.evaluate_t(tmpl.astype(np.float32).astype(np.float64)[None], N, bl, radius)[0]
    refF = S.load_reference_full(tmpl.astype(np.float32), N, 1.0).astype(np.complex64).astype(np.complex128)
    v64 = vols.astype(np.float64)
    tracker = relion.identity_tracker(B)
    shift = np.zeros((B, 3))
    for it in range(n_alt):
        Mp = relion.quat_to_matrix(tracker)
        rot_vols = np.stack([RS.extract_rotated(v64[b], shift[b], Mp[b], radius) for b in range(B)])
        aV = fle.evaluate_t(rot_vols, N, bl, radius)
        sig = fle.compute_sigma(aT, aV, ccm.fle.ls, ccm.fle.mds, ccm.lmax)
        ra, rb, rg, rs = so3.search_orientations(sig, L, Li, candidates=C, num_steps=5, K=K)
        ga, gb, gg, gns, gsh, gss, gq = (x.cpu().numpy().astype(np.float64) for x in pipe.history[it])
        geo = geodesic_deg(ga, gb, gg, ra, rb, rg)
        assert geo.max() < 0.1, (it, geo)
        # the tracker composes the angles the device search returned (FP32)
        tracker, _ = relion.update_rotation_estimate(ga.astype(np.float32), gb.astype(np.float32),
                                                     gg.astype(np.float32), tracker)
        np.testing.assert_allclose(gq, tracker, atol=1e-4, err_msg=f"tracker, alternation {it}")
        Mr = np.swapaxes(relion.quat_to_matrix(tracker), 1, 2)
        ds, _ = S.estimate_shift(refF, v64, shift[:, None, :], Mr, N=N, radius=2)
        shift = shift + ds
        assert np.array_equal(gsh, shift), (it, gsh, shift)
