"""Uneven batches on the GPU: graph replays reproduce the eager pipeline, a
batch of mixed march ends gives each particle the result of a matcher built for
its own end, and a small two-box, two-schedule set with missing wedges is
aligned."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload(gpu_device):
    from matcha_amd.synthetic import UnevenWorkload
    wl = UnevenWorkload(boxes=(32, 48), ratio=(1, 1), L_ends=(8, 12), seed=3, device=gpu_device,
                        tilt_range=(60.0, 60.0), snr_range=(6.0, 6.0), max_shift=2)
    return wl


def _aligner(wl, dev, graphs):
    from matcha_amd.batching import UnevenAligner
    tmpl = {N: wl.work[N].template for N in wl.boxes}
    return UnevenAligner(tmpl, dev, radius=wl.radius, bandlimit=16, batch=4, L_ends=wl.L_ends, Li=(8,),
                         K=2, alternations=2, precision="bf16x3", graphs=graphs, max_shift=2)


def test_graph_replays_match_eager(gpu_device, workload):
    wl = workload
    N = 48
    vols, wedges, R, t, snr, tilt = wl.particles(N, 4)
    eager = _aligner(wl, gpu_device, graphs=False)
    ends = [12, 8, 12, 8]
    ref = eager.group(N, 8).run(vols, wedges, ends)
    ref = [x.clone() for x in ref]
    graph = _aligner(wl, gpu_device, graphs=True)
    bk = graph.group(N, 8)
    other, ow, *_ = wl.particles(N, 4)
    for rep in range(4):
        got = bk.run(vols, wedges, ends)
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            np.testing.assert_allclose(g.cpu().numpy(), r.cpu().numpy(), rtol=1e-6, atol=1e-6,
                                       err_msg=f"replay {rep}")
        # eager work of the program between replays, and a replay on other data
        eager.group(N, 8).run(other, ow, [8, 8, 12, 12])
        bk.run(other, ow, [8, 8, 12, 12])
        torch.cuda.synchronize()
    assert bk.graph is not None


def test_uneven_set_is_aligned(gpu_device, workload):
    from matcha_amd.synthetic import euler_zyz_to_matrix, geodesic_deg
    wl = workload
    al = _aligner(wl, gpu_device, graphs=True)
    specs = wl.draw(18)
    data = {}
    for N in wl.boxes:
        idx = [i for i, s in enumerate(specs) if s[0] == N]
        if idx:
            data[N] = (idx,) + tuple(wl.particles(N, len(idx)))
    particles = [None] * len(specs)
    ctfs = [None] * len(specs)
    Rt = np.zeros((len(specs), 3, 3))
    tt = np.zeros((len(specs), 3), dtype=int)
    for N, (idx, vols, wedges, R, t, snr, tilt) in data.items():
        for j, i in enumerate(idx):
            particles[i], ctfs[i], Rt[i], tt[i] = vols[j], wedges[j], R[j], t[j]
    res = al.align(specs, particles, ctfs)
    err = geodesic_deg(Rt, euler_zyz_to_matrix(res["alpha"].astype(np.float64), res["beta"].astype(np.float64),
                                               res["gamma"].astype(np.float64)))
    assert np.median(err) < 3.0, err
    assert np.mean(np.all(res["shift"].astype(int) == tt[:, ::-1], axis=1)) >= 0.8
    # every group that had a full batch was replayed from a graph
    assert any(b.graph is not None for b in al.groups.values())


def test_benchmark_boxes_and_schedules_are_aligned(gpu_device):
    # the batching workload's shapes: boxes 64 / 96 / 128 (each its own
    # template), march ends 16 / 24 / 32, missing wedges of +-60 deg, 0 dB,
    # shifts in [-4, 4]^3; every (box, L) pair gets particles (one group per
    # box, its batches mixing the three march ends), and the planted rotations
    # and integer shifts are recovered
    from matcha_amd.batching import UnevenAligner
    from matcha_amd.synthetic import UnevenWorkload, euler_zyz_to_matrix, geodesic_deg
    wl = UnevenWorkload(seed=21, device=gpu_device, tilt_range=(60.0, 60.0), snr_range=(0.0, 0.0))
    per = 8
    specs = [(N, L) for N in wl.boxes for L in wl.L_ends for _ in range(per)]
    n = len(specs)
    particles, ctfs = [None] * n, [None] * n
    Rt = np.zeros((n, 3, 3))
    tt = np.zeros((n, 3), dtype=int)
    for N in wl.boxes:
        idx = [i for i, s in enumerate(specs) if s[0] == N]
        vols, wedges, R, t, snr, tilt = wl.particles(N, len(idx))
        for j, i in enumerate(idx):
            particles[i], ctfs[i], Rt[i], tt[i] = vols[j], wedges[j], R[j], t[j]
    al = UnevenAligner({N: wl.work[N].template for N in wl.boxes}, gpu_device, radius=wl.radius, bandlimit=32,
                       batch=16, L_ends=wl.L_ends, precision="bf16x3", graphs=True, max_shift=wl.max_shift)
    res = al.align(specs, particles, ctfs)
    err = geodesic_deg(Rt, euler_zyz_to_matrix(res["alpha"].astype(np.float64), res["beta"].astype(np.float64),
                                               res["gamma"].astype(np.float64)))
    ok = np.all(res["shift"].astype(int) == tt[:, ::-1], axis=1)
    assert len(al.groups) == 3
    for N in wl.boxes:
        m = np.array([s[0] == N for s in specs])
        assert np.median(err[m]) < 1.5, (N, np.sort(err[m]))
        assert ok[m].mean() >= 0.9, (N, ok[m].mean())


def test_mixed_march_ends_equal_one_matcher_per_end(gpu_device, workload):
    # one group's batch mixes march ends 8 and 12: every particle's result is,
    # bit for bit, the one of the same alternation run with a matcher built
    # for its own end (the reference's one-Matcha-per-schedule run)
    # (the per-end groups share the box's projection and shift plans, so the
    # GEMM runs the same tile configuration)
    from matcha_amd.batching import _Group
    wl = workload
    N = 48
    vols, wedges, *_ = wl.particles(N, 4)
    ends = [12, 8, 8, 12]
    al = _aligner(wl, gpu_device, graphs=False)
    mixed = [x.cpu().numpy() for x in al.group(N, 8).run(vols, wedges, ends)]
    for L in (8, 12):
        sel = [i for i, e in enumerate(ends) if e == L]
        one = _Group(al.boxes[N], L, al.batch, al.device, al.cfg, al.alternations, graphs=False)
        assert one.matcher.Lmax == L
        ref = [x.cpu().numpy() for x in one.run(vols[sel], wedges[sel], [L] * len(sel))]
        for k, (m, r) in enumerate(zip(mixed, ref)):
            np.testing.assert_array_equal(m[sel], r, err_msg=f"output {k}, end {L}")
