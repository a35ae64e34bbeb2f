"""Benchmark: aligned volumes/s for Matcha's orientation-search hot path on one MI355X.

Headline (the metric's own definition): the tuned path -- batched (1024
particles per batch, consecutive batches alternating over two streams so that
independent batches overlap on the device) with the projection GEMM on 16-bit
hi/lo split operands (three MFMA products, FP32 accumulation) and every other
stage in FP32 (the shift search keeps complex64 storage, so its integer shifts
are the reference's) -- against the plain FP32 path at batch 64 on one
stream, both measured in this run on the same synthetic particles, each
required to reach <= 1 deg median geodesic error.  ``value`` is the tuned
path's throughput; the plain path's throughput and the ratio are in
``config``; ``dtype`` names the precision of every stage that is not FP32.

A step = one batch of particles aligned on the GPU by two rotation/shift
alternations (matcha_amd.align.Alternation), each:
  CrossCorrelationMatcher.get_sigma  (ball-harmonic projection + sigma)
  Matcha.search_orientations         (SOFFT grid + frequency-marched Newton)
  ShiftMatcher.search_shifts         (rotated-reference FFT cross-correlation)
with the particles re-extracted at the running shift before the second one.
Every input is already resident in HBM.  Prints one JSON line.

Workload (box 96, L = 32): 96^3 template of 120 Gaussian blobs (sigma
U(1,5) px, amplitude U(0.5,1), centres in a ball of 0.35*96), particles =
Haar-random rotation + integer shift in [-6,6]^3 + white noise at 0 dB over
the radius-45 mask; FLE bandlimit 32 (ne 8252, lmax 42), radius 45; march
Li = 8, 16, 24 -> L_max 32; SOFFT grid 64^3 (L0 = 8, K = 4); 10 candidates;
5 Newton steps; plain path batch 64, tuned path batch 1024 (2048 particles
cycled).

Three kernels take 17-24% of the step each: the tuned projection GEMM
(``roofline_gemm16``: HBM bound by its algorithmic operands, ``frac`` on
those bytes, ``mfma_frac`` on the one-product FLOPs over the nonzero operand,
the issue rate of the three split products in ``mfma_issue_frac``), the fused
Newton march (``roofline_march``: FP32 vector bound) and the shift search's
column pass (``roofline_shift_z``: HBM bound); each is timed live
with HIP events on its stream.  ``roofline`` repeats the one with the largest
share of the step in this run.  ``cpu_baseline`` is a torch-CPU FP32
restatement of the reference's algorithm (NUFFT projection, SHT, radial
interpolation, sigma, SOFFT, Newton march, FFT shift search) timed on this
host on a slice of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

# HBM bytes the tuned projection GEMM moves per launch, measured with
# rocprofv3 --pmc FETCH_SIZE (scaled by 2.0: FETCH_SIZE of a device copy of
# 512 MiB reads 256 MiB) and WRITE_SIZE (the K-split partials of its busy work
# items); keyed by (workload, batch, waves per workgroup, 128-column tiles per
# workgroup, K splits) of the configuration the autotuner chose (8 waves is a
# 256-row tile).
GEMM16_TRAFFIC_BYTES = {("box96_L32", 64, 4, 1, 20): 587089 * 1024 * 2 + 47360 * 1024,
                        ("box96_L32", 1024, 8, 2, 2): 1812678 * 1024 * 2 + 67235 * 1024}
# HBM bytes (FETCH_SIZE x 2.0 + WRITE_SIZE) per launch of the shift search's
# column pass, keyed by (workload, batch, reference table precision); measured
# like the GEMM's.
SHIFT_Z_TRAFFIC_BYTES = {("box96_L32", 512, "fp32"): 2979462 * 1024 * 2 + 1953582 * 1024,
                         ("box96_L32", 768, "fp32"): 4466747 * 1024 * 2 + 2928934 * 1024,
                         ("box96_L32", 1024, "fp32"): 5957119 * 1024 * 2 + 3903117 * 1024}
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


WORKLOADS = {
    "box96_L32": dict(N=96, n_blobs=120, sigma=(1.0, 5.0), bandlimit=32, radius=45, L_max=32, tuned_batch=1024,
                      Li=[8, 16, 24], K=4, C=10, steps=5, batch=64, max_shift=6),
    "box64_L16": dict(N=64, n_blobs=40, sigma=(1.5, 4.0), bandlimit=16, radius=28, L_max=16,
                      Li=[8, 12], K=2, C=10, steps=5, batch=64, max_shift=4),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="box96_L32", choices=sorted(WORKLOADS))
    p.add_argument("--batch", type=int, default=None, help="plain-path batch (default: the workload's)")
    p.add_argument("--tuned-batch", type=int, default=None, help="tuned-path batch (default: the workload's)")
    p.add_argument("--tuned-precision", default="bf16x3", choices=["bf16x3", "fp16x3"])
    p.add_argument("--tuned-shift-precision", default="fp32", choices=["fp32", "reduced"],
                   help="storage precision of the tuned path's shift search (fp32: complex64 as the reference; "
                        "reduced: FP16 reference table and BF16 work spectrum, lags of near-tied "
                        "correlations may differ by one voxel)")
    p.add_argument("--tuned-streams", type=int, default=2,
                   help="streams the tuned path's batches alternate over (one plan set each)")
    p.add_argument("--no-plain", action="store_true", help="skip the plain FP32 path")
    p.add_argument("--second-snr", type=float, default=-10.0,
                   help="noise level of an untimed accuracy check of both paths (the workload's "
                        "second SNR); pass nan to skip")
    p.add_argument("--snr", type=float, default=0.0)
    p.add_argument("--pool", type=int, default=2048, help="distinct particles cycled through")
    p.add_argument("--alternations", type=int, default=2)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--diff", action="store_true",
                   help="also measure the differentiable score (256 particles, box 128, L = 48)")
    p.add_argument("--diff-batch", type=int, default=256)
    p.add_argument("--batching", action="store_true",
                   help="also measure uneven batches (4096 particles, boxes 64/96/128, wedges, L 16/24/32)")
    p.add_argument("--batching-particles", type=int, default=4096)
    p.add_argument("--batching-batch", type=int, default=512)
    p.add_argument("--diff-steps", type=int, default=3)
    p.add_argument("--precision-study", action="store_true",
                   help="also run the precision study (box 128, L = 48: FP32 / BF16x3 / FP16x3 on the same particles)")
    p.add_argument("--precision-particles", type=int, default=512)
    p.add_argument("--precision-batch", type=int, default=128)
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="untimed device copy of this size before the run (HBM counter calibration)")
    p.add_argument("--cpu-particles", type=int, default=64,
                   help="particles of the pool the CPU baseline aligns (timed in full)")
    return p.parse_args()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, args, wl, pool_vols, pool_R, pool_t):
    """The reference's algorithm restated for torch on the CPU in FP32
    (oracle/torch_cpu.py: NUFFT projection, SHT, radial interpolation, sigma,
    SOFFT, Newton march, FFT shift search), two global alternations on the
    first ``--cpu-particles`` particles of the same pool, timed in full on the
    host cores (tables and the Fourier reference built beforehand, untimed)."""
    import torch
    from matcha_amd.synthetic import euler_zyz_to_matrix, geodesic_deg
    from oracle import torch_cpu as TC
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    torch.set_num_threads(cores)
    P = min(args.cpu_particles, pool_vols.shape[0])
    vols = pool_vols[:P].float().cpu()
    t0 = time.perf_counter()
    pipe = TC.CpuPipeline(wl.template.cpu().numpy(), w["N"], w["bandlimit"], w["radius"], w["L_max"], w["Li"],
                          w["K"], w["C"], w["steps"])
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    a, b, g, s, sh, ss = pipe.align(vols, alternations=args.alternations)
    el = time.perf_counter() - t0
    err = geodesic_deg(pool_R[:P], euler_zyz_to_matrix(a.double().numpy(), b.double().numpy(), g.double().numpy()))
    ok = np.all(sh.numpy().astype(int) == pool_t[:P, ::-1], axis=1)
    return {
        "value": P / el, "unit": "volumes/s", "cores": cores, "kind": "port",
        "sample": (f"{P} particles of the same pool, {args.alternations} alternations in full (not extrapolated): "
                   f"torch-CPU FP32 restatement of the reference's algorithm (type-2 NUFFT width 5 / 2x "
                   f"oversampling, SHT, radial interpolation, sigma, SOFFT grid, Newton march with "
                   f"materialised D and the 15-column bank, FFT shift search); {el:.1f} s on {cores} torch "
                   f"threads, {_cpu_model()}; setup {setup:.1f} s untimed"),
        "median_err_deg": float(np.median(err)), "shift_exact_frac": float(ok.mean()),
        "torch_threads": torch.get_num_threads(), "cpu_model": _cpu_model(),
    }


def run_path(name, torch, dev, w, args, pool_vols, pool_R, pool_t, precision, B, plan_sets):
    """Throughput, accuracy and per-stage times of one path.

    plan_sets: one (correlator, matcher, shift matcher) per stream; batch k of
    the timed steps runs on stream k mod len(plan_sets), so that independent
    batches overlap on the device (each stream owns its plans' buffers)."""
    from matcha_amd.align import Alternation
    from matcha_amd.core.ShiftMatcher import euler_to_shift_rotations
    from matcha_amd.synthetic import euler_zyz_to_matrix, geodesic_deg
    for c, _, _ in plan_sets:
        c.fle.set_compute_precision(precision)
    pipes = [Alternation(c, m, s, num_alternations=args.alternations, mode="global") for c, m, s in plan_sets]
    ccm, matcher, sm = plan_sets[0]
    pipe = pipes[0]
    nstreams = len(pipes)
    streams = [torch.cuda.current_stream(dev)] + [torch.cuda.Stream(device=dev) for _ in range(nstreams - 1)]
    P = pool_vols.shape[0]
    nb = max(1, P // B)
    batches = [pool_vols[i * B:(i + 1) * B] for i in range(nb)]

    def step(k):
        with torch.cuda.stream(streams[k % nstreams]):
            pipes[k % nstreams].run(batches[k % nb])

    for i in range(max(args.warmup, nstreams)):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    vps = args.steps * B / elapsed
    # accuracy over the pool (untimed)
    errs, shift_ok = [], []
    for i in range(nb):
        a, b, g, ns, sh, ss = pipe.run(batches[i])
        a, b, g = (x.cpu().numpy().astype(np.float64) for x in (a, b, g))
        R = pool_R[i * B:(i + 1) * B]
        t = pool_t[i * B:(i + 1) * B]
        errs.append(geodesic_deg(R, euler_zyz_to_matrix(a, b, g)))
        shift_ok.append(np.all(sh.cpu().numpy() == t[:, ::-1], axis=1))
    errs = np.concatenate(errs)
    shift_ok = np.concatenate(shift_ok)
    # per-stage time of one alternation (HIP events on the compute stream)
    stream = torch.cuda.current_stream(dev)
    vols0 = batches[0]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    sigma, norms = ccm.get_sigma(vols0)
    ev[1].record(stream)
    a0, b0, g0, s0 = matcher.search_orientations(sigma[:vols0.shape[0]])
    ev[2].record(stream)
    sm.search_shifts(subtomos=vols0, rotation_tracker=euler_to_shift_rotations(a0, b0, g0))
    ev[3].record(stream)
    torch.cuda.synchronize()
    return {
        "precision": precision, "shift_precision": sm.precision, "batch": B, "streams": nstreams,
        "volumes_per_s": vps,
        "ms_per_step": elapsed / args.steps * 1e3,
        "median_err_deg": float(np.median(errs)), "p90_err_deg": float(np.percentile(errs, 90)),
        "shift_exact_frac": float(shift_ok.mean()), "particles_checked": int(errs.size),
        "stage_ms_per_batch": {"get_sigma": ev[0].elapsed_time(ev[1]),
                               "search_orientations": ev[1].elapsed_time(ev[2]),
                               "search_shifts": ev[2].elapsed_time(ev[3])},
    }


def accuracy_at(torch, plan_sets, precision, args, vols, R, t):
    """Median / p90 geodesic error and exact-shift fraction of one path on a
    pool of particles (untimed): the same alternations as the timed run."""
    from matcha_amd.align import Alternation
    from matcha_amd.synthetic import euler_zyz_to_matrix, geodesic_deg
    c, m, s = plan_sets[0]
    c.fle.set_compute_precision(precision)
    pipe = Alternation(c, m, s, num_alternations=args.alternations, mode="global")
    B = m.batchsize
    errs, ok = [], []
    for i in range(0, vols.shape[0], B):
        a, b, g, ns, sh, ss = pipe.run(vols[i:i + B])
        a, b, g = (x.cpu().numpy().astype(np.float64) for x in (a, b, g))
        errs.append(geodesic_deg(R[i:i + B], euler_zyz_to_matrix(a, b, g)))
        ok.append(np.all(sh.cpu().numpy() == t[i:i + B][:, ::-1], axis=1))
    errs, ok = np.concatenate(errs), np.concatenate(ok)
    return {"median_err_deg": float(np.median(errs)), "p90_err_deg": float(np.percentile(errs, 90)),
            "shift_exact_frac": float(ok.mean()), "particles_checked": int(errs.size)}


def time_gemm(torch, dev, ccm, vols, precision, reps=20):
    """Average duration of the projection GEMM launch (HIP events on its stream)."""
    from matcha_amd import _lib
    fle_ = ccm.fle
    fle_.set_compute_precision(precision)
    fle_.evaluate_t(vols)
    B = vols.shape[0]
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        _lib.check(_lib.lib().mch_fle_gemm_only(fle_._plan, B, _lib.stream_handle(dev)), "gemm")
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def time_march(torch, dev, ccm, matcher, vols, reps=5):
    """Average duration of one Newton march over the grid candidates of `vols`
    (`Matcha.march`: the sigma repack and the march kernel, HIP events on the
    current stream)."""
    sigma, _ = ccm.get_sigma(vols, max_degree=matcher.Lmax)
    a, b, g, _ = matcher.grid_candidates(sigma[:vols.shape[0]])
    matcher.march(sigma[:vols.shape[0]], a, b, g)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        matcher.march(sigma[:vols.shape[0]], a, b, g)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def march_flops(levels, num_steps, candidates, batch):
    """FP32 operations of the fused march's shell work: per fundamental pair
    (m' >= |m|) and shell l >= m', the three-term recurrences of d, d', d''
    (29) and the twelve accumulating products with sigma (24) for the
    evaluations whose derivatives are used (num_steps per level), and the d
    recurrence (14) with its four products (8) for the last trial of a level,
    whose score alone is used, for every candidate."""
    pair_shells = lambda L: sum((2 * m + 1) * (L - m + 1) for m in range(L + 1))
    per = 53.0 * num_steps + (22.0 if num_steps > 0 else 53.0)
    return per * sum(pair_shells(L) for L in levels) * candidates * batch


def time_shift_kernels(torch, dev, sm, vols, reps=10):
    """Per-kernel ms of the shift search (HIP events around its launches on its
    stream): forward (x, y), column pass (z FFTs + cross power), inverse (y, x),
    final."""
    import ctypes
    from matcha_amd import _lib
    from matcha_amd.core.ShiftMatcher import euler_to_shift_rotations
    B = vols.shape[0]
    g = torch.Generator(device="cpu").manual_seed(5)
    rb = euler_to_shift_rotations(*[(torch.rand(B, generator=g) * s).to(dev) for s in (6.283, 3.1416, 6.283)])
    lib = _lib.lib()
    _lib.check(lib.mch_shift_plan_set_timing(sm._plan, 1), "mch_shift_plan_set_timing")
    buf = (ctypes.c_float * 4)()
    acc = np.zeros(4)
    try:
        sm.search_shifts(subtomos=vols, rotation_tracker=rb)
        for _ in range(reps):
            sm.search_shifts(subtomos=vols, rotation_tracker=rb)
            _lib.check(lib.mch_shift_kernel_ms(sm._plan, buf), "mch_shift_kernel_ms")
            acc += np.array(buf[:], dtype=np.float64)
    finally:
        lib.mch_shift_plan_set_timing(sm._plan, 0)
    return acc / reps


def _lib_depth(fle_):
    from matcha_amd import _lib
    return _lib.lib().mch_fle_tuned_depth(fle_._plan)


def diff_workload(torch, dev, args):
    """Differentiable score at box 128, bandlimit 48, L = 48 with FP32
    coefficients: time of score + gradients in (omega, t, aT) for the batch,
    and a check of the rotation/shift gradient against FP32 central
    differences (1e-3 rad / 1e-2 px) on a few particles; the same with the
    projection in BF16x3 (and its agreement with the FP32 run)."""
    from matcha_amd.diff import ScoreEngine
    from matcha_amd.synthetic import Workload
    from scipy.spatial.transform import Rotation
    N, bl, radius, L, B = 128, 48, 60, 48, args.diff_batch
    t0 = time.perf_counter()
    wl = Workload(N, 200, (1.0, 5.0), args.seed + 11, dev, mask_radius=radius)
    eng = ScoreEngine(N, bl, radius, L, B, dev)
    aT = eng.template_coefficients(wl.template)
    chunks = [wl.particles(min(64, B - i), 6, -6.0) for i in range(0, B, 64)]
    vols = torch.cat([c[0] for c in chunks])
    R = np.concatenate([c[1] for c in chunks])
    tt = np.concatenate([c[2] for c in chunks])
    rng = np.random.default_rng(args.seed + 12)
    # poses near the truth (a few degrees, a fraction of a voxel off)
    omega_np = np.stack([(Rotation.from_rotvec(rng.normal(0, 0.03, 3)) * Rotation.from_matrix(r)).as_rotvec()
                         for r in R])
    shift_np = tt[:, ::-1] + rng.uniform(-0.4, 0.4, (B, 3))
    omega = torch.tensor(omega_np, dtype=torch.float32, device=dev, requires_grad=True)
    shift = torch.tensor(shift_np, dtype=torch.float32, device=dev, requires_grad=True)
    a = aT.clone().requires_grad_(True)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0

    def step():
        omega.grad = shift.grad = a.grad = None
        s = eng.score(omega, shift, a, vols)
        s.sum().backward()
        return s

    def fd_check(nchk):
        """Worst relative error of the (omega, t) gradient against FP32 central
        differences of the score (1e-3 rad / 1e-2 px) on the first particles."""
        with torch.no_grad():
            base, _ = eng.score_and_gradients(omega[:nchk], shift[:nchk], aT, vols[:nchk])
            fd = torch.zeros((nchk, 6), device=dev, dtype=torch.float64)
            for j, h in ((0, 1e-3), (1, 1e-2)):
                for i in range(3):
                    e = torch.zeros(3, device=dev)
                    e[i] = h
                    args_p = [omega[:nchk] + (e if j == 0 else 0), shift[:nchk] + (e if j == 1 else 0)]
                    args_m = [omega[:nchk] - (e if j == 0 else 0), shift[:nchk] - (e if j == 1 else 0)]
                    sp = eng.score_and_gradients(args_p[0], args_p[1], aT, vols[:nchk])[0][:, 0].double()
                    sm = eng.score_and_gradients(args_m[0], args_m[1], aT, vols[:nchk])[0][:, 0].double()
                    fd[:, 3 * j + i] = (sp - sm) / (2 * h)
            an = base[:, 1:].double()
            return ((an - fd).abs().amax(dim=1) / fd.abs().amax(dim=1)).max().item(), base

    # the FP32 projection (the configuration's FP32 coefficients), then the
    # BF16x3 split GEMM for the projection of the 4B patch and tangent volumes
    nchk = min(4, B)
    runs = {}
    for prec in ("fp32", "bf16x3"):
        eng.ccm.fle.set_compute_precision(prec)
        step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.diff_steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) / args.diff_steps * 1e3
        rel, base = fd_check(nchk)
        runs[prec] = {"ms_per_step": ms, "particles_per_s": B / ms * 1e3,
                      "fd_check": {"particles": nchk, "steps": "1e-3 rad / 1e-2 px", "max_rel_err": rel},
                      "_base": base}
    b32, b16 = runs["fp32"].pop("_base").double(), runs["bf16x3"].pop("_base").double()
    runs["bf16x3"]["vs_fp32"] = {
        "score_max_rel": ((b16[:, 0] - b32[:, 0]).abs() / b32[:, 0].abs()).max().item(),
        "gradient_max_rel": ((b16[:, 1:] - b32[:, 1:]).abs().amax(dim=1) / b32[:, 1:].abs().amax(dim=1)).max().item()}
    eng.ccm.fle.set_compute_precision("fp32")
    return {
        "workload": f"{B} particles, box {N}, FLE bandlimit {bl} radius {radius}, L = {L}, -6 dB, FP32 coefficients",
        "ms_per_step": runs["fp32"]["ms_per_step"], "particles_per_s": runs["fp32"]["particles_per_s"],
        "step": "score + d/d(axis-angle, shift, template coefficients), forward and backward",
        "fd_check": runs["fp32"]["fd_check"],
        "precisions": runs,
        "setup_s": setup,
    }


def precision_workload(torch, dev, args):
    """Precision study: particles at box 128 (200-blob template, sigma U(1,5) px),
    FLE bandlimit 48 radius 60, march 8..40 -> L = 48, -6 dB, integer shifts in
    [-6,6]^3.  The same particles go through two alternations with the
    projection GEMM in FP32, BF16x3 and FP16x3 (16-bit hi/lo split operands,
    three MFMA products, FP32 accumulation; the Wigner-d recurrences, quadrature
    nodes and the 3x3 Newton solve stay FP64 in every mode).  Per precision:
    volumes/s, error against the planted poses, and agreement with the FP32
    result (geodesic angle, integer shifts, normalised scores, coefficients)."""
    from matcha_amd.align import Alternation
    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from matcha_amd.synthetic import Workload, euler_zyz_to_matrix, geodesic_deg
    N, bl, radius, L, Li, shift_r = 128, 48, 60, 48, [8, 16, 24, 32, 40], 6
    P, B = args.precision_particles, args.precision_batch
    t0 = time.perf_counter()
    wl = Workload(N, 200, (1.0, 5.0), args.seed + 31, dev, mask_radius=radius)
    ccm = CrossCorrelationMatcher(N, dev, batchsize=B, bandlimit=bl, radius=radius, compute_precision="fp32")
    ccm.set_template(wl.template)
    m = Matcha(B, dev, L, {"Li": Li, "candidates": 10, "num_steps": 5, "oversampling_factor_K": 4})
    sh = ShiftMatcher({"N": N, "num_subtomograms_per_batch": B, "shift_upsample_factor": 1,
                       "shift_search_radius": shift_r, "voxel_size": 1.0}, dev)
    sh.set_reference(wl.template.cpu().numpy())
    chunks = [wl.particles(min(64, P - i), shift_r, -6.0) for i in range(0, P, 64)]
    vols = torch.cat([c[0] for c in chunks])
    R = np.concatenate([c[1] for c in chunks])
    tt = np.concatenate([c[2] for c in chunks])
    pipe = Alternation(ccm, m, sh, num_alternations=2, mode="global")
    batches = [vols[i:i + B] for i in range(0, P, B)]
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    res = {}
    for mode in ("fp32", "bf16x3", "fp16x3"):
        ccm.fle.set_compute_precision(mode)
        pipe.run(batches[0])   # the 16-bit tables are built and tuned on first use
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        outs = []
        for vb in batches:
            a, b, g, ns, s, _ = pipe.run(vb)
            outs.append(torch.cat([torch.stack([a, b, g, ns], 1).float(), s.float()], 1))
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        o = torch.cat(outs).double().cpu().numpy()
        coeffs = ccm.fle.evaluate_t(batches[0]).clone()
        res[mode] = {"out": o, "s": el, "coeffs": coeffs}
    ccm.fle.set_compute_precision("fp32")
    ref = res["fp32"]
    Rref = euler_zyz_to_matrix(*ref["out"][:, :3].T)
    report = {}
    for mode, r in res.items():
        o = r["out"]
        Rm = euler_zyz_to_matrix(*o[:, :3].T)
        err = geodesic_deg(R, Rm)
        ok = np.all(o[:, 4:7].astype(int) == tt[:, ::-1], axis=1)
        d = {"volumes_per_s": P / r["s"], "median_err_deg": float(np.median(err)),
             "p90_err_deg": float(np.percentile(err, 90)), "shift_exact_frac": float(ok.mean())}
        if mode != "fp32":
            dis = geodesic_deg(Rref, Rm)
            c32 = ref["coeffs"]
            d["vs_fp32"] = {
                "median_deg": float(np.median(dis)), "p99_deg": float(np.percentile(dis, 99)),
                "max_deg": float(dis.max()), "within_0.1deg_frac": float((dis <= 0.1).mean()),
                "shifts_equal_frac": float(np.all(o[:, 4:7] == ref["out"][:, 4:7], axis=1).mean()),
                "score_max_rel": float(np.max(np.abs(o[:, 3] - ref["out"][:, 3]) / np.abs(ref["out"][:, 3]))),
                "coeff_max_err_rel": float(((r["coeffs"] - c32).abs().max() / c32.abs().max()).item()),
            }
        report[mode] = d
    return {
        "workload": (f"{P} particles, box {N}, 200 blobs, FLE bandlimit {bl} radius {radius}, march "
                     f"{Li}->{L}, grid 64^3, -6 dB, shifts in [-{shift_r},{shift_r}]^3, 2 alternations, "
                     f"batch {B}"),
        "ne": int(ccm.fle.ne), "lmax": int(ccm.lmax), "precisions": report, "setup_s": setup,
    }


def batching_workload(torch, dev, args):
    """Uneven batches: particles of boxes 64/96/128 (2:2:1, one blob template per
    box), missing wedges of +-U(45, 65) deg about a randomly offset y axis,
    SNR U(-12, 0) dB, march ends L in {16, 24, 32}; grouped per box, each
    batch mixing the march ends in one search launch, eager launches per batch
    (bf16x3 projection, FLE bandlimit 32); a HIP graph replay of one full batch
    is timed beside its eager launch."""
    from matcha_amd.batching import UnevenAligner
    from matcha_amd.synthetic import UnevenWorkload, euler_zyz_to_matrix, geodesic_deg
    n = args.batching_particles
    t0 = time.perf_counter()
    wl = UnevenWorkload(seed=args.seed + 21, device=dev)
    specs = wl.draw(n)
    particles, ctfs = [None] * n, [None] * n
    Rt = np.zeros((n, 3, 3))
    tt = np.zeros((n, 3), dtype=int)
    for N in wl.boxes:
        idx = [i for i, sp in enumerate(specs) if sp[0] == N]
        for c in range(0, len(idx), 64):
            sel = idx[c:c + 64]
            vols, wedges, R, t, snr, tilt = wl.particles(N, len(sel))
            for j, i in enumerate(sel):
                particles[i], ctfs[i], Rt[i], tt[i] = vols[j], wedges[j], R[j], t[j]
    templates = {N: wl.work[N].template for N in wl.boxes}
    al = UnevenAligner(templates, dev, radius=wl.radius, bandlimit=32, batch=args.batching_batch,
                       L_ends=wl.L_ends, precision=args.tuned_precision, graphs=False, max_shift=wl.max_shift)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    # warm every group: plan set-up and autotune stay out of the timed pass
    from matcha_amd.batching import plan_groups
    for (N, L0), chunks in plan_groups(specs, al.batch, al.cfg["Li"]).items():
        full = [c for c in chunks if len(c) == al.batch]
        if full:
            al.group(N, L0).run([particles[i] for i in full[0]], [ctfs[i] for i in full[0]],
                                [specs[i][1] for i in full[0]])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = al.align(specs, particles, ctfs)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    est = euler_zyz_to_matrix(res["alpha"].astype(np.float64), res["beta"].astype(np.float64),
                              res["gamma"].astype(np.float64))
    err = geodesic_deg(Rt, est)
    ok = np.all(res["shift"].astype(int) == tt[:, ::-1], axis=1)
    per_box = {}
    for N in wl.boxes:
        m = np.array([sp[0] == N for sp in specs])
        per_box[str(N)] = {"particles": int(m.sum()), "median_err_deg": float(np.median(err[m])),
                           "shift_exact_frac": float(ok[m].mean())}
    # graph replay against eager launches of the same full batch (box 96, mixed ends)
    L0 = min(al.cfg["Li"])
    bk = al.group(96, L0)
    sel = [i for i, sp in enumerate(specs) if sp[0] == 96][:al.batch]
    # both modes take the particles as a list, as align() passes them: eager
    # launches stack them, the graph stacks them into its static buffer
    vb, cb = [particles[i] for i in sel], [ctfs[i] for i in sel]
    eb = [specs[i][1] for i in sel]
    reps = 10
    times = {}
    for mode in ("graph", "eager"):
        bk.graphs = mode == "graph"
        bk.run(vb, cb, eb)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(reps):
            bk.run(vb, cb, eb)
        torch.cuda.synchronize()
        times[mode] = (time.perf_counter() - t2) / reps * 1e3
    bk.graphs = False
    return {
        "workload": (f"{n} particles, boxes 64/96/128 at 2:2:1 (one template each), missing wedge "
                     f"+-U(45,65) deg, SNR U(-12,0) dB, march end L in {{16,24,32}}, 2 alternations, "
                     f"batch {al.batch} per box with mixed march ends, {args.tuned_precision}, eager launches; box 64: 40 blobs sigma "
                     f"U(1.5,4), box 96: 120 and box 128: 200 blobs sigma U(1,5)"),
        "particles_per_s": n / el, "seconds": el, "groups": len(al.groups),
        "median_err_deg": float(np.median(err)), "shift_exact_frac": float(ok.mean()), "per_box": per_box,
        "graph_vs_eager_ms_per_batch_96_mixed_ends": times, "setup_s": t_data,
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X; "
              "multi-GPU runs (--gpus > 1) are not supported", file=sys.stderr)
        sys.exit(2)
    import torch

    from matcha_amd.core.CrossCorrelationMatcher import CrossCorrelationMatcher
    from matcha_amd.core.Matcha import Matcha
    from matcha_amd.core.ShiftMatcher import ShiftMatcher
    from matcha_amd.synthetic import Workload

    w = dict(WORKLOADS[args.workload])
    Bp = args.batch or w["batch"]
    Bt = args.tuned_batch or w.get("tuned_batch", w["batch"])
    Bmax = max(Bp, Bt)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration_mib > 0:
        src = torch.empty(args.copy_calibration_mib << 18, dtype=torch.float32, device=dev).normal_()
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t_setup = time.perf_counter()
    wl = Workload(w["N"], w["n_blobs"], w["sigma"], args.seed, dev, mask_radius=w["radius"])

    def plans(B, precision, shift_precision):
        """Projection, search and shift plans sized for one path's batch."""
        c = CrossCorrelationMatcher(w["N"], dev, batchsize=B, bandlimit=w["bandlimit"], radius=w["radius"],
                                    compute_precision=precision)
        c.set_template(wl.template)
        m = Matcha(B, dev, w["L_max"], {"Li": w["Li"], "candidates": w["C"], "num_steps": w["steps"],
                                         "oversampling_factor_K": w["K"]})
        sh = ShiftMatcher({"N": w["N"], "num_subtomograms_per_batch": B, "shift_upsample_factor": 1,
                           "shift_search_radius": w["max_shift"], "voxel_size": 1.0,
                           "shift_precision": shift_precision}, dev)
        sh.set_reference(wl.template.cpu().numpy())
        return c, m, sh

    # the 16-bit tables are built (and the GEMM tuned) here, one plan set per stream
    tuned_sets = [plans(Bt, args.tuned_precision, args.tuned_shift_precision)
                  for _ in range(max(1, args.tuned_streams))]
    ccm, matcher, sm = tuned_sets[0]
    plain_plans = None if args.no_plain else plans(Bp, "fp32", "fp32")
    P = max(args.pool, Bmax)
    chunks = [wl.particles(min(64, P - i), w["max_shift"], args.snr) for i in range(0, P, 64)]
    pool_vols = torch.cat([c[0] for c in chunks])
    pool_R = np.concatenate([c[1] for c in chunks])
    pool_t = np.concatenate([c[2] for c in chunks])
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    tuned = run_path("tuned", torch, dev, w, args, pool_vols, pool_R, pool_t, args.tuned_precision, Bt,
                     tuned_sets)
    plain = None
    if not args.no_plain:
        plain = run_path("plain", torch, dev, w, args, pool_vols, pool_R, pool_t, "fp32", Bp, [plain_plans])
    # accuracy of both paths at the workload's second noise level (untimed)
    snr2 = None
    if args.second_snr is not None and not math.isnan(args.second_snr):
        n2 = max(Bt, Bp)
        ch2 = [wl.particles(min(64, n2 - i), w["max_shift"], args.second_snr) for i in range(0, n2, 64)]
        v2 = torch.cat([c[0] for c in ch2])
        R2 = np.concatenate([c[1] for c in ch2])
        t2 = np.concatenate([c[2] for c in ch2])
        snr2 = {"snr_db": args.second_snr,
                "tuned": accuracy_at(torch, tuned_sets, args.tuned_precision, args, v2, R2, t2)}
        if plain_plans is not None:
            snr2["plain"] = accuracy_at(torch, [plain_plans], "fp32", args, v2, R2, t2)
        del v2

    # the tuned path's projection GEMM: algorithmic bytes = W^T hi + lo (4 B per
    # element) + the nonzero part of the hi/lo X operand (4 B per element), each
    # read once; algorithmic flops = one product over the nonzero K of each
    # class (V0 for m = 0, 2 mod 4, 2 V0 for m = 1, 3 mod 4); the GEMM issues
    # three (hi hi + hi lo + lo hi), reported as mfma_issue_frac
    fle_ = ccm.fle
    gemm_t = time_gemm(torch, dev, ccm, pool_vols[:Bt], args.tuned_precision)
    march_t = time_march(torch, dev, ccm, matcher, pool_vols[:Bt])
    march_fl = march_flops(matcher.levels, matcher.num_steps, matcher.candidates, Bt)
    gemm_plain = time_gemm(torch, dev, plain_plans[0], pool_vols[:Bp], "fp32") if plain else None
    w_bytes = 4.0 * fle_.Ntot * fle_.Kpad          # W^T hi + lo, 2 + 2 B per element
    x_bytes = Bt * fle_.operand_bytes_per_volume   # X hi + lo, nonzero halves of the 6 slabs
    alg_bytes = w_bytes + x_bytes
    achieved = alg_bytes / gemm_t / 1e9
    flops_alg = Bt * fle_.flops_per_volume                               # one product
    flops16 = 3 * flops_alg                                              # the three split products issued
    fle_.set_compute_precision(args.tuned_precision)
    depth = int(_lib_depth(fle_))
    tiles = {"k_splits": depth % 10000 // 100, "waves_per_workgroup": depth % 100 // 10,
             "lds_stages": depth % 10, "column_tiles_per_workgroup": depth // 10000 + 1}
    traffic = GEMM16_TRAFFIC_BYTES.get((args.workload, Bt, tiles["waves_per_workgroup"],
                                        tiles["column_tiles_per_workgroup"], tiles["k_splits"]))
    gemm_hbm = achieved / 8000.0
    gemm_mfma = flops_alg / gemm_t / 1e12 / 2500.0
    gemm_issue = flops16 / gemm_t / 1e12 / 2500.0

    # the shift search's column pass (one block of 4 kx x
    # 8 ky columns: forward z FFT, rotated-reference sampling, cross power,
    # inverse z FFT).  Algorithmic bytes: the half spectrum read and written
    # (B H N^2 elements each way, 8 B complex FP32 or 4 B complex BF16 in the
    # reduced mode) + the background spectrum XMy and the reference table once
    # ((N + 2)^2 (N + 1) entries of 16 B, FP32 pairs or FP16 quads);
    # algorithmic flops: 2 passes of H N length-N FFTs
    # (5 N log2 N each) + 38 per sampled frequency (trilinear complex sample,
    # norms, product)
    shift_ms = time_shift_kernels(torch, dev, sm, pool_vols[:Bt])
    N, H = w["N"], w["N"] // 2 + 1
    z_t = shift_ms[1] * 1e-3
    z_red = args.tuned_shift_precision == "reduced"
    z_bytes = 2.0 * Bt * H * N * N * (4 if z_red else 8) + H * N * N * 8 + (N + 2) ** 2 * (N + 1) * 16
    z_flops = Bt * H * (2 * N * 5.0 * N * math.log2(N) + N * N * 38.0)
    z_gbs = z_bytes / z_t / 1e9

    # the fused Newton march, bound by FP32 vector issue: achieved = the fused
    # algorithm's shell flops per launch (march_flops) / its live duration;
    # peak = the published FP32 rate, which counts packed FP32 (frac_nonpacked:
    # against half of it, the rate without packed instructions, which this
    # build does not use)
    roof_march = {"bound": "fp32_vector", "achieved": march_fl / march_t / 1e12, "peak": 157.3,
                  "unit": "TFLOP/s", "frac": march_fl / march_t / 1e12 / 157.3, "traffic": None,
                  "frac_nonpacked": march_fl / march_t / 1e12 / 78.65,
                  "kernel": "so3_march_kernel (with so3_repack_kernel)", "avg_ms": march_t * 1e3,
                  "flops_per_launch": march_fl, "share_of_step": 2 * march_t * 1e3 / tuned["ms_per_step"]}
    dtype = {"bf16x3": "bf16x3+f32acc", "fp16x3": "fp16x3+f32acc"}[args.tuned_precision] + " projection GEMM"
    if args.tuned_shift_precision == "reduced":
        dtype += ", fp16 reference table + bf16 spectrum in the shift search"
    dtype += ", fp32 elsewhere (fp64 Wigner/Bessel tables and Newton solve)"
    out = {
        "metric": "aligned volumes/s, box 96 L=32, at <=1 deg median error: tuned vs plain FP32",
        "value": tuned["volumes_per_s"],
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tuned["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (Gaussian-blob template, Haar rotations, integer shifts, AWGN)",
        "config": {
            "workload": (f"{args.workload}: box {w['N']}, {w['n_blobs']} blobs, FLE bandlimit "
                         f"{w['bandlimit']} radius {w['radius']}, march {w['Li']}->{w['L_max']}, "
                         f"grid {2 * min(w['Li']) * w['K']}^3, {w['C']} candidates, {w['steps']} Newton "
                         f"steps, SNR {args.snr} dB, {P} particles cycled"),
            "path": (f"{args.alternations} x (get_sigma + search_orientations + search_shifts); tuned batches "
                     f"over {max(1, args.tuned_streams)} streams"),
            "tuned": tuned,
            "plain": plain,
            "tuned_over_plain": (tuned["volumes_per_s"] / plain["volumes_per_s"]) if plain else None,
            "meets_accuracy": tuned["median_err_deg"] <= 1.0 and (plain is None or plain["median_err_deg"] <= 1.0),
            "accuracy_second_snr": snr2,
            "gemm_ms": {"tuned": gemm_t * 1e3, "plain_fp32": gemm_plain * 1e3 if gemm_plain else None},
            "gemm_tiles": tiles,
            "setup_s": t_setup,
        },
        "roofline_march": roof_march,
        "roofline_shift_z": {"bound": "hbm", "achieved": z_gbs, "peak": 8000.0, "unit": "GB/s",
                             "frac": z_gbs / 8000.0,
                             "traffic": SHIFT_Z_TRAFFIC_BYTES.get((args.workload, Bt, args.tuned_shift_precision)),
                             "kernel": f"shift_z_kernel<{N}, {'true' if z_red else 'false'}>", "avg_ms": z_t * 1e3,
                             "bytes_per_launch": z_bytes, "flops_per_launch": z_flops,
                             "fp32_vector_tflops": z_flops / z_t / 1e12, "fp32_vector_frac": z_flops / z_t / 157.3e12,
                             "share_of_step": 2 * z_t * 1e3 / tuned["ms_per_step"],
                             "shift_kernels_ms": dict(zip(("fwd_xy", "z", "inv_xy", "final"), map(float, shift_ms)))},
        "roofline_gemm16": {"bound": "mfma" if gemm_mfma > gemm_hbm else "hbm",
                            "achieved": flops_alg / gemm_t / 1e12 if gemm_mfma > gemm_hbm else achieved,
                            "peak": 2500.0 if gemm_mfma > gemm_hbm else 8000.0,
                            "unit": "TFLOP/s" if gemm_mfma > gemm_hbm else "GB/s",
                            "frac": max(gemm_mfma, gemm_hbm), "hbm_frac": gemm_hbm, "mfma_frac": gemm_mfma,
                            "mfma_issue_frac": gemm_issue, "split_products": 3,
                            "traffic": traffic,
                            "kernel": (("fle_gemm16w_kernel" if tiles["column_tiles_per_workgroup"] == 2
                                        else "fle_gemm16_kernel") + f" ({args.tuned_precision})"),
                            "avg_ms": gemm_t * 1e3, "bytes_per_launch": alg_bytes,
                            "flops_per_launch": flops_alg, "flops_issued_per_launch": flops16,
                            "share_of_step": 2 * gemm_t * 1e3 / tuned["ms_per_step"]},
    }
    # `roofline`: the dominant kernel of this run, the one of the three with the
    # largest share of the step (the march, the GEMM and the shift search's
    # column pass take 18-21% each; which leads can change from box to box)
    lead = max(("roofline_march", "roofline_gemm16", "roofline_shift_z"), key=lambda k: out[k]["share_of_step"])
    out["roofline"] = dict(out[lead], entry=lead)
    if args.diff:
        out["differentiation"] = diff_workload(torch, dev, args)
    if args.batching:
        out["batching"] = batching_workload(torch, dev, args)
    if args.precision_study:
        out["precision_study"] = precision_workload(torch, dev, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args, wl, pool_vols, pool_R, pool_t)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
