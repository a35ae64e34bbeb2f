"""Synthetic workloads for the benchmark and tests (GPU, torch).

Template: sum of isotropic Gaussian blobs, centres uniform in a ball of
radius 0.35*N, sigma and amplitude uniform.  Particles: the same blobs with
centres rotated by a Haar-uniform rotation R and moved by an integer shift t,
i.e. V(x) = T(R^-1 (x - t)) evaluated exactly (no interpolation), plus white
Gaussian noise at a given SNR with the signal power taken over a spherical
mask (as ``add_awgn`` of the reference harness, ``src/matcha/example.py:93-124``).

Coordinates are v = index - N/2 with axis 0 = x, 1 = y, 2 = z; rotations
act on (x, y, z).  The orientation search returns (alpha, beta, gamma) with
R = Rz(alpha) Ry(beta) Rz(gamma) for such a particle.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def blob_params(rng: np.random.Generator, N: int, n_blobs: int, sigma_range, amp_range=(0.5, 1.0),
                ball_frac=0.35):
    R = ball_frac * N
    cs = []
    while len(cs) < n_blobs:
        c = rng.uniform(-R, R, 3)
        if np.linalg.norm(c) <= R:
            cs.append(c)
    sig = rng.uniform(*sigma_range, n_blobs)
    amp = rng.uniform(*amp_range, n_blobs)
    return np.asarray(cs), sig, amp


def haar_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)
    return R


def euler_zyz_to_matrix(alpha, beta, gamma) -> np.ndarray:
    """Rz(alpha) Ry(beta) Rz(gamma) acting on (x, y, z) column vectors."""
    a, b, g = (np.asarray(t, dtype=np.float64) for t in (alpha, beta, gamma))

    def rz(t):
        c, s = np.cos(t), np.sin(t)
        o, z = np.ones_like(t), np.zeros_like(t)
        return np.stack([np.stack([c, -s, z], -1), np.stack([s, c, z], -1), np.stack([z, z, o], -1)], -2)

    def ry(t):
        c, s = np.cos(t), np.sin(t)
        o, z = np.ones_like(t), np.zeros_like(t)
        return np.stack([np.stack([c, z, s], -1), np.stack([z, o, z], -1), np.stack([-s, z, c], -1)], -2)

    return rz(a) @ ry(b) @ rz(g)


def geodesic_deg(R1: np.ndarray, R2: np.ndarray) -> np.ndarray:
    M = np.einsum('...ji,...jk->...ik', R1, R2)
    tr = np.trace(M, axis1=-2, axis2=-1)
    return np.degrees(np.arccos(np.clip((tr - 1) / 2, -1.0, 1.0)))


def render_blobs(N: int, centres: torch.Tensor, sig: torch.Tensor, amp: torch.Tensor,
                 device) -> torch.Tensor:
    """Volumes (P, N, N, N) float32 from per-volume blob centres (P, nb, 3)."""
    ax = torch.arange(N, device=device, dtype=torch.float32) - N // 2
    P, nb, _ = centres.shape
    out = torch.zeros((P, N, N, N), dtype=torch.float32, device=device)
    for j in range(nb):
        c = centres[:, j]  # (P, 3)
        inv = (-0.5 / (sig[j] * sig[j])).item()
        gx = torch.exp(inv * (ax[None, :] - c[:, 0:1]) ** 2)  # (P, N)
        gy = torch.exp(inv * (ax[None, :] - c[:, 1:2]) ** 2)
        gz = torch.exp(inv * (ax[None, :] - c[:, 2:3]) ** 2)
        out += amp[j].item() * gx[:, :, None, None] * gy[:, None, :, None] * gz[:, None, None, :]
    return out


def spherical_mask(N: int, radius: float, device) -> torch.Tensor:
    ax = torch.arange(N, device=device, dtype=torch.float32) - N // 2
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    return (torch.sqrt(X * X + Y * Y + Z * Z) <= radius)


class Workload:
    """Template + particle generator for one box size."""

    def __init__(self, N: int, n_blobs: int, sigma_range, seed: int, device, mask_radius: float):
        self.N = N
        self.device = torch.device(device)
        self.rng = np.random.default_rng(seed)
        c, s, a = blob_params(self.rng, N, n_blobs, sigma_range)
        self.centres = c
        self.sig = torch.tensor(s, dtype=torch.float32)
        self.amp = torch.tensor(a, dtype=torch.float32)
        self.mask = spherical_mask(N, mask_radius, self.device)
        self.template = render_blobs(N, torch.tensor(c[None], dtype=torch.float32, device=self.device),
                                     self.sig, self.amp, self.device)[0]
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed + 1)

    def particles(self, P: int, max_shift: int, snr_db: float):
        """(vols (P,N,N,N) float32 GPU, R (P,3,3), shifts (P,3) int)."""
        R = haar_rotations(self.rng, P)
        t = self.rng.integers(-max_shift, max_shift + 1, (P, 3))
        cen = np.einsum('pij,bj->pbi', R, self.centres) + t[:, None, :]
        vols = render_blobs(self.N, torch.tensor(cen, dtype=torch.float32, device=self.device),
                            self.sig, self.amp, self.device)
        if snr_db is not None and math.isfinite(snr_db):
            m = self.mask
            p = (vols[:, m] ** 2).mean(dim=1)
            std = torch.sqrt(p / (10.0 ** (snr_db / 10.0)))
            noise = torch.randn(vols.shape, device=self.device, generator=self.gen)
            vols = vols + noise * std[:, None, None, None]
        return vols.contiguous(), R, t


def wedge_mask_half(N: int, tilt_max_deg: float, axis_offset_deg: float, device) -> torch.Tensor:
    """Missing-wedge weights (N, N, N//2 + 1) float32 in rfftn order of a (z, y, x)
    volume (x the half axis): a tilt series about the in-plane axis y rotated by
    ``axis_offset_deg`` about z, beam along z, tilts within +-``tilt_max_deg``,
    samples the frequencies with |k_z| <= tan(tilt_max) |k_perp|, k_perp the
    component perpendicular to the tilt axis in the x-y plane."""
    kz = torch.fft.fftfreq(N, device=device)[:, None, None]
    ky = torch.fft.fftfreq(N, device=device)[None, :, None]
    kx = torch.fft.rfftfreq(N, device=device)[None, None, :]
    psi = math.radians(axis_offset_deg)
    kp = torch.abs(math.cos(psi) * kx - math.sin(psi) * ky)
    keep = torch.abs(kz) <= math.tan(math.radians(tilt_max_deg)) * kp + 1e-12
    keep = keep | ((kz == 0) & (kp == 0))
    return keep.to(torch.float32).expand(N, N, N // 2 + 1).contiguous()


# Blob templates per box size, as the single-box workloads of that size: box 64
# 40 blobs with sigma U(1.5, 4) px, box 96 120 blobs and box 128 200 blobs with
# sigma U(1, 5) px (other sizes: 120 blobs scaled by the box area, sigma U(1, 5))
BOX_TEMPLATES = {64: (40, (1.5, 4.0)), 96: (120, (1.0, 5.0)), 128: (200, (1.0, 5.0))}


def box_template_spec(N: int):
    return BOX_TEMPLATES.get(int(N), (max(20, int(120 * (N / 96) ** 2)), (1.0, 5.0)))


class UnevenWorkload:
    """Particles of several box sizes, each box with its own blob template
    (``BOX_TEMPLATES`` unless ``n_blobs`` / ``sigma_range`` are given); per
    particle a Haar rotation, an integer shift, a missing wedge (tilt range
    +-U(tilt_range) about y offset by U(-90, 90) degrees), white noise at
    U(snr_range) dB (signal power over the box's spherical mask, noise added
    before the wedge) and a march end L drawn from ``L_ends``."""

    def __init__(self, boxes=(64, 96, 128), ratio=(2, 2, 1), L_ends=(16, 24, 32), seed: int = 0, device="cuda",
                 n_blobs=None, sigma_range=None, tilt_range=(45.0, 65.0), snr_range=(-12.0, 0.0),
                 max_shift: int = 4):
        self.device = torch.device(device)
        self.boxes = tuple(int(b) for b in boxes)
        self.ratio = np.asarray(ratio, dtype=np.float64) / float(np.sum(ratio))
        self.L_ends = tuple(int(x) for x in L_ends)
        self.rng = np.random.default_rng(seed)
        self.tilt_range, self.snr_range, self.max_shift = tilt_range, snr_range, int(max_shift)
        self.radius = {N: float(min(N // 2, int(0.95 * N / 2)) - 2) for N in self.boxes}
        nb = n_blobs or {N: box_template_spec(N)[0] for N in self.boxes}
        sr = {N: sigma_range or box_template_spec(N)[1] for N in self.boxes}
        self.work = {N: Workload(N, nb[N], sr[N], seed + 100 + N, self.device, mask_radius=self.radius[N])
                     for N in self.boxes}

    def draw(self, n: int):
        """Per-particle specs: list of (box, L_end), in draw order."""
        box = self.rng.choice(len(self.boxes), size=n, p=self.ratio)
        le = self.rng.integers(0, len(self.L_ends), size=n)
        return [(self.boxes[b], self.L_ends[l]) for b, l in zip(box, le)]

    def particles(self, N: int, P: int):
        """(vols (P,N,N,N), wedges (P,N,N,N//2+1), R (P,3,3), shifts (P,3), snr (P,), tilt (P,))."""
        w = self.work[N]
        R = haar_rotations(w.rng, P)
        t = w.rng.integers(-self.max_shift, self.max_shift + 1, (P, 3))
        cen = np.einsum('pij,bj->pbi', R, w.centres) + t[:, None, :]
        vols = render_blobs(N, torch.tensor(cen, dtype=torch.float32, device=self.device), w.sig, w.amp,
                            self.device)
        snr = self.rng.uniform(*self.snr_range, P)
        tilt = self.rng.uniform(*self.tilt_range, P)
        off = self.rng.uniform(-90.0, 90.0, P)
        m = w.mask
        p = (vols[:, m] ** 2).mean(dim=1)
        std = torch.sqrt(p / torch.tensor(10.0 ** (snr / 10.0), dtype=torch.float32, device=self.device))
        vols = vols + torch.randn(vols.shape, device=self.device, generator=w.gen) * std[:, None, None, None]
        wedges = torch.stack([wedge_mask_half(N, tilt[i], off[i], self.device) for i in range(P)])
        vols = torch.fft.irfftn(torch.fft.rfftn(vols, dim=(1, 2, 3)) * wedges, s=(N, N, N), dim=(1, 2, 3))
        return vols.contiguous(), wedges, R, t, snr, tilt
