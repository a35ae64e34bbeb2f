"""Rotation/shift alternation over one batch (the hot path's caller).

Mirrors ``BatchProcessor._run_alternations`` (``src/matcha/align_subtomograms.py:
316-361``) for one batch already resident on the device.  Two modes:

* ``"reference"`` (default) -- the reference's loop: each alternation builds
  the particles with ``_extract_rotated_batch`` (the patch at the accumulated
  shift, normalised over the spherical mask, masked, rotated by the running
  rotation tracker; :240-252), searches the orientation of that batch
  (``get_sigma`` + ``search_orientations``), composes it into the tracker
  (``update_rotation_estimate``, :254-266), then searches the shift of the
  original particles at the accumulated shift against the reference rotated by
  the tracker (:268-285) and accumulates it (:350).
* ``"global"`` -- the particles are re-extracted at the accumulated shift but
  neither normalised nor rotated; every orientation search is global, so each
  alternation's orientation is the full estimate.  Same recovered orientations
  on well-posed data, one resampling pass less per alternation.
"""
from __future__ import annotations

import torch

from . import _lib
from .core.ShiftMatcher import euler_to_shift_rotations


def extract_rotated_batch(subtomos: torch.Tensor, shift: torch.Tensor | None, rot: torch.Tensor,
                          radius: float, N: int | None = None, out: torch.Tensor | None = None,
                          work: torch.Tensor | None = None, partial: torch.Tensor | None = None):
    """``_extract_rotated_batch`` on the device: (B, N, N, N) float32 particles.

    subtomos (B, Nv, Nv, Nv); shift (B, 3) displacement in the reference's
    (x, y, z) order or None; rot (B, 3, 3) matrices R(tracker) acting on
    (x, y, z) index coordinates; radius of the binary normalisation mask."""
    _lib.require_gpu_tensor(subtomos, "subtomos")
    vols = subtomos.float().contiguous()
    B, Nv = vols.shape[0], vols.shape[1]
    N = Nv if N is None else int(N)
    dev = vols.device
    sh = None
    if shift is not None:
        sh = shift.reshape(B, -1, 3)[:, 0, :].to(device=dev, dtype=torch.float32).contiguous()
    rot = rot.reshape(B, 9).to(device=dev, dtype=torch.float32).contiguous()
    if out is None:
        out = torch.empty((B, N, N, N), dtype=torch.float32, device=dev)
    if sh is not None and (work is None or work.numel() < B * N ** 3):
        work = torch.empty((B, N, N, N), dtype=torch.float32, device=dev)
    if partial is None or partial.numel() < B * N * 4:
        partial = torch.empty((B, N, 4), dtype=torch.float32, device=dev)
    rc = _lib.lib().mch_extract_rotated(vols.data_ptr(), vols.stride(0), Nv, B, N,
                                        sh.data_ptr() if sh is not None else None, rot.data_ptr(),
                                        float(radius), work.data_ptr() if work is not None else None,
                                        partial.data_ptr(), out.data_ptr(), out.stride(0),
                                        _lib.stream_handle(dev))
    _lib.check(rc, "mch_extract_rotated")
    return out


class RotationTracker:
    """Per-particle rotation trackers (B, 4) float32 (w, x, y, z) on the device,
    identity or ``prior`` (e.g. ``relion.prior_rotation_tracker``), with the
    rotation matrices of the particle resampler (R) and of the shift search
    (R^T) kept current."""

    def __init__(self, B: int, device, prior=None):
        self.device = torch.device(device)
        if prior is None:
            q = torch.zeros((B, 4), dtype=torch.float32, device=self.device)
            q[:, 0] = 1.0
        else:
            q = torch.as_tensor(prior, dtype=torch.float32).reshape(B, 4).to(self.device).contiguous()
        self.q = q
        self.rot_particle = torch.empty((B, 3, 3), dtype=torch.float32, device=self.device)
        self.rot_reference = torch.empty((B, 3, 3), dtype=torch.float32, device=self.device)
        self._call(None)

    def _call(self, abg):
        B = self.q.shape[0]
        ptrs = [None, None, None] if abg is None else [t.data_ptr() for t in abg]
        rc = _lib.lib().mch_tracker_update(*ptrs, B, self.q.data_ptr(), self.rot_particle.data_ptr(),
                                           self.rot_reference.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "mch_tracker_update")

    def update(self, alphas, betas, gammas):
        """``update_rotation_estimate`` with this search's angles."""
        self._call([t.reshape(-1).float().contiguous() for t in (alphas, betas, gammas)])


class Alternation:
    def __init__(self, correlator, rotation_matcher, shift_matcher, num_alternations: int = 2,
                 final_shift: bool = True, mode: str = "reference", mask_radius: float | None = None):
        if mode not in ("reference", "global"):
            raise ValueError("mode must be 'reference' or 'global'")
        self.ccm = correlator
        self.rot = rotation_matcher
        self.sm = shift_matcher
        self.num_alternations = int(num_alternations)
        self.final_shift = final_shift
        self.mode = mode
        self.mask_radius = mask_radius
        self._patches = None
        self._work = None
        self._shift0 = None
        self.history = []
        self.tracker = None
        self._ends = None

    @torch.no_grad()
    def run(self, subtomos: torch.Tensor, prior_tracker=None, ctfs: torch.Tensor | None = None,
            march_ends=None):
        """(alpha, beta, gamma, norm_score, shift (B,3), shift_score) of the last
        alternation.

        ``self.history`` keeps that tuple for every alternation, in order: the
        per-alternation records the STAR output is joined from.  In reference
        mode the angles are the last search's (relative to the particles rotated
        by the previous tracker) and ``self.tracker.q`` holds the composed
        rotation trackers, the reference's ``rotation_tracker``; in global mode
        the angles are the full estimate.  ``ctfs`` (B, N, N, N//2+1): CTF or
        missing-wedge weights of the shift search (``do_ctf_correction``).
        ``march_ends``: per-particle march ends of the orientation searches
        (``Matcha.search_orientations``)."""
        self._ends = march_ends
        if self.mode == "reference":
            return self._run_reference(subtomos, prior_tracker, ctfs)
        if prior_tracker is not None:
            raise ValueError("a prior rotation needs mode='reference'")
        return self._run_global(subtomos, ctfs)

    def _search(self, sigma):
        if self._ends is None:
            return self.rot.search_orientations(sigma)
        return self.rot.search_orientations(sigma, march_ends=self._ends)

    def _buffers(self, subtomos, patches: bool):
        """Running-shift buffer, and (reference mode) the rotated-patch buffers."""
        if self._shift0 is None or self._shift0.shape[0] != subtomos.shape[0]:
            self._shift0 = torch.empty((subtomos.shape[0], 3), dtype=torch.float32, device=subtomos.device)
        if patches and (self._patches is None or self._patches.shape != subtomos.shape):
            self._patches = torch.empty_like(subtomos, dtype=torch.float32)
            self._work = torch.empty_like(subtomos, dtype=torch.float32)

    def _run_global(self, subtomos, ctfs=None):
        B = subtomos.shape[0]
        self._buffers(subtomos, patches=False)
        shift = self._shift0.zero_()
        out = None
        self.history = []
        for it in range(self.num_alternations):
            # alternations after the first see the patches at the running shift;
            # the projection and the shift search read them straight from the
            # particles (the values extract_shifted_patches would copy out)
            sh = shift if it > 0 else None
            sigma, norms = self.ccm.get_sigma(subtomos, max_degree=self.rot.Lmax, shift_zyx=sh)
            a, b, g, s = self._search(sigma[:B])
            nscore = s / (norms[:B] * self.ccm.bh_norm_template).clamp_min(1e-12)
            last = it == self.num_alternations - 1
            sscore = None
            if not last or self.final_shift:
                rb = euler_to_shift_rotations(a, b, g)
                ds, sscore = self.sm.search_shifts(subtomos=subtomos, ctfs=ctfs,
                                                   shift_zyx=sh, rotation_tracker=rb)
                shift = shift + ds
            out = (a, b, g, nscore, shift, sscore)
            self.history.append(out)
        return out

    def _run_reference(self, subtomos, prior_tracker, ctfs=None):
        B, N = subtomos.shape[0], subtomos.shape[1]
        dev = subtomos.device
        radius = self.mask_radius if self.mask_radius is not None else getattr(self.ccm, "radius", N // 2)
        self._buffers(subtomos, patches=True)
        shift = self._shift0.zero_()
        self.tracker = RotationTracker(B, dev, prior_tracker)
        out = None
        self.history = []
        for it in range(self.num_alternations):
            vols = extract_rotated_batch(subtomos, shift, self.tracker.rot_particle, radius,
                                         out=self._patches, work=self._work)
            sigma, norms = self.ccm.get_sigma(vols, max_degree=self.rot.Lmax)
            a, b, g, s = self._search(sigma[:B])
            nscore = s / (norms[:B] * self.ccm.bh_norm_template).clamp_min(1e-12)
            self.tracker.update(a, b, g)
            last = it == self.num_alternations - 1
            sscore = None
            if not last or self.final_shift:
                ds, sscore = self.sm.search_shifts(subtomos=subtomos, ctfs=ctfs, shift_zyx=shift,
                                                   rotation_tracker=self.tracker.rot_reference)
                shift = shift + ds
            out = (a, b, g, nscore, shift, sscore)
            self.history.append(out + (self.tracker.q.clone(),))
        return out
