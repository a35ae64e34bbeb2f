"""Rotation/shift alternation over one batch (the hot path's caller).

Mirrors ``BatchProcessor._run_alternations`` (``src/matcha/align_subtomograms.py:
316-361``) for one batch already resident on the device: each alternation
re-extracts the particles at the accumulated shift, computes sigma and the
orientation (``get_sigma`` + ``search_orientations``), then searches the shift
against the reference rotated by that orientation.  The particles are not
pre-rotated by the running estimate (every orientation search is global), so
the per-alternation rotation is the full estimate.
"""
from __future__ import annotations

import torch

from .core.ShiftMatcher import euler_to_shift_rotations, extract_shifted_patches


class Alternation:
    def __init__(self, correlator, rotation_matcher, shift_matcher, num_alternations: int = 2,
                 final_shift: bool = True):
        self.ccm = correlator
        self.rot = rotation_matcher
        self.sm = shift_matcher
        self.num_alternations = int(num_alternations)
        self.final_shift = final_shift
        self._patches = None
        self.history = []

    @torch.no_grad()
    def run(self, subtomos: torch.Tensor):
        """(alpha, beta, gamma, norm_score, shift_zyx (B,3), shift_score) for the batch.

        ``self.history`` keeps that tuple for every alternation, in order: the
        per-alternation records the STAR output is joined from."""
        B = subtomos.shape[0]
        dev = subtomos.device
        shift = torch.zeros((B, 3), dtype=torch.float32, device=dev)
        if self._patches is None or self._patches.shape != subtomos.shape:
            self._patches = torch.empty_like(subtomos, dtype=torch.float32)
        vols = subtomos
        out = None
        self.history = []
        for it in range(self.num_alternations):
            if it > 0:
                vols = extract_shifted_patches(subtomos, shift, out=self._patches)
            sigma, norms = self.ccm.get_sigma(vols)
            a, b, g, s = self.rot.search_orientations(sigma)
            nscore = s / (norms[:B] * self.ccm.bh_norm_template).clamp_min(1e-12)
            last = it == self.num_alternations - 1
            sscore = None
            if not last or self.final_shift:
                rb = euler_to_shift_rotations(a, b, g)
                ds, sscore = self.sm.search_shifts(subtomos=vols, ctfs=None,
                                                   shift_zyx=None, rotation_tracker=rb)
                shift = shift + ds
            out = (a, b, g, nscore, shift, sscore)
            self.history.append(out)
        return out
