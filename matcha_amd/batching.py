"""Uneven batches: particles of several box sizes and march schedules.

The reference aligns one box size with one march schedule per run
(``CrossCorrelationMatcher``/``FLEBasis3D`` are built for one N, ``Matcha`` for
one ``Li``; ``src/matcha/align_subtomograms.py:61-134``).  Here a stream of
particles with per-particle (box, march end L) is aligned in one pass:

* particles are bucketed by (box, L); each box owns one set of projection,
  sigma and shift-search plans (the projection GEMM's tile configuration is
  autotuned per box when its tuned precision is enabled), each bucket its own
  ``Matcha`` (march schedule ``Li`` truncated below L, then L) and alternation
  harness;
* by default a bucket launches its kernels eagerly per batch (about twenty
  long kernels per batch: launch overhead is not the bound, and a graph
  replay measured slower for the copies into its static buffers); with
  ``graphs=True`` (for callers that must replay a captured graph) a bucket runs
  full batches from static device buffers: one alternation pass over a full
  batch (projection, sigma, SOFFT grid, the fused Newton march -- every level
  and step in one launch -- and the shift search, twice) is captured once
  into a HIP graph and replayed per batch, the ragged last batch of a bucket
  padded with empty particles, as the reference pads its last batch;
* missing-wedge or CTF weights go to the shift search per particle, as the
  reference's ``ctfs`` (``ShiftMatcher.py:637-638``).
"""
from __future__ import annotations

import numpy as np
import torch

from .align import Alternation
from .core.CrossCorrelationMatcher import CrossCorrelationMatcher
from .core.Matcha import Matcha
from .core.ShiftMatcher import ShiftMatcher


def march_levels(L_end: int, Li=(8, 16, 24)):
    """The march levels below L_end (the reference appends L_max to sorted Li)."""
    lv = [int(x) for x in sorted(Li) if int(x) < int(L_end)]
    return lv or [int(L_end)]


def plan_buckets(specs, batch: int):
    """{(box, L): [chunks of particle indices of at most ``batch``]} in first-seen order."""
    buckets: dict = {}
    for i, key in enumerate(specs):
        buckets.setdefault((int(key[0]), int(key[1])), []).append(i)
    return {k: [v[j:j + batch] for j in range(0, len(v), batch)] for k, v in buckets.items()}


class _Box:
    def __init__(self, N, template, radius, bandlimit, batch, device, precision, max_shift):
        self.N = N
        self.ccm = CrossCorrelationMatcher(N, device, batchsize=batch, bandlimit=bandlimit, radius=radius,
                                           compute_precision=precision)
        self.ccm.set_template(template)
        self.sm = ShiftMatcher({"N": N, "num_subtomograms_per_batch": batch, "shift_upsample_factor": 1,
                                "shift_search_radius": max_shift, "voxel_size": 1.0,
                                "do_ctf_correction": True}, device)
        self.sm.set_reference(template.detach().cpu().numpy())


class _Bucket:
    def __init__(self, box: _Box, L_end, batch, device, cfg, alternations, graphs):
        self.box, self.L, self.batch, self.device = box, int(L_end), int(batch), device
        mcfg = {"Li": march_levels(L_end, cfg["Li"]), "candidates": cfg["candidates"],
                "num_steps": cfg["num_steps"], "oversampling_factor_K": cfg["K"]}
        self.matcher = Matcha(batch, device, self.L, mcfg)
        self.pipe = Alternation(box.ccm, self.matcher, box.sm, num_alternations=alternations, mode="global")
        N = box.N
        self.vin = torch.zeros((batch, N, N, N), dtype=torch.float32, device=device)
        self.cin = torch.ones((batch, N, N, N // 2 + 1), dtype=torch.float32, device=device)
        self.graphs = graphs
        self.graph = None
        self.gout = None

    def _capture(self):
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            for _ in range(2):
                self.pipe.run(self.vin, ctfs=self.cin)
        torch.cuda.current_stream(self.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.gout = self.pipe.run(self.vin, ctfs=self.cin)
        self.graph = g

    def run(self, vols, ctfs):
        """(alpha, beta, gamma, score, shift (B, 3), shift score) for one chunk of
        B <= batch particles; with graphs a ragged chunk is padded with empty
        particles (as the reference pads its last batch, ``setup_utils.pad_data``)."""
        B = vols.shape[0]
        if not self.graphs:
            return self.pipe.run(vols.float().contiguous(), ctfs=ctfs)
        self.vin[:B].copy_(vols)
        if B < self.batch:
            self.vin[B:].zero_()
        if ctfs is not None:
            self.cin[:B].copy_(ctfs)
            if B < self.batch:
                self.cin[B:].fill_(1.0)
        else:
            self.cin.fill_(1.0)
        if self.graph is None:
            self._capture()
        self.graph.replay()
        return tuple(t[:B].clone() if t is not None else None for t in self.gout)


class UnevenAligner:
    """Align particles of several box sizes and march ends.

    templates: {N: template volume (N, N, N)}; radius: {N: FLE / mask radius};
    L_ends: the march ends in use (each <= the expansion's lmax)."""

    def __init__(self, templates: dict, device, radius: dict | None = None, bandlimit: int = 32, batch: int = 64,
                 L_ends=(16, 24, 32), Li=(8, 16, 24), K: int = 4, candidates: int = 10, num_steps: int = 5,
                 alternations: int = 2, precision: str = "bf16x3", graphs: bool = False, max_shift: int = 4):
        self.device = torch.device(device)
        self.batch = int(batch)
        self.cfg = {"Li": tuple(Li), "K": int(K), "candidates": int(candidates), "num_steps": int(num_steps)}
        self.alternations = int(alternations)
        self.graphs = bool(graphs)
        self.boxes = {}
        for N, tmpl in templates.items():
            r = (radius or {}).get(N, min(N // 2, int(0.95 * N / 2)))
            self.boxes[int(N)] = _Box(int(N), tmpl.to(self.device, torch.float32), r, bandlimit, self.batch,
                                      self.device, precision, max_shift)
        self.buckets = {}
        for N in self.boxes:
            for L in L_ends:
                if L > self.boxes[N].ccm.lmax:
                    raise ValueError(f"L = {L} exceeds the expansion's lmax {self.boxes[N].ccm.lmax} at box {N}")

    def bucket(self, N: int, L: int) -> _Bucket:
        key = (int(N), int(L))
        if key not in self.buckets:
            self.buckets[key] = _Bucket(self.boxes[key[0]], key[1], self.batch, self.device, self.cfg,
                                        self.alternations, self.graphs)
        return self.buckets[key]

    @torch.no_grad()
    def align(self, specs, particles, ctfs=None):
        """specs: per particle (box, L_end); particles[i]: (N, N, N) device volume of
        that box (any indexable: list or per-box tensors); ctfs[i]: its
        (N, N, N//2+1) weights or None.  Returns per-particle float32 arrays
        alpha, beta, gamma, score, shift (n, 3), shift_score on the host."""
        n = len(specs)
        out = {k: torch.empty(n, dtype=torch.float32, device=self.device)
               for k in ("alpha", "beta", "gamma", "score", "shift_score")}
        out["shift"] = torch.empty((n, 3), dtype=torch.float32, device=self.device)
        for (N, L), chunks in plan_buckets(specs, self.batch).items():
            bk = self.bucket(N, L)
            for idx in chunks:
                vols = torch.stack([particles[i] for i in idx])
                cw = torch.stack([ctfs[i] for i in idx]) if ctfs is not None else None
                a, b, g, s, sh, ss = bk.run(vols, cw)
                it = torch.tensor(idx, device=self.device)
                for k, v in (("alpha", a), ("beta", b), ("gamma", g), ("score", s), ("shift", sh),
                             ("shift_score", ss)):
                    out[k].index_copy_(0, it, v)
        return {k: v.cpu().numpy() for k, v in out.items()}
