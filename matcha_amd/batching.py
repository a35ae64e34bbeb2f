"""Uneven batches: particles of several box sizes and march schedules.

The reference aligns one box size with one march schedule per run
(``CrossCorrelationMatcher``/``FLEBasis3D`` are built for one N, ``Matcha`` for
one ``Li``; ``src/matcha/align_subtomograms.py:61-134``).  Here a stream of
particles with per-particle (box, march end L) is aligned in one pass:

* particles are grouped by box (and by the SOFFT grid's bandlimit L0 = the
  smallest level of their schedule, which is min(Li) for every end above it);
  each box owns one set of projection, sigma and shift-search plans (the
  projection GEMM's tile configuration is autotuned per box when its tuned
  precision is enabled);
* a group's batches mix march ends: one ``Matcha`` of L_max = the largest end
  runs the grid, the fused Newton march and the candidate reduction once per
  batch, each particle marching the levels of ``Li`` below its own end and
  then that end (``mch_search_orientations_ends``): the result a ``Matcha``
  built for that end gives, in one launch of each kernel for the whole batch;
* by default a group launches its kernels eagerly per batch; with
  ``graphs=True`` it runs full batches from static device buffers (the
  particles are stacked straight into them, the march ends copied into a
  device table): one alternation pass over a full batch (projection, sigma,
  SOFFT grid, the fused march -- every level and step of every schedule in one
  launch -- and the shift search, twice) is captured once into a HIP graph and
  replayed per batch, the ragged last batch padded with empty particles, as
  the reference pads its last batch;
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


def grid_degree(L_end: int, Li=(8, 16, 24)) -> int:
    """L0 of a ``Matcha`` built for L_max = L_end: min(Li + [L_end])."""
    return min([int(x) for x in Li] + [int(L_end)])


def plan_buckets(specs, batch: int):
    """{(box, L): [chunks of particle indices of at most ``batch``]} in first-seen order."""
    buckets: dict = {}
    for i, key in enumerate(specs):
        buckets.setdefault((int(key[0]), int(key[1])), []).append(i)
    return {k: [v[j:j + batch] for j in range(0, len(v), batch)] for k, v in buckets.items()}


def plan_groups(specs, batch: int, Li=(8, 16, 24)):
    """{(box, L0): [chunks of particle indices of at most ``batch``]}, first-seen
    order: the particles of one box whose schedules share the grid bandlimit L0,
    whatever their march ends."""
    groups: dict = {}
    for i, key in enumerate(specs):
        groups.setdefault((int(key[0]), grid_degree(key[1], Li)), []).append(i)
    return {k: [v[j:j + batch] for j in range(0, len(v), batch)] for k, v in groups.items()}


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


def _fill(dst: torch.Tensor, src, pad: float):
    """dst[:B] = the B volumes of src (a stacked tensor or a list), dst[B:] = pad."""
    B = len(src)
    if isinstance(src, torch.Tensor):
        dst[:B].copy_(src)
    else:
        torch.stack(src, out=dst[:B])
    if B < dst.shape[0]:
        dst[B:].fill_(pad)


class _Group:
    """The particles of one box whose schedules share one grid bandlimit."""

    def __init__(self, box: _Box, L_top: int, batch, device, cfg, alternations, graphs):
        self.box, self.L, self.batch, self.device = box, int(L_top), int(batch), device
        li = [int(x) for x in cfg["Li"] if int(x) < self.L]
        mcfg = {"Li": li or [self.L], "candidates": cfg["candidates"], "num_steps": cfg["num_steps"],
                "oversampling_factor_K": cfg["K"]}
        self.matcher = Matcha(batch, device, self.L, mcfg)
        self.pipe = Alternation(box.ccm, self.matcher, box.sm, num_alternations=alternations, mode="global")
        self.graphs = graphs
        self.graph = None
        self.gout = None
        self.vin = None

    def _static(self):
        N, dev = self.box.N, self.device
        self.vin = torch.zeros((self.batch, N, N, N), dtype=torch.float32, device=dev)
        self.cin = torch.ones((self.batch, N, N, N // 2 + 1), dtype=torch.float32, device=dev)
        self.ends = torch.full((self.batch,), self.L, dtype=torch.int32, device=dev)

    def _capture(self):
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            for _ in range(2):
                self.pipe.run(self.vin, ctfs=self.cin, march_ends=self.ends)
        torch.cuda.current_stream(self.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.gout = self.pipe.run(self.vin, ctfs=self.cin, march_ends=self.ends)
        self.graph = g

    def run(self, vols, ctfs, ends):
        """(alpha, beta, gamma, score, shift (B, 3), shift score) for one chunk of
        B <= batch particles; ``vols`` a (B, N, N, N) device tensor or a list of
        (N, N, N) volumes, ``ctfs`` the same for the weights or None, ``ends``
        their march ends (host ints; None: all at the group's top end).  With
        graphs a ragged chunk is padded with empty particles (as the reference
        pads its last batch, ``setup_utils.pad_data``)."""
        B = len(vols)
        ends = [self.L] * B if ends is None else [int(x) for x in ends]
        if not self.graphs:
            v = vols if isinstance(vols, torch.Tensor) else torch.stack(vols)
            v = v.float().contiguous()
            cw = None
            if ctfs is not None:
                cw = (ctfs if isinstance(ctfs, torch.Tensor) else torch.stack(ctfs)).float().contiguous()
            e = None if all(x == self.L for x in ends) else self.matcher.march_ends_tensor(ends)
            return self.pipe.run(v, ctfs=cw, march_ends=e)
        if self.vin is None:
            self._static()
        _fill(self.vin, vols, 0.0)
        if ctfs is not None:
            _fill(self.cin, ctfs, 1.0)
        else:
            self.cin.fill_(1.0)
        e = self.matcher.march_ends_tensor(ends + [self.L] * (self.batch - B))
        self.ends.copy_(e)
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
        self.cfg = {"Li": tuple(int(x) for x in Li), "K": int(K), "candidates": int(candidates),
                    "num_steps": int(num_steps)}
        self.L_ends = tuple(int(x) for x in L_ends)
        self.alternations = int(alternations)
        self.graphs = bool(graphs)
        self.boxes = {}
        for N, tmpl in templates.items():
            r = (radius or {}).get(N, min(N // 2, int(0.95 * N / 2)))
            self.boxes[int(N)] = _Box(int(N), tmpl.to(self.device, torch.float32), r, bandlimit, self.batch,
                                      self.device, precision, max_shift)
        self.groups = {}
        for N in self.boxes:
            for L in self.L_ends:
                if L > self.boxes[N].ccm.lmax:
                    raise ValueError(f"L = {L} exceeds the expansion's lmax {self.boxes[N].ccm.lmax} at box {N}")

    def group(self, N: int, L0: int) -> _Group:
        key = (int(N), int(L0))
        if key not in self.groups:
            top = max(L for L in self.L_ends if grid_degree(L, self.cfg["Li"]) == key[1])
            self.groups[key] = _Group(self.boxes[key[0]], top, self.batch, self.device, self.cfg,
                                      self.alternations, self.graphs)
        return self.groups[key]

    @torch.no_grad()
    def align(self, specs, particles, ctfs=None):
        """specs: per particle (box, L_end), L_end one of ``L_ends``; particles[i]:
        (N, N, N) device volume of that box (any indexable: list or per-box
        tensors); ctfs[i]: its (N, N, N//2+1) weights or None.  Returns
        per-particle float32 arrays alpha, beta, gamma, score, shift (n, 3),
        shift_score on the host."""
        n = len(specs)
        for key in specs:
            if int(key[1]) not in self.L_ends:
                raise ValueError(f"march end {key[1]} is not one of L_ends {self.L_ends}")
        out = {k: torch.empty(n, dtype=torch.float32, device=self.device)
               for k in ("alpha", "beta", "gamma", "score", "shift_score")}
        out["shift"] = torch.empty((n, 3), dtype=torch.float32, device=self.device)
        for (N, L0), chunks in plan_groups(specs, self.batch, self.cfg["Li"]).items():
            grp = self.group(N, L0)
            for idx in chunks:
                vols = [particles[i] for i in idx]
                cw = [ctfs[i] for i in idx] if ctfs is not None else None
                a, b, g, s, sh, ss = grp.run(vols, cw, [int(specs[i][1]) for i in idx])
                it = torch.tensor(idx, device=self.device)
                for k, v in (("alpha", a), ("beta", b), ("gamma", g), ("score", s), ("shift", sh),
                             ("shift_score", ss)):
                    out[k].index_copy_(0, it, v)
        return {k: v.cpu().numpy() for k, v in out.items()}
