"""Frequency-marched Newton orientation search on the MI355X library.

Drop-in for ``src/matcha/core/Matcha.py`` (class ``Matcha``): same
constructor ``Matcha(batchsize, device, L_max, matcha_config)``, same config
keys and defaults (``_assert_config``, ``Matcha.py:120-148``), same
``search_orientations(sigma) -> (alphas, betas, gammas, scores)`` contract
(``Matcha.py:369-403``): radians, float32 device tensors of shape (B,), score
= Re sum D sigma (unnormalised).

The whole search is three HIP launches per call (grid top-C, fused Newton
march over every level, candidate reduction) instead of the reference's
per-step D/bank materialisation.  ``stop_early`` (a batch-global early exit,
``Matcha.py:198-199``: a level stops after the first step that accepted no
candidate of the batch) runs the march one launch per (level, step); each
launch reads the accept counts of the previous ones on the device, so unlike
the reference no host synchronisation is needed.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..config import ConfigDict, cfg_get
from .SOFFT import SOFFT, _check_sigma
from .wigner_index import Dsize


_DEFAULTS = {
    "candidates": 10,
    "reinits": 1,
    "num_steps": 5,
    "step_type": "newton",
    "stop_early": False,
    "do_random_sampling": False,
    "reinits_iterations": 0,
    "oversampling_factor_K": 2,
}


def march_schedule(Li, L_max):
    """Levels actually optimised: ``sorted(Li) + [L_max]`` minus levels >= L_max
    that are not last (``Matcha.py:76,381-386``)."""
    sched = sorted(int(x) for x in Li) + [int(L_max)]
    return [l for i, l in enumerate(sched) if not (l >= L_max and i != len(sched) - 1)]


class Matcha:
    def __init__(self, batchsize: int, device: torch.device, L_max: int,
                 matcha_config=None):
        cfg = self._assert_config(matcha_config)
        self.device = torch.device(device)
        self.Lmax = int(L_max)
        self.Li = sorted(cfg_get(cfg, "Li")) + [self.Lmax]
        self.L0 = min(self.Li)
        self.do_random_sampling = cfg_get(cfg, "do_random_sampling")
        self.batchsize = int(batchsize)
        self.candidates = int(cfg_get(cfg, "candidates"))
        self.reinits = int(cfg_get(cfg, "reinits"))
        self.reinits_iterations = cfg_get(cfg, "reinits_iterations")
        self.num_steps = int(cfg_get(cfg, "num_steps"))
        self.stop_early = bool(cfg_get(cfg, "stop_early"))
        self.num_templates = 1
        self.step_type = cfg_get(cfg, "step_type")
        if self.step_type not in ("newton", "gradient"):
            raise ValueError(f"step_type must be 'newton' or 'gradient', got {self.step_type!r}")
        self.base_lr = 1 if self.step_type == "newton" else 5e-7
        if self.reinits != 1:
            # the reference's D buffers are sized for candidates*reinits rotations
            # but the grid search only yields `candidates`; only reinits == 1 runs there
            raise ValueError("reinits != 1 is not supported (the reference only runs reinits=1)")
        if self.L0 < 1:
            raise ValueError("min(Li) must be >= 1")
        self.levels = march_schedule(cfg_get(cfg, "Li"), self.Lmax)
        self.SOFFT = SOFFT(L=self.L0, device=self.device, batchsize=self.batchsize,
                           oversampling_factor=int(cfg_get(cfg, "oversampling_factor_K")),
                           candidates=self.candidates)
        # the plan of the march shares the grid plan's geometry but needs L_max
        from .SOFFT import So3Plan
        self.plan = So3Plan(L_max=self.Lmax, L0=self.L0,
                            K=int(cfg_get(cfg, "oversampling_factor_K")),
                            batchsize=self.batchsize, candidates=self.candidates,
                            device=self.device)
        if self.stop_early:
            if self.num_steps > 63:
                raise ValueError("stop_early supports at most 63 num_steps")
            _lib.check(_lib.lib().mch_so3_set_stop_early(self.plan.handle, 1), "mch_so3_set_stop_early")
        self._levels_c = (ctypes.c_int * len(self.levels))(*self.levels)
        self._li_sorted = sorted(int(x) for x in cfg_get(cfg, "Li"))
        self._li_c = (ctypes.c_int * len(self._li_sorted))(*self._li_sorted)
        self._truncate_dsize_cache = {int(l): Dsize(int(l)) for l in set(self.Li)}

    def _assert_config(self, matcha_config):
        if matcha_config is None:
            matcha_config = ConfigDict()
        if cfg_get(matcha_config, "Li") is None:
            raise ValueError("Matcha config missing required key: 'Li'")
        for key, default in _DEFAULTS.items():
            if cfg_get(matcha_config, key) is None:
                try:
                    matcha_config[key] = default
                except TypeError:
                    setattr(matcha_config, key, default)
        return matcha_config

    # -- the three stages, exposed for tests ---------------------------------
    def grid_candidates(self, sigma: torch.Tensor):
        """Top-C SOFFT grid candidates: (alpha, beta, gamma, score), each (B, C)."""
        sig = _check_sigma(sigma, self.device, Dsize(self.L0 - 1), self.batchsize)
        return self.SOFFT.topk(sig, self.candidates)

    def march(self, sigma: torch.Tensor, alphas, betas, gammas, levels=None):
        """Newton-march (B, C) candidates over ``levels`` (default: the schedule)."""
        sig = _check_sigma(sigma, self.device, Dsize(self.Lmax), self.batchsize)
        B, C = alphas.shape
        a = alphas.contiguous().clone().float()
        b = betas.contiguous().clone().float()
        g = gammas.contiguous().clone().float()
        s = torch.empty_like(a)
        lv = self.levels if levels is None else [int(x) for x in levels]
        lv_c = (ctypes.c_int * len(lv))(*lv)
        rc = _lib.lib().mch_newton_march(
            self.plan.handle, sig.data_ptr(), sig.stride(0), B, C, lv_c, len(lv), self.num_steps,
            1 if self.step_type == "newton" else 0, float(self.base_lr), a.data_ptr(),
            b.data_ptr(), g.data_ptr(), s.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "mch_newton_march")
        return a, b, g, s

    def march_ends_tensor(self, march_ends) -> torch.Tensor:
        """Per-particle march ends as the (B,) int32 device tensor that
        ``search_orientations(sigma, march_ends=...)`` reads, checked on the host:
        each end in [min(Li), L_max], so that every particle shares this
        matcher's SOFFT grid (L0 = min(Li)) and marches the schedule of a
        ``Matcha`` built with ``L_max`` = its end."""
        ends = torch.as_tensor(march_ends).reshape(-1).to(torch.int64)
        if ends.is_cuda:
            raise ValueError("march_ends_tensor checks host values; pass a list, array or CPU tensor")
        lo = self._li_sorted[0] if self._li_sorted else 0
        if ends.numel() and (int(ends.min()) < max(lo, 1) or int(ends.max()) > self.Lmax):
            raise ValueError(f"march ends must lie in [{max(lo, 1)}, {self.Lmax}], got "
                             f"[{int(ends.min())}, {int(ends.max())}]")
        return ends.to(torch.int32).to(self.device)

    @torch.no_grad()
    def search_orientations(self, sigma: torch.Tensor, march_ends=None):
        """``Matcha.search_orientations`` (``Matcha.py:369-403``).

        ``march_ends`` (optional, beyond the reference): per-particle march
        ends -- a (B,) int32 device tensor from ``march_ends_tensor`` (read on
        the device, so a graph can capture the call), or host values checked
        here.  Particle b then gets the result a ``Matcha`` of ``L_max`` =
        its end would give, in one launch for the whole batch."""
        sig = _check_sigma(sigma, self.device, Dsize(self.Lmax), self.batchsize)
        B = sig.shape[0]
        outs = [torch.empty(B, dtype=torch.float32, device=self.device) for _ in range(4)]
        if march_ends is not None:
            if self.stop_early:
                raise ValueError("per-particle march ends cannot be combined with stop_early")
            ends = march_ends
            if not (isinstance(ends, torch.Tensor) and ends.is_cuda):
                ends = self.march_ends_tensor(ends)
            if ends.dtype != torch.int32 or ends.device != self.device or ends.numel() < B:
                raise ValueError("march_ends must be a device int32 tensor with one end per particle")
            ends = ends.contiguous()
            rc = _lib.lib().mch_search_orientations_ends(
                self.plan.handle, sig.data_ptr(), sig.stride(0), B, self.candidates, self._li_c,
                len(self._li_sorted), ends.data_ptr(), self.num_steps,
                1 if self.step_type == "newton" else 0, float(self.base_lr),
                *[o.data_ptr() for o in outs], _lib.stream_handle(self.device))
            _lib.check(rc, "mch_search_orientations_ends")
            return tuple(outs)
        rc = _lib.lib().mch_search_orientations(
            self.plan.handle, sig.data_ptr(), sig.stride(0), B, self.candidates, self._levels_c,
            len(self.levels), self.num_steps, 1 if self.step_type == "newton" else 0,
            float(self.base_lr), *[o.data_ptr() for o in outs], _lib.stream_handle(self.device))
        _lib.check(rc, "mch_search_orientations")
        return tuple(outs)


def so3_evaluate(matcher: "Matcha", sigma: torch.Tensor, alphas, betas, gammas, L: int):
    """Score/gradient/Hessian at (B, C) angles through the library; (B, C, 10)."""
    sig = _check_sigma(sigma, matcher.device, Dsize(matcher.Lmax), matcher.batchsize)
    B, C = alphas.shape
    a = alphas.contiguous().float()
    b = betas.contiguous().float()
    g = gammas.contiguous().float()
    out = torch.empty((B, C, 10), dtype=torch.float32, device=matcher.device)
    rc = _lib.lib().mch_so3_evaluate(matcher.plan.handle, sig.data_ptr(), sig.stride(0), B, C,
                                     int(L), a.data_ptr(), b.data_ptr(), g.data_ptr(),
                                     out.data_ptr(), _lib.stream_handle(matcher.device))
    _lib.check(rc, "mch_so3_evaluate")
    return out
