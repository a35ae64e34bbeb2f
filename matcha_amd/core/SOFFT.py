"""SO(3) FFT grid on the MI355X library.

Mirrors ``src/matcha/core/SOFFT.py:6-124`` (class ``SOFFT``: ``eval``,
``ids_to_angles``) and ``SOFFTMatcher`` of ``src/matcha/example.py:42-67``.
The grid values, the top-C selection and the argmax run in HIP
(``sofft_slice_kernel``/``sofft_merge_kernel``); only index-to-angle glue is
host-side.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .wigner_index import compute_size


class So3Plan:
    """Owns the library plan (tables + workspaces) for one (L_max, L0, K, batch, C)."""

    def __init__(self, L_max: int, L0: int, K: int, batchsize: int, candidates: int,
                 device: torch.device):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError(f"matcha_amd runs on a GPU device only (got {self.device})")
        self.L_max, self.L0, self.K = int(L_max), int(L0), int(K)
        self.batchsize, self.candidates = int(batchsize), int(candidates)
        self.n = 2 * self.L0 * self.K
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = _lib.lib().mch_so3_plan_create(self.L_max, self.L0, self.K, self.batchsize,
                                                self.candidates, ctypes.byref(handle))
        _lib.check(rc, "mch_so3_plan_create")
        self.handle = handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().mch_so3_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _check_sigma(sigma: torch.Tensor, device, min_cols: int, max_rows: int) -> torch.Tensor:
    _lib.require_gpu_tensor(sigma, "sigma", dtype=torch.complex64, device=device)
    if sigma.ndim != 2:
        raise ValueError(f"sigma must be 2-D (B, Dsize), got shape {tuple(sigma.shape)}")
    if sigma.shape[1] < min_cols:
        raise ValueError(f"sigma has {sigma.shape[1]} coefficients, need at least {min_cols}")
    if sigma.shape[0] < 1 or sigma.shape[0] > max_rows:
        raise ValueError(f"sigma batch {sigma.shape[0]} outside [1, {max_rows}]")
    if sigma.stride(1) != 1:
        sigma = sigma.contiguous()
    return sigma


class SOFFT:
    """``SOFFT(L, device, batchsize, oversampling_factor)`` with the same grid."""

    def __init__(self, L: int, device: torch.device, batchsize: int, oversampling_factor=1,
                 candidates: int = 1):
        self.num_betas = 2 * L * oversampling_factor
        self.num_alphas = self.num_betas
        self.num_gammas = self.num_betas
        self.batch_size = batchsize
        self.device = torch.device(device)
        self.L = L
        self.Dsize = compute_size(L + 1)
        self.plan = So3Plan(L_max=max(L, 1), L0=L, K=oversampling_factor, batchsize=batchsize,
                            candidates=max(1, candidates), device=self.device)

    def eval(self, flmn_coeffs_t: torch.Tensor) -> torch.Tensor:
        """Score grid (B, n_beta, n_alpha, n_gamma), float32 (``SOFFT.eval``)."""
        assert flmn_coeffs_t.ndim == 2
        sig = _check_sigma(flmn_coeffs_t, self.device, compute_size(self.L), self.batch_size)
        B = sig.shape[0]
        n = self.num_betas
        out = torch.empty((B, n, n, n), dtype=torch.float32, device=self.device)
        rc = _lib.lib().mch_sofft_eval(self.plan.handle, sig.data_ptr(), sig.stride(0), B,
                                       out.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "mch_sofft_eval")
        return out

    def topk(self, sigma: torch.Tensor, k: int):
        """Top-k grid points as Euler angles (B, k) each + scores (B, k)."""
        sig = _check_sigma(sigma, self.device, compute_size(self.L), self.batch_size)
        B = sig.shape[0]
        outs = [torch.empty((B, k), dtype=torch.float32, device=self.device) for _ in range(4)]
        rc = _lib.lib().mch_sofft_topk(self.plan.handle, sig.data_ptr(), sig.stride(0), B, k,
                                       *[o.data_ptr() for o in outs],
                                       _lib.stream_handle(self.device))
        _lib.check(rc, "mch_sofft_topk")
        return tuple(outs)

    def ids_to_angles(self, ids: torch.Tensor, shape):
        """(i_beta, i_alpha, i_gamma) -> (alpha, beta, gamma) (``SOFFT.ids_to_angles``)."""
        num_betas, num_alphas, num_gammas = shape
        betas = ids[:, :, 0] * (np.pi / (num_betas - 1))
        alphas = ids[:, :, 1] * (2 * np.pi / num_alphas)
        gammas = ids[:, :, 2] * (2 * np.pi / num_gammas)
        return alphas, betas, gammas


class SOFFTMatcher:
    """``search_orientations`` by the grid argmax only (``method: "sofft"``)."""

    def __init__(self, batchsize: int, device, L_max: int, oversampling_factor_K: int = 2):
        self.sofft = SOFFT(L=L_max, device=device, batchsize=batchsize,
                           oversampling_factor=oversampling_factor_K, candidates=1)

    def search_orientations(self, sigma):
        a, b, g, s = self.sofft.topk(sigma, 1)
        return a[:, 0], b[:, 0], g[:, 0], s[:, 0]
