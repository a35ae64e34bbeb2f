"""Template/particle ball-harmonic coefficients -> SO(3) correlation coefficients.

Drop-in for ``src/matcha/core/CrossCorrelationMatcher.py`` (same constructor,
``set_template``, ``compute_sigma``, ``get_sigma``; attributes ``lmax``,
``bh_norm_template``, ``sigma``, ``bh_norms``, ``fle``).  The expansion runs
in ``FLEBasis3D`` (HIP GEMMs) and sigma in ``sigma_kernel``:

    sigma[Dindex(l, u, v)] = (-1)^{u+v} sum_k a^T_{l,k,-v} conj(a^V_{l,k,-u})

(``CrossCorrelationMatcher.py:133-212``), with k running over the radial
orders present for both m values (the reference pads missing ones with a
zero coefficient).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .FLEBasis3D import FLEBasis3D
from .wigner_index import compute_size


def _to_complex(dtype):
    if dtype == torch.float32:
        return torch.complex64
    raise ValueError(f"Unsupported dtype {dtype}: the MI355X path computes in float32")


def get_sht_index(m: int) -> int:
    """Order m -> position in 0, -1, 1, -2, 2, ... (CrossCorrelationMatcher.py:19-34)."""
    if m == 0:
        return 0
    if m > 0:
        return 2 * m
    return -2 * m - 1


class CrossCorrelationMatcher:
    def __init__(self, N: int, device: torch.device, expansion_epsilon: float = 1e-4,
                 batchsize: int = 1, reduce_memory: bool = True, bandlimit: int = None,
                 micro_batch_split: int = 2, dtype=torch.float32, radius: int = None,
                 jl_zeros_path: str = None, cs_path: str = None, precision_mode: str = "accurate",
                 compute_precision: str = "fp32"):
        self.N = N
        self.fle_basis_eps = expansion_epsilon
        self.batchsize = batchsize
        self.reduce_memory = reduce_memory
        self.bandlimit = bandlimit
        self.micro_batch_split = int(micro_batch_split)
        if self.micro_batch_split <= 0:
            raise ValueError("micro_batch_split must be a positive integer.")
        if self.batchsize == 1:
            self.micro_batch_split = 1
        # memory is not the constraint on a 288 GB device: one pass per batch
        self.micro_batch_size = self.batchsize
        self.device = torch.device(device)
        self.dtype_real = dtype
        self.dtype = _to_complex(dtype)
        # the reference passes radius=None through to FLEBasis3D, whose
        # h = radius ** -1.5 then fails; its caller clamps the radius to
        # min(N // 2, int(N // 2 * 0.95)) (align_subtomograms.py:94), the default here
        self.radius = radius if radius is not None else min(N // 2, int(N // 2 * 0.95))
        if bandlimit is None:
            raise ValueError("bandlimit must be given")
        self.fle = FLEBasis3D(N=N, bandlimit=bandlimit, eps=expansion_epsilon,
                              reduce_memory=reduce_memory, batchsize=self.batchsize,
                              radius=self.radius, device=self.device, dtype=self.dtype,
                              jl_zeros_path=jl_zeros_path, cs_path=cs_path,
                              precision_mode=precision_mode, compute_precision=compute_precision)
        self.lmax = self.fle.lmax
        self._make_corr_plan()
        self.sigma = torch.zeros((self.batchsize, compute_size(self.lmax + 1)), dtype=self.dtype,
                                 device=self.device)
        self.bh_norms = torch.zeros((self.batchsize,), dtype=self.dtype_real, device=self.device)
        self.bh_volumes_buffer = torch.zeros((self.batchsize, self.fle.ne), dtype=self.dtype,
                                             device=self.device)
        self.template_coefficients_alphas_torch = None
        self.bh_norm_template = None

    def _make_corr_plan(self):
        lmax = self.lmax
        kdim = max(len(self.fle.idlm_list[l][md]) for l in range(lmax + 1) for md in range(2 * l + 1))
        idx = np.full((lmax + 1, 2 * lmax + 1, kdim), -1, dtype=np.int32)
        for l in range(lmax + 1):
            for m in range(-l, l + 1):
                lst = self.fle.idlm_list[l][get_sht_index(m)]
                idx[l, m + l, :len(lst)] = lst
        self._kdim = kdim
        self._idx = idx
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = _lib.lib().mch_corr_plan_create(lmax, kdim, idx.ctypes.data_as(ctypes.c_void_p),
                                                 self.batchsize, ctypes.byref(h))
        _lib.check(rc, "mch_corr_plan_create")
        self._corr = h

    def __del__(self):
        h = getattr(self, "_corr", None)
        if h is not None and h.value:
            try:
                _lib.lib().mch_corr_plan_destroy(h)
            except Exception:
                pass

    def set_template(self, template_data, mask: torch.Tensor = None):
        """Expand the template; keeps a^T (padded with one zero) and its norm."""
        if isinstance(template_data, np.ndarray):
            t = torch.from_numpy(np.ascontiguousarray(template_data)).to(self.device, torch.float32)
        else:
            t = template_data.to(self.device, torch.float32)
        a = self.fle.evaluate_t(t.reshape(1, self.N, self.N, self.N))[0]
        self._aT = a.contiguous()
        self.template_coefficients_alphas_torch = torch.nn.functional.pad(a, (0, 1))
        self.bh_norm_template = torch.linalg.norm(a)

    def compute_sigma(self, bh_volumes: torch.Tensor, i_from: int, i_to: int, max_degree: int = None):
        """sigma rows [i_from, i_to) from particle coefficients (B, ne); with
        ``max_degree`` only the shells l <= max_degree are written."""
        if self.template_coefficients_alphas_torch is None:
            raise RuntimeError("set_template must be called before compute_sigma")
        _lib.require_gpu_tensor(bh_volumes, "bh_volumes", dtype=self.dtype, device=self.device)
        bv = bh_volumes.contiguous()
        rows = self.sigma[i_from:i_to]
        norms = self.bh_norms[i_from:i_to]
        L = self.lmax if max_degree is None else int(max_degree)
        rc = _lib.lib().mch_compute_sigma_degree(self._corr, self._aT.data_ptr(), bv.data_ptr(), bv.stride(0),
                                                 self.fle.ne, bv.shape[0], L, rows.data_ptr(), rows.stride(0),
                                                 norms.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "mch_compute_sigma")

    def get_sigma(self, volumes: torch.Tensor, max_degree: int = None, shift_zyx: torch.Tensor = None):
        """(sigma (batchsize, Dsize(lmax)) complex64, bh_norms (batchsize,)).

        Returns the persistent buffers (overwritten by the next call), as the
        reference does; rows beyond the input batch are zero.  ``max_degree``
        (not in the reference) writes only the shells l <= max_degree, all a
        search up to that L_max reads; ``shift_zyx`` (not in the reference)
        projects the patches extract_shifted_patches would cut at those shifts."""
        num = volumes.shape[0]
        if num > self.batchsize:
            raise ValueError(
                f"Input batch size {num} exceeds configured batchsize {self.batchsize}.")
        if num < self.batchsize:
            self.sigma[num:].zero_()
            self.bh_norms[num:].zero_()
        a = self.fle.evaluate_t(volumes, out=self.bh_volumes_buffer[:num], shift_zyx=shift_zyx)
        self.compute_sigma(a, 0, num, max_degree)
        return self.sigma, self.bh_norms
