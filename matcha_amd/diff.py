"""Differentiable correlation score: a torch.autograd.Function over HIP kernels.

The score of the reference's orientation search at a given pose,

    s(omega, t, aT) = Re sum_{l <= L} D^l(R(omega)) sigma^l(aT, aV(t)),

with R(omega) = exp([omega]x) acting on the volumes' (axis 0, 1, 2)
coordinates (the frame of ``search_orientations``' R = Rz(a) Ry(b) Rz(g)), aV(t)
the ball-harmonic coefficients of the particle patch at shift t (trilinear,
patch[a] = V[a + flip(t)], as ``extract_subtomogram_patch_batch``,
``src/matcha/utils/io_utils.py:113-178``), sigma as
``CrossCorrelationMatcher.compute_sigma`` (``CrossCorrelationMatcher.py:
183-249``) and D, the score, as ``Matcha.optimize`` (``Matcha.py:184,288``).
The reference differentiates the score only in Euler angles
(``WignerDGradients.py:247-302``).

Forward: the patch and its three exact t-derivatives (``mch_shift_tangent``)
go through one projection of 4B volumes and one sigma pass; then
``mch_score_gradients`` re-marches d^l(beta) per shell and evaluates, in one
pass over each shell, the score, the body-frame rotation derivatives (scores
of sigma X_i^T, X_i the generators of D; chained with the right Jacobian of
exp, so no Euler singularity), the shift derivatives and the per-particle
template adjoint.  Backward scales these by the upstream gradient
(``mch_score_backward``).  The template gradient follows PyTorch's complex
convention (d/dRe + i d/dIm).
"""
from __future__ import annotations

import torch

from . import _lib
from .core.CrossCorrelationMatcher import CrossCorrelationMatcher


class ScoreEngine:
    """Plans of the differentiable score for particles of one box size.

    N, bandlimit, radius: the ball-harmonic expansion (as CrossCorrelationMatcher);
    L: the highest Wigner order of the score (the search's L_max, <= lmax);
    batchsize: particles per call."""

    def __init__(self, N: int, bandlimit: int, radius: float, L: int, batchsize: int, device,
                 compute_precision: str = "fp32"):
        self.device = torch.device(device)
        self.N = int(N)
        self.batchsize = int(batchsize)
        self.ccm = CrossCorrelationMatcher(N, self.device, batchsize=4 * self.batchsize, bandlimit=bandlimit,
                                           radius=radius, compute_precision=compute_precision)
        self.lmax = self.ccm.lmax
        if not 0 <= int(L) <= self.lmax:
            raise ValueError(f"L must be in [0, {self.lmax}] (the expansion's lmax)")
        self.L = int(L)
        self.ne = self.ccm.fle.ne
        self.kdim = self.ccm._kdim
        self.idx = torch.tensor(self.ccm._idx, dtype=torch.int32, device=self.device).contiguous()

    def template_coefficients(self, template: torch.Tensor) -> torch.Tensor:
        """Ball-harmonic coefficients (ne,) complex64 of a template volume."""
        return self.ccm.fle.evaluate_t(template.reshape(1, self.N, self.N, self.N).to(self.device))[0]

    def tangent_coefficients(self, vols: torch.Tensor, t: torch.Tensor, fused: bool = None) -> torch.Tensor:
        """(4B, ne) coefficients of the patches at shift t (B, 3) and of their three
        t-derivatives.  fused (default: the FP32 and BF16x3 projections) forms
        the four volumes inside the projection's operand build
        (mch_fle_project_tangents); otherwise they go through HBM
        (mch_shift_tangent, then the projection)."""
        B, Nv, N = vols.shape[0], vols.shape[-1], self.N
        fle = self.ccm.fle
        stream = _lib.stream_handle(self.device)
        if fused is None:
            fused = fle.compute_precision in ("fp32", "bf16x3")
        if fused:
            coeffs = torch.empty((4 * B, self.ne), dtype=torch.complex64, device=self.device)
            _lib.check(_lib.lib().mch_fle_project_tangents(fle._plan, vols.data_ptr(), vols.stride(0), Nv, B,
                                                           t.data_ptr(), coeffs.data_ptr(), coeffs.stride(0),
                                                           stream), "mch_fle_project_tangents")
            return coeffs
        tang = torch.empty((4 * B, N, N, N), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().mch_shift_tangent(vols.data_ptr(), vols.stride(0), Nv, B, N, t.data_ptr(),
                                                tang.data_ptr(), tang.stride(0), stream), "mch_shift_tangent")
        return fle.evaluate_t(tang)

    def _prepare(self, omega, shift, aT, particles):
        """Inputs of mch_score_gradients: (B, angles, jr, coeffs, sigma)."""
        B = omega.shape[0]
        if B > self.batchsize:
            raise ValueError(f"batch {B} exceeds batchsize {self.batchsize}")
        _lib.require_gpu_tensor(particles, "particles", device=self.device)
        vols = particles.float().contiguous()
        Nv = vols.shape[-1]
        if vols.shape[0] != B or tuple(vols.shape[1:]) != (Nv, Nv, Nv) or Nv < self.N:
            raise ValueError(f"particles must be (B, Nv, Nv, Nv) with Nv >= {self.N}")
        om = omega.detach().to(self.device, torch.float32).contiguous()
        t = shift.detach().to(self.device, torch.float32).reshape(B, 3).contiguous()
        a_t = aT.detach().to(self.device, torch.complex64).reshape(-1).contiguous()
        if a_t.numel() != self.ne:
            raise ValueError(f"aT must hold {self.ne} coefficients")
        stream = _lib.stream_handle(self.device)
        lib = _lib.lib()
        angles = torch.empty((B, 3), dtype=torch.float64, device=self.device)
        jr = torch.empty((B, 9), dtype=torch.float64, device=self.device)
        _lib.check(lib.mch_score_pose(om.data_ptr(), B, angles.data_ptr(), jr.data_ptr(), stream), "mch_score_pose")
        coeffs = self.tangent_coefficients(vols, t)
        dsz = self.ccm.sigma.shape[1]
        sigma = torch.empty((4 * B, dsz), dtype=torch.complex64, device=self.device)
        _lib.check(lib.mch_compute_sigma_degree(self.ccm._corr, a_t.data_ptr(), coeffs.data_ptr(),
                                                coeffs.stride(0), self.ne, 4 * B, self.L, sigma.data_ptr(),
                                                sigma.stride(0), None, stream), "mch_compute_sigma")
        return B, angles, jr, coeffs, sigma

    def _forward(self, omega, shift, aT, particles, want_template: bool):
        B, angles, jr, coeffs, sigma = self._prepare(omega, shift, aT, particles)
        stream = _lib.stream_handle(self.device)
        lib = _lib.lib()
        partial = torch.empty((B, self.L + 1, 8), dtype=torch.float32, device=self.device)
        out = torch.empty((B, 7), dtype=torch.float32, device=self.device)
        G = torch.empty((B, self.ne), dtype=torch.complex64, device=self.device) if want_template else None
        _lib.check(lib.mch_score_gradients(self.idx.data_ptr(), self.lmax, self.kdim, sigma.data_ptr(),
                                           sigma.stride(0), B, coeffs.data_ptr(), coeffs.stride(0), self.L,
                                           angles.data_ptr(), jr.data_ptr(), partial.data_ptr(), out.data_ptr(),
                                           G.data_ptr() if G is not None else None,
                                           G.stride(0) if G is not None else 0, stream), "mch_score_gradients")
        return out, G

    def score_and_gradients(self, omega, shift, aT, particles, template_gradient: bool = False):
        """(out (B, 7): s, ds/domega, ds/dt; G (B, ne) per-particle template adjoint or None)."""
        return self._forward(omega, shift, aT, particles, template_gradient)

    def score(self, omega, shift, aT, particles):
        """Scores (B,) differentiable in omega (B, 3), shift (B, 3) and aT (ne,)."""
        return _ScoreFunction.apply(omega, shift, aT, particles, self)


class _ScoreFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, omega, shift, aT, particles, engine):
        want_t = bool(ctx.needs_input_grad[2])
        out, G = engine._forward(omega, shift, aT, particles, want_t)
        ctx.engine = engine
        ctx.want = (bool(ctx.needs_input_grad[0]), bool(ctx.needs_input_grad[1]), want_t)
        ctx.dtypes = (omega.dtype, shift.dtype, aT.dtype)
        ctx.save_for_backward(out, G if G is not None else torch.empty(0, device=out.device))
        return out[:, 0].to(omega.dtype if omega.dtype.is_floating_point else torch.float32)

    @staticmethod
    def backward(ctx, grad_s):
        out, G = ctx.saved_tensors
        eng = ctx.engine
        B = out.shape[0]
        w = grad_s.detach().to(torch.float32).contiguous()
        d_om = torch.empty((B, 3), dtype=torch.float32, device=out.device) if ctx.want[0] else None
        d_t = torch.empty((B, 3), dtype=torch.float32, device=out.device) if ctx.want[1] else None
        d_aT = torch.empty((eng.ne,), dtype=torch.complex64, device=out.device) if ctx.want[2] else None
        rc = _lib.lib().mch_score_backward(w.data_ptr(), out.data_ptr(), B,
                                           G.data_ptr() if ctx.want[2] else None,
                                           G.stride(0) if ctx.want[2] else 0, eng.ne,
                                           d_om.data_ptr() if d_om is not None else None,
                                           d_t.data_ptr() if d_t is not None else None,
                                           d_aT.data_ptr() if d_aT is not None else None,
                                           _lib.stream_handle(out.device))
        _lib.check(rc, "mch_score_backward")
        return (d_om.to(ctx.dtypes[0]) if d_om is not None else None,
                d_t.to(ctx.dtypes[1]) if d_t is not None else None,
                d_aT.to(ctx.dtypes[2]) if d_aT is not None else None, None, None)
