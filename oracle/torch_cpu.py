"""Torch-CPU FP32 restatement of the reference's orientation-search path.

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule): the
``cpu_baseline`` leg of ``bench.py`` times it on the host cores, and
``tests/test_torch_cpu_baseline.py`` checks it against the FP64 oracle.  The
product (``matcha_amd``) never imports it.

The reference has no CPU path (its FLE runs cufinufft plans and its Wigner
code is numba.cuda); this is its algorithm written for torch on the CPU in
FP32 / complex64, stage by stage (reference = swing-research/Matcha,
``src/matcha/core/`` unless noted):

* ``TorchFLE.evaluate_t`` -- ``FLEBasis3D.evaluate_t`` (``FLEBasis3D.py:
  945-976``): the ball mask (``:502-515``), then
  step 1 (``:1026-1038``): a type-2 NUFFT of the volume at the points
  lambda_q omega(theta_j, phi_k) / radius over half of the phi circle, the
  other half by conjugation with theta flipped.  The grid sizes (n_radial,
  n_theta, n_phi) follow ``precomp`` (``:159-208``), the radial nodes are the
  Chebyshev points of ``:375-378`` mapped to [lambda_min, lambda_max], the
  angular grid is Clenshaw-Curtis in cos(theta) and uniform in phi
  (``:233-262``).  The NUFFT is the published exponential-of-semicircle
  method (Barnett, Magland, af Klinteberg 2019) with the tolerance of the
  shipped configs (``precision_mode: "fast"``: 1e-4, kernel width 5, 2x
  oversampling): deconvolution by the kernel's Fourier transform, an FFT of
  the oversampled grid, interpolation with w^3 = 125 taps per point;
  step 2 (``:1040-1060``): the spherical harmonic transform of the samples
  (FFT over phi, Legendre matmul over theta with the quadrature weights),
  times i^l / (4 pi);
  step 3 (``:1233-1267``): interpolation in lambda from the n_radial
  Chebyshev nodes to lambda_lk (by the nodes' Chebyshev interpolant; the
  reference resamples it to 2 n_radial nodes by DCT and interpolates
  those), times c_lk and h = radius^-1.5.
* ``compute_sigma`` -- ``CrossCorrelationMatcher.compute_sigma``
  (``CrossCorrelationMatcher.py:183-249``), one matmul per shell.
* ``sofft_topk`` -- ``SOFFT.eval`` + ``sparse_topk_per_batch_jit`` +
  ``ids_to_angles`` (``SOFFT.py:30-124``, ``Matcha.py:27-58``) with
  ``torch.fft.ifft2`` and ``torch.topk``.
* ``march`` -- ``Matcha.optimize`` / ``_take_step`` (``Matcha.py:150-367``):
  D materialised per step (d^l by the three-term l-recurrence), the
  15-column coefficient bank of ``WignerDGradients.set_so3_coeffs`` /
  ``get_derivatives`` (``WignerDGradients.py:155-302``), Newton steps by
  ``torch.linalg.solve_ex`` with the singular-Hessian skip, the lr rule.
* ``ShiftSearch.search`` -- ``estimate_shift_like_relion`` +
  ``compute_shift`` (``ShiftMatcher.py:94-282,491-674``) for integer lags:
  the rotated Fourier reference by ``grid_sample`` (align_corners=True), the
  cosine-masked patch, ``rfftn`` / ``ifftn``, first argmax.
* ``align`` -- the benchmark's two global alternations
  (``align_subtomograms.py:316-361``).

Precomputation (grids, kernel tables, Wigner tables, the Fourier reference)
happens in the constructors and is not part of the timed work.

Parity pins (tests/test_torch_cpu_baseline.py): the projection against the
FP64 brute-force projection (``oracle.fle.evaluate_t``) within the NUFFT
tolerance; d^l against ``oracle.wigner.wigner_d``; sigma against
``oracle.fle.compute_sigma``; the search against ``oracle.so3`` on a peaked
sigma; the shift search against ``oracle.shift.estimate_shift``; and the
grid sizes against the reference's values at boxes 64 and 96.
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F
from scipy import special

from . import fle as FLE
from . import shift as SH
from . import so3 as SO3
from . import wigner as W

# ---------------------------------------------------------------------------
# FLE projection
# ---------------------------------------------------------------------------


def grid_sizes(N: int, eps: float, lmd_last: float, lmax: int, reduce_memory: bool = True):
    """(n_radial, n_theta, n_phi) of ``FLEBasis3D.precomp`` (FLEBasis3D.py:159-208,
    torch-harmonics branch)."""
    N = N + N % 2
    epsdis = eps / 4 / (np.pi ** 2 * (3 / 2) ** (1 / 4) * (3 + np.pi / 2 * np.log(5.3 * N)))
    Q = int(np.ceil(max(5.3 * N, np.log2(1 / epsdis))))
    tmp = 1 / np.sqrt(4 * np.pi)
    Q2 = 1
    for Q2 in range(1, Q):
        tmp = tmp / Q2 * ((np.sqrt(3) * np.pi / 16) ** (2 / 3) * (N + 1))
        if tmp < epsdis:
            break
    n_radial = int(Q2)
    if reduce_memory:
        n_radial = int(np.ceil(int(Q2) * 0.65))
    S = int(max(np.ceil(2 * np.exp(1) * 6 ** (1 / 3) * np.pi ** (2 / 3) * (N // 2)), 4 * np.log2(27.6 / epsdis)))
    S2 = 1
    for S2 in range(1, S):
        t = np.exp(1) * lmd_last / (2 * (S2 // 2 + 1) + 3)
        t2 = 28 / 27 * np.sqrt(2 * lmax + 1) * (np.exp(1) * lmd_last) ** 1.5 * (t ** (S2 // 2 - 1 / 2)) / (1 - t)
        if t2 < epsdis and t < 1:
            break
    S = max(S2, 2 * lmax, 18)
    n_phi = S + 1
    if n_phi % 2 == 1:
        n_phi += 1
    return n_radial, S, n_phi


def clenshaw_curtis(n: int):
    """Nodes cos(theta_j), theta_j = pi j / (n - 1), and weights of n-point
    Clenshaw-Curtis quadrature on [-1, 1] (Waldvogel 2006)."""
    N = n - 1
    theta = np.pi * np.arange(n) / N
    w = np.zeros(n)
    for j in range(n):
        s = 0.0
        for k in range(1, N // 2 + 1):
            b = 1.0 if 2 * k == N else 2.0
            s += b / (4 * k * k - 1) * np.cos(2 * k * theta[j])
        c = 1.0 if j in (0, N) else 2.0
        w[j] = c / N * (1.0 - s)
    return theta, w


def _es_kernel(z, beta):
    return np.where(np.abs(z) < 1, np.exp(beta * (np.sqrt(np.maximum(0.0, 1 - z * z)) - 1)), 0.0)


class TorchFLE:
    """The reference's NUFFT / SHT / radial-interpolation projection for one
    (N, bandlimit, radius), torch CPU, FP32."""

    W_KERNEL = 5  # spreading width at the shipped tolerance 1e-4 (2x oversampling)

    def __init__(self, N: int, bandlimit: int, radius: float, eps: float = 1e-4, micro_batch: int = 8):
        self.N, self.radius, self.micro_batch = int(N), float(radius), int(micro_batch)
        ls, ks, ms, mds, lmds, cs, ne = FLE.lap_eig_ball(N, bandlimit)
        self.ls, self.ks, self.ms, self.mds, self.lmds, self.cs, self.ne = ls, ks, ms, mds, lmds, cs, ne
        self.lmax = lmax = int(ls.max())
        nr, nt, nphi = grid_sizes(N, eps, float(lmds[-1]), lmax)
        self.n_radial, self.n_theta, self.n_phi = nr, nt, nphi
        # radial Chebyshev nodes in lambda (FLEBasis3D.py:370-378)
        l0, l1 = float(lmds.min()), float(lmds.max())
        xq = np.cos(np.pi * (1 - (2 * np.arange(nr) + 1) / (2 * nr)))
        pts = (xq + 1) / 2 * (l1 - l0) + l0
        theta, wcc = clenshaw_curtis(nt)
        phi = 2 * np.pi * np.arange(nphi // 2) / nphi
        st, ct = np.sin(theta), np.cos(theta)
        dirs = np.stack([np.cos(phi)[None, :] * st[:, None], np.sin(phi)[None, :] * st[:, None],
                         np.broadcast_to(ct[:, None], (nt, nphi // 2))], -1)          # (nt, nphi/2, 3)
        xyz = (pts[:, None, None, None] / self.radius) * dirs[None]                     # (nr, nt, nphi/2, 3)
        self.M = xyz.shape[0] * xyz.shape[1] * xyz.shape[2]
        # NUFFT tables: oversampled grid, kernel and its Fourier transform
        w = self.W_KERNEL
        self.nf = nf = 2 * N
        h = 2 * np.pi / nf
        beta = 2.30 * w
        alpha = w * h / 2
        zg, wg = np.polynomial.legendre.leggauss(64)
        k = np.arange(-N // 2, N - N // 2)
        psihat = alpha * (wg[None, :] * _es_kernel(zg[None, :], beta) * np.cos(k[:, None] * alpha * zg[None, :])).sum(1)
        deconv = (1.0 / psihat).astype(np.float32)
        self.deconv = (torch.tensor(deconv)[:, None, None] * torch.tensor(deconv)[None, :, None]
                       * torch.tensor(deconv)[None, None, :])
        t = xyz.reshape(-1, 3) / h                                   # continuous grid index
        base = np.ceil(t - w / 2).astype(np.int64)                   # (M, 3)
        offs = np.arange(w)
        z = (t[:, :, None] - (base[:, :, None] + offs[None, None, :])) * (2.0 / w)
        wts = _es_kernel(z, beta) * h                                # (M, 3, w), h per dimension
        # the points need the oversampled grid only within |index| <= R: that
        # sub-box (centred), stored point-major so that a tap reads the batch's
        # values from one cache line
        R = int(np.abs(base).max()) + w
        if 2 * R + 1 > nf:
            raise ValueError("NUFFT points beyond the oversampled grid")
        self.sub = (nf // 2 - R, nf // 2 + R + 1)
        self.sub_n = 2 * R + 1
        self.base = torch.tensor(base + R, dtype=torch.int64)       # index into the sub-box
        self.wts = torch.tensor(wts, dtype=torch.float32)
        self.offs = torch.arange(w)
        # ball mask (FLEBasis3D.py:502-515)
        ax = np.arange(N) - N // 2
        X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
        self.inside = torch.tensor(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) / self.radius <= 1 + 1e-13)
        # SHT: weighted orthonormal Legendre tables per m (theta ascending)
        self.leg = []
        for m in range(lmax + 1):
            P = np.stack([special.sph_harm_y(l, m, theta, 0.0).real for l in range(m, lmax + 1)])  # (lmax+1-m, nt)
            self.leg.append(torch.tensor((P * wcc[None, :]).T, dtype=torch.complex64))           # (nt, L-m+1)
        # radial interpolation in lambda with h c_lk i^l / (4 pi) folded in
        xnode = xq
        bw = (-1.0) ** np.arange(nr) * np.sin((2 * np.arange(nr) + 1) * np.pi / (2 * nr))
        hh = self.radius ** -1.5
        self.radial = []
        for l in range(lmax + 1):
            sel = np.nonzero(ls == l)[0]
            kk = np.unique(ks[sel])
            xt = 2 * (np.array([lmds[sel[ks[sel] == kv][0]] for kv in kk]) - l0) / (l1 - l0) - 1
            A = np.zeros((len(kk), nr))
            for r, x in enumerate(xt):
                d = x - xnode
                hit = np.nonzero(d == 0)[0]
                if hit.size:
                    A[r, hit[0]] = 1.0
                else:
                    q = bw / d
                    A[r] = q / q.sum()
            c_l = np.array([cs[sel[ks[sel] == kv][0]] for kv in kk])
            A = A * (hh * c_l / (4 * np.pi))[:, None] * (1j ** l)
            idx = np.full((len(kk), 2 * l + 1), -1, dtype=np.int64)
            for i in sel:
                idx[np.searchsorted(kk, ks[i]), ms[i] + l] = i
            flat = idx.reshape(-1)
            ok = np.nonzero(flat >= 0)[0]
            self.radial.append((torch.tensor(A, dtype=torch.complex64), torch.tensor(flat[ok]), torch.tensor(ok)))

    def _nufft(self, f):
        """c_j = sum_v f(v) e^{-i xi_j . v} (v = index - N/2) at the half-circle points; (B, M)."""
        B, N, nf, w = f.shape[0], self.N, self.nf, self.W_KERNEL
        g = f.to(torch.complex64) * self.deconv
        fine = torch.zeros((B, nf, nf, nf), dtype=torch.complex64)
        lo = nf // 2 - N // 2
        fine[:, lo:lo + N, lo:lo + N, lo:lo + N] = g
        G = torch.fft.fftn(torch.fft.ifftshift(fine, dim=(1, 2, 3)), dim=(1, 2, 3))
        # grid index l -> position l + nf/2 after fftshift; keep the points' sub-box
        a, b_ = self.sub
        G = torch.fft.fftshift(G, dim=(1, 2, 3))[:, a:b_, a:b_, a:b_]
        S = self.sub_n
        # point-major (S^3, B) as real pairs: a tap reads the batch's values
        # from one cache line; the 125 taps of a point are one weighted sum (bmm)
        GT = torch.view_as_real(G.reshape(B, -1).transpose(0, 1).contiguous()).reshape(S ** 3, 2 * B)
        out = torch.empty((self.M, 2 * B), dtype=torch.float32)
        chunk = 16384
        for s in range(0, self.M, chunk):
            b = self.base[s:s + chunk]
            wt = self.wts[s:s + chunk]
            n = len(b)
            ix = b[:, 0:1] + self.offs
            iy = b[:, 1:2] + self.offs
            iz = b[:, 2:3] + self.offs
            flat = ((ix[:, :, None, None] * S + iy[:, None, :, None]) * S + iz[:, None, None, :]).reshape(-1)
            vals = GT.index_select(0, flat).reshape(n, w ** 3, 2 * B)
            wgt = (wt[:, 0, :, None, None] * wt[:, 1, None, :, None] * wt[:, 2, None, None, :]).reshape(n, 1, -1)
            out[s:s + chunk] = torch.bmm(wgt, vals)[:, 0]
        out = torch.view_as_complex(out.reshape(self.M, B, 2))
        return out.transpose(0, 1)

    def _project(self, f):
        B = f.shape[0]
        nr, nt, nphi = self.n_radial, self.n_theta, self.n_phi
        f = torch.where(self.inside, f, torch.zeros((), dtype=f.dtype))
        z0 = self._nufft(f).reshape(B, nr, nt, nphi // 2)
        z = torch.cat([z0, torch.flip(z0, dims=[2]).conj()], dim=3)                   # (B, nr, nt, nphi)
        Zm = torch.fft.fft(z, dim=3) * (2 * np.pi / nphi)
        L = self.lmax
        b = torch.zeros((B, nr, L + 1, 2 * L + 1), dtype=torch.complex64)
        for m in range(L + 1):
            P = self.leg[m]                                                            # (nt, L-m+1)
            b[:, :, m:, L + m] = Zm[..., m] @ P
            if m:
                b[:, :, m:, L - m] = ((-1) ** m) * (Zm[..., nphi - m] @ P)
        a = torch.zeros((B, self.ne), dtype=torch.complex64)
        for l in range(L + 1):
            A, dst, src = self.radial[l]
            out = torch.einsum("kq,bqm->bkm", A, b[:, :, l, L - l:L + l + 1]).reshape(B, -1)
            a[:, dst] = out[:, src]
        return a

    def evaluate_t(self, vols: torch.Tensor) -> torch.Tensor:
        """a (B, ne) complex64 of volumes (B, N, N, N) float32 on the CPU."""
        outs = [self._project(vols[i:i + self.micro_batch]) for i in range(0, vols.shape[0], self.micro_batch)]
        return torch.cat(outs)


# ---------------------------------------------------------------------------
# sigma
# ---------------------------------------------------------------------------

class TorchSigma:
    """sigma[Dindex(l, u, v)] = (-1)^{u+v} sum_k aT[l, k, -v] conj(aV[l, k, -u])."""

    def __init__(self, fle: TorchFLE, L: int):
        self.L = int(L)
        ls, ks, ms = fle.ls, fle.ks, fle.ms
        self.shells = []
        for l in range(self.L + 1):
            sel = np.nonzero(ls == l)[0]
            K = int(ks[sel].max()) if sel.size else 0
            idx = np.full((K, 2 * l + 1), -1, dtype=np.int64)
            for i in sel:
                idx[ks[i] - 1, ms[i] + l] = i
            # column j holds m = l - j, so that column u + l is m = -u
            idx = idx[:, ::-1].copy()
            sign = np.array([[(-1.0) ** (u + v) for v in range(-l, l + 1)] for u in range(-l, l + 1)])
            self.shells.append((torch.tensor(np.where(idx >= 0, idx, 0)), torch.tensor(idx >= 0),
                                torch.tensor(sign, dtype=torch.float32)))
        self.dsize = W.dsize(self.L)

    def __call__(self, aT: torch.Tensor, aV: torch.Tensor) -> torch.Tensor:
        B = aV.shape[0]
        out = torch.empty((B, self.dsize), dtype=torch.complex64)
        zero = torch.zeros((), dtype=torch.complex64)
        for l, (idx, ok, sign) in enumerate(self.shells):
            T = torch.where(ok, aT[idx], zero)                  # (K, 2l+1): [k, v + l] = aT[l, k, -v]
            V = torch.where(ok[None], aV[:, idx], zero)         # (B, K, 2l+1): [., k, u + l] = aV[l, k, -u]
            s = torch.einsum("bku,kv->buv", V.conj(), T) * sign
            out[:, W.dsize(l - 1):W.dsize(l)] = s.reshape(B, -1)
        return out


# ---------------------------------------------------------------------------
# Wigner d (three-term l-recurrence), the SOFFT grid and the Newton march
# ---------------------------------------------------------------------------

def wigner_d_torch(beta: torch.Tensor, L: int) -> torch.Tensor:
    """d^l_{m'm}(beta) for l <= L in D layout, (nb, dsize(L)) float32: seeds at
    l = max(|m'|, |m|) from Wigner's sum's closed-form edges, then the
    three-term recurrence in l."""
    beta = beta.reshape(-1).to(torch.float32)
    nb, Wd = beta.shape[0], 2 * L + 1
    x = torch.cos(beta)[:, None, None]
    c = torch.cos(0.5 * beta)[:, None, None]
    s = torch.sin(0.5 * beta)[:, None, None]
    out = torch.empty((nb, W.dsize(L)), dtype=torch.float32)
    d1 = torch.zeros((nb, Wd, Wd))
    d2 = torch.zeros((nb, Wd, Wd))
    mpg = torch.arange(-L, L + 1, dtype=torch.float64)[:, None].expand(Wd, Wd)
    mg = torch.arange(-L, L + 1, dtype=torch.float64)[None, :].expand(Wd, Wd)
    for l in range(L + 1):
        lo, hi = L - l, L + l + 1
        dn = torch.zeros((nb, Wd, Wd))
        if l >= 1:
            # interior |m'|, |m| <= l - 1
            a, b_ = lo + 1, hi - 1
            mp, m = mpg[a:b_, a:b_], mg[a:b_, a:b_]
            P = (l * l - m * m) * (l * l - mp * mp)
            A = (l * (2 * l - 1) / torch.sqrt(P)).float()
            Bc = (m * mp / ((l - 1) * l)).float() if l > 1 else torch.zeros_like(A)
            Pp = ((l - 1) ** 2 - m * m) * ((l - 1) ** 2 - mp * mp)
            Cp = (torch.sqrt(Pp.clamp_min(0)) / ((l - 1) * (2 * l - 1))).float() if l > 1 else torch.zeros_like(A)
            dn[:, a:b_, a:b_] = A * ((x - Bc) * d1[:, a:b_, a:b_] - Cp * d2[:, a:b_, a:b_])
        # the border of the window: closed forms at l0 = l
        j = l
        lf = lambda n: math.lgamma(n + 1.0)
        mm = torch.arange(-l, l + 1)
        K = torch.tensor([math.exp(0.5 * (lf(2 * j) - lf(j + q) - lf(j - q))) for q in range(-l, l + 1)],
                         dtype=torch.float32)
        par = lambda n: 1.0 - 2.0 * (n.remainder(2)).float()
        # m' = l: (-1)^{l-m} K(l+m) c^{l+m} s^{l-m}
        dn[:, hi - 1, lo:hi] = par(l - mm) * K * torch.pow(c[:, 0], (j + mm)) * torch.pow(s[:, 0], (j - mm))
        # m' = -l: K(l+m) c^{l-m} s^{l+m}
        dn[:, lo, lo:hi] = K * torch.pow(c[:, 0], (j - mm)) * torch.pow(s[:, 0], (j + mm))
        # m = l: K(l+m') c^{l+m'} s^{l-m'}
        dn[:, lo:hi, hi - 1] = K * torch.pow(c[:, :, 0], (j + mm)) * torch.pow(s[:, :, 0], (j - mm))
        # m = -l: (-1)^{l+m'} K(l-m') c^{l-m'} s^{l+m'}
        dn[:, lo:hi, lo] = par(l + mm) * K.flip(0) * torch.pow(c[:, :, 0], (j - mm)) * torch.pow(s[:, :, 0], (j + mm))
        out[:, W.dsize(l - 1):W.dsize(l)] = dn[:, lo:hi, lo:hi].reshape(nb, -1)
        d2, d1 = d1, dn
    return out


class TorchSearch:
    """``Matcha.search_orientations`` (grid top-C, march, reduction), torch CPU FP32."""

    def __init__(self, L_max: int, Li, K: int, candidates: int = 10, num_steps: int = 5):
        self.L_max, self.C, self.num_steps = int(L_max), int(candidates), int(num_steps)
        self.levels = SO3.march_schedule(Li, L_max)
        self.L0 = min(self.levels)
        self.n = 2 * self.L0 * K
        n = self.n
        betas = np.linspace(0.0, np.pi, n, endpoint=True)
        self.dgrid = torch.tensor(W.wigner_d(betas, self.L0 - 1), dtype=torch.float32)  # (n, nsz)
        _, mps, ms = W.d_layout(self.L0 - 1)
        self.lin = torch.tensor((mps % n) * n + (ms % n))
        self.tabs = {L: self._tables(L) for L in set(self.levels)}

    @staticmethod
    def _tables(L):
        ls, mA, mG = W.d_layout(L)
        n = len(ls)
        lf, m, mp = ls.astype(np.float64), mA.astype(np.float64), mG.astype(np.float64)
        fm = np.where((lf + m) * (lf - m + 1) >= 0, -0.5 * np.sqrt(np.maximum((lf + m) * (lf - m + 1), 0)), 0.0)
        fp = np.where((lf - m) * (lf + m + 1) >= 0, 0.5 * np.sqrt(np.maximum((lf - m) * (lf + m + 1), 0)), 0.0)
        tn2 = np.sqrt(np.maximum((lf + m) * (lf - m + 1) * (lf + m - 1) * (lf - m + 2), 0)) / 4
        t0 = -((lf + m) * (lf - m + 1) + (lf + m + 1) * (lf - m)) / 4
        tp2 = np.sqrt(np.maximum((lf + m + 1) * (lf - m) * (lf + m + 2) * (lf - m - 1), 0)) / 4
        off = np.array([W.dsize(int(l) - 1) for l in range(L + 1)])[ls]

        def shifted(dm):
            ok = np.abs(mA + dm) <= ls
            dst = off + (mA + dm + ls) * (2 * ls + 1) + (mG + ls)
            return torch.tensor(np.nonzero(ok)[0]), torch.tensor(dst[ok])

        t = lambda a: torch.tensor(a, dtype=torch.float32)
        return dict(n=n, m=t(m), mp=t(mp), fm=t(fm), fp=t(fp), tn2=t(tn2), t0=t(t0), tp2=t(tp2),
                    mp1=shifted(+1), mn1=shifted(-1), mp2=shifted(+2), mn2=shifted(-2),
                    mpi=torch.tensor(mA), mi=torch.tensor(mG))

    def sofft_topk(self, sigma: torch.Tensor):
        B, n, C = sigma.shape[0], self.n, self.C
        nsz = self.dgrid.shape[1]
        temp = self.dgrid[None] * sigma[:, None, :nsz]                     # (B, n, nsz) complex
        Fg = torch.zeros((B, n, n * n), dtype=torch.complex64)
        Fg.index_add_(2, self.lin, temp)
        f = torch.fft.ifft2(Fg.reshape(B, n, n, n), dim=(2, 3)).real
        vals, idx = torch.topk(f.reshape(B, -1), C, dim=1)
        ib, ia, ig = idx // (n * n), (idx // n) % n, idx % n
        return (ia.float() * (2 * np.pi / n), ib.float() * (np.pi / (n - 1)), ig.float() * (2 * np.pi / n))

    def _D(self, a, b, g, L):
        tb = self.tabs[L]
        d = wigner_d_torch(b.reshape(-1), L)
        ang = a.reshape(-1, 1) * tb["mpi"][None].float() + g.reshape(-1, 1) * tb["mi"][None].float()
        ph = torch.polar(torch.ones_like(ang), ang)
        return (d * ph).reshape(a.shape + (tb["n"],))

    def _bank(self, sig, L):
        tb = self.tabs[L]
        c = sig[:, :tb["n"]]
        im, imp = 1j * tb["m"], 1j * tb["mp"]
        cols = [c, im * c, imp * c, (im * im) * c, (imp * imp) * c, (im * imp) * c]
        bm = torch.zeros_like(c)
        src, dst = tb["mn1"]
        bm[:, dst] = c[:, src] * tb["fm"][src]
        bp = torch.zeros_like(c)
        src, dst = tb["mp1"]
        bp[:, dst] = c[:, src] * tb["fp"][src]
        bbm = torch.zeros_like(c)
        src, dst = tb["mn2"]
        bbm[:, dst] = c[:, src] * tb["tn2"][src]
        bbp = torch.zeros_like(c)
        src, dst = tb["mp2"]
        bbp[:, dst] = c[:, src] * tb["tp2"][src]
        cols += [bm, bp, im * bm, im * bp, imp * bm, imp * bp, bbm, tb["t0"] * c, bbp]
        return torch.stack(cols, dim=1)                                     # (B, 15, n)

    @staticmethod
    def _derivatives(Y, a):
        pp, pn = torch.polar(torch.ones_like(a), a), torch.polar(torch.ones_like(a), -a)
        p2p, p2n = pp * pp, pn * pn
        Ybm, Ybp, Yabm, Yabp = Y[..., 6], Y[..., 7], Y[..., 8], Y[..., 9]
        gb = pp * Ybm + pn * Ybp
        gbg = pp * Y[..., 10] + pn * Y[..., 11]
        ab = pp * (1j * Ybm + Yabm) + pn * (-1j * Ybp + Yabp)
        bb = Y[..., 13] + p2p * Y[..., 12] + p2n * Y[..., 14]
        grad = torch.stack([Y[..., 1], gb, Y[..., 2]], -1).real
        H = torch.stack([torch.stack([Y[..., 3], ab, Y[..., 5]], -1),
                         torch.stack([ab, bb, gbg], -1),
                         torch.stack([Y[..., 5], gbg, Y[..., 4]], -1)], -2).real
        return grad, H

    def optimize(self, a, b, g, sigma, L):
        n = self.tabs[L]["n"]
        sig = sigma[:, :n]
        D = self._D(a, b, g, L)
        best = torch.einsum("bcn,bn->bc", D, sig).real
        ba, bb, bg = a.clone(), b.clone(), g.clone()
        bank = self._bank(sig, L)
        lr = torch.ones_like(a)
        for _ in range(self.num_steps):
            Y = torch.bmm(D, bank.transpose(1, 2))                             # (B, C, 15)
            grad, H = self._derivatives(Y, a)
            step, info = torch.linalg.solve_ex(H, grad.unsqueeze(-1))
            step = torch.where((info == 0)[..., None, None], step, torch.zeros_like(step))[..., 0] * lr[..., None]
            ta, tb, tg = a - step[..., 0], b - step[..., 1], g - step[..., 2]
            Dt = self._D(ta, tb, tg, L)
            ev = torch.einsum("bcn,bn->bc", Dt, sig).real
            upd = ev > best
            lr = torch.where(upd, lr * 1.1, lr * 0.5)
            ta, tb, tg = torch.where(upd, ta, a), torch.where(upd, tb, b), torch.where(upd, tg, g)
            Dt = torch.where(upd[..., None], Dt, D)
            best = torch.where(upd, ev, best)
            ba, bb, bg = torch.where(upd, ta, ba), torch.where(upd, tb, bb), torch.where(upd, tg, bg)
            a, b, g, D = ta, tb, tg, Dt
        return ba, bb, bg, best

    def search(self, sigma: torch.Tensor):
        a, b, g = self.sofft_topk(sigma)
        s = None
        for L in self.levels:
            a, b, g, s = self.optimize(a, b, g, sigma, L)
        i = torch.argmax(s, dim=1, keepdim=True)
        pick = lambda t: torch.gather(t, 1, i)[:, 0]
        return pick(a), pick(b), pick(g), pick(s)


# ---------------------------------------------------------------------------
# shift search (integer lags)
# ---------------------------------------------------------------------------

class ShiftSearch:
    """``estimate_shift_like_relion`` with ``shift_upsample_factor`` 1, torch CPU FP32."""

    def __init__(self, template: np.ndarray, N: int, voxel_size: float = 1.0):
        self.N = N
        ref = SH.load_reference_full(np.asarray(template, dtype=np.float64), N, voxel_size)
        self.ref = torch.tensor(np.stack([ref.real, ref.imag]), dtype=torch.float32)[None]   # (1, 2, N, N, N)
        outer, cos_, raised = SH.compute_masks((N, N, N), N / (2.0 * voxel_size), 5.0)
        self.outer = torch.tensor(outer)
        self.cos_ = torch.tensor(cos_)
        self.raised = torch.tensor(raised, dtype=torch.float32)
        self.wsum = float(outer.sum() + raised.sum())
        self.mask_full = torch.tensor(SH.get_mask_full((N, N, N // 2 + 1)))
        d = N
        de = d + ((d + 1) % 2)
        cz = (de - 1) / 2
        zz, yy, xx = torch.meshgrid(torch.arange(d), torch.arange(d), torch.arange(d + 1), indexing="ij")
        self.base = torch.stack([xx - cz, yy - cz, zz - cz], -1).reshape(-1, 3).double()    # (x, y, z)
        self.c, self.scale = cz, 2.0 / (de - 2)

    def _rotated(self, rot):
        B, N = rot.shape[0], self.N
        p = torch.einsum("bij,nj->bni", rot.double(), self.base) + self.c
        grid = (p * self.scale - 1.0).float().reshape(B, N, N, N + 1, 3)
        r = F.grid_sample(self.ref.expand(B, -1, -1, -1, -1), grid, mode="bilinear", padding_mode="zeros",
                          align_corners=True)
        rr = torch.complex(r[:, 0], r[:, 1])[..., N // 2:]
        return torch.fft.ifftshift(rr, dim=(1, 2)) * self.mask_full

    def search(self, subtomos: torch.Tensor, rot: torch.Tensor, shift_zyx: torch.Tensor):
        """Integer lags (B, 3) (flipped order) and scores; subtomos (B, N, N, N)."""
        B, N = subtomos.shape[0], self.N
        rr = self._rotated(rot)
        s_arr = shift_zyx.flip(1).round().long()
        patch = torch.zeros_like(subtomos)
        for i in range(B):
            sz, sy, sx = (int(v) for v in s_arr[i])
            src = subtomos[i, max(sz, 0):N + min(sz, 0), max(sy, 0):N + min(sy, 0), max(sx, 0):N + min(sx, 0)]
            patch[i, max(-sz, 0):N - max(sz, 0), max(-sy, 0):N - max(sy, 0), max(-sx, 0):N - max(sx, 0)] = src
        bg = (patch[:, self.outer].sum(1) + (self.raised * patch[:, self.cos_]).sum(1)) / self.wsum
        patch[:, self.outer] = bg[:, None]
        patch[:, self.cos_] = (1 - self.raised) * patch[:, self.cos_] + self.raised * bg[:, None]
        sub = torch.fft.rfftn(torch.roll(patch, (-(N // 2),) * 3, dims=(1, 2, 3)), dim=(1, 2, 3), norm="forward")
        prod = rr.conj() * sub
        full = torch.zeros((B, N, N, N), dtype=torch.complex64)
        full[..., :N // 2 + 1] = prod
        cc = torch.fft.ifftn(full, dim=(1, 2, 3), norm="forward").real
        nrm = torch.linalg.vector_norm(sub.reshape(B, -1), dim=1) * torch.linalg.vector_norm(rr.reshape(B, -1), dim=1)
        cc = cc.reshape(B, -1) / nrm[:, None]
        idx = cc.argmax(dim=1)                                                # first maximum
        sc = cc.gather(1, idx[:, None])[:, 0]
        m = torch.stack([idx // (N * N), (idx // N) % N, idx % N], 1)
        shift = torch.where(m > N // 2, m - N, m).float()
        return shift.flip(1), sc


# ---------------------------------------------------------------------------
# the benchmark's alternations
# ---------------------------------------------------------------------------

def shift_rotations(a, b, g):
    """P R^T P with R = Rz(a) Ry(b) Rz(g) and P the axis reversal (the shift
    search's reference rotation from search_orientations' angles)."""
    def rz(t):
        c, s = torch.cos(t), torch.sin(t)
        o, z = torch.ones_like(t), torch.zeros_like(t)
        return torch.stack([torch.stack([c, -s, z], -1), torch.stack([s, c, z], -1), torch.stack([z, z, o], -1)], -2)

    def ry(t):
        c, s = torch.cos(t), torch.sin(t)
        o, z = torch.ones_like(t), torch.zeros_like(t)
        return torch.stack([torch.stack([c, z, s], -1), torch.stack([z, o, z], -1), torch.stack([-s, z, c], -1)], -2)

    R = rz(a.double()) @ ry(b.double()) @ rz(g.double())
    P = torch.eye(3, dtype=torch.float64).flip(0)
    return P @ R.transpose(1, 2) @ P


class CpuPipeline:
    """Two global alternations (get_sigma, search_orientations, search_shifts)."""

    def __init__(self, template: np.ndarray, N: int, bandlimit: int, radius: float, L_max: int, Li, K: int,
                 candidates: int = 10, num_steps: int = 5):
        self.fle = TorchFLE(N, bandlimit, radius)
        self.sigma = TorchSigma(self.fle, L_max)
        self.search = TorchSearch(L_max, Li, K, candidates, num_steps)
        self.shift = ShiftSearch(template, N)
        self.aT = self.fle.evaluate_t(torch.tensor(template, dtype=torch.float32)[None])[0]
        self.N = N

    def patches(self, vols, shift_zyx):
        B, N = vols.shape[0], self.N
        s_arr = shift_zyx.flip(1).round().long()
        out = torch.zeros_like(vols)
        for i in range(B):
            sz, sy, sx = (int(v) for v in s_arr[i])
            src = vols[i, max(sz, 0):N + min(sz, 0), max(sy, 0):N + min(sy, 0), max(sx, 0):N + min(sx, 0)]
            out[i, max(-sz, 0):N - max(sz, 0), max(-sy, 0):N - max(sy, 0), max(-sx, 0):N - max(sx, 0)] = src
        return out

    def align(self, vols: torch.Tensor, alternations: int = 2):
        B = vols.shape[0]
        shift = torch.zeros((B, 3))
        out = None
        for it in range(alternations):
            pv = vols if it == 0 else self.patches(vols, shift)
            aV = self.fle.evaluate_t(pv)
            sig = self.sigma(self.aT, aV)
            a, b, g, s = self.search.search(sig)
            ds, ss = self.shift.search(vols, shift_rotations(a, b, g), shift)
            shift = shift + ds
            out = (a, b, g, s, shift, ss)
        return out
