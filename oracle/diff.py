"""FP64 reference of the differentiable correlation score.

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule).

The reference differentiates the score only in Euler angles
(``WignerDGradients.get_derivatives``, ``src/matcha/core/WignerDGradients.py:
247-302``); the gradient with respect to an axis-angle rotation, the shift and
the template coefficients is new.  The score is the reference's pipeline for
one particle at a given pose:

    s(omega, t, aT) = Re sum_{l <= L} D^l(R(omega)) sigma^l(aT, aV(t))

* R(omega) = exp([omega]x), acting on the volumes' (axis 0, 1, 2)
  coordinates; D in the reference's convention at the z-y-z Euler angles of R
  (``Matcha.py:184,288``, ``oracle.so3.score``);
* aV(t) = FLE of the particle patch at shift t, extracted as
  ``extract_subtomogram_patch_batch`` does (``oracle.resample.extract_patch``:
  trilinear, zeros outside, patch[a] = V[a + flip(t)]);
* sigma = ``CrossCorrelationMatcher.compute_sigma`` (``oracle.fle.compute_sigma``).

Every piece is an oracle function pinned elsewhere; the reference gradient is
the FP64 central difference of this composition (``fd_gradient``).  The helpers
below (Rodrigues, the right Jacobian of exp, the body-frame generators of D)
are the closed forms the GPU adjoint uses, pinned in
``tests/test_diff_host.py`` against scipy and against FP64 differences.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial.transform import Rotation

from . import fle, resample, so3
from . import wigner as W


def rotvec_to_matrix(omega) -> np.ndarray:
    """exp([omega]x) by Rodrigues' formula, (..., 3) -> (..., 3, 3)."""
    w = np.asarray(omega, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1)[..., None, None]
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2], K[..., 1, 2] = -w[..., 2], w[..., 1], -w[..., 0]
    K[..., 1, 0], K[..., 2, 0], K[..., 2, 1] = w[..., 2], -w[..., 1], w[..., 0]
    small = th < 1e-8
    ths = np.where(small, 1.0, th)
    a = np.where(small, 1.0 - th ** 2 / 6, np.sin(ths) / ths)
    b = np.where(small, 0.5 - th ** 2 / 24, (1 - np.cos(ths)) / ths ** 2)
    return np.eye(3) + a * K + b * (K @ K)


def right_jacobian(omega) -> np.ndarray:
    """J_r with exp([omega + d]x) = exp([omega]x) exp([J_r d]x) + O(d^2):
    I - (1 - cos t)/t^2 [w]x + (t - sin t)/t^3 [w]x^2."""
    w = np.asarray(omega, dtype=np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th < 1e-6:
        a, b = 0.5 - th ** 2 / 24, 1.0 / 6 - th ** 2 / 120
    else:
        a, b = (1 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
    return np.eye(3) - a * K + b * (K @ K)


def euler_zyz(R: np.ndarray):
    """(alpha, beta, gamma) with R = Rz(alpha) Ry(beta) Rz(gamma)."""
    return Rotation.from_matrix(R).as_euler("ZYZ")


def body_sigmas(sigma: np.ndarray, L: int):
    """sigma'_i (i = x, y, z) with Re sum D(R) sigma'_i = d/de s(R exp(e [e_i]x)) at 0:
    sigma'_{m'n} = sum_m X_i[n, m] sigma_{m'm}, X_z = diag(i m),
    X_y[n, n+1] = C(n) = -X_y[n+1, n], X_x[n, n+-1] = i C(min), C(n) = sqrt((l-n)(l+n+1))/2."""
    sigma = np.asarray(sigma, dtype=np.complex128)
    out = np.zeros((3, W.dsize(L)), dtype=np.complex128)
    for l in range(L + 1):
        for mp in range(-l, l + 1):
            for n in range(-l, l + 1):
                j = W.dindex(l, mp, n)
                up = sigma[W.dindex(l, mp, n + 1)] if n < l else 0.0
                dn = sigma[W.dindex(l, mp, n - 1)] if n > -l else 0.0
                cu = 0.5 * np.sqrt((l - n) * (l + n + 1.0))
                cd = 0.5 * np.sqrt((l - n + 1.0) * (l + n))
                out[0, j] = 1j * (cu * up + cd * dn)
                out[1, j] = cu * up - cd * dn
                out[2, j] = 1j * n * sigma[j]
    return out


class Problem:
    """One particle volume and a template's coefficients at a FLE size."""

    def __init__(self, vol: np.ndarray, aT: np.ndarray, N: int, bandlimit: int, radius: float, L: int):
        self.vol = np.asarray(vol, dtype=np.float64)
        self.aT = np.asarray(aT, dtype=np.complex128)
        self.N, self.bl, self.radius, self.L = N, bandlimit, radius, L
        ls, ks, ms, mds, lmds, cs, ne = fle.lap_eig_ball(N, bandlimit)
        self.ls, self.mds, self.ne = ls, mds, ne
        self.lmax = int(ls.max())
        self.Bmat, self.flat = fle.basis_matrix(N, bandlimit, radius)

    def coeffs(self, vol: np.ndarray) -> np.ndarray:
        return vol.reshape(-1)[self.flat] @ np.conj(self.Bmat)

    def sigma(self, t, aT=None) -> np.ndarray:
        aT = self.aT if aT is None else aT
        aV = self.coeffs(resample.extract_patch(self.vol, t))
        return fle.compute_sigma(aT, aV[None], self.ls, self.mds, self.lmax)[0]

    def score(self, omega, t, aT=None) -> float:
        a, b, g = euler_zyz(rotvec_to_matrix(omega))
        return float(so3.score(a, b, g, self.sigma(t, aT), self.L)[0])

    def fd_gradient(self, omega, t, h_rot=1e-6, h_shift=1e-6):
        """FP64 central differences in omega (3) and t (3)."""
        omega = np.asarray(omega, dtype=np.float64)
        t = np.asarray(t, dtype=np.float64)
        g = np.zeros(6)
        for i in range(3):
            e = np.zeros(3)
            e[i] = h_rot
            g[i] = (self.score(omega + e, t) - self.score(omega - e, t)) / (2 * h_rot)
            e[i] = h_shift
            g[3 + i] = (self.score(omega, t + e) - self.score(omega, t - e)) / (2 * h_shift)
        return g

    def template_directional(self, omega, t, delta: np.ndarray) -> float:
        """s(aT + delta) - s(aT): exact (s is linear in aT)."""
        return self.score(omega, t, self.aT + delta) - self.score(omega, t)
