"""FP64 CPU restatement of the reference's Wigner-D machinery.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product (``matcha_amd``) never
imports anything from ``oracle/``.

What it restates (reference = swing-research/Matcha, paths relative to its
repository root):

* ``dsize``/``dindex``          -- ``src/matcha/core/WignerDMatrices.py:48-97,
  969-1012`` (D layout ``[(l, m', m) for l for m' for m]``).
* ``hindex``                    -- ``WignerDMatrices.py:146-239`` (H wedge).
* ``h_wedge``                   -- the four-step H recurrence of
  ``WignerDMatrices.py:673-801`` with the coefficient tables of
  ``:803-965,1096-1107`` (Gumerov-Duraiswami recurrence as in moble/spherical),
  vectorised over beta, in FP64.
* ``wigner_D``                  -- ``WignerDMatrices.D`` (``:1291-1375``) with the
  fill of ``:281-306``: ``D[l,m',m] = eps(m') eps(-m) H(beta) e^{i m' a} e^{i m g}``
  and ``eps`` of ``:662-669``; truncation at ``l <= truncate``.
* ``wigner_d_explicit``         -- independent closed form (Wigner's sum), used
  only to pin the convention of the restatement; not a reference function.

Parity status: ``h_wedge``/``wigner_D`` are pinned by (a) Wigner's explicit
sum for l <= 8 (``wigner_d_explicit``), (b) unitarity of D(R) for l <= 32,
(c) the group law D(R1 R2) = D(R1) D(R2), (d) d(0) = identity.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


# --------------------------------------------------------------------------
# index algebra
# --------------------------------------------------------------------------

def dsize(ell_max: int) -> int:
    """Number of D entries for 0 <= l <= ell_max (full m' range).

    Closed form of ``WignerDsize(0, ell_max, ell_max)``
    (WignerDMatrices.py:48-97): sum_l (2l+1)^2.
    """
    if ell_max < 0:
        return 0
    L = ell_max
    return (L + 1) * (2 * L + 1) * (2 * L + 3) // 3


def dindex(ell: int, mp: int, m: int) -> int:
    """``WignerDindex(ell, mp, m)`` with full m' range (WignerDMatrices.py:969-1012)."""
    return dsize(ell - 1) + (mp + ell) * (2 * ell + 1) + (m + ell)


def compute_size(n: int) -> int:
    """``setup_utils.compute_size`` (setup_utils.py:35-37) == dsize(n-1)."""
    n = n - 1
    return (2 * n * (n + 1) * (2 * n + 1)) // 3 + 2 * n * (n + 1) + n + 1


def hsize(ell_max: int) -> int:
    """``WignerHsize(mp_max=ell_max, ell_max)`` (WignerDMatrices.py:181-213)."""
    if ell_max < 0:
        return 0
    return (ell_max + 1) * (ell_max + 2) * (2 * ell_max + 3) // 6


def _hindex_raw(ell: int, mp: int, m: int) -> int:
    """``_WignerHindex`` with mp_max >= ell (WignerDMatrices.py:216-227)."""
    mp_max = ell
    i = hsize(ell - 1)
    if mp < 1:
        i += (mp_max + mp) * (2 * ell - mp_max + mp + 1) // 2
    else:
        i += (mp_max + 1) * (2 * ell - mp_max + 2) // 2
        i += (mp - 1) * (2 * ell - mp + 2) // 2
    i += m - abs(mp)
    return i


def hindex(ell: int, mp: int, m: int) -> int:
    """``WignerHindex`` (WignerDMatrices.py:146-178): symmetric wedge lookup."""
    if ell == 0:
        return 0
    if m < -mp:
        if m < mp:
            return _hindex_raw(ell, -mp, -m)
        return _hindex_raw(ell, -m, -mp)
    if m < mp:
        return _hindex_raw(ell, m, mp)
    return _hindex_raw(ell, mp, m)


def eps(m: int) -> int:
    """``ϵ(m)`` (WignerDMatrices.py:662-669)."""
    if m <= 0:
        return 1
    return -1 if (m % 2) else 1


@lru_cache(maxsize=None)
def d_layout(ell_max: int):
    """(l, m', m) integer arrays in D order for l <= ell_max."""
    ls, mps, ms = [], [], []
    for ell in range(ell_max + 1):
        for mp in range(-ell, ell + 1):
            for m in range(-ell, ell + 1):
                ls.append(ell)
                mps.append(mp)
                ms.append(m)
    return (np.array(ls, dtype=np.int64), np.array(mps, dtype=np.int64),
            np.array(ms, dtype=np.int64))


# --------------------------------------------------------------------------
# H recurrence (restated, FP64)
# --------------------------------------------------------------------------

def _nm(n, m):
    return m + n * (n + 1)


def _coef_tables(ell_max: int):
    """a, b, d, g, h of WignerDMatrices.py:1096-1107 as dict lookups."""
    def a(n, absm):
        return math.sqrt((n + 1 + absm) * (n + 1 - absm) / ((2 * n + 1) * (2 * n + 3)))

    def b(n, m):
        v = math.sqrt((n - m - 1) * (n - m) / ((2 * n - 1) * (2 * n + 1)))
        return -v if m < 0 else v

    def d(n, m):
        v = 0.5 * math.sqrt((n - m) * (n + m + 1))
        return -v if m < 0 else v

    def g(n, m):
        return 2 * (m + 1) / math.sqrt((n - m) * (n + m + 1))

    def h(n, m):
        return math.sqrt((n + m + 2) * (n - m - 1) / ((n - m) * (n + m + 1)))

    return a, b, d, g, h


def h_wedge(beta, ell_max: int):
    """H^l_{m',m}(beta) for l <= ell_max, wedge layout, shape (nb, hsize).

    Restates steps 1-5 of WignerDMatrices.H (``:673-801`` kernels with the
    index/coefficient precompute of ``:803-965``) in FP64; the "extra"
    row n = ell_max+1 of ``step_1_2_cuda`` is kept in a separate array.
    """
    beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    nb = beta.shape[0]
    cb = np.cos(beta)
    sb = np.sin(beta)
    sb2 = sb * sb
    N = ell_max
    a, b, d, g, h = _coef_tables(N)

    H = np.zeros((nb, hsize(N)), dtype=np.float64)
    # column m'=0 for n = 0..N+1, stored per n as (nb, n+1) arrays
    col0 = [None] * (N + 2)
    H0 = np.ones(nb)
    col0[0] = H0[:, None].copy()
    if N >= 0:
        H[:, 0] = 1.0
    if N > 0:
        # step 2 (step_1_2_cuda, :673-741)
        corner_un = [None] * (N + 2)  # un-normalised H^n_{0,n}
        c1 = np.zeros((nb, 2))
        c1[:, 1] = math.sqrt(3.0)
        # H^1_{0,0} = g(1,0) cos(beta) / sqrt(2), written directly (not from the
        # un-normalised corner) at WignerDMatrices.py:697.
        c1[:, 0] = g(1, 0) * cb / math.sqrt(2.0)
        corner_un[1] = c1[:, 1].copy()
        col0[1] = c1
        for n in range(2, N + 2):
            col = np.zeros((nb, n + 1))
            col[:, n] = math.sqrt(1.0 + 0.5 / n) * corner_un[n - 1]
            corner_un[n] = col[:, n].copy()
            col[:, n - 1] = g(n, n - 1) * cb * col[:, n]
            for i in range(2, n):
                mm = n - i
                col[:, mm] = g(n, mm) * cb * col[:, mm + 1] - h(n, mm) * sb2 * col[:, mm + 2]
            const = 1.0 / math.sqrt(4 * n + 2)
            col[:, 0] = (g(n, 0) * cb * col[:, 1] - h(n, 0) * sb2 * col[:, 2]) * const
            pref = np.full(nb, const)
            for i in range(1, n):
                pref = pref * sb
                col[:, i] = col[:, i] * pref
            col0[n] = col
        # corner normalisation (:735-741)
        pref = np.ones(nb)
        for n in range(1, N + 2):
            pref = pref * sb
            col0[n][:, n] = col0[n][:, n] * pref / math.sqrt(4 * n + 2)
        for n in range(1, N + 1):
            base = _hindex_raw(n, 0, 0)
            H[:, base:base + n + 1] = col0[n]
        # step 3 (:745-774 with :803-841): H^n_{1,m}, m = 1..n from H^{n+1}_{0,.}
        for n in range(1, N + 1):
            i1 = _hindex_raw(n, 1, 1)
            inv_b5 = 1.0 / b(n + 1, 0)
            nxt = col0[n + 1]
            for i in range(n):
                b6 = b(n + 1, -i - 2) * inv_b5 * 0.5
                b7 = b(n + 1, i) * inv_b5 * 0.5
                a8 = a(n, 1 + i) * inv_b5
                H[:, i1 + i] = (b6 * (1 - cb) * nxt[:, i + 2]
                                - b7 * (1 + cb) * nxt[:, i]
                                - a8 * sb * nxt[:, i + 1])
        # step 4 (:776-789 with :843-892): m' > 0
        for n in range(2, N + 1):
            for mp in range(1, n):
                d5 = d(n, mp)
                for m in range(mp + 1, n + 1):
                    v = d(n, mp - 1) * H[:, hindex(n, mp - 1, m)] - d(n, m - 1) * H[:, hindex(n, mp, m - 1)]
                    if m < n:
                        v = v + d(n, m) * H[:, hindex(n, mp, m + 1)]
                    H[:, _hindex_raw(n, mp + 1, m)] = v / d5
        # step 5 (:790-801 with :907-965): m' < 0
        for n in range(0, N + 1):
            for mp in range(0, -n, -1):
                d5 = d(n, mp - 1)
                for m in range(-mp + 1, n + 1):
                    v = d(n, mp) * H[:, hindex(n, mp + 1, m)] + d(n, m - 1) * H[:, hindex(n, mp, m - 1)]
                    if m < n:
                        v = v - d(n, m) * H[:, hindex(n, mp, m + 1)]
                    H[:, _hindex_raw(n, mp - 1, m)] = v / d5
    return H


def wigner_d(beta, ell_max: int):
    """d^l_{m'm}(beta) = eps(m') eps(-m) H^l_{m'm}(beta) in D layout, (nb, dsize)."""
    H = h_wedge(beta, ell_max)
    ls, mps, ms = d_layout(ell_max)
    hidx = np.array([hindex(int(l), int(mp), int(m)) for l, mp, m in zip(ls, mps, ms)], dtype=np.int64)
    res = np.array([eps(int(mp)) * eps(int(-m)) for mp, m in zip(mps, ms)], dtype=np.float64)
    return H[:, hidx] * res[None, :]


def wigner_D(alpha, beta, gamma, ell_max: int):
    """D^l_{m'm}(alpha, beta, gamma) = d^l_{m'm}(beta) e^{i m' alpha} e^{i m gamma}.

    ``WignerDMatrices.D`` (WignerDMatrices.py:1291-1375, fill :281-306).
    Returns complex128 array (nb, dsize(ell_max)).
    """
    alpha = np.atleast_1d(np.asarray(alpha, dtype=np.float64))
    gamma = np.atleast_1d(np.asarray(gamma, dtype=np.float64))
    dd = wigner_d(beta, ell_max)
    _, mps, ms = d_layout(ell_max)
    phase = np.exp(1j * (mps[None, :] * alpha[:, None] + ms[None, :] * gamma[:, None]))
    return dd * phase


# --------------------------------------------------------------------------
# independent closed form (for pinning only)
# --------------------------------------------------------------------------

def wigner_d_explicit(beta: float, ell: int, mp: int, m: int) -> float:
    """Wigner's explicit sum for d^j_{m'm}(beta) (textbook convention)."""
    j = ell
    f = math.factorial
    pref = math.sqrt(f(j + mp) * f(j - mp) * f(j + m) * f(j - m))
    c = math.cos(beta / 2.0)
    s = math.sin(beta / 2.0)
    tot = 0.0
    for k in range(0, 2 * j + 1):
        a1 = j + m - k
        a2 = mp - m + k
        a3 = j - mp - k
        if a1 < 0 or a2 < 0 or a3 < 0:
            continue
        term = ((-1) ** (mp - m + k)) / (f(a1) * f(k) * f(a2) * f(a3))
        term *= c ** (2 * j + m - mp - 2 * k) * s ** (mp - m + 2 * k)
        tot += term
    return pref * tot
