"""FP64 CPU restatement of the reference's ball-harmonic ("FLE") expansion and
of the correlation coefficients sigma.

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule).

Restated reference functions (swing-research/Matcha):

* ``lap_eig_ball`` -- ``FLEBasis3D.lap_eig_ball`` (``src/matcha/core/
  FLEBasis3D.py:700-799``): enumerate (k, l, m) with m ordered 0,-1,1,-2,2,...,
  stable sort by lambda_lk, truncate at lambda < bandlimit*pi/2 with the
  reference's ``searchsorted(...) - 1`` and the +-m pair rule.  The roots
  lambda_lk of j_l come from the reference's missing table
  ``jl_zeros_l=3000_k=2500.mat``; here they are recomputed in FP64
  (bracket + Brent on scipy's spherical_jn).  c_lk = sqrt(2)/|j_{l+1}(lambda_lk)|
  (unit-ball L2 normalisation; the ``cs_l=3000_k=2500.mat`` table is missing).
* ``evaluate_t`` -- the quantity ``FLEBasis3D.evaluate_t`` approximates
  (``:945-976`` via NUFFT/SHT/interpolation, ``:1026-1267``): the adjoint of
  the dense basis of ``create_denseB`` (``:570-698``) with the radius scaling of
  ``precomp`` (``:383,390-409,502-515``):
      a_i = h c_i sum_{|v| <= radius} f(v) j_l(lambda_i |v|/radius) conj(Y_lm(v/|v|)),
  h = radius^-1.5, v = index - N/2 with axis 0 = x, 1 = y, 2 = z; Y_lm
  orthonormal with the Condon-Shortley phase (scipy's sph_harm_y).  Computed
  by brute force.
* ``compute_sigma`` -- ``CrossCorrelationMatcher.compute_sigma`` /
  ``_precompute_indices`` (``src/matcha/core/CrossCorrelationMatcher.py:
  133-212``): sigma[Dindex(l,u,v)] = (-1)^{u+v} sum_k a^T_{l,k,-v} conj(a^V_{l,k,-u}).

Parity pins: roots against mpmath's besseljzero and the closed form
lambda_{0,k} = k pi; c_lk and the basis by discrete orthonormality on the
voxel grid; evaluate_t by exact rotation covariance under 90-degree grid
rotations (a size-independent property); sigma by the identity
Re sum D(R) sigma = <a^T, rotated a^V>.  The exact values of the missing
.mat tables (last-bit effects at lambda == bandlimit*pi/2 for even
bandlimits) are parity-unpinned.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np
from scipy import optimize, special

from . import wigner as W


# --------------------------------------------------------------------------
# roots of spherical Bessel functions
# --------------------------------------------------------------------------

@lru_cache(maxsize=None)
def jl_roots(l: int, xmax: float) -> tuple:
    """All positive roots of j_l below xmax (FP64)."""
    if l == 0:
        k = np.arange(1, int(xmax / np.pi) + 2)
        r = k * np.pi
        return tuple(r[r < xmax])
    f = lambda x: special.spherical_jn(l, x)
    xs = np.linspace(max(l * 0.5, 1e-3), xmax + 1.0, int((xmax + 1.0) * 40) + 10)
    ys = f(xs)
    roots = []
    for i in range(len(xs) - 1):
        if ys[i] == 0.0:
            roots.append(xs[i])
        elif ys[i] * ys[i + 1] < 0:
            roots.append(optimize.brentq(f, xs[i], xs[i + 1], xtol=1e-15, rtol=8.9e-16, maxiter=200))
    return tuple(r for r in roots if r < xmax)


def lap_eig_ball(N: int, bandlimit: int):
    """(ls, ks, ms, mds, lmds, cs, ne) in the reference's order and truncation."""
    threshold = bandlimit * np.pi / 2
    # every root below the threshold (the reference's loop enumerates them all
    # before its early stop; see module docstring)
    ent_l, ent_k, ent_lmd = [], [], []
    l = 0
    while True:
        rts = jl_roots(l, threshold + 1.0)
        if not rts:
            break
        for k, r in enumerate(rts, start=1):
            ent_l.append(l)
            ent_k.append(k)
            ent_lmd.append(r)
        l += 1
    # insertion order of the reference: k outer, l inner, m in 0,-1,1,...
    order = sorted(range(len(ent_l)), key=lambda i: (ent_k[i], ent_l[i]))
    ls, ks, ms, lmds = [], [], [], []
    for i in order:
        l, k, lmd = ent_l[i], ent_k[i], ent_lmd[i]
        for j in range(2 * l + 1):
            ls.append(l)
            ks.append(k)
            ms.append(((-1) ** j) * ((j + 1) // 2))
            lmds.append(lmd)
    ls, ks, ms, lmds = map(np.array, (ls, ks, ms, lmds))
    idx = np.argsort(lmds, kind="stable")
    ls, ks, ms, lmds = ls[idx], ks[idx], ms[idx], lmds[idx]
    ne = int(np.searchsorted(lmds, threshold, side="left")) - 1
    if ms[ne - 1] < 0:
        ne -= 1
    if ne <= 1:
        ne = 1
    ls, ks, ms, lmds = ls[:ne], ks[:ne], ms[:ne], lmds[:ne]
    mds = 2 * np.abs(ms) - (ms < 0)
    cs = np.sqrt(2.0) / np.abs(special.spherical_jn(ls + 1, lmds))
    return ls, ks, ms, mds, lmds, cs, ne


# --------------------------------------------------------------------------
# exact projection (brute force)
# --------------------------------------------------------------------------

def ball_voxels(N: int, radius: float):
    """Integer coordinates (x, y, z) = index - N/2 inside |v|/radius <= 1 + 1e-13,
    and their flat indices (axis 0 = x)."""
    ax = np.arange(N) - N // 2
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) / radius
    inside = r <= 1 + 1e-13
    flat = np.nonzero(inside.ravel())[0]
    return X.ravel()[flat], Y.ravel()[flat], Z.ravel()[flat], flat


def basis_matrix(N: int, bandlimit: int, radius: float):
    """Dense B (V_ball x ne) complex with h folded in, and the ball's flat indices."""
    ls, ks, ms, mds, lmds, cs, ne = lap_eig_ball(N, bandlimit)
    x, y, z, flat = ball_voxels(N, radius)
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    theta = np.arctan2(np.sqrt(x ** 2 + y ** 2), z)
    phi = np.arctan2(y, x)
    h = radius ** -1.5
    B = np.zeros((len(flat), ne), dtype=np.complex128)
    for i in range(ne):
        jl = special.spherical_jn(ls[i], lmds[i] * r / radius)
        Y = special.sph_harm_y(ls[i], ms[i], theta, phi)
        B[:, i] = h * cs[i] * jl * Y
    return B, flat


def evaluate_t(vols: np.ndarray, N: int, bandlimit: int, radius: float) -> np.ndarray:
    """a = B^H f for a batch of volumes (Bt, N, N, N) -> (Bt, ne) complex128."""
    vols = np.asarray(vols, dtype=np.float64).reshape(-1, N ** 3)
    B, flat = basis_matrix(N, bandlimit, radius)
    return vols[:, flat] @ np.conj(B)


# --------------------------------------------------------------------------
# sigma
# --------------------------------------------------------------------------

def sht_index(m: int) -> int:
    """``get_sht_index`` (CrossCorrelationMatcher.py:19-34)."""
    if m == 0:
        return 0
    return 2 * m if m > 0 else -2 * m - 1


def idlm_lists(ls, mds, lmax):
    out = [[[] for _ in range(2 * lmax + 1)] for _ in range(lmax + 1)]
    for i, (l, md) in enumerate(zip(ls, mds)):
        out[l][md].append(i)
    return out


def compute_sigma(aT: np.ndarray, aV: np.ndarray, ls, mds, lmax: int) -> np.ndarray:
    """sigma (Bt, Dsize(lmax)) from template aT (ne,) and particles aV (Bt, ne)."""
    aV = np.atleast_2d(aV)
    idl = idlm_lists(ls, mds, lmax)
    out = np.zeros((aV.shape[0], W.dsize(lmax)), dtype=np.complex128)
    for l in range(lmax + 1):
        for u in range(-l, l + 1):
            iv = idl[l][sht_index(-u)]
            for v in range(-l, l + 1):
                it = idl[l][sht_index(-v)]
                n = min(len(iv), len(it))
                acc = (aT[it[:n]][None, :] * np.conj(aV[:, iv[:n]])).sum(axis=1)
                out[:, W.dindex(l, u, v)] = ((-1) ** (u + v)) * acc
    return out


# --------------------------------------------------------------------------
# synthetic volumes
# --------------------------------------------------------------------------

def blob_volume(rng, N: int, n_blobs: int, sigma_range=(1.5, 4.0), amp_range=(0.5, 1.0),
                ball_frac=0.35):
    """Sum of isotropic Gaussian blobs, centres uniform in a ball of radius
    ball_frac*N (the synthetic template of the benchmark workloads)."""
    ax = np.arange(N) - N // 2
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    vol = np.zeros((N, N, N))
    R = ball_frac * N
    for _ in range(n_blobs):
        while True:
            c = rng.uniform(-R, R, 3)
            if np.linalg.norm(c) <= R:
                break
        s = rng.uniform(*sigma_range)
        a = rng.uniform(*amp_range)
        vol += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    return vol


def evaluate_t_subset(vols: np.ndarray, N: int, bandlimit: int, radius: float, indices) -> np.ndarray:
    """Selected coefficients a_i (brute force, FP64): (Bt, len(indices))."""
    ls, ks, ms, mds, lmds, cs, ne = lap_eig_ball(N, bandlimit)
    x, y, z, flat = ball_voxels(N, radius)
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    theta = np.arctan2(np.sqrt(x ** 2 + y ** 2), z)
    phi = np.arctan2(y, x)
    h = radius ** -1.5
    vols = np.asarray(vols, dtype=np.float64).reshape(-1, N ** 3)[:, flat]
    out = np.zeros((vols.shape[0], len(indices)), dtype=np.complex128)
    for j, i in enumerate(indices):
        basis = h * cs[i] * special.spherical_jn(ls[i], lmds[i] * r / radius) * \
            special.sph_harm_y(ls[i], ms[i], theta, phi)
        out[:, j] = vols @ np.conj(basis)
    return out


def rotate_y90(vol: np.ndarray) -> np.ndarray:
    """f'(v) = f(R^-1 v) for R: (x, y, z) -> (z, y, -x) (90 degrees about y), on the
    grid v = index - N/2; voxels whose preimage leaves the box are set to 0."""
    N = vol.shape[-1]
    ax = np.arange(N) - N // 2
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    # R^-1 (x, y, z) = (-z, y, x)
    sx, sy, sz = -Z, Y, X
    ok = (sx >= -N // 2) & (sx < N - N // 2)
    out = np.zeros_like(vol)
    ix = np.clip(sx + N // 2, 0, N - 1)
    out[..., ok] = vol[..., ix[ok], (sy + N // 2)[ok], (sz + N // 2)[ok]]
    return out
