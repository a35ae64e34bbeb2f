"""FP64 CPU restatement of the reference's SO(3) search (grid + Newton march).

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule).

Restated reference functions (swing-research/Matcha):

* ``sofft_grid``         -- ``SOFFT._precompute`` + ``SOFFT.eval``
  (``src/matcha/core/SOFFT.py:30-98``): f(b_j, a_a, g_g) =
  Re ifft2 over (alpha, gamma) of sum_{l < L0} d^l(beta_j) sigma, beta_j =
  linspace(0, pi, n) with endpoint, alpha/gamma bins m mod n.
* ``topk_grid``          -- ``sparse_topk_per_batch_jit`` (``Matcha.py:27-58``)
  + ``SOFFT.ids_to_angles`` (``SOFFT.py:100-124``).  Tie order among equal
  values is implementation-defined in torch; here ties break by lower flat
  index (documented as unpinned).
* ``coeff_bank``/``derivatives`` -- ``WignerDGradients.set_so3_coeffs`` and
  ``get_derivatives`` (``WignerDGradients.py:6-152,155-302``): the 15-column
  ladder bank, one contraction with D, alpha phases.
* ``optimize``/``search_orientations`` -- ``Matcha.optimize``,
  ``Matcha._take_step``, ``Matcha.search_orientations`` and
  ``reduce_over_candidates`` (``Matcha.py:11-24,150-205,239-403``), including
  the singular-Hessian per-candidate skip (``:315-343``) and the lr rule
  (x1.1 on accept, x0.5 on reject).

Parity pins: ``derivatives`` vs central finite differences of the score
(closed form, FP64); ``sofft_grid`` vs direct evaluation of Re sum D sigma / n^2
on every grid point; the march vs monotone-score and fixed-point invariants,
and its rules (accept iff s' > best, lr x1.1 / x0.5, per-level reset of lr
and best, the Newton step lr H^-1 g subtracted, the gradient step, the
batch-global stop_early exit) by closed-form trajectories of scores
A cos(beta) + B P2(cos(beta)) and by a Newton step from finite differences.
The accept/reject trajectory of the FP32 reference where the decision rests
on round-off (near convergence) is unpinned.
"""
from __future__ import annotations

import numpy as np

from . import wigner as W


# --------------------------------------------------------------------------
# SOFFT grid (SOFFT.py:30-98)
# --------------------------------------------------------------------------

def sofft_tables(L0: int, K: int):
    n = 2 * L0 * K
    betas = np.linspace(0.0, np.pi, n, endpoint=True)
    d = W.wigner_d(betas, L0 - 1)  # (n, compute_size(L0)) -> l < L0
    _, mps, ms = W.d_layout(L0 - 1)
    return n, betas, d, mps, ms


def sofft_grid(sigma: np.ndarray, L0: int, K: int) -> np.ndarray:
    """Score grid f[b, j(beta), a(alpha), g(gamma)] (real, FP64)."""
    sigma = np.atleast_2d(sigma)
    n, _, d, mps, ms = sofft_tables(L0, K)
    nsz = d.shape[1]
    B = sigma.shape[0]
    out = np.zeros((B, n, n, n))
    for b in range(B):
        temp = d * sigma[b, :nsz][None, :]  # (n_beta, K)
        F = np.zeros((n, n * n), dtype=np.complex128)
        lin = (mps % n) * n + (ms % n)
        for j in range(n):
            np.add.at(F[j], lin, temp[j])
        out[b] = np.fft.ifft2(F.reshape(n, n, n), axes=(1, 2)).real
    return out


def topk_grid(f: np.ndarray, C: int):
    """Top-C per batch over the flattened (beta, alpha, gamma) grid.

    Returns (ids (B, C, 3) int64 as (i_beta, i_alpha, i_gamma), values (B, C)).
    Ties: descending value, then ascending flat index.
    """
    B = f.shape[0]
    X, Y, Z = f.shape[1:]
    flat = f.reshape(B, -1)
    k = min(C, flat.shape[1])
    ids = np.full((B, C, 3), -1, dtype=np.int64)
    vals = np.full((B, C), -1.0)
    for b in range(B):
        order = np.lexsort((np.arange(flat.shape[1]), -flat[b]))[:k]
        vals[b, :k] = flat[b, order]
        ids[b, :k, 0] = order // (Y * Z)
        ids[b, :k, 1] = (order // Z) % Y
        ids[b, :k, 2] = order % Z
    return ids, vals


def ids_to_angles(ids: np.ndarray, n: int):
    betas = ids[..., 0] * (np.pi / (n - 1))
    alphas = ids[..., 1] * (2 * np.pi / n)
    gammas = ids[..., 2] * (2 * np.pi / n)
    return alphas, betas, gammas


# --------------------------------------------------------------------------
# gradient / Hessian via the 15-column bank (WignerDGradients.py)
# --------------------------------------------------------------------------

def _c1(j, m):
    s = (j + m) * (j - m + 1)
    return 0.0 if s < 0 else -0.5 * np.sqrt(s)


def _c2(j, m):
    s = (j - m) * (j + m + 1)
    return 0.0 if s < 0 else 0.5 * np.sqrt(s)


def coeff_bank(sigma_t: np.ndarray, ell: int) -> np.ndarray:
    """15-column bank for sigma truncated to l <= ell; (B, 15, dsize(ell)).

    Restates ``set_so3_coeffs`` (WignerDGradients.py:155-243): first D index
    is "m" there (the alpha index), second is "m_prime" (the gamma index).
    """
    sigma_t = np.atleast_2d(sigma_t)[:, :W.dsize(ell)]
    B, C = sigma_t.shape
    ls, mA, mG = W.d_layout(ell)
    m = mA.astype(np.float64)
    mp = mG.astype(np.float64)
    f_minus = np.array([_c1(int(j), int(a)) for j, a in zip(ls, mA)])
    f_plus = np.array([_c2(int(j), int(a)) for j, a in zip(ls, mA)])
    lf = ls.astype(np.float64)
    tn2 = np.sqrt(np.maximum((lf + m) * (lf - m + 1) * (lf + m - 1) * (lf - m + 2), 0)) / 4
    t0 = -((lf + m) * (lf - m + 1) + (lf + m + 1) * (lf - m)) / 4
    tp2 = np.sqrt(np.maximum((lf + m + 1) * (lf - m) * (lf + m + 2) * (lf - m - 1), 0)) / 4
    # the reference's sqrt((l+m)(l-m+1)(l+m-1)(l-m+2)) is only evaluated for
    # index pairs whose shifted partner exists; invalid ones are dropped by the
    # map below, so clamping the radicand at 0 does not change the bank.
    idx = np.arange(C)
    def shifted(dm):
        out = np.full(C, -1, dtype=np.int64)
        ok = (np.abs(mA + dm) <= ls)
        out[ok] = [W.dindex(int(l), int(a + dm), int(g)) for l, a, g in zip(ls[ok], mA[ok], mG[ok])]
        return out
    mp1, mn1, mp2, mn2 = shifted(+1), shifted(-1), shifted(+2), shifted(-2)
    j = 1j
    c = sigma_t
    c_val = c
    c_a = (j * m) * c
    c_g = (j * mp) * c
    c_aa = (j * m) ** 2 * c
    c_gg = (j * mp) ** 2 * c
    c_ag = (j * m) * (j * mp) * c
    c_bm = np.zeros_like(c)
    c_bp = np.zeros_like(c)
    c_bbm = np.zeros_like(c)
    c_bbp = np.zeros_like(c)
    ok = mn1 >= 0
    np.add.at(c_bm, (slice(None), mn1[ok]), c[:, idx[ok]] * f_minus[idx[ok]])
    ok = mp1 >= 0
    np.add.at(c_bp, (slice(None), mp1[ok]), c[:, idx[ok]] * f_plus[idx[ok]])
    c_bb0 = t0 * c
    ok = mn2 >= 0
    np.add.at(c_bbm, (slice(None), mn2[ok]), c[:, idx[ok]] * tn2[idx[ok]])
    ok = mp2 >= 0
    np.add.at(c_bbp, (slice(None), mp2[ok]), c[:, idx[ok]] * tp2[idx[ok]])
    c_abm = (j * m) * c_bm
    c_abp = (j * m) * c_bp
    c_gbm = (j * mp) * c_bm
    c_gbp = (j * mp) * c_bp
    return np.stack([c_val, c_a, c_g, c_aa, c_gg, c_ag, c_bm, c_bp, c_abm, c_abp,
                     c_gbm, c_gbp, c_bbm, c_bb0, c_bbp], axis=1)


def derivatives(D: np.ndarray, bank: np.ndarray, alpha: np.ndarray):
    """``get_derivatives`` (WignerDGradients.py:247-302).

    D: (B, C, n) complex, bank: (B, 15, n), alpha: (B, C).
    Returns grad (B, C, 3) and hessian (B, C, 3, 3), order (alpha, beta, gamma).
    """
    Y = np.einsum('bnk,bck->bnc', D, bank)
    pp = np.exp(1j * alpha)
    pn = np.exp(-1j * alpha)
    p2p = np.exp(2j * alpha)
    p2n = np.exp(-2j * alpha)
    Ybm, Ybp, Yabm, Yabp = Y[..., 6], Y[..., 7], Y[..., 8], Y[..., 9]
    beta_out = pp * Ybm + pn * Ybp
    gb_out = pp * Y[..., 10] + pn * Y[..., 11]
    ab = pp * (1j * Ybm + Yabm) + pn * (-1j * Ybp + Yabp)
    bb = Y[..., 13] + p2p * Y[..., 12] + p2n * Y[..., 14]
    grad = np.stack([Y[..., 1], beta_out, Y[..., 2]], axis=-1).real
    H = np.stack([
        np.stack([Y[..., 3], ab, Y[..., 5]], axis=-1),
        np.stack([ab, bb, gb_out], axis=-1),
        np.stack([Y[..., 5], gb_out, Y[..., 4]], axis=-1)], axis=-2).real
    return grad, H


def score(alpha, beta, gamma, sigma_b, ell):
    """s = Re sum_{l <= ell} D(alpha, beta, gamma) sigma (Matcha.py:184,288)."""
    D = W.wigner_D(alpha, beta, gamma, ell)
    return (D * sigma_b[None, :W.dsize(ell)]).sum(axis=1).real


# --------------------------------------------------------------------------
# Newton march (Matcha.py)
# --------------------------------------------------------------------------

def _solve_or_skip(H: np.ndarray, g: np.ndarray):
    """torch.linalg.solve semantics with the per-candidate singular skip.

    A candidate whose Hessian is exactly singular gets a zero step
    (Matcha.py:315-343).
    """
    flatH = H.reshape(-1, 3, 3)
    flatg = g.reshape(-1, 3)
    out = np.zeros_like(flatg)
    for i in range(flatH.shape[0]):
        try:
            out[i] = np.linalg.solve(flatH[i], flatg[i])
        except np.linalg.LinAlgError:
            out[i] = 0.0
    return out.reshape(g.shape)


def optimize(alphas, betas, gammas, sigma, ell, num_steps, base_lr=1.0, stop_early=False,
             step_type="newton", trace=None):
    """``Matcha.optimize`` + ``_take_step`` for one level ``ell`` (FP64).

    alphas/betas/gammas: (B, C); sigma: (B, >= dsize(ell)).
    Returns refined (alpha, beta, gamma, score) each (B, C).  ``trace`` (a
    list, optional) receives one dict per step taken: the trial point, its
    score, the accept mask and the learning rate after the step.
    """
    alphas = np.array(alphas, dtype=np.float64)
    betas = np.array(betas, dtype=np.float64)
    gammas = np.array(gammas, dtype=np.float64)
    B, C = alphas.shape
    n = W.dsize(ell)
    sig = np.asarray(sigma)[:, :n]
    bank = coeff_bank(sig, ell)

    def D_of(a, b_, g):
        return W.wigner_D(a.reshape(-1), b_.reshape(-1), g.reshape(-1), ell).reshape(B, C, n)

    D = D_of(alphas, betas, gammas)
    best = np.einsum('bcn,bn->bc', D, sig).real
    best_a, best_b, best_g = alphas.copy(), betas.copy(), gammas.copy()
    lr = base_lr * np.ones((B, C, 3))
    for _ in range(num_steps):
        g, H = derivatives(D, bank, alphas)
        if step_type == "newton":
            step = lr * _solve_or_skip(H, g)
        else:
            step = -lr * g
        ta = alphas - step[..., 0]
        tb = betas - step[..., 1]
        tg = gammas - step[..., 2]
        Dt = D_of(ta, tb, tg)
        ev = np.einsum('bcn,bn->bc', Dt, sig).real
        upd = ev > best
        lr[~upd] *= 0.5
        lr[upd] *= 1.1
        if trace is not None:
            trace.append({"trial": (ta.copy(), tb.copy(), tg.copy()), "score": ev.copy(),
                          "accept": upd.copy(), "lr": lr[..., 0].copy()})
        ta = np.where(upd, ta, alphas)
        tb = np.where(upd, tb, betas)
        tg = np.where(upd, tg, gammas)
        Dt[~upd] = D[~upd]
        if stop_early and upd.sum() == 0:
            alphas, betas, gammas, D = ta, tb, tg, Dt
            break
        best = np.where(upd, ev, best)
        best_a = np.where(upd, ta, best_a)
        best_b = np.where(upd, tb, best_b)
        best_g = np.where(upd, tg, best_g)
        alphas, betas, gammas, D = ta, tb, tg, Dt
    return best_a, best_b, best_g, best


def march_schedule(Li, L_max):
    """``self.Li = sorted(Li) + [L_max]`` with the skip rule (Matcha.py:76,381-386)."""
    sched = sorted(int(x) for x in Li) + [int(L_max)]
    out = []
    for i, l in enumerate(sched):
        if l >= L_max and i != len(sched) - 1:
            continue
        out.append(l)
    return out


def march(alphas, betas, gammas, sigma, levels, num_steps=5, base_lr=1.0, stop_early=False,
          step_type="newton"):
    """``optimize`` over a level schedule, each level from the previous one's
    best points (Matcha.py:381-398, before the candidate reduction)."""
    a, b, g = alphas, betas, gammas
    s = None
    for ell in levels:
        a, b, g, s = optimize(a, b, g, sigma, ell, num_steps, base_lr=base_lr, stop_early=stop_early,
                              step_type=step_type)
    return a, b, g, s


def search_orientations(sigma, L_max, Li, candidates=10, num_steps=5, K=2, step_type="newton",
                        stop_early=False):
    """``Matcha.search_orientations`` (Matcha.py:369-403) in FP64."""
    sigma = np.atleast_2d(sigma)
    sched = march_schedule(Li, L_max)
    L0 = min(sched)
    f = sofft_grid(sigma, L0, K)
    ids, _ = topk_grid(f, candidates)
    a, b, g = ids_to_angles(ids, f.shape[1])
    s = None
    base_lr = 1.0 if step_type == "newton" else 5e-7
    for ell in sched:
        a, b, g, s = optimize(a, b, g, sigma, ell, num_steps, base_lr=base_lr, step_type=step_type,
                              stop_early=stop_early)
    idx = np.argmax(s, axis=1)
    r = np.arange(s.shape[0])
    return a[r, idx], b[r, idx], g[r, idx], s[r, idx]


def sofft_search(sigma, L_max, K):
    """``SOFFTMatcher.search_orientations`` (example.py:42-67): grid argmax."""
    f = sofft_grid(sigma, L_max, K)
    B, n = f.shape[0], f.shape[1]
    flat = f.reshape(B, -1)
    idx = np.argmax(flat, axis=1)
    bi = idx // (n * n)
    rem = idx % (n * n)
    ai = rem // n
    gi = rem % n
    ids = np.stack([bi, ai, gi], axis=1)[:, None, :]
    a, b, g = ids_to_angles(ids, n)
    return a[:, 0], b[:, 0], g[:, 0], flat[np.arange(B), idx]
