"""Shared synthetic-sigma helpers for the SO(3) tests (CPU, FP64, oracle-based)."""
import numpy as np

from oracle import wigner as W


def peaked_sigma(rng, L, alpha0, beta0, gamma0, noise=0.0, decay=0.15):
    """sigma_l = w_l conj(D^l(R0)) (+ noise): Re sum D(R) sigma = sum_l w_l chi_l(R R0^-1),
    whose global maximum is R = R0."""
    D0 = W.wigner_D([alpha0], [beta0], [gamma0], L)[0]
    ls, _, _ = W.d_layout(L)
    w = np.exp(-decay * ls)
    sig = w * np.conj(D0)
    if noise:
        sig = sig + noise * w * (rng.standard_normal(sig.shape) + 1j * rng.standard_normal(sig.shape))
    return sig


def random_sigma(rng, B, L, decay=0.1):
    ls, _, _ = W.d_layout(L)
    w = np.exp(-decay * ls)
    return w[None, :] * (rng.standard_normal((B, W.dsize(L))) + 1j * rng.standard_normal((B, W.dsize(L))))


def geodesic_deg(a1, b1, g1, a2, b2, g2):
    """Rotation angle between two Euler triples from the l=1 character (convention-free)."""
    D1 = W.wigner_D(np.atleast_1d(a1), np.atleast_1d(b1), np.atleast_1d(g1), 1)[:, 1:]
    D2 = W.wigner_D(np.atleast_1d(a2), np.atleast_1d(b2), np.atleast_1d(g2), 1)[:, 1:]
    out = []
    for x, y in zip(D1, D2):
        M1 = x.reshape(3, 3)
        M2 = y.reshape(3, 3)
        chi = np.trace(M1 @ M2.conj().T).real
        out.append(np.degrees(np.arccos(np.clip((chi - 1) / 2, -1, 1))))
    return np.array(out)
