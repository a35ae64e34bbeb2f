"""Pins of the FP64 Wigner restatement (oracle/wigner.py) by closed forms and invariants."""
import numpy as np
import pytest

from oracle import wigner as W


def test_sizes_match_reference_formulas():
    # Dsize(L) = (L+1)(2L+1)(2L+3)/3 ; compute_size(n) = Dsize(n-1)
    for L in range(0, 40):
        assert W.dsize(L) == sum((2 * l + 1) ** 2 for l in range(L + 1))
        assert W.compute_size(L + 1) == W.dsize(L)
    assert [W.dsize(L) for L in (8, 12, 16, 24, 32, 48)] == [969, 2925, 6545, 20825, 47905, 156849]
    assert W.hsize(32) == 12529


def test_dindex_is_row_major_layout():
    k = 0
    for l in range(6):
        for mp in range(-l, l + 1):
            for m in range(-l, l + 1):
                assert W.dindex(l, mp, m) == k
                k += 1


@pytest.mark.parametrize("beta", [0.0, 0.37, 1.3, np.pi / 2, 2.9, np.pi, -0.6, 4.1])
def test_recurrence_matches_wigner_explicit_sum(beta):
    L = 8
    d = W.wigner_d(np.array([beta]), L)[0]
    ls, mps, ms = W.d_layout(L)
    ref = np.array([W.wigner_d_explicit(beta, int(l), int(mp), int(m)) for l, mp, m in zip(ls, mps, ms)])
    assert np.max(np.abs(d - ref)) < 1e-12


def test_known_low_order_values():
    b = 0.7
    d = W.wigner_d(np.array([b]), 1)[0]
    # d^1_{0,0} = cos b ; d^1_{1,0} = -sin b / sqrt 2 ; d^1_{1,1} = (1 + cos b)/2
    assert d[W.dindex(1, 0, 0)] == pytest.approx(np.cos(b), abs=1e-14)
    assert d[W.dindex(1, 1, 0)] == pytest.approx(-np.sin(b) / np.sqrt(2), abs=1e-14)
    assert d[W.dindex(1, 1, 1)] == pytest.approx((1 + np.cos(b)) / 2, abs=1e-14)


def test_unitarity_and_identity():
    L = 32
    rng = np.random.default_rng(1)
    a, b, g = rng.uniform(0, 2 * np.pi), rng.uniform(0, np.pi), rng.uniform(0, 2 * np.pi)
    D = W.wigner_D([a], [b], [g], L)[0]
    d0 = W.wigner_d(np.array([0.0]), L)[0]
    for l in (0, 1, 5, 17, 32):
        o = W.dsize(l - 1)
        M = D[o:o + (2 * l + 1) ** 2].reshape(2 * l + 1, 2 * l + 1)
        assert np.max(np.abs(M @ M.conj().T - np.eye(2 * l + 1))) < 1e-10
        I = d0[o:o + (2 * l + 1) ** 2].reshape(2 * l + 1, 2 * l + 1)
        assert np.max(np.abs(I - np.eye(2 * l + 1))) < 1e-13


def test_group_law_under_composition():
    # D(a,b,g) = Dz(a) d(b) Dz(g) as matrices; the composite Dz(a) d(b) Dz(g)
    # must equal the product of its factors (checks phase placement m' <-> alpha).
    L = 6
    a, b, g = 0.4, 1.1, 2.3
    D = W.wigner_D([a], [b], [g], L)[0]
    Da = W.wigner_D([a], [0.0], [0.0], L)[0]
    Db = W.wigner_D([0.0], [b], [0.0], L)[0]
    Dg = W.wigner_D([0.0], [0.0], [g], L)[0]
    for l in range(L + 1):
        o, n = W.dsize(l - 1), 2 * l + 1
        M = lambda X: X[o:o + n * n].reshape(n, n)
        assert np.max(np.abs(M(Da) @ M(Db) @ M(Dg) - M(D))) < 1e-12


def test_projection_rotation_covariance_oracle():
    # pins the oracle's projection conventions: a 90-degree grid rotation about y
    # maps a_{lk,.} to d^l(pi/2) a_{lk,.}
    from oracle import fle
    N, bl, radius = 24, 10, 11
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((N, N, N))
    a = fle.evaluate_t(np.stack([vol, fle.rotate_y90(vol)]), N, bl, radius)
    ls, ks, ms, mds, lmds, cs, ne = fle.lap_eig_ball(N, bl)
    for l in range(int(ls.max()) + 1):
        d = W.wigner_d(np.array([np.pi / 2]), l)[0][W.dsize(l - 1):W.dsize(l)].reshape(2 * l + 1, 2 * l + 1)
        for k in range(1, ks.max() + 1):
            idx = {int(m): i for i, (ll, kk, m) in enumerate(zip(ls, ks, ms)) if ll == l and kk == k}
            if len(idx) != 2 * l + 1:
                continue
            v0 = np.array([a[0, idx[m]] for m in range(-l, l + 1)])
            v1 = np.array([a[1, idx[m]] for m in range(-l, l + 1)])
            assert np.max(np.abs(d @ v0 - v1)) < 1e-12 * max(1.0, np.linalg.norm(v0))
