"""Pins of the FP64 SO(3) search restatement (oracle/so3.py)."""
import numpy as np
import pytest

from oracle import so3, wigner as W
from tests.so3_helpers import geodesic_deg, peaked_sigma, random_sigma


def test_bank_derivatives_match_finite_differences():
    rng = np.random.default_rng(3)
    L = 7
    sig = random_sigma(rng, 2, L)
    a = rng.uniform(0, 2 * np.pi, (2, 3))
    b = rng.uniform(0.2, 3.0, (2, 3))
    g = rng.uniform(0, 2 * np.pi, (2, 3))
    D = W.wigner_D(a.ravel(), b.ravel(), g.ravel(), L).reshape(2, 3, -1)
    grad, H = so3.derivatives(D, so3.coeff_bank(sig, L), a)

    def f(bi, x):
        return so3.score(np.array([x[0]]), np.array([x[1]]), np.array([x[2]]), sig[bi], L)[0]

    for bi in range(2):
        for c in range(3):
            x0 = np.array([a[bi, c], b[bi, c], g[bi, c]])
            h = 1e-5
            fd = np.array([(f(bi, x0 + h * e) - f(bi, x0 - h * e)) / (2 * h) for e in np.eye(3)])
            assert np.max(np.abs(fd - grad[bi, c])) < 1e-7 * max(1.0, np.abs(fd).max())
            h2 = 1e-4
            fh = np.zeros((3, 3))
            for i, ei in enumerate(np.eye(3)):
                for k, ek in enumerate(np.eye(3)):
                    fh[i, k] = (f(bi, x0 + h2 * ei + h2 * ek) - f(bi, x0 + h2 * ei - h2 * ek)
                                - f(bi, x0 - h2 * ei + h2 * ek) + f(bi, x0 - h2 * ei - h2 * ek)) / (4 * h2 * h2)
            assert np.max(np.abs(fh - H[bi, c])) < 1e-5 * max(1.0, np.abs(fh).max())


def test_sofft_grid_equals_direct_sum():
    rng = np.random.default_rng(4)
    L0, K = 4, 2
    sig = random_sigma(rng, 1, L0)
    f = so3.sofft_grid(sig, L0, K)
    n = 2 * L0 * K
    for (j, ai, gi) in [(0, 0, 0), (3, 5, 7), (n - 1, 2, 9), (7, 15, 1)]:
        beta = np.pi * j / (n - 1)
        direct = so3.score(np.array([2 * np.pi * ai / n]), np.array([beta]),
                           np.array([2 * np.pi * gi / n]), sig[0], L0 - 1)[0] / n ** 2
        assert f[0, j, ai, gi] == pytest.approx(direct, abs=1e-12)


def test_march_recovers_peak_and_is_monotone():
    rng = np.random.default_rng(5)
    a0, b0, g0 = 1.234, 0.987, 4.321
    sig = peaked_sigma(rng, 12, a0, b0, g0, noise=0.02)[None, :]
    f = so3.sofft_grid(sig, 4, 2)
    ids, vals = so3.topk_grid(f, 4)
    a, b, g = so3.ids_to_angles(ids, f.shape[1])
    s_prev = so3.score(a.ravel(), b.ravel(), g.ravel(), sig[0], 8)
    ra, rb, rg, rs = so3.optimize(a, b, g, sig, 8, 5)
    assert np.all(rs.ravel() >= s_prev - 1e-12)  # best score never decreases
    A, B_, G, S = so3.search_orientations(sig, 12, [4, 8], candidates=4, num_steps=5, K=2)
    assert geodesic_deg(A, B_, G, a0, b0, g0)[0] < 1.0


def test_fixed_point_at_exact_peak():
    a0, b0, g0 = 0.5, 1.2, 2.0
    sig = peaked_sigma(None, 10, a0, b0, g0)[None, :]
    ra, rb, rg, rs = so3.optimize(np.array([[a0]]), np.array([[b0]]), np.array([[g0]]), sig, 10, 3)
    assert geodesic_deg(ra[0], rb[0], rg[0], a0, b0, g0)[0] < 1e-6


def test_march_schedule_skip_rule():
    assert so3.march_schedule([8, 16, 24], 32) == [8, 16, 24, 32]
    assert so3.march_schedule([8, 16, 24, 32], 32) == [8, 16, 24, 32]
    assert so3.march_schedule([30, 40, 65], 40) == [30, 40]


# --- the march's rules, pinned by closed-form trajectories -------------------
# sigma with only the (l, m', m) = (1, 0, 0) and (2, 0, 0) entries A, B gives
# Re sum_{l <= L} D sigma = A cos(beta) (L = 1) and A cos(beta) + B P2(cos beta)
# (L >= 2), independent of alpha and gamma (d^1_00 = cos, d^2_00 = P2).  In
# gradient mode the step is beta' = beta + lr ds/dbeta (Matcha.py:14-15,
# 166-168), so the whole accept/reject sequence follows from cos and sin.

def _cos_sigma(pairs, L=2):
    sig = np.zeros((len(pairs), W.dsize(L)), dtype=complex)
    for i, (A, B) in enumerate(pairs):
        sig[i, W.dindex(1, 0, 0)] = A
        sig[i, W.dindex(2, 0, 0)] = B
    return sig


def _closed_form(pairs, betas, levels, steps, lr0, up=1.1, down=0.5, reset_lr=True, reset_best=True,
                 stop_early=False):
    """Trajectory of the rules for a batch of candidates (one per particle)."""
    def s(b, A, B, L):
        return A * np.cos(b) + (B * (3 * np.cos(b) ** 2 - 1) / 2 if L >= 2 else 0.0)

    def ds(b, A, B, L):
        return -A * np.sin(b) - (3 * B * np.cos(b) * np.sin(b) if L >= 2 else 0.0)

    n = len(pairs)
    beta = list(betas)
    lr = [lr0] * n
    best = [None] * n
    log = []
    for L in levels:
        for i, (A, B) in enumerate(pairs):
            if reset_best or best[i] is None:
                best[i] = s(beta[i], A, B, L)
            if reset_lr:
                lr[i] = lr0
        for _ in range(steps):
            step = []
            for i, (A, B) in enumerate(pairs):
                t = beta[i] + lr[i] * ds(beta[i], A, B, L)
                st = s(t, A, B, L)
                acc = st > best[i]
                if acc:
                    beta[i], best[i] = t, st
                lr[i] *= up if acc else down
                step.append((t, acc, lr[i]))
            log.append((L, step))
            if stop_early and not any(a for _, a, _ in step):
                break
    return log, beta, best


def _oracle_log(pairs, betas, levels, steps, lr0, stop_early=False):
    sig = _cos_sigma(pairs)
    n = len(pairs)
    a = np.full((n, 1), 0.3)
    b = np.array(betas, dtype=float).reshape(n, 1)
    g = np.full((n, 1), 1.7)
    log = []
    for L in levels:
        tr = []
        a, b, g, s = so3.optimize(a, b, g, sig, L, steps, base_lr=lr0, step_type="gradient",
                                  stop_early=stop_early, trace=tr)
        log += [(L, [(t["trial"][1][i, 0], bool(t["accept"][i, 0]), t["lr"][i, 0]) for i in range(n)])
                for t in tr]
        assert np.all(a == 0.3) and np.all(g == 1.7)  # no alpha / gamma dependence
    return log, b[:, 0], s[:, 0]


def _same_log(x, y):
    return len(x) == len(y) and all(
        lx == ly and all(ax == ay and abs(tx - ty) < 1e-12 and abs(rx - ry) < 1e-12
                         for (tx, ax, rx), (ty, ay, ry) in zip(sx, sy))
        for (lx, sx), (ly, sy) in zip(x, y))


def test_march_rules_follow_closed_form_trajectory():
    pairs, betas, levels, steps, lr0 = [(1.0, 0.3)], [1.0], [1, 2], 3, 3.0
    want, wb, ws = _closed_form(pairs, betas, levels, steps, lr0)
    got, gb, gs = _oracle_log(pairs, betas, levels, steps, lr0)
    assert _same_log(got, want), (got, want)
    assert abs(gb[0] - wb[0]) < 1e-12 and abs(gs[0] - ws[0]) < 1e-12
    # the case exercises both branches on both levels ...
    acc = [[a for _, a, _ in st][0] for _, st in want]
    assert acc[:3].count(True) and acc[:3].count(False) and acc[3:].count(True) and acc[3:].count(False)
    # ... and a swapped lr factor, a carried-over lr or best score, or a
    # non-strict accept would each change it
    for variant in (dict(up=0.5, down=1.1), dict(reset_lr=False), dict(reset_best=False),
                    dict(up=1.1, down=0.25)):
        assert not _same_log(_closed_form(pairs, betas, levels, steps, lr0, **variant)[0], want), variant


def test_march_accepts_only_strict_improvements():
    # at the exact maximum beta = 0 every trial point equals the start: s' == best
    # is rejected (lr halves each step), as update_mask = evaled > curr_max_val
    log, b, s = _oracle_log([(1.0, 0.0)], [0.0], [1], 3, 1.0)
    assert [st[0][1] for _, st in log] == [False] * 3
    assert [st[0][2] for _, st in log] == [0.5, 0.25, 0.125]


def test_stop_early_is_batch_global():
    levels, steps, lr0 = [1, 2], 3, 3.0
    # both particles reject their first step: each level stops after one step
    pairs, betas = [(1.0, 0.3), (1.0, 0.3)], [1.0, 1.4]
    want, wb, ws = _closed_form(pairs, betas, levels, steps, lr0, stop_early=True)
    got, gb, gs = _oracle_log(pairs, betas, levels, steps, lr0, stop_early=True)
    assert _same_log(got, want)
    assert [L for L, _ in want] == [1, 2] and np.allclose(gb, betas)
    # one particle accepting keeps the whole batch marching
    pairs2, betas2 = [(1.0, 0.3), (0.5, 0.2)], [1.0, 2.2]
    want2, _, _ = _closed_form(pairs2, betas2, levels, steps, lr0, stop_early=True)
    got2, gb2, _ = _oracle_log(pairs2, betas2, levels, steps, lr0, stop_early=True)
    assert _same_log(got2, want2)
    assert len(want2) > 2
    free, fb, _ = _closed_form(pairs2, betas2, levels, steps, lr0)
    assert len(free) > len(want2)  # and stopping did cut steps


def test_newton_step_pinned_by_finite_differences():
    # the first trial point is theta - lr H^-1 g (Matcha.py:11-12,310-311,
    # 163-168) with g, H of the score itself (central differences, FP64)
    rng = np.random.default_rng(12)
    L = 3
    sig = random_sigma(rng, 1, L)
    x0 = np.array([0.7, 1.3, 2.1])
    lr0 = 0.8
    tr = []
    so3.optimize(x0[0].reshape(1, 1), x0[1].reshape(1, 1), x0[2].reshape(1, 1), sig, L, 1, base_lr=lr0,
                 trace=tr)

    def f(x):
        return so3.score(np.array([x[0]]), np.array([x[1]]), np.array([x[2]]), sig[0], L)[0]

    h = 1e-4
    E = np.eye(3)
    g = np.array([(f(x0 + h * e) - f(x0 - h * e)) / (2 * h) for e in E])
    H = np.array([[(f(x0 + h * ei + h * ek) - f(x0 + h * ei - h * ek) - f(x0 - h * ei + h * ek)
                    + f(x0 - h * ei - h * ek)) / (4 * h * h) for ek in E] for ei in E])
    want = x0 - lr0 * np.linalg.solve(H, g)
    got = np.array([tr[0]["trial"][k][0, 0] for k in range(3)])
    assert np.max(np.abs(got - want)) < 1e-5
