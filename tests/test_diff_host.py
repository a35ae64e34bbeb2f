"""Closed forms behind the score adjoint, pinned in FP64 on the CPU: Rodrigues
against scipy, the right Jacobian of exp and the body-frame generators of the
reference's Wigner D against central differences, and the chain rule of the
whole score (rotation, shift, template) against differences of the oracle."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import diff as DF
from oracle import fle, so3, resample
from oracle import wigner as W

RNG = np.random.default_rng(21)


def test_rodrigues_matches_scipy():
    for w in [np.zeros(3), np.array([1e-9, 0, 0]), RNG.standard_normal(3), 3.0 * RNG.standard_normal(3)]:
        np.testing.assert_allclose(DF.rotvec_to_matrix(w), Rotation.from_rotvec(w).as_matrix(), atol=1e-13)


@pytest.mark.parametrize("scale", [1e-7, 0.3, 2.5])
def test_right_jacobian_against_differences(scale):
    w = scale * RNG.standard_normal(3)
    J = DF.right_jacobian(w)
    h = 1e-6
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        d = Rotation.from_matrix(DF.rotvec_to_matrix(w).T @ DF.rotvec_to_matrix(w + e)).as_rotvec()
        d2 = Rotation.from_matrix(DF.rotvec_to_matrix(w).T @ DF.rotvec_to_matrix(w - e)).as_rotvec()
        np.testing.assert_allclose((d - d2) / (2 * h), J[:, i], atol=1e-8)


def _random_sigma(L):
    return RNG.standard_normal(W.dsize(L)) + 1j * RNG.standard_normal(W.dsize(L))


def _s_of_R(R, sigma, L):
    a, b, g = DF.euler_zyz(R)
    return so3.score(a, b, g, sigma, L)[0]


@pytest.mark.parametrize("beta_kind", ["generic", "near_zero", "near_pi"])
def test_body_generators_against_differences(beta_kind):
    L = 5
    sigma = _random_sigma(L)
    if beta_kind == "generic":
        R = Rotation.random(random_state=2).as_matrix()
    else:
        b = 1e-4 if beta_kind == "near_zero" else np.pi - 1e-4
        R = Rotation.from_euler("ZYZ", [0.7, b, -1.9]).as_matrix()
    sp = DF.body_sigmas(sigma, L)
    a, b, g = DF.euler_zyz(R)
    h = 1e-6
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        fd = (_s_of_R(R @ DF.rotvec_to_matrix(e), sigma, L) - _s_of_R(R @ DF.rotvec_to_matrix(-e), sigma, L)) / (2 * h)
        an = so3.score(a, b, g, sp[i], L)[0]
        assert abs(an - fd) < 1e-7 * (1 + abs(fd)), (i, an, fd)
    # the z generator is the gamma derivative of the Euler score
    np.testing.assert_allclose(so3.score(a, b, g, sp[2], L)[0],
                               (so3.score(a, b, g + h, sigma, L) - so3.score(a, b, g - h, sigma, L))[0] / (2 * h),
                               atol=1e-6)


def _trilinear_derivatives(vol, t):
    """d/dt_i of the patch V[a + flip(t)] (trilinear), exact inside each cell."""
    N = vol.shape[0]
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    px, py, pz = x + t[0], y + t[1], z + t[2]
    x0, y0, z0 = np.floor(px).astype(int), np.floor(py).astype(int), np.floor(pz).astype(int)
    fx, fy, fz = px - x0, py - y0, pz - z0
    out = np.zeros((3,) + vol.shape)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                xi, yi, zi = x0 + dx, y0 + dy, z0 + dz
                ok = (xi >= 0) & (xi < N) & (yi >= 0) & (yi < N) & (zi >= 0) & (zi < N)
                v = np.where(ok, vol[np.clip(zi, 0, N - 1), np.clip(yi, 0, N - 1), np.clip(xi, 0, N - 1)], 0.0)
                wx, wy, wz = (fx if dx else 1 - fx), (fy if dy else 1 - fy), (fz if dz else 1 - fz)
                sx, sy, sz = (1 if dx else -1), (1 if dy else -1), (1 if dz else -1)
                out[0] += sx * wy * wz * v
                out[1] += wx * sy * wz * v
                out[2] += wx * wy * sz * v
    return out


def test_score_chain_rule_against_fp64_differences():
    N, bl, radius, L = 16, 6, 7.0, 4
    vol = fle.blob_volume(RNG, N, 6, sigma_range=(1.0, 2.0))
    tmpl = fle.blob_volume(RNG, N, 6, sigma_range=(1.0, 2.0))
    pr = DF.Problem(vol, np.zeros(1), N, bl, radius, L)
    pr.aT = pr.coeffs(tmpl)
    omega = np.array([0.4, -1.1, 0.7])
    t = np.array([0.3, -0.45, 0.2])
    fd = pr.fd_gradient(omega, t)
    # analytic: rotation through the body-frame generators and J_r, shift through the
    # FLE of the trilinear derivative volumes (sigma is linear in aV)
    R = DF.rotvec_to_matrix(omega)
    a, b, g = DF.euler_zyz(R)
    sig = pr.sigma(t)
    gb = np.array([so3.score(a, b, g, s, L)[0] for s in DF.body_sigmas(sig, L)])
    grot = DF.right_jacobian(omega).T @ gb
    dV = _trilinear_derivatives(pr.vol, t)
    gsh = []
    for i in range(3):
        aVd = pr.coeffs(dV[i])
        sd = fle.compute_sigma(pr.aT, aVd[None], pr.ls, pr.mds, pr.lmax)[0]
        gsh.append(so3.score(a, b, g, sd, L)[0])
    an = np.concatenate([grot, gsh])
    np.testing.assert_allclose(an, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
    # template: s is linear in aT, so the directional change is exact
    delta = RNG.standard_normal(pr.ne) + 1j * RNG.standard_normal(pr.ne)
    ds = pr.template_directional(omega, t, delta)
    # gradient (d/dRe + i d/dIm): sum_u conj(D_uv) (-1)^{u+v} aV_{l k, -u} at index (l, k, -v)
    D = W.wigner_D(a, b, g, L)[0]
    aV = pr.coeffs(resample.extract_patch(pr.vol, t))
    idl = fle.idlm_lists(pr.ls, pr.mds, pr.lmax)
    grad = np.zeros(pr.ne, dtype=np.complex128)
    for l in range(L + 1):
        for v in range(-l, l + 1):
            it = idl[l][fle.sht_index(-v)]
            for u in range(-l, l + 1):
                iv = idl[l][fle.sht_index(-u)]
                n = min(len(iv), len(it))
                grad[it[:n]] += np.conj(D[W.dindex(l, u, v)]) * (-1) ** (u + v) * aV[iv[:n]]
    assert abs(np.real(np.vdot(grad, delta)) - ds) < 1e-9 * (1 + abs(ds))
