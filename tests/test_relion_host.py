"""RELION STAR output: rotation-tracker conventions, STAR text round trip and
the join into the particles table (io_utils.join_data, rotation_ops.py:23-34).

The quaternion conventions are pinned against scipy's rotation matrices (the
reference's z-y-z quaternion is R = Rz(alpha) Ry(beta) Rz(gamma), scipy's
intrinsic "ZYZ"); the whole search -> STAR angle chain is pinned by the closed
form Rz(psi) Ry(tilt) Rz(rot) = Rx(pi) P R(alpha, beta, gamma) P Rx(pi), P the
x<->z axis exchange, which follows from the tracker's component swap, the
conjugate and the reference's extrinsic-"zyz" -> inverse -> intrinsic-"ZYZ"
conversion, and fails for a dropped swap, conjugate or inverse or for
exchanged psi/rot columns.
"""
import numpy as np
import pandas as pd
import pytest
from scipy.spatial.transform import Rotation

from matcha_amd import relion

RNG = np.random.default_rng(7)
P = np.array([[0.0, 0, 1], [0, 1, 0], [1, 0, 0]])
RX_PI = np.diag([1.0, -1.0, -1.0])


def random_angles(n):
    return np.stack([RNG.uniform(-np.pi, np.pi, n), RNG.uniform(0.05, np.pi - 0.05, n),
                     RNG.uniform(-np.pi, np.pi, n)], axis=1)


def test_quaternion_matches_scipy_intrinsic_zyz():
    ang = random_angles(64)
    q = relion.quat_from_euler_zyz(ang)
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(relion.quat_to_matrix(q), Rotation.from_euler("ZYZ", ang).as_matrix(),
                               atol=1e-12)


def test_quaternion_product_is_matrix_product_and_conj_is_transpose():
    p = relion.quat_from_euler_zyz(random_angles(32))
    q = relion.quat_from_euler_zyz(random_angles(32))
    Rp, Rq = relion.quat_to_matrix(p), relion.quat_to_matrix(q)
    np.testing.assert_allclose(relion.quat_to_matrix(relion.quat_mul(p, q)), Rp @ Rq, atol=1e-12)
    np.testing.assert_allclose(relion.quat_to_matrix(relion.quat_conj(p)), np.swapaxes(Rp, 1, 2), atol=1e-12)


def test_euler_round_trip_through_rotation():
    ang = random_angles(64)
    back = relion.quat_to_euler_zyz(relion.quat_from_euler_zyz(ang))
    assert np.all(back[:, 1] >= 0) and np.all(back[:, 1] <= np.pi)
    np.testing.assert_allclose(Rotation.from_euler("ZYZ", back).as_matrix(),
                               Rotation.from_euler("ZYZ", ang).as_matrix(), atol=1e-12)
    # special cases: identity and beta = 0 (gimbal lock keeps the total z rotation)
    np.testing.assert_allclose(relion.quat_to_euler_zyz([1.0, 0, 0, 0]), [0, 0, 0], atol=1e-15)
    e = relion.quat_to_euler_zyz(relion.quat_from_euler_zyz([0.3, 0.0, 0.4]))
    assert abs(e[1]) < 1e-12 and abs(e[0] + e[2] - 0.7) < 1e-12


def test_tracker_update_is_axis_swapped_search_rotation():
    ang = random_angles(48)
    tr, cur = relion.update_rotation_estimate(ang[:, 0], ang[:, 1], ang[:, 2], relion.identity_tracker(48))
    want = P @ Rotation.from_euler("ZYZ", ang).as_matrix() @ P
    np.testing.assert_allclose(relion.quat_to_matrix(tr), want, atol=1e-12)
    # composition: a second update right-multiplies by the new swapped rotation
    ang2 = random_angles(48)
    tr2, _ = relion.update_rotation_estimate(ang2[:, 0], ang2[:, 1], ang2[:, 2], tr)
    want2 = want @ (P @ Rotation.from_euler("ZYZ", ang2).as_matrix() @ P)
    np.testing.assert_allclose(relion.quat_to_matrix(tr2), want2, atol=1e-12)


def test_prior_tracker_is_inverse_relion_rotation():
    df = pd.DataFrame({"rlnAngleRot": [10.0, -70.0], "rlnAngleTilt": [30.0, 120.0], "rlnAnglePsi": [5.0, 44.0]})
    tr = relion.prior_rotation_tracker(df)
    want = Rotation.from_euler("ZYZ", df[["rlnAngleRot", "rlnAngleTilt", "rlnAnglePsi"]].to_numpy(),
                               degrees=True).as_matrix()
    np.testing.assert_allclose(relion.quat_to_matrix(tr), np.swapaxes(want, 1, 2), atol=1e-12)
    with pytest.raises(ValueError):
        relion.prior_rotation_tracker(df.drop(columns=["rlnAnglePsi"]))


@pytest.mark.parametrize("path,token", [
    ("3@/data/tomo1/particle_5_data.mrc", "tomo1/particle_5"),
    ("tomo1/particle_5_subtomo.mrc", "tomo1/particle_5_"),
    ("C:\\x\\TS_01\\p7_weights.mrc", "TS_01/p7"),
    ("p9_ctf.mrc", "p9"),
    ("  TS_02/p3  ", "TS_02/p3"),
])
def test_particle_token(path, token):
    assert relion.particle_token(path) == token


def make_particles_star(tmp_path, names):
    n = len(names)
    blocks = {
        "general": {"rlnTomoSubTomosAre2DStacks": 0},
        "optics": pd.DataFrame({"rlnOpticsGroup": [1], "rlnOpticsGroupName": ["opticsGroup1"],
                                "rlnImagePixelSize": [2.5], "rlnVoltage": [300.0]}),
        "particles": pd.DataFrame({"rlnTomoName": ["TS_01"] * n, "rlnTomoParticleName": names,
                                   "rlnCoordinateX": np.arange(n, dtype=np.float64) * 10.0,
                                   "rlnAngleRot": np.zeros(n), "rlnAngleTilt": np.zeros(n),
                                   "rlnAnglePsi": np.zeros(n), "rlnOpticsGroup": np.ones(n, dtype=np.int64)}),
    }
    path = tmp_path / "particles.star"
    relion.write_star(blocks, path)
    return path, blocks


def test_star_round_trip(tmp_path):
    path, blocks = make_particles_star(tmp_path, ["TS_01/1", "TS_01/2", "TS_01/3"])
    back = relion.read_star(path)
    assert list(back) == ["general", "optics", "particles"]
    assert back["general"] == {"rlnTomoSubTomosAre2DStacks": 0}
    pd.testing.assert_frame_equal(back["optics"], blocks["optics"], check_dtype=False)
    pd.testing.assert_frame_equal(back["particles"], blocks["particles"], check_dtype=False)
    assert back["particles"]["rlnOpticsGroup"].dtype == np.int64


def test_join_angles_origins_and_flags(tmp_path):
    names = [f"TS_01/{i}" for i in range(6)]
    star, _ = make_particles_star(tmp_path, names)
    paths = [f"/data/TS_01/{i}_data.mrc" for i in range(4)]          # particles 4, 5 unaligned
    ang = random_angles(4)
    shift = RNG.uniform(-3, 3, (4, 3)).astype(np.float32)
    hist = [(ang[:, 0] + 0.2, ang[:, 1], ang[:, 2], np.full(4, 0.1), shift * 0.5, np.full(4, 0.2)),
            (ang[:, 0], ang[:, 1], ang[:, 2], np.linspace(0.5, 0.8, 4), shift, np.full(4, 0.9))]
    rec = relion.records_from_history(paths[:2], [tuple(v[:2] for v in h) for h in hist], half="half1")
    rec2 = relion.records_from_history(paths[2:], [tuple(v[2:] for v in h) for h in hist], half="half2")
    out = tmp_path / "aligned.star"
    assert relion.join_to_star(star, [rec, rec2], out, num_alternations=2, random_half_split=True)
    got = relion.read_star(out)
    assert list(got) == ["general", "optics", "particles"]
    df = got["particles"]
    assert len(df) == 6
    assert list(df["fastAlignment"]) == [1, 1, 1, 1, 0, 0]
    assert list(df["rlnRandomSubset"]) == [1, 1, 2, 2, 0, 0]
    for c in relion.RECORD_COLUMNS:
        if c != "rlnTomoParticleName":
            assert c not in df.columns
    # closed form of the whole search -> STAR chain
    R_star = Rotation.from_euler("ZYZ", df[["rlnAnglePsi", "rlnAngleTilt", "rlnAngleRot"]].to_numpy()[:4],
                                 degrees=True).as_matrix()
    want = RX_PI @ P @ Rotation.from_euler("ZYZ", ang).as_matrix() @ P @ RX_PI
    np.testing.assert_allclose(R_star, want, atol=2e-7)
    # origins: -translation * pixel size (final alternation's accumulated shift)
    org = df[["rlnOriginXAngst", "rlnOriginYAngst", "rlnOriginZAngst"]].to_numpy()
    np.testing.assert_allclose(org[:4], -shift.astype(np.float64) * 2.5, atol=1e-6)
    assert np.all(np.isnan(org[4:]))
    # untouched particles keep their input angles
    np.testing.assert_array_equal(df["rlnAngleRot"].to_numpy()[4:], 0.0)


def test_join_keeps_only_final_alternation(tmp_path):
    names = ["TS_01/0", "TS_01/1"]
    star, _ = make_particles_star(tmp_path, names)
    a = np.array([0.1, 0.2])
    hist = [(a, a, a, a, np.zeros((2, 3)), a)]
    rec = relion.records_from_history(["TS_01/0", "TS_01/1"], hist)
    out = tmp_path / "o.star"
    # only alternation 0 stored while two were configured: the matched rows drop out
    assert relion.join_to_star(star, rec, out, num_alternations=2)
    assert len(relion.read_star(out)["particles"]) == 0
    assert relion.join_to_star(star, rec, out, num_alternations=1)
    df = relion.read_star(out)["particles"]
    assert len(df) == 2 and "rlnRandomSubset" not in df.columns
    assert not relion.join_to_star(star, [], out)
