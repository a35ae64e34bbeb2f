"""RELION STAR output of the alignment (the drop-in's output contract).

Host-side restatement of the reference's result path:

* ``update_rotation_estimate`` (``src/matcha/utils/rotation_ops.py:23-34``): the
  per-particle rotation tracker, tracker <- tracker * conj(q), q the z-y-z
  quaternion of (alpha, beta, gamma) with its x and z components exchanged;
* ``store_alignment_parameters`` (``src/matcha/utils/io_utils.py:181-225``): one
  record per particle and alternation;
* ``_particle_token`` (``io_utils.py:14-25``) and ``join_data``
  (``io_utils.py:256-368``): merge into the input particles table by token,
  ``rlnAnglePsi, rlnAngleTilt, rlnAngleRot`` = ``Rotation.from_euler("zyz",
  tracker Euler angles).inv().as_euler("ZYZ", degrees=True)``,
  ``rlnOrigin{X,Y,Z}Angst = -translation * rlnImagePixelSize``, the final
  alternation only, ``half1/2`` -> ``rlnRandomSubset``, STAR write.

The reference relies on quaternionic 1.0.17 (``from_euler_angles``,
``to_euler_angles``, ``conj``, products, ``normalized``) and starfile 0.5.13
(STAR read/write), which are not installed here; their published conventions
are restated below: q(alpha, beta, gamma) = exp(alpha z/2) exp(beta y/2)
exp(gamma z/2) as (w, x, y, z), Hamilton products, and the STAR 3.1 text format
(``data_`` blocks holding ``loop_`` tables or ``_name value`` pairs).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pandas as pd

# ---------------------------------------------------------------------------
# quaternions (w, x, y, z), Hamilton convention
# ---------------------------------------------------------------------------


def quat_from_euler_zyz(angles) -> np.ndarray:
    """exp(alpha z/2) exp(beta y/2) exp(gamma z/2) for angles (..., 3) in radians."""
    a = np.asarray(angles, dtype=np.float64)
    al, be, ga = a[..., 0], a[..., 1], a[..., 2]
    cb, sb = np.cos(be / 2), np.sin(be / 2)
    return np.stack([cb * np.cos((al + ga) / 2), -sb * np.sin((al - ga) / 2),
                     sb * np.cos((al - ga) / 2), cb * np.sin((al + ga) / 2)], axis=-1)


def quat_to_euler_zyz(q) -> np.ndarray:
    """Inverse of quat_from_euler_zyz (alpha, gamma in (-2pi, 2pi], beta in [0, pi])."""
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    s = np.arctan2(z, w)
    d = np.arctan2(-x, y)
    beta = 2.0 * np.arctan2(np.sqrt(x * x + y * y), np.sqrt(w * w + z * z))
    return np.stack([s + d, beta, s - d], axis=-1)


def quat_mul(p, q) -> np.ndarray:
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    pw, px, py, pz = (p[..., i] for i in range(4))
    qw, qx, qy, qz = (q[..., i] for i in range(4))
    return np.stack([pw * qw - px * qx - py * qy - pz * qz,
                     pw * qx + px * qw + py * qz - pz * qy,
                     pw * qy - px * qz + py * qw + pz * qx,
                     pw * qz + px * qy - py * qx + pz * qw], axis=-1)


def quat_conj(q) -> np.ndarray:
    q = np.array(q, dtype=np.float64, copy=True)
    q[..., 1:] *= -1.0
    return q


def quat_normalized(q) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def quat_to_matrix(q) -> np.ndarray:
    """Rotation matrix of unit quaternions (..., 4) -> (..., 3, 3)."""
    q = quat_normalized(q)
    w, x, y, z = (q[..., i] for i in range(4))
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=-1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=-1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=-1)], axis=-2)


def identity_tracker(n: int) -> np.ndarray:
    """The reference's tracker without prior rotations (setup_utils.get_rotation_tracker)."""
    return np.tile(np.array([[1.0, 0.0, 0.0, 0.0]]), (n, 1))


def prior_rotation_tracker(particles: pd.DataFrame) -> np.ndarray:
    """conj(q(rot, tilt, psi)) from the particles' RELION angles in degrees
    (setup_utils.get_rotation_tracker with use_prior_rotation)."""
    cols = ["rlnAngleRot", "rlnAngleTilt", "rlnAnglePsi"]
    if not all(c in particles.columns for c in cols):
        raise ValueError("Missing Euler angle columns in particles table for prior rotation.")
    return quat_conj(quat_from_euler_zyz(np.deg2rad(particles[cols].to_numpy(dtype=np.float64))))


def update_rotation_estimate(alphas, betas, gammas, rotation_tracker):
    """(tracker * conj(q~)).normalized with q~ the z-y-z quaternion of the
    angles with x and z exchanged (rotation_ops.py:23-34).  Returns
    (tracker, q~) as (n, 4) arrays."""
    ang = np.stack([np.asarray(a, dtype=np.float64).reshape(-1) for a in (alphas, betas, gammas)], axis=1)
    cur = quat_from_euler_zyz(ang)
    cur[:, [1, 3]] = cur[:, [3, 1]]
    tracker = quat_normalized(quat_mul(np.asarray(rotation_tracker, dtype=np.float64), quat_conj(cur)))
    return tracker, cur


# ---------------------------------------------------------------------------
# per-particle records and the STAR join
# ---------------------------------------------------------------------------


def particle_token(path_like) -> str:
    """Normalised particle key (io_utils._particle_token)."""
    p = str(path_like).strip()
    if "@" in p:
        p = p.split("@", 1)[1]
    p = p.replace("\\", "/")
    parts = Path(p).parts
    token = "/".join(parts[-2:]) if len(parts) >= 2 else parts[-1]
    token = token.replace("_weights.mrc", "").replace("_ctf.mrc", "")
    token = token.replace("_data.mrc", "").replace(".mrc", "")
    token = token.replace("_subtomo", "_")
    return token.replace("//", "/")


RECORD_COLUMNS = ["path", "rotation_score", "rotation", "translation", "file_name", "prior_shift",
                  "alternation_index", "shift_score", "rlnTomoParticleName", "half"]


class AlignmentRecords:
    """Per-particle results in the reference's record layout
    (store_alignment_parameters, one candidate per particle)."""

    def __init__(self):
        self.frame = pd.DataFrame(columns=RECORD_COLUMNS)

    def add(self, paths, rotation_scores, shift_scores, local_shifts, rotation_tracker, prior_shifts=None,
            half: str = "half1", alternation_index: int = 0):
        rs = np.asarray(_to_numpy(rotation_scores), dtype=np.float64).reshape(-1)
        ss = np.asarray(_to_numpy(shift_scores), dtype=np.float64).reshape(-1)
        sh = np.asarray(_to_numpy(local_shifts), dtype=np.float64).reshape(len(paths), -1, 3)[:, 0]
        tr = np.asarray(rotation_tracker, dtype=np.float64).reshape(len(paths), -1, 4)[:, 0]
        ps = (np.zeros_like(sh) if prior_shifts is None else
              np.asarray(_to_numpy(prior_shifts), dtype=np.float64).reshape(len(paths), -1, 3)[:, 0])
        eul = quat_to_euler_zyz(tr)
        rows = []
        for i, path in enumerate(paths):
            rows.append({"path": path, "rotation_score": float(rs[i]), "rotation": eul[i],
                         "translation": sh[i].astype(np.float32), "file_name": str(path).split("/")[-1],
                         "prior_shift": ps[i].astype(np.float32), "alternation_index": int(alternation_index),
                         "shift_score": float(ss[i]), "rlnTomoParticleName": particle_token(path), "half": half})
        new = pd.DataFrame(rows, columns=RECORD_COLUMNS)
        self.frame = new if self.frame.empty else pd.concat([self.frame, new], ignore_index=True)
        return self


def records_from_history(paths, history, half: str = "half1", rotation_tracker=None,
                         records: AlignmentRecords = None) -> AlignmentRecords:
    """One record per particle and alternation from an alternation history of
    (alpha, beta, gamma, norm_score, shift, shift_score[, tracker]) tuples
    (align_subtomograms.py:330-361).  With a 7th entry (the reference-mode
    harness) that composed tracker is stored; without one (global searches) the
    tracker is the prior tracker (identity by default) updated by that
    alternation's angles."""
    records = AlignmentRecords() if records is None else records
    prior = identity_tracker(len(paths)) if rotation_tracker is None else np.asarray(rotation_tracker)
    for k, entry in enumerate(history):
        a, b, g, nscore, shift, sscore = entry[:6]
        if len(entry) > 6:
            tracker = np.asarray(_to_numpy(entry[6]), dtype=np.float64).reshape(len(paths), 4)
        else:
            tracker, _ = update_rotation_estimate(_to_numpy(a), _to_numpy(b), _to_numpy(g), prior)
        ss = np.full(len(paths), np.nan) if sscore is None else _to_numpy(sscore)
        records.add(paths, nscore, ss, shift, tracker, half=half, alternation_index=k)
    return records


def _to_numpy(x):
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def euler_zyz_to_relion(rotation) -> np.ndarray:
    """(rlnAnglePsi, rlnAngleTilt, rlnAngleRot) in degrees from the tracker's
    z-y-z angles: Rotation.from_euler("zyz", angles).inv().as_euler("ZYZ")
    (io_utils.py:326-329)."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_euler("zyz", np.asarray(rotation, dtype=np.float64), degrees=False).inv().as_euler(
        "ZYZ", degrees=True)


def join_to_star(particles_star, records, out_path, num_alternations: int = 1,
                 random_half_split: bool = False) -> bool:
    """Merge alignment records into the input particles STAR and write
    ``out_path`` (io_utils.join_data, with the records in memory)."""
    raw = read_star(particles_star) if not isinstance(particles_star, dict) else particles_star
    df_base = raw.get("particles", list(raw.values())[-1]).copy()
    optics = raw.get("optics")
    general = raw.get("general")
    parts = [r.frame if isinstance(r, AlignmentRecords) else r for r in
             (records if isinstance(records, (list, tuple)) else [records])]
    parts = [p for p in parts if len(p)]
    if not parts:
        return False
    pixel_size = float(optics["rlnImagePixelSize"].iloc[0])
    df_result = pd.concat(parts, ignore_index=True)
    base_col = "rlnTomoParticleName" if "rlnTomoParticleName" in df_base.columns else "rlnImageName"
    if base_col not in df_base.columns:
        raise KeyError("Base particles table must contain 'rlnTomoParticleName' or 'rlnImageName'.")
    res_col = "rlnTomoParticleName" if "rlnTomoParticleName" in df_result.columns else "path"
    df_base["_match_token"] = df_base[base_col].astype(str).map(particle_token)
    df_result = df_result.copy()
    df_result["_match_token"] = df_result[res_col].astype(str).map(particle_token)
    df_result = df_result.drop_duplicates(subset="_match_token", keep="last")
    dfm = df_base.merge(df_result, on="_match_token", how="left", suffixes=("", "_res"), indicator=True)
    dfm["fastAlignment"] = (dfm["_merge"] == "both").astype(np.int8)
    dfm = dfm.drop(columns=["_merge"])
    if "rotation" in dfm.columns:
        mask = dfm["rotation"].notna()
        if mask.any():
            eul = euler_zyz_to_relion(np.stack(dfm.loc[mask, "rotation"].to_numpy()))
            for j, col in enumerate(["rlnAnglePsi", "rlnAngleTilt", "rlnAngleRot"]):
                if col not in dfm.columns:
                    dfm[col] = np.nan
                dfm.loc[mask, col] = eul[:, j]
    if "translation" in dfm.columns:
        mask = dfm["translation"].notna()
        if mask.any():
            t = -np.stack(dfm.loc[mask, "translation"].to_numpy())
            for j, col in enumerate(["rlnOriginXAngst", "rlnOriginYAngst", "rlnOriginZAngst"]):
                if col not in dfm.columns:
                    dfm[col] = np.nan
                dfm.loc[mask, col] = t[:, j] * pixel_size
    if "alternation_index" in dfm.columns:
        final = int(num_alternations) - 1
        dfm = dfm[dfm["alternation_index"].isna() | (dfm["alternation_index"] == final)]
    if "half" in dfm.columns:
        mapped = dfm["half"].map({"half1": 1, "half2": 2})
        if "rlnRandomSubset" in dfm.columns:
            dfm.loc[mapped.notna(), "rlnRandomSubset"] = mapped[mapped.notna()].astype(np.int32)
        elif random_half_split:
            dfm["rlnRandomSubset"] = mapped.fillna(0).astype(np.int32)
    drop = [c for c in df_result.columns if c != "rlnTomoParticleName"]
    drop += [f"{c}_res" for c in drop]
    dfm = dfm.drop(columns=drop, errors="ignore")
    dfm = dfm.drop_duplicates(subset="rlnTomoParticleName" if "rlnTomoParticleName" in dfm.columns
                              else "_match_token")
    dfm = dfm.drop(columns=["_match_token"], errors="ignore").reset_index(drop=True)
    out = {}
    if general is not None:
        out["general"] = general
    if optics is not None:
        out["optics"] = optics
    out["particles"] = dfm
    write_star(out, out_path)
    return True


# ---------------------------------------------------------------------------
# STAR files (text): data_ blocks with loop_ tables or _name value pairs
# ---------------------------------------------------------------------------


def _convert(tok: str):
    try:
        return int(tok)
    except ValueError:
        pass
    try:
        return float(tok)
    except ValueError:
        return tok


def read_star(path) -> dict:
    """{block name: DataFrame (loop blocks) or dict (name/value blocks)}."""
    blocks: dict = {}
    name = None
    cols: list = []
    rows: list = []
    pairs: dict = {}
    mode = None

    def close():
        nonlocal cols, rows, pairs, mode
        if name is None:
            return
        if mode == "loop":
            frame = pd.DataFrame(rows, columns=cols)
            for c in frame.columns:
                frame[c] = frame[c].map(_convert)
                if all(isinstance(v, (int, np.integer)) for v in frame[c]):
                    frame[c] = frame[c].astype(np.int64)
                elif all(isinstance(v, (int, float, np.integer, np.floating)) for v in frame[c]):
                    frame[c] = frame[c].astype(np.float64)
            blocks[name] = frame
        else:
            blocks[name] = dict(pairs)
        cols, rows, pairs, mode = [], [], {}, None

    for line in Path(path).read_text().splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        if s.startswith("data_"):
            close()
            name = s[5:]
            continue
        if s == "loop_":
            mode = "loop"
            continue
        if s.startswith("_"):
            if mode == "loop" and not rows:
                cols.append(s.split()[0][1:])
            else:
                mode = "pairs"
                key, _, val = s.partition(" ")
                pairs[key[1:]] = _convert(val.strip())
            continue
        if mode == "loop":
            toks = s.split()
            if len(toks) != len(cols):
                raise ValueError(f"STAR row with {len(toks)} fields, expected {len(cols)}: {line!r}")
            rows.append(toks)
    close()
    return blocks


def _fmt(v) -> str:
    if isinstance(v, (bool, np.bool_)):
        return str(int(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return "nan" if math.isnan(v) else f"{float(v):.6f}"
    return str(v)


def write_star(blocks: dict, path) -> None:
    out = ["", "# version 30001", ""]
    for name, data in blocks.items():
        out.append(f"data_{name}")
        out.append("")
        if isinstance(data, pd.DataFrame):
            out.append("loop_")
            for j, c in enumerate(data.columns):
                out.append(f"_{c} #{j + 1}")
            for row in data.itertuples(index=False):
                out.append(" ".join(_fmt(v) for v in row))
        else:
            for k, v in data.items():
                out.append(f"_{k} {_fmt(v)}")
        out.append("")
    Path(path).write_text("\n".join(out) + "\n")
