"""FP64 CPU restatement of the reference's alternation resampler.

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule).

``BatchProcessor._extract_rotated_batch`` (swing-research/Matcha,
``src/matcha/align_subtomograms.py:240-252``) builds each alternation's
particles in three steps:

* ``extract_patch`` -- ``io_utils.extract_subtomogram_patch_batch``
  (``src/matcha/utils/io_utils.py:94-178``, normalize=True,
  set_coords_to_zero=False): patch[z, y, x] = V at (x + s0, y + s1, z + s2)
  by ``grid_sample`` (trilinear, align_corners=True, zero padding), the
  patch centred on the volume's ``shape // 2``;
* ``normalise_masked`` -- ``volume_ops.normalise`` (``src/matcha/utils/
  volume_ops.py:17-35``) times the mask: (P - mean) / std * mask with the
  masked mean and sqrt(E[P^2] - mean^2); the mask is ``get_spherical_mask``
  (``volume_ops.py:6-14``, sigma 0): distance to ``shape // 2`` < radius
  (``setup_utils.setup_mask`` :272-280);
* ``rotate_volume`` -- ``volume_rotation.rotate_volumes_in_batches``
  (``src/matcha/utils/volume_rotation.py:52-125``): zero-pad even axes by one,
  coordinates ``linspace(-1, 1)`` with x = last axis, rotated by M acting on
  (x, y, z), ``grid_sample`` trilinear with align_corners=True, crop.  In index
  units: out[i] = P at c + M (i - c), c = (padded size - 1) / 2 = N // 2.
  ``rotate_volumes`` (:9-49) takes M = R(tracker) via the tracker's z-y-z
  Euler angles and scipy's intrinsic "ZYZ".

Parity pins (``tests/test_resample_host.py``): extraction against scipy's
``map_coordinates(order=1, mode="grid-constant")``; a 90-degree rotation about
z as an exact voxel permutation (fixes the centre, the axis order and the
direction); exactness of trilinear sampling on affine volumes under arbitrary
rotations; the masked mean 0 / standard deviation 1 of the normalised patch
and its invariance to affine intensity changes.
"""
from __future__ import annotations

import numpy as np


def spherical_mask(N: int, radius: float) -> np.ndarray:
    """Binary sphere: distance to N // 2 (per axis) below ``radius``."""
    c = N // 2
    z, y, x = np.ogrid[:N, :N, :N]
    return (np.sqrt((z - c) ** 2 + (y - c) ** 2 + (x - c) ** 2) < radius).astype(np.float64)


def trilinear_xyz(vol: np.ndarray, px, py, pz):
    """grid_sample(align_corners=True, zeros) of a (D, H, W) volume at index
    coordinates (x along W, y along H, z along D)."""
    D, H, W = vol.shape
    x0, y0, z0 = np.floor(px).astype(np.int64), np.floor(py).astype(np.int64), np.floor(pz).astype(np.int64)
    fx, fy, fz = px - x0, py - y0, pz - z0
    out = np.zeros(np.broadcast(px, py, pz).shape)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                xi, yi, zi = x0 + dx, y0 + dy, z0 + dz
                w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                ok = (xi >= 0) & (xi < W) & (yi >= 0) & (yi < H) & (zi >= 0) & (zi < D)
                out += w * np.where(ok, vol[np.clip(zi, 0, D - 1), np.clip(yi, 0, H - 1), np.clip(xi, 0, W - 1)], 0.0)
    return out


def extract_patch(vol: np.ndarray, shift, N: int | None = None) -> np.ndarray:
    """N^3 patch of ``vol`` centred on shape // 2, displaced by shift = (sx, sy, sz)."""
    N = vol.shape[0] if N is None else N
    cz, cy, cx = (s // 2 for s in vol.shape)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    s = np.asarray(shift, dtype=np.float64)
    return trilinear_xyz(vol, x + cx - N // 2 + s[0], y + cy - N // 2 + s[1], z + cz - N // 2 + s[2])


def normalise_masked(patch: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """(patch - mean) / std * mask with masked mean and population std."""
    n = mask.sum()
    if n == 0:
        return patch * mask
    mean = (patch * mask).sum() / n
    std = np.sqrt((patch * patch * mask).sum() / n - mean * mean)
    return (patch - mean) / std * mask


def rotate_volume(vol: np.ndarray, M: np.ndarray) -> np.ndarray:
    """out[i] = vol at c + M (i - c) in (x, y, z) index units, c = N // 2."""
    N = vol.shape[0]
    c = N // 2
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    d = np.stack([x - c, y - c, z - c]).reshape(3, -1).astype(np.float64)
    p = np.asarray(M, dtype=np.float64) @ d + c
    return trilinear_xyz(vol, p[0], p[1], p[2]).reshape(N, N, N)


def extract_rotated(vol: np.ndarray, shift, M: np.ndarray, radius: float) -> np.ndarray:
    """One particle of ``_extract_rotated_batch``."""
    N = vol.shape[0]
    return rotate_volume(normalise_masked(extract_patch(vol, shift), spherical_mask(N, radius)), M)
