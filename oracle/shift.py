"""FP64 CPU restatement of the reference's shift search.

TEST INFRASTRUCTURE ONLY (see ``oracle/wigner.py`` for the import rule).

Restated reference functions (swing-research/Matcha,
``src/matcha/core/ShiftMatcher.py`` unless noted):

* ``apply_soft_mask`` (:377-408), ``gridding_correct`` (:411-439),
  ``get_mask_full`` (:442-453, = ``setup_utils._get_mask_full`` :239-249),
  ``load_reference_full`` (:456-472): FP64 Fourier reference.
* ``compute_masks`` -- ``setup_utils._compute_masks`` (:252-269): outer /
  cosine-edge masks and raised-cosine weights of the background fill.
* ``rotate_complex_volume`` (:491-537): trilinear sampling (grid_sample,
  align_corners=True, zero padding) of the centred reference at rotated
  coordinates, x = last axis, centre (d+1 odd)/2, optional extra x sample.
* ``extract_shifted_patch`` (:558-616): the patch at v[a + s] (s = the
  flipped shift, array-axis order), trilinear for fractional shifts with the
  border slice zeroed (set_coords_to_zero).
* ``apply_cosine_mask`` (:540-555).
* ``compute_shift`` (:94-282) for space="complex", normalization=None,
  score_normalization=True, mode="local": half-spectrum cross power,
  ifftn with s = (N, N, N) (the half axis zero-padded), real part, first
  argmax, wrap to (-N/2, N/2]; for upsample > 1 the local matrix-DFT
  refinement around zero with fftfreq(N/2+1) on the half axis.
* ``estimate_shift_like_relion`` (:619-674) composes the above.

Parity pins: an integer circular shift of a volume against itself gives that
shift exactly (closed form); the cross-correlation equals a direct lag sum
on small grids; rotation by the identity reproduces the half spectrum;
``apply_soft_mask`` scales single Fourier modes inside, on and beyond its
cosine band by the closed-form factors; ``gridding_correct`` returns deltas
divided by sinc^2 at known radii (1, pi^2/8, pi^2/4); ``get_mask_full``
equals a brute-force count of wrapped frequencies; the upsampled refinement
peaks at the closed-form Dirichlet-kernel maximum.
"""
from __future__ import annotations

import numpy as np


def apply_soft_mask(vol, ori_size, pixel_size, ini_high, width_fmask_edge=6.0):
    if ini_high <= 0:
        return vol
    radius = ori_size * pixel_size / ini_high
    radius -= width_fmask_edge / 2.0
    radius_p = radius + width_fmask_edge
    ft = np.fft.fftshift(np.fft.fftn(vol, norm="forward"))
    zd, yd, xd = vol.shape
    zz, yy, xx = np.indices(vol.shape)
    r = np.sqrt((xx - xd // 2) ** 2 + (yy - yd // 2) ** 2 + (zz - zd // 2) ** 2)
    mask = np.ones_like(r)
    mask[r > radius_p] = 0
    tz = (r >= radius) & (r <= radius_p)
    mask[tz] = 0.5 - 0.5 * np.cos(np.pi * (radius_p - r[tz]) / width_fmask_edge)
    ft *= mask
    return np.fft.ifftn(np.fft.ifftshift(ft), norm="forward").real


def gridding_correct(vol, ori_size, padding_factor=1.0):
    z, y, x = vol.shape
    kk, ii, jj = np.meshgrid(np.arange(-z // 2, z // 2), np.arange(-y // 2, y // 2),
                             np.arange(-x // 2, x // 2), indexing="ij")
    rval = np.sqrt(kk ** 2 + ii ** 2 + jj ** 2) / (ori_size * padding_factor)
    sinc = np.ones_like(rval)
    nz = rval > 0
    sinc[nz] = np.sin(np.pi * rval[nz]) / (np.pi * rval[nz])
    corr = sinc ** 2
    corr[corr == 0] = 1.0
    return vol / corr


def get_mask_full(shape):
    nz, ny, nx = shape
    cz, cy = nz // 2, ny // 2
    Z, Y, X = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    kp = np.where(Z < cz, Z, Z - nz)
    ip = np.where(Y < cy, Y, Y - ny)
    jp = np.where(X < cy, X, X - nx)
    return (kp ** 2 + ip ** 2 + jp ** 2) <= nz // 2 * nz // 2


def load_reference_full(ref, volume_size, voxel_size):
    ref = ref.astype(np.float64)
    f = apply_soft_mask(ref, volume_size, voxel_size, 4.0, 2.0)
    f = gridding_correct(f, volume_size, 1.0)
    h = int(volume_size // 2)
    faux = np.fft.fftn(np.roll(f, (-h, -h, -h), axis=(0, 1, 2)), norm="forward")
    mask = get_mask_full(faux.shape)
    return np.fft.fftshift(faux) * np.fft.fftshift(mask)


def compute_masks(shape, radius, cosine_width=5.0):
    z, y, x = shape
    if radius < 0:
        radius = x / 2.0
    radius_p = radius + cosine_width
    zz, yy, xx = np.meshgrid(np.arange(-z // 2, z // 2), np.arange(-y // 2, y // 2),
                             np.arange(-x // 2, x // 2), indexing="ij")
    r = np.sqrt(xx ** 2 + yy ** 2 + zz ** 2)
    inner = r < radius
    outer = r > radius_p
    cos_ = (~inner) & (~outer)
    raised = 0.5 + 0.5 * np.cos(np.pi * (radius_p - r[cos_]) / cosine_width)
    return outer, cos_, raised


def trilinear(vol, cz, cy, cx):
    """grid_sample(align_corners=True, padding zeros) at index coordinates."""
    D, H, W = vol.shape[-3:]
    z0 = np.floor(cz).astype(np.int64)
    y0 = np.floor(cy).astype(np.int64)
    x0 = np.floor(cx).astype(np.int64)
    fz, fy, fx = cz - z0, cy - y0, cx - x0
    out = 0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                zi, yi, xi = z0 + dz, y0 + dy, x0 + dx
                w = (fz if dz else 1 - fz) * (fy if dy else 1 - fy) * (fx if dx else 1 - fx)
                ok = (zi >= 0) & (zi < D) & (yi >= 0) & (yi < H) & (xi >= 0) & (xi < W)
                v = np.where(ok, vol[np.clip(zi, 0, D - 1), np.clip(yi, 0, H - 1), np.clip(xi, 0, W - 1)], 0)
                out = out + w * v
    return out


def rotate_complex_volume(vol, rotmat, pad_for_rfft_slicing=True):
    """One (D, H, W) complex volume rotated by rotmat (acting on (x, y, z) = (w, h, d))."""
    d, h, w = vol.shape
    de, he, we = d + ((d + 1) % 2), h + ((h + 1) % 2), w + ((w + 1) % 2)
    if pad_for_rfft_slicing:
        w += 1
    cz, cy, cx = (de - 1) / 2, (he - 1) / 2, (we - 1) / 2
    zz, yy, xx = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    base = np.stack([xx - cx, yy - cy, zz - cz]).reshape(3, -1)
    rot = rotmat @ base + np.array([[cx], [cy], [cz]])
    # normalised by (size_e - 2) then mapped back by align_corners onto input index units
    gx = rot[0] * (2.0 / (we - 2)) - 1.0
    gy = rot[1] * (2.0 / (he - 2)) - 1.0
    gz = rot[2] * (2.0 / (de - 2)) - 1.0
    ix = (gx + 1) / 2 * (vol.shape[2] - 1)
    iy = (gy + 1) / 2 * (vol.shape[1] - 1)
    iz = (gz + 1) / 2 * (vol.shape[0] - 1)
    out = trilinear(vol.real, iz, iy, ix) + 1j * trilinear(vol.imag, iz, iy, ix)
    return out.reshape(d, h, w)


def extract_shifted_patch(vol, s_arr, set_coords_to_zero=True):
    """patch[a] = vol[a + s_arr] (trilinear, zeros outside), border slice zeroed for
    fractional components (extract_shifted_patch_batch with normalize=False)."""
    N = vol.shape[0]
    a0, a1, a2 = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    patch = trilinear(vol, a0 + s_arr[0], a1 + s_arr[1], a2 + s_arr[2])
    if set_coords_to_zero:
        s = np.asarray(s_arr, dtype=np.float64)
        ab = np.abs(s)
        fl = np.floor(ab).astype(int)
        ce = np.ceil(ab).astype(int)
        for dim in range(3):
            if ab[dim] - fl[dim] > 0:
                idx = [slice(None)] * 3
                idx[dim] = fl[dim] if np.sign(s[dim]) < 0 else -ce[dim]
                patch[tuple(idx)] = 0
    return patch


def apply_cosine_mask(vol, outer, cos_, raised):
    s = vol[outer].sum() + (raised * vol[cos_]).sum()
    wsum = outer.sum() + raised.sum()
    bg = s / wsum if wsum > 0 else 0.0
    out = vol.copy()
    out[outer] = bg
    out[cos_] = (1 - raised) * vol[cos_] + raised * bg
    return out


def compute_shift_local(ref_half, sub_half, upsample=1, window=5):
    """(shift (3,), score) of compute_shift(space='complex', normalization=None,
    score_normalization=True, mode='local') for one volume."""
    N = ref_half.shape[0]
    prod = np.conj(ref_half) * sub_half
    full = np.zeros((N, N, N), dtype=np.complex128)
    full[:, :, :prod.shape[2]] = prod
    cc = np.fft.ifftn(full, norm="forward").real
    nrm = np.linalg.norm(sub_half.ravel()) * np.linalg.norm(ref_half.ravel())
    cc = cc / nrm
    idx = int(np.argmax(cc))
    m = np.array(np.unravel_index(idx, cc.shape), dtype=np.float64)
    shape = np.array([N, N, N], dtype=np.float64)
    shift = np.where(m > np.floor(shape / 2), m - shape, m)
    if upsample == 1:
        return shift, float(cc.max())
    shift = shift * 0.0
    up = float(upsample)
    region = int(np.ceil(up * window))
    dftshift = np.floor(region / 2.0)
    off = dftshift - shift * up
    data = np.conj(prod)
    for dim in reversed(range(3)):
        n = data.shape[dim]
        rows = np.arange(region)[:, None] - off[dim]
        freqs = np.fft.fftfreq(n, d=up)[None, :]
        kern = np.exp(-2j * np.pi * rows * freqs)
        data = np.moveaxis(np.tensordot(kern, np.moveaxis(data, dim, 0), axes=(1, 0)), 0, dim)
    ref_cc = np.conj(data).real / nrm
    k = int(np.argmax(ref_cc))
    mm = np.array(np.unravel_index(k, ref_cc.shape), dtype=np.float64) - dftshift
    return shift + mm / up, float(ref_cc.max())


def quat_to_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def estimate_shift(ref_F, subtomos, shift_zyx, rotmats, *, N, box_size=None, voxel_size=1.0,
                   upsample=1, radius=2, ctfs=None):
    """``estimate_shift_like_relion`` for a batch; rotmats = the matrices the
    reference builds from the conjugated tracker (R(q)^T).  Returns (B, 3), (B,)."""
    box_size = N if box_size is None else box_size
    outer, cos_, raised = compute_masks((N, N, N), box_size / (2.0 * voxel_size), 5.0)
    mask_full = get_mask_full((N, N, N // 2 + 1))
    shifts, scores = [], []
    for b in range(subtomos.shape[0]):
        rr = rotate_complex_volume(ref_F, rotmats[b], pad_for_rfft_slicing=True)[:, :, N // 2:]
        rr = np.fft.ifftshift(rr, axes=(0, 1)) * mask_full
        if ctfs is not None:
            rr = rr * ctfs[b]
        s_arr = np.asarray(shift_zyx[b]).reshape(-1, 3)[0][::-1]
        patch = extract_shifted_patch(np.asarray(subtomos[b], dtype=np.float64), s_arr)
        patch = apply_cosine_mask(patch, outer, cos_, raised)
        sub = np.fft.rfftn(np.roll(patch, (-(N // 2),) * 3, axis=(0, 1, 2)), norm="forward")
        sh, sc = compute_shift_local(rr, sub, upsample, 2 * radius + 1)
        shifts.append(sh[::-1])
        scores.append(sc)
    return np.array(shifts), np.array(scores)
