"""FFT shift match on the MI355X library.

Drop-in for ``src/matcha/core/ShiftMatcher.py`` (class ``ShiftMatcher``:
``set_reference``, ``get_base_shifts``, ``search_shifts``).  The Fourier
reference is prepared on the host in FP64 exactly as the reference does
(``load_reference_full``: 4 A cosine low-pass, sinc^2 gridding correction,
roll, FFT (norm "forward"), sphere mask, fftshift; ``:377-472``) and held on
the device as complex64; the per-batch search (rotation of the Fourier
reference, shifted-patch extraction, cosine-masked background, rFFT, cross
power, inverse FFT, argmax) runs in ``shift.hip``.

``shift_upsample_factor`` 1 gives integer lags over the whole box; a factor
up > 1 gives the reference's local sub-voxel refinement (``compute_shift``,
mode "local": the coarse lag is zeroed and a matrix DFT is evaluated on
ceil(up * (2 * shift_search_radius + 1))^3 points around zero, 1/up apart).

``shift_precision`` (an extension, no reference counterpart): "fp32"
(default) keeps the rotated reference and the work spectra in complex FP32,
the reference's complex64; "reduced" samples the reference from FP16 values
scaled by a power of two (FP16 trilinear weights, FP32 accumulation) and
stores the integer-lag path's work spectrum in BF16 between passes, all
arithmetic in FP32 -- for throughput: lags agree where the correlation has a
clear peak, scores to about 2e-3.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from ..config import cfg_get


def _soft_lowpass(vol, ori_size, pixel_size, ini_high, width):
    radius = ori_size * pixel_size / ini_high - width / 2.0
    radius_p = radius + width
    ft = np.fft.fftshift(np.fft.fftn(vol, norm="forward"))
    idx = np.indices(vol.shape)
    c = [s // 2 for s in vol.shape]
    r = np.sqrt(sum((idx[d] - c[d]) ** 2 for d in range(3)))
    m = np.where(r > radius_p, 0.0, 1.0)
    edge = (r >= radius) & (r <= radius_p)
    m[edge] = 0.5 - 0.5 * np.cos(np.pi * (radius_p - r[edge]) / width)
    return np.fft.ifftn(np.fft.ifftshift(ft * m), norm="forward").real


def fourier_reference(template: np.ndarray, volume_size: int, voxel_size: float) -> np.ndarray:
    """Centred complex half-cube-free Fourier reference (N, N, N) complex128."""
    vol = _soft_lowpass(np.asarray(template, dtype=np.float64), volume_size, voxel_size, 4.0, 2.0)
    n = vol.shape
    grids = np.meshgrid(*[np.arange(-s // 2, s // 2) for s in n], indexing="ij")
    rv = np.sqrt(sum(g ** 2 for g in grids)) / (volume_size * 1.0)
    sinc = np.ones_like(rv)
    nz = rv > 0
    sinc[nz] = np.sin(np.pi * rv[nz]) / (np.pi * rv[nz])
    corr = sinc ** 2
    corr[corr == 0] = 1.0
    vol = vol / corr
    h = int(volume_size // 2)
    F = np.fft.fftn(np.roll(vol, (-h, -h, -h), axis=(0, 1, 2)), norm="forward")
    nz_, ny_, nx_ = F.shape
    Z, Y, X = np.meshgrid(np.arange(nz_), np.arange(ny_), np.arange(nx_), indexing="ij")
    kp = np.where(Z < nz_ // 2, Z, Z - nz_)
    ip = np.where(Y < ny_ // 2, Y, Y - ny_)
    jp = np.where(X < ny_ // 2, X, X - nx_)
    mask = (kp ** 2 + ip ** 2 + jp ** 2) <= nz_ // 2 * nz_ // 2
    return np.fft.fftshift(F) * np.fft.fftshift(mask)


def tracker_to_matrices(rotation_tracker) -> np.ndarray:
    """Matrices the reference builds from conj(tracker) via ZYZ Euler angles
    (``search_shifts`` :913 + ``rotate_complex_volume`` :503-504), i.e. R(q)^T
    for unit quaternions q = (w, x, y, z); (B, 3, 3) matrices pass through."""
    arr = getattr(rotation_tracker, "ndarray", rotation_tracker)
    if isinstance(arr, torch.Tensor):
        arr = arr.detach().cpu().numpy()
    arr = np.asarray(arr, dtype=np.float64)
    if arr.ndim == 3 and arr.shape[1:] == (3, 3):
        return arr
    q = arr.reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)
    return np.transpose(R, (0, 2, 1))


class ShiftMatcher:
    def __init__(self, config, device: torch.device, dtype: torch.dtype = torch.complex64,
                 micro_batch_size: int = None):
        self.config = config
        self.device = torch.device(device)
        self.dtype = dtype
        self.N = int(cfg_get(config, "N"))
        nb = cfg_get(config, "num_subtomograms_per_batch", 1)
        try:
            nb = int(nb)
        except (TypeError, ValueError):
            nb = 1
        self.max_batch = max(1, nb)
        self.micro_batch_size = int(micro_batch_size) if micro_batch_size is not None else max(1, nb // 2)
        self.upsample = int(cfg_get(config, "shift_upsample_factor", 16))
        self.radius = int(cfg_get(config, "shift_search_radius", 2))
        self.do_ctf = bool(cfg_get(config, "do_ctf_correction", True))
        self.box_size = float(cfg_get(config, "box_size", self.N))
        self.voxel_size = float(cfg_get(config, "voxel_size", 1.0))
        self.precision = str(cfg_get(config, "shift_precision", "fp32"))
        if self.precision not in ("fp32", "reduced"):
            raise ValueError(f"shift_precision must be 'fp32' or 'reduced', got {self.precision!r}")
        self.reference_data_fourier = None
        self._plan = None

    def set_reference(self, template_data: np.ndarray):
        ref = fourier_reference(template_data, self.N, self.voxel_size).astype(np.complex64)
        self.reference_data_fourier = torch.from_numpy(ref).unsqueeze(0).to(self.device)
        if self._plan is not None:
            _lib.lib().mch_shift_plan_destroy(self._plan)
            self._plan = None
        h = ctypes.c_void_p()
        ref = np.ascontiguousarray(ref)
        with torch.cuda.device(self.device):
            rc = _lib.lib().mch_shift_plan_create(self.N, ref.ctypes.data_as(ctypes.c_void_p),
                                                  float(self.box_size / (2.0 * self.voxel_size)),
                                                  self.max_batch, ctypes.byref(h))
        _lib.check(rc, "mch_shift_plan_create")
        self._plan = h
        rc = _lib.lib().mch_shift_plan_set_upsample(self._plan, self.upsample, self.radius)
        _lib.check(rc, "mch_shift_plan_set_upsample")
        self.set_precision(self.precision)

    def set_precision(self, mode: str):
        """Storage precision of the searches: "fp32" (the reference's complex64
        values) or "reduced" (FP16 reference table, BF16 work spectrum, FP32
        arithmetic)."""
        if mode not in ("fp32", "reduced"):
            raise ValueError(f"shift precision must be 'fp32' or 'reduced', got {mode!r}")
        self.precision = mode
        if self._plan is not None:
            rc = _lib.lib().mch_shift_set_precision(self._plan, 1 if mode == "reduced" else 0)
            _lib.check(rc, "mch_shift_set_precision")

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and h.value:
            try:
                _lib.lib().mch_shift_plan_destroy(h)
            except Exception:
                pass

    def get_base_shifts(self, num_subtomograms_per_batch: int, num_base_shifts: int, dtype_real):
        execution = cfg_get(self.config, "execution", None)
        grid = cfg_get(execution, "translational_search_grid", None) if execution is not None else None
        if grid is None:
            base = torch.zeros((num_base_shifts, 3), device=self.device, dtype=dtype_real)
        else:
            base = torch.as_tensor(grid[0], device=self.device, dtype=dtype_real)
            if base.ndim == 1:
                base = base.unsqueeze(0)
            if base.shape[0] > num_base_shifts:
                base = base[:num_base_shifts]
            elif base.shape[0] < num_base_shifts:
                pad = torch.zeros((num_base_shifts - base.shape[0], 3), device=self.device, dtype=dtype_real)
                base = torch.cat((base, pad), dim=0)
        return base.unsqueeze(0).expand(num_subtomograms_per_batch, -1, -1)

    @torch.no_grad()
    def search_shifts(self, *, subtomos: torch.Tensor, ctfs: torch.Tensor = None, shift_zyx=None,
                      rotation_tracker=None, micro_batch_size: int = None):
        """(shift (B, 3), scores (B,)) float32 on the device."""
        if self._plan is None:
            raise RuntimeError("ShiftMatcher reference is not initialized. Call `set_reference` first.")
        _lib.require_gpu_tensor(subtomos, "subtomos", device=self.device)
        vols = subtomos.float().contiguous()
        B = vols.shape[0]
        if tuple(vols.shape[1:]) != (self.N,) * 3:
            raise ValueError(f"subtomos must be (B, {self.N}, {self.N}, {self.N})")
        if B > self.max_batch:
            raise ValueError(f"batch {B} exceeds num_subtomograms_per_batch {self.max_batch}")
        if isinstance(rotation_tracker, torch.Tensor) and rotation_tracker.is_cuda and \
                rotation_tracker.numel() == 9 * B:
            # device-resident sampling matrices rb (B, 3, 3) (e.g. from euler_to_shift_rotations)
            rot_t = rotation_tracker.reshape(B, 9).to(torch.float32).contiguous()
        else:
            if rotation_tracker is None:
                rot = np.tile(np.eye(3), (B, 1, 1))
            else:
                rot = tracker_to_matrices(rotation_tracker)
            if rot.shape[0] != B:
                raise ValueError("rotation_tracker does not match the batch")
            rot_t = torch.tensor(rot.reshape(B, 9), dtype=torch.float32, device=self.device)
        sh = None
        if shift_zyx is not None:
            sh = torch.as_tensor(shift_zyx, device=self.device, dtype=torch.float32).reshape(B, -1, 3)[:, 0, :]
            sh = sh.contiguous()
        ctf = None
        if self.do_ctf and ctfs is not None:
            ctf = torch.as_tensor(ctfs, device=self.device, dtype=torch.float32).reshape(
                B, self.N, self.N, self.N // 2 + 1).contiguous()
        out_shift = torch.empty((B, 3), dtype=torch.float32, device=self.device)
        out_score = torch.empty((B,), dtype=torch.float32, device=self.device)
        rc = _lib.lib().mch_shift_search(self._plan, vols.data_ptr(), vols.stride(0), B,
                                         sh.data_ptr() if sh is not None else None, rot_t.data_ptr(),
                                         ctf.data_ptr() if ctf is not None else None,
                                         out_shift.data_ptr(), out_score.data_ptr(),
                                         _lib.stream_handle(self.device))
        _lib.check(rc, "mch_shift_search")
        return out_shift, out_score


def euler_to_shift_rotations(alphas: torch.Tensor, betas: torch.Tensor, gammas: torch.Tensor):
    """(B, 3, 3) device matrices for ``search_shifts`` from search_orientations'
    angles (the x<->z frame change of ``update_rotation_estimate``, on the GPU)."""
    B = alphas.shape[0]
    a, b, g = (t.contiguous().float() for t in (alphas, betas, gammas))
    out = torch.empty((B, 3, 3), dtype=torch.float32, device=a.device)
    rc = _lib.lib().mch_euler_to_shift_rot(a.data_ptr(), b.data_ptr(), g.data_ptr(), B, out.data_ptr(),
                                           _lib.stream_handle(a.device))
    _lib.check(rc, "mch_euler_to_shift_rot")
    return out


def extract_shifted_patches(subtomos: torch.Tensor, shift_zyx: torch.Tensor, out: torch.Tensor = None):
    """patch[a] = vol[a + flip(shift_zyx)] (zeros outside; trilinear for fractional
    shifts): ``extract_shifted_patch_batch(..., normalize=False)``."""
    _lib.require_gpu_tensor(subtomos, "subtomos")
    vols = subtomos.float().contiguous()
    B, N = vols.shape[0], vols.shape[1]
    sh = shift_zyx.reshape(B, -1, 3)[:, 0, :].float().contiguous() if shift_zyx is not None else None
    if out is None:
        out = torch.empty_like(vols)
    rc = _lib.lib().mch_extract_patches(vols.data_ptr(), vols.stride(0), B, N,
                                        sh.data_ptr() if sh is not None else None, out.data_ptr(),
                                        out.stride(0), _lib.stream_handle(vols.device))
    _lib.check(rc, "mch_extract_patches")
    return out
