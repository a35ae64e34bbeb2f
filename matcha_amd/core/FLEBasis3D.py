"""Ball-harmonic expansion a = B^H f on the MI355X library.

Drop-in for ``src/matcha/core/FLEBasis3D.py`` (class ``FLEBasis3D``, same
constructor, attributes ``ne, lmax, kmax, ls, ks, ms, mds, lmds, cs,
idlm_list, h``; ``evaluate_t``).  The reference approximates B^H f to ~eps by
a type-2 NUFFT, two real SHTs and barycentric radial interpolation
(``:945-976,1026-1267``).  Here the projection is computed exactly, as MFMA
GEMMs, with the 16-element symmetry group D4h of the voxel grid folded in:

* every ball voxel v is g*v0 for one fundamental voxel v0 (x >= 0, 0 <= y <= x,
  z >= 0) and g = R^p My^t Mz^q (R: 90 degrees about z, My: y -> -y,
  Mz: z -> -z);
* under g, P_lm(cos theta) picks up (-1)^{q(l+m)} and e^{-i m phi} picks up
  (-i)^{m p} and phi -> -phi for t = 1, so the orbit sum of f * conj(Y_lm)
  collapses to P_lm(x0) [cos(m phi0) X1 + sin(m phi0) X2] (real part) and
  [cos X3 + sin X4] (imaginary part), where X1..X4 are fixed combinations of
  the 16 orbit values that depend only on the class ((l+m) mod 2, m mod 4);
* a_{lkm} (m >= 0) = sum_{v0} X . W[v0, (l,k,m)] with
  W = h c_lk w(v0) j_l(lambda_lk r0/radius) P_lm(x0) (cos | sin)(m phi0),
  w = 1/|stabiliser(v0)|; a_{lk,-m} = (-1)^m conj(a_{lkm}) (f is real);
* the class m = 3 mod 4 is the conjugate of m = 1 mod 4 (its X2, X3 are
  negated), so it reuses that operand with its sin rows of W negated and
  its imaginary output negated: 6 operand slabs instead of 8.

The basis therefore spans ~V_ball/16 voxels and only m >= 0, i.e. about 1/8
of the dense complex basis.  Tables are built here in FP64 and stored FP32.
Orbit images outside the array (coordinate +N/2, which a ball of radius N/2
reaches) enter the orbit sums as zeros.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch
from scipy import optimize, special

from .. import _lib

N_CLASSES = 8  # (l+m) mod 2 in {0,1} x m mod 4 in {0,1,2,3}
COL_PAD = 128  # class column counts are padded to the GEMM column tile


# ---------------------------------------------------------------------------
# spectrum: roots of j_l by interlacing, the reference's ordering and cut
# ---------------------------------------------------------------------------

def _sph_jn_roots_interlaced(xmax: float):
    """roots[l] = increasing roots of j_l below xmax, l = 0, 1, ...

    Roots of j_{l+1} interlace those of j_l (lambda_{l,k} < lambda_{l+1,k} <
    lambda_{l,k+1}), so each is bracketed by two consecutive roots of j_l."""
    kmax0 = int(xmax / math.pi) + 2
    prev = [math.pi * k for k in range(1, kmax0 + 1)]  # one beyond xmax
    out = [[r for r in prev if r < xmax]]
    l = 0
    while True:
        f = lambda x, n=l + 1: special.spherical_jn(n, x)
        cur = []
        for a, b in zip(prev[:-1], prev[1:]):
            cur.append(optimize.brentq(f, a, b, xtol=1e-15, rtol=8.9e-16, maxiter=300))
            if cur[-1] >= xmax:
                break
        # keep one root beyond xmax (bracket for the next l)
        if not cur or cur[0] >= xmax:
            break
        if cur[-1] < xmax:
            # need one more root beyond the last bracket
            a = prev[-1]
            b = a + math.pi
            while f(a) * f(b) > 0:
                b += math.pi / 2
            cur.append(optimize.brentq(f, a, b, xtol=1e-15, rtol=8.9e-16, maxiter=300))
        out.append([r for r in cur if r < xmax])
        prev = cur
        l += 1
    return out


def ball_spectrum(bandlimit: int):
    """Basis enumeration with the reference's order and truncation
    (``FLEBasis3D.lap_eig_ball``, FLEBasis3D.py:700-799): entries (k, l, m) with
    m in 0,-1,1,-2,2,..., stable-sorted by lambda_lk, cut at
    searchsorted(lambda, bandlimit*pi/2) - 1 and then at a trailing negative m."""
    threshold = bandlimit * np.pi / 2
    roots = _sph_jn_roots_interlaced(threshold + 1.0)
    groups = []  # (k, l, lambda) in the reference's insertion order (k outer, l inner)
    kmax = max(len(r) for r in roots)
    for k in range(1, kmax + 1):
        for l, rl in enumerate(roots):
            if len(rl) >= k:
                groups.append((k, l, rl[k - 1]))
    ls, ks, ms, lmds = [], [], [], []
    for k, l, lam in groups:
        for j in range(2 * l + 1):
            ls.append(l)
            ks.append(k)
            ms.append(((-1) ** j) * ((j + 1) // 2))
            lmds.append(lam)
    ls, ks, ms, lmds = (np.asarray(x) for x in (ls, ks, ms, lmds))
    order = np.argsort(lmds, kind="stable")
    ls, ks, ms, lmds = ls[order], ks[order], ms[order], lmds[order]
    ne = int(np.searchsorted(lmds, threshold, side="left")) - 1
    if ms[ne - 1] < 0:
        ne -= 1
    ne = max(ne, 1)
    ls, ks, ms, lmds = ls[:ne], ks[:ne], ms[:ne], lmds[:ne]
    mds = 2 * np.abs(ms) - (ms < 0)
    cs = np.sqrt(2.0) / np.abs(special.spherical_jn(ls + 1, lmds))
    return ls.astype(np.int64), ks.astype(np.int64), ms.astype(np.int64), mds.astype(np.int64), \
        lmds.astype(np.float64), cs.astype(np.float64), ne


# ---------------------------------------------------------------------------
# geometry: fundamental voxels of D4h and their orbits
# ---------------------------------------------------------------------------

def _orbit(v):
    """The 16 images R^p My^t Mz^q v, ordered p + 4 q + 8 t."""
    out = []
    for t in (0, 1):
        for q in (0, 1):
            for p in range(4):
                x, y, z = v
                if q:
                    z = -z
                if t:
                    y = -y
                for _ in range(p):
                    x, y = -y, x
                out.append((x, y, z))
    return out


def fundamental_voxels(N: int, radius: float):
    """Fundamental voxels of the ball |v| <= radius (1 + 1e-13), v = index - N/2.

    The array holds v in [-N/2, N/2 - 1] per axis and the ball is its part
    within the radius, as the reference's mask (``np.arange(-R, R + N % 2)``,
    FLEBasis3D.py:502-515): at radius N/2 the -N/2 boundary voxels are in the
    ball and their +N/2 mirror images are not in the array.  Such images get
    the index -1 (read as 0); orbits with no image in the array are dropped.

    Returns coords (V0, 3) int (x >= y >= 0, z >= 0, ordered by x, then y,
    then z), orbit flat indices (V0, 16) int32 (image p + 4 q + 8 t of
    _orbit, axis 0 = x, -1 outside the array), weights 1/|stabiliser| (V0,)."""
    if not radius > 0:
        raise ValueError(f"radius must be positive, got {radius}")
    R = int(math.floor(radius * (1 + 1e-13)))
    c = N // 2
    lim = radius * (1 + 1e-13)
    x, y, z = np.meshgrid(np.arange(R + 1), np.arange(R + 1), np.arange(R + 1), indexing="ij")
    keep = (y <= x) & (np.sqrt((x * x + y * y + z * z).astype(np.float64)) <= lim)
    x, y, z = x[keep], y[keep], z[keep]
    # the 16 images R^p My^t Mz^q v, in the order p + 4 q + 8 t
    imgs = np.empty((x.size, 16, 3), dtype=np.int64)
    for t in (0, 1):
        for q in (0, 1):
            yt = -y if t else y
            zq = -z if q else z
            ax, ay = x, yt
            for p in range(4):
                imgs[:, p + 4 * q + 8 * t] = np.stack([ax, ay, zq], axis=1)
                ax, ay = -ay, ax
    inside = np.all((imgs >= -c) & (imgs < N - c), axis=2)
    g = imgs + c
    ids = np.where(inside, (g[..., 0] * N + g[..., 1]) * N + g[..., 2], -1)
    any_in = ids.max(axis=1) >= 0
    base = np.stack([x, y, z], axis=1)[:, None, :]
    stab = np.all(imgs == base, axis=2).sum(axis=1)
    coords = np.stack([x, y, z], axis=1)[any_in]
    return (coords.astype(np.int64), ids[any_in].astype(np.int32),
            1.0 / stab[any_in].astype(np.float64))


def _legendre_orthonormal(lmax: int, x: np.ndarray):
    """P[l][m] (m <= l): orthonormal associated Legendre with Condon-Shortley phase,
    so that Y_lm(theta, phi) = P[l][m](cos theta) e^{i m phi}."""
    x = np.asarray(x, dtype=np.float64)
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = [[None] * (l + 1) for l in range(lmax + 1)]
    P[0][0] = np.full_like(x, 1.0 / math.sqrt(4 * math.pi))
    for m in range(1, lmax + 1):
        P[m][m] = -math.sqrt((2 * m + 1) / (2.0 * m)) * s * P[m - 1][m - 1]
    for m in range(0, lmax):
        P[m + 1][m] = math.sqrt(2 * m + 3) * x * P[m][m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1) ** 2 - 1))
            P[l][m] = a * (x * P[l - 1][m] - b * P[l - 2][m])
    return P


def projection_layout(N: int, radius: float, spectrum, lmax: int, BK: int = 16):
    """Layout of the symmetry-adapted projection GEMM and the inputs of its
    table, which the library builds on the device (``mch_fle_build_basis``).

    Returns dict with V0, Kpad, Ntot, class_cols, class_colpad, orbits (V0, 16)
    int32, chan_plus/chan_minus/chan_m (Ntot,) int32, and the table's inputs:
    vox (V0, 4) FP64 = (r, cos theta, phi, h w) per fundamental voxel,
    chan_lm (Ntot, 2) int32 = (l, m) per column (-1 for padding), chan_val
    (Ntot, 3) FP64 = (lambda_lk, c_lk, sign of the sin half) per column."""
    ls, ks, ms, lmds, cs = spectrum
    coords, orbits, w = fundamental_voxels(N, radius)
    V0 = len(coords)
    x, y, z = (coords[:, i].astype(np.float64) for i in range(3))
    r = np.sqrt(x * x + y * y + z * z)
    x0 = np.where(r > 0, z / np.where(r > 0, r, 1), 1.0)
    phi0 = np.arctan2(y, x)
    index = {(int(l), int(k), int(m)): i for i, (l, k, m) in enumerate(zip(ls, ks, ms))}
    classes = [[] for _ in range(N_CLASSES)]
    for (l, k, m), i in index.items():
        if m < 0:
            continue
        cls = ((l + m) % 2) * 4 + (m % 4)
        classes[cls].append((l, k, m, i, index.get((l, k, -m), -1)))
    for c in classes:
        c.sort()
    # K order: the cos-weighted operand of fundamental voxel v at k = v, the
    # sin-weighted one at k = Kpad/2 + v (classes m = 0, 2 mod 4 have a zero
    # sin half in their real rows and a zero cos half in their imaginary rows,
    # which the GEMM then skips)
    KH = ((V0 + BK - 1) // BK) * BK
    Kpad = 2 * KH
    cols = [len(c) for c in classes]
    colpad = [((n + COL_PAD - 1) // COL_PAD) * COL_PAD for n in cols]
    Ntot = int(sum(colpad))
    chan_plus = np.full(Ntot, -1, dtype=np.int32)
    chan_minus = np.full(Ntot, -1, dtype=np.int32)
    chan_m = np.zeros(Ntot, dtype=np.int32)
    chan_lm = np.full((Ntot, 2), -1, dtype=np.int32)
    chan_val = np.zeros((Ntot, 3), dtype=np.float64)
    h = radius ** -1.5
    off = 0
    for cls in range(N_CLASSES):
        # classes (eps, 3) read the (eps, 1) operand, which has X2 and X3
        # negated: their sin rows are negated here and their imaginary
        # output in the kernel's finalize
        ssign = -1.0 if cls % 4 == 3 else 1.0
        for j, (l, k, m, ip, im) in enumerate(classes[cls]):
            chan_plus[off + j] = ip
            chan_minus[off + j] = im if m > 0 else -1
            chan_m[off + j] = m
            chan_lm[off + j] = (l, m)
            chan_val[off + j] = (lmds[ip], cs[ip], ssign)
        off += colpad[cls]
    vox = np.stack([r, x0, phi0, h * w], axis=1)
    return dict(V0=V0, Kpad=Kpad, Ntot=Ntot, class_cols=cols, class_colpad=colpad,
                orbits=np.ascontiguousarray(orbits), chan_plus=chan_plus, chan_minus=chan_minus,
                chan_m=chan_m, vox=np.ascontiguousarray(vox), chan_lm=chan_lm, chan_val=chan_val,
                lmax=int(lmax), radius=float(radius))


def host_basis(layout) -> np.ndarray:
    """The table W (Kpad, Ntot) float32 of a layout on the host in FP64 (SciPy's
    spherical Bessel functions, the host Legendre recurrence): what the
    library's device build computes, for tests and small sizes."""
    V0, Kpad, Ntot = layout["V0"], layout["Kpad"], layout["Ntot"]
    KH = Kpad // 2
    r, x0, phi0, hw = (layout["vox"][:, i] for i in range(4))
    P = _legendre_orthonormal(layout["lmax"], x0)
    Wb = np.zeros((Kpad, Ntot), dtype=np.float32)
    jl_cache = {}
    for col in range(Ntot):
        l, m = (int(v) for v in layout["chan_lm"][col])
        if l < 0:
            continue
        lam, c, ssign = layout["chan_val"][col]
        if (l, lam) not in jl_cache:
            jl_cache[(l, lam)] = special.spherical_jn(l, lam * r / layout["radius"])
        rad = hw * jl_cache[(l, lam)] * P[l][m] * c
        ang = phi0 * m
        Wb[0:V0, col] = rad * np.cos(ang)
        Wb[KH:KH + V0, col] = ssign * rad * np.sin(ang)
    return Wb


def projection_tables(N: int, radius: float, spectrum, lmax: int, BK: int = 16):
    """projection_layout with the host-built table as ``W`` (tests)."""
    t = projection_layout(N, radius, spectrum, lmax, BK)
    t["W"] = host_basis(t)
    return t


# compute precision of the projection GEMM (mch_fle_set_precision)
COMPUTE_PRECISIONS = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}


class FLEBasis3D:
    """Same constructor as the reference (FLEBasis3D.py:39-109), plus
    ``compute_precision``: "fp32" (plain FP32 MFMA GEMM), "bf16x3" or "fp16x3"
    (16-bit hi/lo split operands, three MFMA products, FP32 accumulation)."""

    BK = 64  # K padding of the GEMM (the 16-bit path splits W^T in 64 x 64 tiles)

    def __init__(self, N, bandlimit, eps, maxitr=None, maxfun=None, max_l=None, mode="complex",
                 sph_harm_solver="nvidia_torch", reduce_memory=True, batchsize=1, radius=1,
                 device="cuda", dtype=torch.complex128, jl_zeros_path=None, cs_path=None,
                 precision_mode="accurate", compute_precision="fp32"):
        if mode != "complex":
            raise ValueError("only mode='complex' is supported")
        if maxfun or max_l:
            raise ValueError("maxfun/max_l truncations are not supported")
        self.batchsize = int(batchsize)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("matcha_amd runs on a GPU device only")
        self.complexmode = True
        self.precision_mode = precision_mode
        self.reduce_memory = reduce_memory
        self.eps = eps
        self.N1 = int(N)
        self.N = self.N1 + (self.N1 % 2)
        self.radius = float(radius)
        self.dtype = torch.complex64
        self.h = self.radius ** -1.5
        self.R = self.N // 2
        ls, ks, ms, mds, lmds, cs, ne = ball_spectrum(bandlimit)
        self.ls, self.ks, self.ms, self.mds, self.lmds, self.cs, self.ne = ls, ks, ms, mds, lmds, cs, ne
        self.lmax = int(ls.max())
        self.kmax = int(ks.max())
        self.mmax = int(np.abs(ms).max())
        self.idlm_list = [[[] for _ in range(2 * self.lmax + 1)] for _ in range(self.lmax + 1)]
        for i in range(ne):
            self.idlm_list[ls[i]][mds[i]].append(i)
        self._build_tables()
        self._plan = None
        self._make_plan()
        self.compute_precision = "fp32"
        self.set_compute_precision(compute_precision)

    def set_gemm_tiles(self, stages: int, waves: int, column_tiles: int, k_splits: int):
        """Pin the 16-bit GEMM to one autotune candidate (tuning/testing aid;
        the plan must have switched to a 16-bit mode)."""
        _lib.check(_lib.lib().mch_fle_set_gemm_tiles(self._plan, int(stages), int(waves), int(column_tiles),
                                                     int(k_splits)), "mch_fle_set_gemm_tiles")

    def gemm_tiles(self) -> dict:
        """The 16-bit GEMM's tile configuration (autotuned or pinned)."""
        d = int(_lib.lib().mch_fle_tuned_depth(self._plan))
        return {"column_tiles": d // 10000 + 1, "k_splits": d % 10000 // 100, "waves": d % 100 // 10,
                "stages": d % 10}

    def set_compute_precision(self, mode: str):
        """Switch the projection GEMM between "fp32", "bf16x3" and "fp16x3"."""
        if mode not in COMPUTE_PRECISIONS:
            raise ValueError(f"compute_precision must be one of {sorted(COMPUTE_PRECISIONS)}, got {mode!r}")
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().mch_fle_set_precision(self._plan, COMPUTE_PRECISIONS[mode]),
                       "mch_fle_set_precision")
        self.compute_precision = mode

    # -- tables ---------------------------------------------------------------
    def _build_tables(self):
        if self.N != self.N1:
            raise ValueError("odd box sizes are not supported")
        t = projection_layout(self.N, self.radius, (self.ls, self.ks, self.ms, self.lmds, self.cs),
                              self.lmax, self.BK)
        self._layout = t
        self.V0, self.Kpad, self.Ntot = t["V0"], t["Kpad"], t["Ntot"]
        self.class_cols, self.class_colpad = t["class_cols"], t["class_colpad"]
        self._orbits = t["orbits"]
        self._chan_plus, self._chan_minus, self._chan_m = t["chan_plus"], t["chan_minus"], t["chan_m"]

    def _make_plan(self):
        handle = ctypes.c_void_p()
        cc = (ctypes.c_int * N_CLASSES)(*self.class_colpad)
        with torch.cuda.device(self.device):
            rc = _lib.lib().mch_fle_plan_create(
                self.N, self.V0, self.Kpad, self.Ntot, cc,
                self._orbits.ctypes.data_as(ctypes.c_void_p), None,
                self._chan_plus.ctypes.data_as(ctypes.c_void_p),
                self._chan_minus.ctypes.data_as(ctypes.c_void_p),
                self._chan_m.ctypes.data_as(ctypes.c_void_p),
                self.ne, self.batchsize, ctypes.byref(handle))
            _lib.check(rc, "mch_fle_plan_create")
            self._plan = handle
            # the table itself is built on the device, in FP64, from the layout
            t = self._layout
            rc = _lib.lib().mch_fle_build_basis(
                handle, t["vox"].ctypes.data_as(ctypes.c_void_p), t["chan_lm"].ctypes.data_as(ctypes.c_void_p),
                t["chan_val"].ctypes.data_as(ctypes.c_void_p), float(self.radius))
            _lib.check(rc, "mch_fle_build_basis")

    def basis_table(self) -> torch.Tensor:
        """The projection table W (Kpad, Ntot) float32 as the library holds it
        (a device copy; tests)."""
        W = torch.empty((self.Kpad, self.Ntot), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().mch_fle_copy_basis(self._plan, W.data_ptr(), _lib.stream_handle(self.device)),
                       "mch_fle_copy_basis")
        return W

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and h.value:
            try:
                _lib.lib().mch_fle_plan_destroy(h)
            except Exception:
                pass

    # -- transform ---------------------------------------------------------------
    def evaluate_t(self, f: torch.Tensor, out: torch.Tensor | None = None,
                   shift_zyx: torch.Tensor | None = None) -> torch.Tensor:
        """a = B^H f for a batch (Bt, N, N, N) (or one (N, N, N) volume) of real volumes.

        Returns (Bt, ne) complex64.  Unlike the reference, the caller's tensor is
        not modified (the reference zeroes its out-of-ball voxels in place).
        ``shift_zyx`` (Bt, 3) (not in the reference): project the patches
        ``extract_shifted_patches(f, shift_zyx)`` would cut, read straight from
        ``f`` (no patch copy)."""
        if f.ndim == 3:
            f = f.unsqueeze(0)
        _lib.require_gpu_tensor(f, "volumes", device=self.device)
        if f.dtype != torch.float32:
            f = f.float()
        if tuple(f.shape[1:]) != (self.N1, self.N1, self.N1):
            raise ValueError(f"volumes must be (B, {self.N1}, {self.N1}, {self.N1}), got {tuple(f.shape)}")
        Bt = f.shape[0]
        if Bt > self.batchsize:
            raise ValueError(f"batch {Bt} exceeds the configured batchsize {self.batchsize}")
        f = f.contiguous()
        if out is None:
            out = torch.empty((Bt, self.ne), dtype=torch.complex64, device=self.device)
        if shift_zyx is None:
            rc = _lib.lib().mch_fle_project(self._plan, f.data_ptr(), self.N1 ** 3, Bt, out.data_ptr(),
                                            out.stride(0), _lib.stream_handle(self.device))
        else:
            _lib.require_gpu_tensor(shift_zyx, "shift_zyx", device=self.device)
            sh = shift_zyx.to(torch.float32).reshape(-1, 3).contiguous()
            if sh.shape[0] != Bt:
                raise ValueError(f"shift_zyx must be ({Bt}, 3), got {tuple(shift_zyx.shape)}")
            rc = _lib.lib().mch_fle_project_shifted(self._plan, f.data_ptr(), self.N1 ** 3, Bt, sh.data_ptr(),
                                                    out.data_ptr(), out.stride(0), _lib.stream_handle(self.device))
        _lib.check(rc, "mch_fle_project")
        return out

    @property
    def flops_per_volume(self) -> float:
        """Algorithmic FLOPs of the projection GEMM per volume: two operand rows
        (real and imaginary) against every basis column, over V0 voxels for the
        classes m = 0, 2 mod 4 (one of the cos / sin halves of their rows is
        zero) and 2 V0 for the classes m = 1, 3 mod 4."""
        even = sum(c for i, c in enumerate(self.class_cols) if i % 2 == 0)
        odd = sum(c for i, c in enumerate(self.class_cols) if i % 2 == 1)
        return 2.0 * 2 * self.V0 * (even + 2 * odd)

    @property
    def operand_bytes_per_volume(self) -> float:
        """Algorithmic bytes of the projection operand X per volume at 4 B per
        element (FP32, or BF16/FP16 hi + lo): the six stored class slabs (the
        classes m = 3 mod 4 read their m = 1 mod 4 slab), two rows each, the
        zero half of the m = 0, 2 mod 4 slabs not read."""
        return 4.0 * 2 * (4 * self.Kpad / 2 + 2 * self.Kpad)
