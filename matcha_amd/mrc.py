"""MRC volumes and RELION5 subtomogram / weight files (host I/O).

The reference reads particles and CTF weights with mrcfile 1.5.4
(``src/matcha/utils/io_utils.py:229-254``), which is not installed here; the
MRC2014 format is restated: a 1024-byte header (NX, NY, NZ, MODE at words 0-3,
CELLA at 40, NSYMBT at 92, MAP at 208, MACHST at 212), NSYMBT bytes of extended
header, then the data with x fastest, i.e. a (NZ, NY, NX) array.  Modes 0
(int8), 1 (int16), 2 (float32), 6 (uint16) and 12 (float16); the byte order
comes from MACHST (0x44 0x44: little endian, 0x11 0x11: big).

* ``load_subtomogram`` = ``load_subtomogram_cpu`` (:229-238): the volume as
  C-ordered float32.
* ``load_ctf_relion5`` = ``load_ctf_relion5_cpu`` (:240-254): float32,
  optionally transposed (2, 1, 0), then the slice ``[:n0 // 2, :, :n0 // 2 + 1]``
  of the reference, n0 the first dimension (the RELION5 weights file of a
  particle, ``<name>_weights.mrc`` beside ``<name>_data.mrc``,
  ``run_utils.py:476-479``).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

HEADER_BYTES = 1024
_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
_MODE_OF = {np.dtype(v).str[1:]: k for k, v in _MODES.items()}


class MrcHeader:
    def __init__(self, shape, mode, voxel_size, nsymbt, endian, origin=(0.0, 0.0, 0.0)):
        self.shape = tuple(int(s) for s in shape)  # (nz, ny, nx)
        self.mode = int(mode)
        self.voxel_size = tuple(float(v) for v in voxel_size)  # (x, y, z) in Angstrom
        self.nsymbt = int(nsymbt)
        self.endian = endian
        self.origin = tuple(float(v) for v in origin)

    @property
    def data_offset(self) -> int:
        return HEADER_BYTES + self.nsymbt


def read_header(path) -> MrcHeader:
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES:
        raise ValueError(f"{path}: shorter than an MRC header")
    mach = raw[212:214]
    endian = ">" if mach[0] == 0x11 else "<"
    nx, ny, nz, mode = struct.unpack(endian + "4i", raw[0:16])
    if not (0 < nx and 0 < ny and 0 < nz) or mode not in _MODES:
        # MACHST absent in some writers: try the other byte order once
        endian = ">" if endian == "<" else "<"
        nx, ny, nz, mode = struct.unpack(endian + "4i", raw[0:16])
        if not (0 < nx and 0 < ny and 0 < nz) or mode not in _MODES:
            raise ValueError(f"{path}: unsupported MRC header (mode {mode}, size {nx}x{ny}x{nz})")
    mx, my, mz = struct.unpack(endian + "3i", raw[28:40])
    cx, cy, cz = struct.unpack(endian + "3f", raw[40:52])
    nsymbt = struct.unpack(endian + "i", raw[92:96])[0]
    origin = struct.unpack(endian + "3f", raw[196:208])
    vs = tuple(c / m if m > 0 else 0.0 for c, m in ((cx, mx), (cy, my), (cz, mz)))
    return MrcHeader((nz, ny, nx), mode, vs, max(nsymbt, 0), endian, origin)


def read_mrc(path, mmap: bool = True):
    """(data (NZ, NY, NX) array, header); memory-mapped read-only by default."""
    h = read_header(path)
    dt = np.dtype(_MODES[h.mode]).newbyteorder(h.endian)
    count = int(np.prod(h.shape))
    size = Path(path).stat().st_size
    if size < h.data_offset + count * dt.itemsize:
        raise ValueError(f"{path}: truncated data ({size} bytes)")
    if mmap:
        data = np.memmap(path, dtype=dt, mode="r", offset=h.data_offset, shape=h.shape)
    else:
        data = np.fromfile(path, dtype=dt, count=count, offset=h.data_offset).reshape(h.shape)
    return data, h


def write_mrc(path, data, voxel_size: float = 1.0) -> None:
    """MRC2014 file of a 3-D array (mode from its dtype: int8, int16, float32, uint16, float16)."""
    a = np.ascontiguousarray(data)
    if a.ndim != 3:
        raise ValueError("write_mrc expects a 3-D array (nz, ny, nx)")
    key = a.dtype.str[1:]
    if key not in _MODE_OF:
        a = a.astype(np.float32)
        key = "f4"
    a = a.astype(a.dtype.newbyteorder("<"), copy=False)
    nz, ny, nx = a.shape
    hdr = bytearray(HEADER_BYTES)
    struct.pack_into("<4i", hdr, 0, nx, ny, nz, _MODE_OF[key])
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)
    struct.pack_into("<3f", hdr, 40, nx * voxel_size, ny * voxel_size, nz * voxel_size)
    struct.pack_into("<3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)
    f = a.astype(np.float64) if a.size else np.zeros(1)
    struct.pack_into("<3f", hdr, 76, float(f.min()), float(f.max()), float(f.mean()))
    struct.pack_into("<i", hdr, 92, 0)
    struct.pack_into("<i", hdr, 108, 20140)
    hdr[208:212] = b"MAP "
    hdr[212:216] = bytes([0x44, 0x44, 0, 0])
    struct.pack_into("<f", hdr, 216, float(f.std()))
    with open(path, "wb") as fh:
        fh.write(hdr)
        fh.write(a.tobytes())


def load_subtomogram(path, dtype=np.float32) -> np.ndarray:
    """``load_subtomogram_cpu``: the volume as a C-ordered array of ``dtype``."""
    data, _ = read_mrc(path)
    return np.array(data, dtype=dtype, order="C", copy=True)


def load_ctf_relion5(path, transpose: bool = False) -> np.ndarray:
    """``load_ctf_relion5_cpu``: float32 weights, optionally transposed (2, 1, 0),
    sliced ``[:n0 // 2, :, :n0 // 2 + 1]``."""
    data, _ = read_mrc(path)
    ctf = np.asarray(data, dtype=np.float32, order="C")
    if transpose:
        ctf = np.transpose(ctf, (2, 1, 0))
    return np.ascontiguousarray(ctf[: ctf.shape[0] // 2, :, : ctf.shape[0] // 2 + 1])


def weights_path(data_path: str, base: str = "_data.", replace: str = "_weights.") -> str:
    """The weights file of a particle (``run_utils.py:35,476-479``)."""
    return str(data_path).replace(base, replace)
