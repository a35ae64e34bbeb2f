"""Host pipeline: threaded MRC reads into pinned staging, H2D on a copy stream.

Restates the reference's feeding path for one GPU: ``cpu_reader`` reads
particles and their RELION5 weights with a thread pool, capturing per-file
errors so that one bad file does not stop the run, and flushes batches in
order (``src/matcha/utils/run_utils.py:23-90,429-611``); the consumer pins
them and copies them to the device on a copy stream with events that the
compute stream waits on (``align_subtomograms.py:435-460``).

Here the reads land directly in a ring of pinned staging slots (one host copy
from the memory-mapped file), each full slot goes to a ring of device slots
with a non-blocking copy on a copy stream (default priority), and the
compute stream waits on the copy's event.  Two hazards are ordered with
events: a pinned slot is refilled only after its copy has finished, and a
device slot is overwritten only after the compute work queued on the batch it
held (everything enqueued before the next batch is requested) has finished.

``probe_batch_size`` replaces the reference's trial-allocation probe
(``run_utils.py:294-315,732-741``) with the plans' own sizes against the
device's free memory.
"""
from __future__ import annotations

import queue
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import mrc


class Batch:
    def __init__(self, paths, vols, ctfs, slot):
        self.paths = paths          # the particles' file paths, in read order
        self.vols = vols            # (B, N, N, N) float32 on the device
        self.ctfs = ctfs            # (B, N, N, N//2+1) float32 or None
        self.slot = slot


class ParticleBatches:
    """Iterate over ``paths`` in batches on ``device``.

    ctf: also read each particle's weights file (``mrc.weights_path``).
    workers: reader threads; depth: staging and device slots (>= 2 overlaps
    the reads and copies of batch k+1 with the compute on batch k)."""

    def __init__(self, paths, batch: int, device, ctf: bool = True, workers: int = 8, depth: int = 2,
                 weights_base: str = "_data.", weights_replace: str = "_weights."):
        self.paths = list(paths)
        self.batch = int(batch)
        self.device = torch.device(device)
        self.ctf = bool(ctf)
        self.workers = max(1, int(workers))
        self.depth = max(2, int(depth))
        self.base, self.replace = weights_base, weights_replace
        self.errors = []            # (path, exception) of files that could not be read
        if not self.paths:
            raise ValueError("no particle paths")
        first = mrc.read_header(self.paths[0])
        nz, ny, nx = first.shape
        if not nz == ny == nx:
            raise ValueError(f"{self.paths[0]}: particles must be cubes, got {first.shape}")
        self.N = nz
        self.ctf_shape = (self.N, self.N, self.N // 2 + 1)
        cuda = self.device.type == "cuda"
        self._cuda = cuda
        pin = dict(pin_memory=True) if cuda else {}
        self._host_v = [torch.empty((self.batch, self.N, self.N, self.N), dtype=torch.float32, **pin)
                        for _ in range(self.depth)]
        self._host_c = [torch.empty((self.batch,) + self.ctf_shape, dtype=torch.float32, **pin)
                        if self.ctf else None for _ in range(self.depth)]
        if cuda:
            self._dev_v = [torch.empty_like(h, device=self.device) for h in self._host_v]
            self._dev_c = [torch.empty_like(h, device=self.device) if h is not None else None
                           for h in self._host_c]
            self._copy_stream = torch.cuda.Stream(device=self.device)
            self._copied = [None] * self.depth     # H2D done: the pinned slot may be refilled
            self._consumed = [None] * self.depth   # compute on the device slot done

    # -- reading -------------------------------------------------------------
    def _read(self, path):
        try:
            v = mrc.load_subtomogram(path)
            if v.shape != (self.N,) * 3:
                raise ValueError(f"shape {v.shape}, expected {(self.N,) * 3}")
            c = None
            if self.ctf:
                c = mrc.load_ctf_relion5(mrc.weights_path(path, self.base, self.replace))
                if c.shape != self.ctf_shape:
                    raise ValueError(f"weights shape {c.shape}, expected {self.ctf_shape}")
            return path, v, c, None
        except Exception as e:  # a bad file is reported and skipped, as the reference does
            return path, None, None, e

    def _producer(self, ready: queue.Queue, free: queue.Queue):
        try:
            with ThreadPoolExecutor(max_workers=self.workers) as ex:
                inflight = self.workers * 4
                slot, fill, names = None, 0, []
                for start in range(0, len(self.paths), inflight):
                    for path, v, c, err in ex.map(self._read, self.paths[start:start + inflight]):
                        if err is not None:
                            self.errors.append((path, err))
                            continue
                        if slot is None:
                            slot = free.get()
                            if slot is None:
                                return
                            if self._cuda and self._copied[slot] is not None:
                                self._copied[slot].synchronize()
                            fill, names = 0, []
                        self._host_v[slot][fill].numpy()[...] = v
                        if self.ctf:
                            self._host_c[slot][fill].numpy()[...] = c
                        names.append(path)
                        fill += 1
                        if fill == self.batch:
                            ready.put((slot, fill, names))
                            slot = None
                if slot is not None and fill:
                    ready.put((slot, fill, names))
        except Exception as e:  # surfaced to the consumer
            ready.put(e)
        finally:
            ready.put(None)

    # -- iteration -----------------------------------------------------------
    def __iter__(self):
        ready: queue.Queue = queue.Queue()
        free: queue.Queue = queue.Queue()
        for s in range(self.depth):
            free.put(s)
        th = threading.Thread(target=self._producer, args=(ready, free), daemon=True)
        th.start()
        prev = None
        try:
            while True:
                item = ready.get()
                if prev is not None:
                    self._release(prev, free)
                    prev = None
                if item is None:
                    break
                if isinstance(item, Exception):
                    raise item
                slot, n, names = item
                prev = slot
                yield self._stage(slot, n, names)
        finally:
            free.put(None)
            th.join()

    def _stage(self, slot, n, names):
        hv, hc = self._host_v[slot], self._host_c[slot]
        if not self._cuda:
            # host batches (reader tests): views of the staging slot, valid until
            # the next batch is requested
            return Batch(names, hv[:n], hc[:n] if hc is not None else None, slot)
        comp = torch.cuda.current_stream(self.device)
        dv, dc = self._dev_v[slot], self._dev_c[slot]
        with torch.cuda.stream(self._copy_stream):
            if self._consumed[slot] is not None:
                self._copy_stream.wait_event(self._consumed[slot])
            dv[:n].copy_(hv[:n], non_blocking=True)
            if hc is not None:
                dc[:n].copy_(hc[:n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self._copy_stream)
        self._copied[slot] = ev
        comp.wait_event(ev)
        return Batch(names, dv[:n], dc[:n] if dc is not None else None, slot)

    def _release(self, slot, free):
        if self._cuda:
            ev = torch.cuda.Event()
            ev.record(torch.cuda.current_stream(self.device))
            self._consumed[slot] = ev
        free.put(slot)


def particle_bytes(N: int, ne: int, lmax: int, L_max: int, candidates: int = 10, n_grid: int = 64,
                   ctf: bool = True) -> int:
    """Device bytes per particle of one batch slot of the alignment plans:
    the particle, its patch, its weights, the projection operand (8 classes, 2
    rows, 2 V0 ~ N^3 pi / 12 columns, FP32 and 16-bit hi/lo), coefficients,
    sigma (Dsize(lmax)) and its repack, the SOFFT partials and the shift
    search's half spectrum."""
    dsize = lambda L: (L + 1) * (2 * L + 1) * (2 * L + 3) // 3
    vol = 4 * N ** 3
    half = 8 * N * N * (N // 2 + 1)
    kpad = int(np.ceil(np.pi * N ** 3 / 12.0 / 64.0)) * 64
    proj = 8 * 2 * kpad * (4 + 4)
    return int(vol * 2 + (4 * N * N * (N // 2 + 1) if ctf else 0) + proj + 8 * ne + 8 * dsize(lmax)
               + 4 * dsize(L_max) + 8 * n_grid * candidates * 2 + half)


def probe_batch_size(N: int, ne: int, lmax: int, L_max: int, device, fraction: float = 0.5,
                     limit: int = 4096, **kw) -> int:
    """Largest power-of-two batch whose per-particle plan memory fits in
    ``fraction`` of the device's free memory (the shared tables, a few GB,
    stay in the other part)."""
    free, _ = torch.cuda.mem_get_info(torch.device(device))
    per = particle_bytes(N, ne, lmax, L_max, **kw)
    b = 1
    while b * 2 <= limit and (b * 2) * per <= fraction * free:
        b *= 2
    return b


def align_paths(paths, correlator, rotation_matcher, shift_matcher, batch: int, num_alternations: int = 2,
                half: str = "half1", mode: str = "reference", records=None, workers: int = 8, ctf: bool = True):
    """Read, align and record particles: the reference's per-GPU loop of reading
    batches, running the alternations and storing every alternation's records
    (``run_utils.py:429-611``, ``align_subtomograms.py:316-413``), feeding
    ``relion.join_to_star``.  Returns the ``relion.AlignmentRecords``."""
    from .align import Alternation
    from .relion import AlignmentRecords, records_from_history
    pipe = Alternation(correlator, rotation_matcher, shift_matcher, num_alternations=num_alternations, mode=mode)
    records = AlignmentRecords() if records is None else records
    for b in ParticleBatches(paths, batch, correlator.device, ctf=ctf, workers=workers):
        pipe.run(b.vols, ctfs=b.ctfs)
        records_from_history(b.paths, pipe.history, half=half, records=records)
    return records
