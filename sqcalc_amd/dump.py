"""LAMMPS dump frames through the library's native reader.

``DumpReader`` mirrors the reference's ``lammps_dump_reader_t``
(src/sqc_dump.f90:72-283): ``next_frame`` returns the next frame or None at
the end of the file and raises ``DumpError`` with the reference reader's
ierr code otherwise.  Positions come back continuous, as the reference's
``frame_t%pos`` does (src/sqc_dump.f90:300-364).  ``read_frames`` fills a
(F, N, 3) batch in the layout ``HipStructureFactor.accumulate`` takes, with the
driver's per-frame checks of atom count and box (app/main.f90:240-247).

``feed`` streams a whole file into a structure factor: the library parses
batch k + 1 into one half of a double buffer on a thread of its own while
batch k runs on the device from the other.
"""

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

END = 1


class DumpError(RuntimeError):
    def __init__(self, code, detail=""):
        text = _lib.load().sqc_dump_error_string(code).decode()
        super().__init__("dump reader error %d: %s%s" % (code, text,
                                                         (" -- " + detail) if detail else ""))
        self.code = code


@dataclass
class Frame:
    natoms: int
    timestep: int
    cell: np.ndarray        # (3, 3), lattice vectors in the columns
    origin: np.ndarray
    pbc: tuple
    triclinic: bool
    scaled: bool
    has_id: bool
    unwrapped_from_dump: bool
    n_ambiguous: int
    pos: np.ndarray = None  # (N, 3) Cartesian, continuous across frames
    id: np.ndarray = None
    type_id: np.ndarray = None


def _frame(h):
    return Frame(natoms=h.natoms, timestep=h.timestep,
                 cell=np.array(h.cell[:]).reshape(3, 3, order="F"),
                 origin=np.array(h.origin[:]), pbc=tuple(bool(v) for v in h.pbc),
                 triclinic=bool(h.triclinic), scaled=bool(h.scaled), has_id=bool(h.has_id),
                 unwrapped_from_dump=bool(h.unwrapped_from_dump), n_ambiguous=h.n_ambiguous)


class DumpReader:
    """Native LAMMPS dump reader; ``nthreads`` <= 0 picks the host's threads."""

    def __init__(self, path, nthreads=0):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        rc = self.lib.sqc_dump_open(ctypes.byref(self.handle), str(path).encode(), int(nthreads))
        if rc:
            raise DumpError(rc, str(path))

    def _raise(self, rc):
        raise DumpError(rc, self.lib.sqc_dump_message(self.handle).decode())

    def next_frame(self):
        """The next frame with its arrays, or None at the end of the file."""
        h = _lib.DumpFrame()
        rc = self.lib.sqc_dump_next(self.handle, ctypes.byref(h))
        if rc == END:
            return None
        if rc:
            self._raise(rc)
        fr = _frame(h)
        n = h.natoms
        fr.pos = np.empty((n, 3))
        fr.id = np.empty(n, dtype=np.int64)
        fr.type_id = np.empty(n, dtype=np.int32)
        rc = self.lib.sqc_dump_copy(self.handle, fr.pos.ctypes.data, fr.id.ctypes.data,
                                    fr.type_id.ctypes.data)
        if rc:
            self._raise(rc)
        return fr

    def __iter__(self):
        while True:
            fr = self.next_frame()
            if fr is None:
                return
            yield fr

    def read_frames(self, out, natoms, cell=None):
        """Fill out[:k] (an (F, natoms, 3) float64 C-contiguous array or pinned
        tensor) with the next k <= F frames; returns (k, header of the last)."""
        if hasattr(out, "data_ptr"):           # torch tensor (e.g. pinned host memory)
            import torch
            if out.dtype != torch.float64 or not out.is_contiguous() or out.device.type != "cpu":
                raise ValueError("out must be a contiguous float64 host tensor")
            ptr = out.data_ptr()
        else:
            if out.dtype != np.float64 or not out.flags.c_contiguous or not out.flags.writeable:
                raise ValueError("out must be a writeable C-contiguous float64 array")
            ptr = out.ctypes.data
        if out.ndim != 3 or tuple(out.shape[1:]) != (natoms, 3):
            raise ValueError("out must be (F, natoms, 3)")
        cbuf = None
        if cell is not None:
            cbuf = np.asfortranarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        n = ctypes.c_int64(0)
        h = _lib.DumpFrame()
        rc = self.lib.sqc_dump_read_frames(self.handle, out.shape[0], natoms,
                                           cbuf.ctypes.data if cbuf is not None else None,
                                           ptr, ctypes.byref(n), ctypes.byref(h))
        if rc:
            self._raise(rc)
        return n.value, (_frame(h) if n.value else None)

    def read_begin(self, out, natoms, cell=None):
        """Start read_frames into out on a library thread and return at once;
        out must stay alive and untouched until read_end."""
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.ndim != 3 \
                or tuple(out.shape[1:]) != (natoms, 3):
            raise ValueError("out must be a C-contiguous float64 (F, natoms, 3) array")
        cbuf = None
        if cell is not None:
            cbuf = np.asfortranarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
        self._pending = out
        rc = self.lib.sqc_dump_read_frames_begin(self.handle, out.shape[0], natoms,
                                                 cbuf.ctypes.data if cbuf is not None else None,
                                                 out.ctypes.data)
        if rc:
            self._raise(rc)

    def read_end(self):
        """Wait for read_begin's batch; returns the number of frames read."""
        n = ctypes.c_int64(0)
        rc = self.lib.sqc_dump_read_frames_end(self.handle, ctypes.byref(n))
        self._pending = None
        if rc:
            self._raise(rc)
        return n.value

    def close(self):
        if self.handle:
            self.lib.sqc_dump_close(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host_buffer(shape):
    """Page-locked host memory when torch has a device (uploads then run at
    link speed), plain memory otherwise."""
    try:
        import torch
        if torch.cuda.is_available():
            return torch.empty(shape, dtype=torch.float64, pin_memory=True).numpy()
    except ImportError:
        pass
    return np.empty(shape)


def feed(sf, path, cell=None, batch=8, nthreads=0, max_frames=None):
    """Stream the frames of a dump into ``sf.accumulate``, batch by batch:
    the library parses batch k + 1 on a thread of its own (read_begin) while
    batch k runs on the device.

    Every frame must keep ``sf.natoms`` atoms and the cell (3, 3) of ``cell``,
    or of the file's first frame when ``cell`` is None.  nthreads <= 0 parses
    with one worker fewer than the host's threads: the thread that drives the
    device waits by spinning, and a parse worker that shares its core holds
    up every frame's join.  Returns the number of frames added.
    """
    if nthreads <= 0:
        import os
        cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        nthreads = max(1, cores - 1)
    with DumpReader(path, nthreads) as rd:
        natoms = sf.natoms
        bufs = [_host_buffer((batch, natoms, 3)), _host_buffer((batch, natoms, 3))]
        # the first batch in the foreground; its first frame fixes the cell
        if cell is None:
            n1, h1 = rd.read_frames(bufs[0][:1], natoms)
            if n1 == 0:
                return 0
            cell = h1.cell
            n = 1 + rd.read_frames(bufs[0][1:], natoms, cell)[0]
        else:
            n = rd.read_frames(bufs[0], natoms, cell)[0]
        done, k = 0, 0
        while True:
            if max_frames is not None:
                n = min(n, max_frames - done)
            more = n == batch and (max_frames is None or done + n < max_frames)
            if more:
                rd.read_begin(bufs[(k + 1) % 2], natoms, cell)
            try:
                if n:
                    sf.accumulate(bufs[k % 2][:n])
                    done += n
            finally:
                nxt = rd.read_end() if more else 0
            if not more or nxt == 0:
                return done
            n = nxt
            k += 1
