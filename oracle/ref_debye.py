"""Runs the reference's own cell list (oracle/_ref/ref_debye, built by
``make -C oracle ref`` from src/sqc_neighbour_list.f90 with this directory's
driver of the Debye pair loop) on a trajectory.  Test infrastructure only.

ref_hist(cell, origin, types, ntypes, pbc, rmax, dr, skin, frames) ->
(hist[a][b][k] float64 ordered pair counts summed over the frames, 0-based
types, nbuilds, npairs of the last list).
"""

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DEBYE = os.path.join(HERE, "_ref", "ref_debye")


def available():
    return os.path.exists(REF_DEBYE)


def ref_hist(cell, origin, types, ntypes, pbc, rmax, dr, skin, frames):
    frames = np.ascontiguousarray(np.asarray(frames, dtype=np.float64))
    if frames.ndim == 2:
        frames = frames[None]
    nf, n, _ = frames.shape
    with tempfile.TemporaryDirectory() as tmp:
        inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(inp, "wb") as fh:
            np.array([n, ntypes, nf] + [1 if p else 0 for p in pbc], dtype=np.int32).tofile(fh)
            np.asarray(cell, dtype=np.float64).reshape(3, 3).ravel(order="F").tofile(fh)
            np.asarray(origin, dtype=np.float64).tofile(fh)
            np.array([rmax, dr, skin], dtype=np.float64).tofile(fh)
            np.asarray(types, dtype=np.int32).tofile(fh)
            frames.tofile(fh)       # (3, natoms, nframes) column-major
        subprocess.run([REF_DEBYE, inp, out], check=True, capture_output=True)
        with open(out, "rb") as fh:
            nbins, nbuilds = np.fromfile(fh, dtype=np.int32, count=2)
            npairs = int(np.fromfile(fh, dtype=np.int64, count=1)[0])
            h = np.fromfile(fh, dtype=np.float64)
    # histogram(ntypes, ntypes, nbins) column-major -> [a][b][k]
    return h.reshape((nbins, ntypes, ntypes)).transpose(2, 1, 0), int(nbuilds), npairs
