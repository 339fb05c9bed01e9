"""Runs the reference's own LAMMPS dump reader (oracle/_ref/ref_dump, built by
``make -C oracle ref`` from src/sqc_dump.f90 and its two modules) and decodes
the records its driver writes.  Test infrastructure only.

ref_read(path) -> (frames, status): frames is a list of dicts with natoms,
timestep, cell (3, 3; lattice vectors in the columns), origin, pbc, pos
(N, 3), id, type_id and n_ambiguous (the reader's running count); status is
the reader's ierr that ended the file (1 = clean end).
"""

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DUMP = os.path.join(HERE, "_ref", "ref_dump")


def available():
    return os.path.exists(REF_DUMP)


def ref_read(path):
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "frames.bin")
        subprocess.run([REF_DUMP, str(path), out], check=True, capture_output=True)
        raw = open(out, "rb").read()
    frames, p = [], 0

    def take(dtype, n):
        nonlocal p
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=p)
        p += a.nbytes
        return a

    while True:
        ierr = int(take("<i4", 1)[0])
        if ierr != 0:
            return frames, ierr
        natoms, timestep = (int(v) for v in take("<i4", 2))
        cell = take("<f8", 9).reshape(3, 3, order="F").copy()
        origin = take("<f8", 3).copy()
        pbc = tuple(bool(v) for v in take("<i4", 3))
        pos = take("<f8", 3 * natoms).reshape(natoms, 3).copy()
        ids = take("<i4", natoms).astype(np.int64)
        types = take("<i4", natoms).astype(np.int32)
        namb = int(take("<i4", 1)[0])
        frames.append(dict(natoms=natoms, timestep=timestep, cell=cell, origin=origin, pbc=pbc,
                           pos=pos, id=ids, type_id=types, n_ambiguous=namb))
