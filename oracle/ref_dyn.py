"""Runs the reference's own separable phase factors (oracle/_ref/ref_dyn,
built by ``make -C oracle ref`` from src/sqc_phase.f90 with this directory's
driver of the dyn amplitude loop) on a case.  Test infrastructure only.

ref_rho(pos, species_of, nspecies, recip, hkl, base=None) -> complex
(nspecies, nmodes), the reference's rho(m, species) of dyn_accumulate.
"""

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DYN = os.path.join(HERE, "_ref", "ref_dyn")


def available():
    return os.path.exists(REF_DYN)


def ref_rho(pos, species_of, nspecies, recip, hkl, base=None):
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    hkl = np.ascontiguousarray(np.asarray(hkl, dtype=np.int32).reshape(-1, 3))
    sp = np.asarray(species_of, dtype=np.int32) + 1
    recip = np.asarray(recip, dtype=np.float64).reshape(3, 3)
    bv = np.zeros(3) if base is None else np.asarray(base, dtype=np.float64)
    with tempfile.TemporaryDirectory() as tmp:
        inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(inp, "wb") as fh:
            np.array([len(sp), nspecies, len(hkl), 0 if base is None else 1],
                     dtype=np.int32).tofile(fh)
            np.asarray(recip.ravel(order="F"), dtype=np.float64).tofile(fh)
            bv.astype(np.float64).tofile(fh)
            pos.ravel().tofile(fh)          # (3, natoms) column-major = (natoms, 3) row-major
            sp.tofile(fh)
            hkl.ravel().tofile(fh)
        subprocess.run([REF_DYN, inp, out], check=True, capture_output=True)
        raw = np.fromfile(out, dtype=np.complex128)
    return raw.reshape(nspecies, len(hkl))
