#!/usr/bin/env python3
"""Write golden S(q) tables with the reference's own numpy reference.

Runs in the build container only (it reads /root/reference, which does not
exist where the tests run).  For each case it writes a small LAMMPS dump with
this project's own seeded generator, runs the reference's independent numpy
reference ``test/ref_sq.py`` on it (the reference uses the same script to
check its Fortran output, test/test_sqcalc.py:179-306), and stores the inputs
(box, types, positions) and the reference's output table as JSON under
tests/golden/.  The fixtures are data: nothing of the reference's source is
kept.

Usage: python tests/golden/make_golden.py [/root/reference]
"""

import importlib.util
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def write_dump(path, frames, types, lo, hi, tilt=None, columns="id type x y z"):
    with open(path, "w") as fh:
        for k, pos in enumerate(frames):
            fh.write("ITEM: TIMESTEP\n%d\n" % k)
            fh.write("ITEM: NUMBER OF ATOMS\n%d\n" % len(types))
            if tilt is None:
                fh.write("ITEM: BOX BOUNDS pp pp pp\n")
                for d in range(3):
                    fh.write("%.10f %.10f\n" % (lo[d], hi[d]))
            else:
                fh.write("ITEM: BOX BOUNDS xy xz yz pp pp pp\n")
                for d in range(3):
                    fh.write("%.10f %.10f %.10f\n" % (lo[d], hi[d], tilt[d]))
            fh.write("ITEM: ATOMS %s\n" % columns)
            for i in range(len(types)):
                vals = []
                for c in columns.split():
                    if c == "id":
                        vals.append("%d" % (i + 1))
                    elif c == "type":
                        vals.append("%d" % types[i])
                    else:
                        vals.append("%.10f" % pos[i, "xyz".index(c[0])])
                fh.write(" ".join(vals) + "\n")


def read_back(path, columns):
    """Positions exactly as written (10 decimals), so the fixture matches the dump."""
    frames = []
    cols = columns.split()
    ix = [cols.index(c) for c in cols if c[0] in "xyz"]
    with open(path) as fh:
        lines = fh.readlines()
    i = 0
    while i < len(lines):
        if lines[i].startswith("ITEM: NUMBER OF ATOMS"):
            n = int(lines[i + 1])
        if lines[i].startswith("ITEM: ATOMS"):
            body = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            frames.append(body[:, ix])
            i += 1 + n
            continue
        i += 1
    return np.array(frames)


def bounds_for(length, tilt):
    xy, xz, yz = tilt
    lo = [0.0 + min(0.0, xy, xz, xy + xz), 0.0 + min(0.0, yz), 0.0]
    hi = [length + max(0.0, xy, xz, xy + xz), length + max(0.0, yz), length]
    return lo, hi


CASES = [
    # name, natoms, length, frames, seed, weight, norm, mapping, tilt, columns
    ("gas_unit_self", 250, 20.0, 4, 7, "unit", "self", "1:Si,2:O", None, "id type x y z"),
    ("gas_neutron_mean", 250, 20.0, 4, 7, "neutron", "mean", "1:Si,2:O", None, "id type x y z"),
    ("gas_xray_mean", 250, 20.0, 4, 7, "xray", "mean", "1:Si,2:O", None, "id type x y z"),
    ("gas_ion_xray", 250, 20.0, 2, 7, "xray", "mean", "1:Si4+,2:O2-", None, "id type x y z"),
    ("tri_unit_self", 150, 18.0, 3, 11, "unit", "self", "1:Si,2:O", (1.5, 0.5, -2.0),
     "id xu yu zu type"),
    ("gas_neutron_n", 300, 16.0, 2, 5, "neutron", "n", "1:Si,2:O", None, "id type x y z"),
]


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    spec = importlib.util.spec_from_file_location("ref_sq", os.path.join(ref_root, "test",
                                                                          "ref_sq.py"))
    ref_sq = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref_sq)
    qmax, nq = 6.0, 120
    for name, natoms, length, nframes, seed, weight, norm, mapping, tilt, columns in CASES:
        rng = np.random.default_rng(seed)
        types = rng.integers(1, 3, natoms)
        frames = [rng.random((natoms, 3)) * length for _ in range(nframes)]
        if tilt is not None:
            # spread the atoms over the sheared cell: r = A s
            A = np.array([[length, 0, 0], [tilt[0], length, 0], [tilt[1], tilt[2], length]]).T
            frames = [f @ A.T / length for f in frames]
            lo, hi = bounds_for(length, tilt)
        else:
            lo, hi = [0.0] * 3, [length] * 3
        with tempfile.TemporaryDirectory() as tmp:
            dump = os.path.join(tmp, "in.dump")
            out = os.path.join(tmp, "out.dat")
            write_dump(dump, frames, types, lo, hi, tilt, columns)
            pos = read_back(dump, columns)
            argv = sys.argv
            sys.argv = ["ref_sq.py", "--input", dump, "--mapping", mapping, "--weight", weight,
                        "--norm", norm, "--qmax", str(qmax), "--nq", str(nq), "--output", out]
            try:
                ref_sq.main()
            finally:
                sys.argv = argv
            table = np.loadtxt(out)
        fixture = {
            "name": name, "weight": weight, "norm": norm, "mapping": mapping,
            "qmin": 0.0, "qmax": qmax, "nq": nq,
            "lo": lo, "hi": hi, "tilt": list(tilt) if tilt else [0.0, 0.0, 0.0],
            "types": [int(t) for t in types],
            "positions": np.round(pos, 10).tolist(),
            "q": table[:, 0].tolist(), "S": table[:, 1].tolist(),
            "source": "reference test/ref_sq.py run on this dump",
        }
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(fixture, fh)
        print("wrote", name, table.shape)


if __name__ == "__main__":
    main()
