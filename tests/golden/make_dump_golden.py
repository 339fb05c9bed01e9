#!/usr/bin/env python3
"""Write dump-reader fixtures with the reference's own reader.

Runs in the build container only: it needs oracle/_ref/ref_dump, which
``make -C oracle ref`` compiles from the reference's src/sqc_dump.f90 (plus
src/sqc_kinds.f90 and src/sqc_cell.f90) and oracle/ref_dump_driver.f90.  Each
case is a small LAMMPS dump written here by a seeded generator; the fixture
keeps the dump text and what the reference reader returned for it (frames
with cell, origin, periodicity, positions, ids, types, the running count of
ambiguous steps, and the status that ended the file).

Usage: python tests/golden/make_dump_golden.py
"""

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.ref_dump import available, ref_read  # noqa: E402


def header(step, n, bounds, flags="pp pp pp", tilt=None, columns="id type x y z"):
    s = "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (step, n)
    if tilt is None:
        s += "ITEM: BOX BOUNDS %s\n" % flags
        s += "".join("%.10f %.10f\n" % tuple(b) for b in bounds)
    else:
        s += "ITEM: BOX BOUNDS xy xz yz %s\n" % flags
        s += "".join("%.10f %.10f %.10f\n" % (b[0], b[1], t) for b, t in zip(bounds, tilt))
    return s + "ITEM: ATOMS %s\n" % columns


def walk(rng, n, frames, L, step):
    pos = rng.random((n, 3)) * L
    out = []
    for _ in range(frames):
        out.append(pos.copy())
        pos = pos + rng.normal(0.0, step, pos.shape)
    return out


def cases():
    rng = np.random.default_rng(20260604)
    L = 10.0
    cub = [(0.0, L)] * 3
    out = {}

    # wrapped x y z with boundary crossings, some steps over 0.4 L (ambiguous)
    traj = walk(rng, 40, 5, L, 1.5)
    traj[3][:4] += 4.6
    txt = ""
    for f, p in enumerate(traj):
        txt += header(f * 10, 40, cub)
        w = np.mod(p, L)
        txt += "".join("%d %d %.10f %.10f %.10f\n" % (i + 1, 1 + i % 3, *w[i]) for i in range(40))
    out["wrapped_crossings"] = txt

    # triclinic, unwrapped xu yu zu and the same trajectory wrapped into the cell
    tilt = (1.5, 0.5, -2.0)
    lo = np.array([0.0, 0.0, 0.0])
    a = np.array([[L, tilt[0], tilt[1]], [0.0, L, tilt[2]], [0.0, 0.0, L]])
    xlo_b = lo[0] + min(0.0, tilt[0], tilt[1], tilt[0] + tilt[1])
    xhi_b = lo[0] + L + max(0.0, tilt[0], tilt[1], tilt[0] + tilt[1])
    ylo_b = lo[1] + min(0.0, tilt[2])
    yhi_b = lo[1] + L + max(0.0, tilt[2])
    bnd = [(xlo_b, xhi_b), (ylo_b, yhi_b), (lo[2], lo[2] + L)]
    s_traj = walk(rng, 30, 4, 1.0, 0.08)
    txt_u, txt_w = "", ""
    for f, s in enumerate(s_traj):
        r = s @ a.T + lo
        rw = np.mod(s, 1.0) @ a.T + lo
        txt_u += header(f, 30, bnd, tilt=tilt, columns="id xu yu zu type")
        txt_u += "".join("%d %.10f %.10f %.10f %d\n" % (i + 1, *r[i], 1 + i % 2) for i in range(30))
        txt_w += header(f, 30, bnd, tilt=tilt, columns="id type x y z")
        txt_w += "".join("%d %d %.10f %.10f %.10f\n" % (i + 1, 1 + i % 2, *rw[i]) for i in range(30))
    out["triclinic_xu"] = txt_u
    out["triclinic_wrapped"] = txt_w

    # number formats: exponents, signs, integers, upper-case column names,
    # extra columns, no id column
    txt = header(0, 6, cub, columns="TYPE vx X Y Z q")
    vals = ["1 0.5 1.25e+00 -2.5E-1 3.0d0 7", "2 -1 +4 5. .75 0", "1 1e-3 9.999999999999999 1D1 2.5 -0",
            "2 0 1.0000000000000002 -0.0 7.123456789012345678 1", "1 0 3.3e0 4.4E+0 5.5e-0 2",
            "2 0 1234567890123456789e-18 0.1 0.2 3"]
    txt += "\n".join(vals) + "\n"
    out["number_formats"] = txt

    # non-periodic y, atoms leaving the box along it
    traj = walk(rng, 20, 3, L, 1.2)
    txt = ""
    for f, p in enumerate(traj):
        txt += header(f, 20, cub, flags="pp fm pp")
        w = p.copy()
        w[:, 0] = np.mod(w[:, 0], L)
        w[:, 2] = np.mod(w[:, 2], L)
        txt += "".join("%d 1 %.10f %.10f %.10f\n" % (i + 1, *w[i]) for i in range(20))
    out["nonperiodic_y"] = txt

    # an orthogonal header whose rows carry a third number (read as a tilt)
    txt = header(0, 3, cub).replace("%.10f %.10f\n" % (0.0, L), "%.10f %.10f 0.5\n" % (0.0, L), 1)
    txt += "1 1 1.0 2.0 3.0\n2 1 4.0 5.0 6.0\n3 2 7.0 8.0 9.0\n"
    out["orthogonal_three_values"] = txt

    # extra ITEM blocks and blanks before the ATOMS line; CRLF line ends
    txt = ("ITEM: TIMESTEP\n5\nITEM: UNITS\nreal\nITEM: NUMBER OF ATOMS\n2\n   ITEM: BOX BOUNDS pp pp pp\n"
           "0 10\n0 10\n0 10\n\nITEM: ATOMS id type x y z\n1 1 1 2 3\n2 2 4 5 6\n")
    out["extra_items"] = txt
    out["crlf"] = txt.replace("\n", "\r\n")

    # the reference's rejection cases (atom count / order changes)
    for name in ("grow", "reorder"):
        txt = ""
        for frame in (0, 1, 2):
            n = 20 if (name == "grow" and frame == 0) else 40
            txt += "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (frame, n)
            txt += "ITEM: BOX BOUNDS pp pp pp\n0 10\n0 10\n0 10\nITEM: ATOMS id type x y z\n"
            order = range(n, 0, -1) if name == "reorder" and frame > 0 else range(1, n + 1)
            txt += "".join("%d 1 %.6f %.6f %.6f\n" % (k, (k % 10) * 1.0, (k % 7) * 1.3, (k % 5) * 1.7)
                           for k in order)
        out[name] = txt

    base = header(0, 3, cub) + "1 1 1.0 2.0 3.0\n2 1 4.0 5.0 6.0\n3 2 7.0 8.0 9.0\n"
    out["trailing_blank_line"] = base + "\n"
    out["bad_token"] = base.replace("5.0", "5.0abc")
    out["overflow"] = base.replace("5.0", "1e400")
    out["repeat_count"] = base.replace("4.0 5.0 6.0", "3*2.5").replace("1 1 1.0", "2*1 1.0")
    out["short_atom_block"] = base.rsplit("3 2", 1)[0]
    out["missing_type"] = base.replace("id type x y z", "id x y z")
    out["no_newline_at_end"] = base.rstrip("\n")
    out["bad_natoms"] = base.replace("ATOMS\n3", "ATOMS\n-3")
    out["bad_box_row"] = base.replace("0.0000000000 10.0000000000\n", "0.0000000000\n", 1)
    out["no_timestep"] = base.replace("ITEM: TIMESTEP", "ITEM: TIMESTEPS").replace("TIMESTEPS", "STEP")
    out["empty_file"] = ""
    return out


def main():
    if not available():
        raise SystemExit("build the reference reader first: make -C oracle ref")
    fixtures = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, text in cases().items():
            path = os.path.join(tmp, name + ".dump")
            with open(path, "w", newline="") as fh:
                fh.write(text)
            frames, status = ref_read(path)
            fixtures[name] = {
                "text": text,
                "status": status,
                "frames": [{
                    "natoms": f["natoms"], "timestep": f["timestep"],
                    "cell": f["cell"].T.ravel().tolist(),   # column-major
                    "origin": f["origin"].tolist(), "pbc": list(f["pbc"]),
                    "pos": f["pos"].ravel().tolist(), "id": f["id"].tolist(),
                    "type_id": f["type_id"].tolist(), "n_ambiguous": f["n_ambiguous"],
                } for f in frames],
            }
            print("%-24s frames %d status %d" % (name, len(frames), status))
    meta = {"source": "the reference reader lammps_dump_reader_t (src/sqc_dump.f90), "
                      "built by make -C oracle ref and run on each dump"}
    os.makedirs(os.path.join(HERE, "dump"), exist_ok=True)
    with open(os.path.join(HERE, "dump", "dump_reader.json"), "w") as fh:
        json.dump({"meta": meta, "cases": fixtures}, fh)


if __name__ == "__main__":
    main()
