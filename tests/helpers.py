"""Shared helpers of the test-suite: golden fixtures, oracle runs, the
comparison metric of the reference suite, and the layout of the library's
canonical half-spectrum (to compare per-mode amplitudes)."""

import glob
import json
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_dev(reference, candidate):
    """max |c - r| / max(|r|, 1): the metric of the reference suite
    (test/test_sqcalc.py:102-110)."""
    reference = np.asarray(reference, dtype=np.float64)
    candidate = np.asarray(candidate, dtype=np.float64)
    scale = np.maximum(np.abs(reference), 1.0)
    return float(np.max(np.abs(candidate - reference) / scale))


def golden_cases():
    return sorted(os.path.splitext(os.path.basename(p))[0]
                  for p in glob.glob(os.path.join(GOLDEN, "*.json")))


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as fh:
        d = json.load(fh)
    cell, origin = O.cell_from_bounds(d["lo"], d["hi"], d["tilt"])
    mapping = {int(k): v for k, v in (s.split(":") for s in d["mapping"].split(","))}
    d["cell"], d["origin"], d["mapping_dict"] = cell, origin, mapping
    d["frames"] = np.array(d["positions"], dtype=np.float64)
    return d


def oracle_run(cell, origin, types, mapping, weight, norm, qmin, qmax, nq, frames,
               partials=True, faber_ziman=False):
    run = O.Run(cell, origin, types, mapping, weight=weight, norm=norm, qmin=qmin, qmax=qmax,
                nq=nq, partials=partials, faber_ziman=faber_ziman)
    for pos in frames:
        run.add_frame(pos)
    return run


def random_case(natoms=300, length=20.0, nframes=2, seed=0, ntypes=2, tilt=None):
    rng = np.random.default_rng(seed)
    types = rng.integers(1, ntypes + 1, natoms)
    if tilt is None:
        cell = np.diag([length] * 3)
    else:
        cell = np.array([[length, tilt[0], tilt[1]], [0.0, length, tilt[2]], [0.0, 0.0, length]])
    origin = np.zeros(3)
    # positions may lie outside the cell: the path wraps them
    frames = [(rng.random((natoms, 3)) * 1.4 - 0.2) @ cell.T for _ in range(nframes)]
    return cell, origin, types, np.array(frames)


def candidate_layout(md):
    """(h1, h2, h3) of every candidate slot of the library's line table.

    Restates the layout rule of the plan (canonical half: h3 > 0; h3 = 0,
    h2 > 0; h3 = h2 = 0, h1 > 0; per h3 the h2 interval holding kept modes,
    per line the h1 interval of its kept modes) from the oracle's mode set.
    Slots that are not kept modes get h = None.
    """
    kept = set(map(tuple, md["hkl"].tolist()))
    M = [(int(m) - 1) // 2 for m in md["m"]]

    def canon(h1, h2, h3):
        return h3 > 0 or (h3 == 0 and (h2 > 0 or (h2 == 0 and h1 > 0)))

    by_line = {}
    for h in kept:
        if canon(*h):
            by_line.setdefault((h[1], h[2]), []).append(h[0])
    slots = []
    for h3 in range(0, M[2] + 1):
        h2s = [l[0] for l in by_line if l[1] == h3]
        if not h2s:
            continue
        for h2 in range(min(h2s), max(h2s) + 1):
            h1s = by_line.get((h2, h3), [])
            if not h1s:
                continue
            for h1 in range(min(h1s), max(h1s) + 1):
                slots.append((h1, h2, h3) if (h1, h2, h3) in kept and canon(h1, h2, h3)
                             else None)
    return slots


def write_dump(path, frames, types, lo, hi, tilt=(0.0, 0.0, 0.0), scaled=False, ids=None,
               natoms_override=None):
    """LAMMPS text dump of Cartesian frames (N, 3): `x y z`, or `xs ys zs`
    when scaled (fractional coordinates of the restricted triclinic cell).
    The format written by test/gen_dump.py:134-142 (%.10f columns)."""
    cell, origin = O.cell_from_bounds(lo, hi, tilt)
    inv = np.linalg.inv(cell)
    n = len(types)
    triclinic = any(abs(t) > 0 for t in tilt)
    with open(path, "w") as fh:
        for k, pos in enumerate(frames):
            fh.write("ITEM: TIMESTEP\n%d\n" % (100 * k))
            fh.write("ITEM: NUMBER OF ATOMS\n%d\n" % (natoms_override or n))
            if triclinic:
                fh.write("ITEM: BOX BOUNDS xy xz yz pp pp pp\n")
                for i in range(3):
                    fh.write("%.10f %.10f %.10f\n" % (lo[i], hi[i], tilt[i]))
            else:
                fh.write("ITEM: BOX BOUNDS pp pp pp\n")
                for i in range(3):
                    fh.write("%.10f %.10f\n" % (lo[i], hi[i]))
            cols = "xs ys zs" if scaled else "x y z"
            fh.write("ITEM: ATOMS id type %s\n" % cols)
            order = ids[k] if ids is not None else np.arange(1, n + 1)
            for j in range(n):
                r = np.asarray(pos[j], dtype=np.float64)
                v = inv @ (r - origin) if scaled else r
                fh.write("%d %d %.10f %.10f %.10f\n" % (order[j], types[j], v[0], v[1], v[2]))


def read_table(text):
    """(header, rows) of a `# q S(q) ...` table."""
    lines = [l for l in text.splitlines() if l.strip()]
    header = lines[0]
    rows = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    return header, rows
