"""The native LAMMPS dump reader (sqc_dump_* in libsqc_hip.so) against the
reference's own reader.

Host code only, so these run without a GPU.  Pins:
  - tests/golden/dump/dump_reader.json: what the reference reader
    (src/sqc_dump.f90, compiled by ``make -C oracle ref``) returned for
    edge-case dumps; the native reader must agree bit for bit, error codes
    included;
  - the live reference reader (when oracle/_ref/ref_dump is built) on larger
    random trajectories that split the atom block over several threads;
  - geometry: a wrapped dump unwraps to the xu yu zu trajectory up to one
    fixed lattice vector per atom (the reference's own check that wrapped and
    unwrapped dumps give the same result, test/test_sqcalc.py:1505-1512), and
    the scaled columns xs ys zs reproduce the Cartesian path.
Malformed tokens that the reference's Fortran runtime reads leniently
(``5.0x``, ``.e1``, null values ``,,`` and ``/`` that leave stale buffer
contents) are errors here; the fixtures hold only inputs both readers define.
"""

import json
import os

import numpy as np
import pytest

from sqcalc_amd.dump import DumpError, DumpReader

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "dump", "dump_reader.json")


def _cases():
    return json.load(open(GOLDEN))["cases"]


def _read_all(path, nthreads=2):
    frames, status = [], 1
    try:
        with DumpReader(path, nthreads) as rd:
            for fr in rd:
                frames.append(fr)
    except DumpError as exc:
        status = exc.code
    return frames, status


def _write(tmp_path, name, text):
    p = tmp_path / (name + ".dump")
    with open(p, "w", newline="") as fh:
        fh.write(text)
    return p


@pytest.mark.parametrize("name", sorted(_cases()))
def test_matches_reference_fixture(tmp_path, name):
    case = _cases()[name]
    frames, status = _read_all(_write(tmp_path, name, case["text"]))
    assert status == case["status"]
    assert len(frames) == len(case["frames"])
    for ref, fr in zip(case["frames"], frames):
        assert fr.natoms == ref["natoms"] and fr.timestep == ref["timestep"]
        assert np.array_equal(fr.cell.ravel(order="F"), ref["cell"])
        assert np.array_equal(fr.origin, ref["origin"])
        assert list(fr.pbc) == ref["pbc"]
        assert np.array_equal(fr.pos.ravel(), ref["pos"])
        assert np.array_equal(fr.id, ref["id"])
        assert np.array_equal(fr.type_id, ref["type_id"])
        assert fr.n_ambiguous == ref["n_ambiguous"]


def _trajectory(rng, n, frames, L, step, tilt=(0.0, 0.0, 0.0)):
    a = np.array([[L, tilt[0], tilt[1]], [0.0, L, tilt[2]], [0.0, 0.0, L]])
    s = rng.random((n, 3))
    out = []
    for _ in range(frames):
        out.append(s.copy())
        s = s + rng.normal(0.0, step / L, s.shape)
    return a, out


def _bounds(a, lo=(0.0, 0.0, 0.0)):
    xy, xz, yz = a[0, 1], a[0, 2], a[1, 2]
    return [(lo[0] + min(0.0, xy, xz, xy + xz), lo[0] + a[0, 0] + max(0.0, xy, xz, xy + xz)),
            (lo[1] + min(0.0, yz), lo[1] + a[1, 1] + max(0.0, yz)), (lo[2], lo[2] + a[2, 2])]


def _dump_text(a, traj, columns, fmt="%.10f", lo=(0.0, 0.0, 0.0), wrap=True, ragged=False):
    """columns: a permutation of id/type and one coordinate family."""
    names = columns.split()
    b = _bounds(a, lo)
    out = []
    n = len(traj[0])
    for f, s in enumerate(traj):
        out.append("ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n" % (f, n))
        out.append("ITEM: BOX BOUNDS xy xz yz pp pp pp\n")
        tilt = (a[0, 1], a[0, 2], a[1, 2])
        out.extend("%.10f %.10f %.10f\n" % (b[d][0], b[d][1], tilt[d]) for d in range(3))
        out.append("ITEM: ATOMS %s\n" % columns)
        sw = np.mod(s, 1.0) if wrap else s
        r = sw @ a.T + np.asarray(lo)
        cols = {}
        for k, ax in enumerate("xyz"):
            cols[ax] = cols[ax + "u"] = r[:, k]
            cols[ax + "s"] = cols[ax + "su"] = sw[:, k]
        lines = []
        for i in range(n):
            vals = []
            for c in names:
                if c == "id":
                    vals.append("%d" % (i + 1))
                elif c == "type":
                    vals.append("%d" % (1 + i % 3))
                else:
                    vals.append(fmt % cols[c][i])
            sep = "  " if (ragged and (i * 7 + f) % 5 == 0) else " "
            lines.append(sep.join(vals))
        out.append("\n".join(lines) + "\n")
    return "".join(out)


@pytest.mark.parametrize("n,frames,nthreads,ragged", [(60000, 3, 7, False), (45000, 4, 16, True)])
def test_matches_live_reference_reader(tmp_path, n, frames, nthreads, ragged):
    from oracle.ref_dump import available, ref_read
    if not available():
        pytest.skip("reference reader not built (make -C oracle ref)")
    rng = np.random.default_rng(n)
    a, traj = _trajectory(rng, n, frames, 40.0, 0.9, tilt=(2.0, -1.0, 0.5))
    p = _write(tmp_path, "big", _dump_text(a, traj, "id type x y z", ragged=ragged))
    ref, ref_status = ref_read(p)
    mine, status = _read_all(p, nthreads)
    assert status == ref_status == 1
    assert len(mine) == len(ref) == frames
    for r, m in zip(ref, mine):
        assert np.array_equal(r["pos"], m.pos)
        assert np.array_equal(r["id"], m.id) and np.array_equal(r["type_id"], m.type_id)
        assert r["n_ambiguous"] == m.n_ambiguous


def test_thread_count_does_not_change_bits(tmp_path):
    rng = np.random.default_rng(5)
    a, traj = _trajectory(rng, 50000, 3, 30.0, 0.5)
    # the second frame's lines are longer, so the first frame's block size
    # undershoots the next window and the reader extends it
    t0 = _dump_text(a, traj[:1], "id type x y z", fmt="%.6f")
    t1 = _dump_text(a, traj[1:], "id type x y z", fmt="%.12f").replace("TIMESTEP\n0", "TIMESTEP\n7")
    p = _write(tmp_path, "mixed", t0 + t1)
    one, s1 = _read_all(p, 1)
    many, s2 = _read_all(p, 12)
    assert s1 == s2 == 1 and len(one) == len(many) == 3
    for x, y in zip(one, many):
        assert np.array_equal(x.pos, y.pos) and np.array_equal(x.id, y.id)


def test_wrapped_unwraps_to_xu_up_to_a_lattice_vector(tmp_path):
    """The reference reader's contract: x y z come back continuous, so a
    wrapped dump equals the xu yu zu dump minus the origin minus the image of
    the first frame (one lattice vector per atom, the same in every frame)."""
    rng = np.random.default_rng(11)
    lo = (-3.0, 2.0, 1.0)
    a, traj = _trajectory(rng, 3000, 5, 25.0, 1.0, tilt=(1.5, 0.5, -2.0))
    traj = [s + 2.0 for s in traj]   # start several cells away from the origin cell
    wrapped, _ = _read_all(_write(tmp_path, "w", _dump_text(a, traj, "id type x y z", lo=lo)))
    unwr, _ = _read_all(_write(tmp_path, "u", _dump_text(a, traj, "id type xu yu zu", lo=lo,
                                                         wrap=False)))
    assert len(wrapped) == len(unwr) == 5
    ainv = np.linalg.inv(wrapped[0].cell)
    shift0 = None
    for w, u in zip(wrapped, unwr):
        assert u.unwrapped_from_dump and not w.unwrapped_from_dump
        d = (u.pos - u.origin - w.pos) @ ainv.T    # lattice images, fractional
        img = np.round(d)
        assert np.abs(d - img).max() < 1e-8
        if shift0 is None:
            shift0 = img
            assert np.abs(img).max() >= 1
        assert np.array_equal(img, shift0)


def test_scaled_columns_reproduce_cartesian(tmp_path):
    rng = np.random.default_rng(3)
    a, traj = _trajectory(rng, 2000, 3, 20.0, 0.8, tilt=(1.0, 0.0, -0.5))
    lo = (1.0, -2.0, 0.5)
    cart, _ = _read_all(_write(tmp_path, "c", _dump_text(a, traj, "id type x y z", lo=lo,
                                                         fmt="%.15e")))
    sc, _ = _read_all(_write(tmp_path, "s", _dump_text(a, traj, "id type xs ys zs", lo=lo,
                                                       fmt="%.15e")))
    scu, _ = _read_all(_write(tmp_path, "su", _dump_text(a, traj, "id type xsu ysu zsu", lo=lo,
                                                         wrap=False, fmt="%.15e")))
    for c, s, su, st in zip(cart, sc, scu, traj):
        assert s.scaled and not c.scaled
        assert np.abs(c.pos - s.pos).max() < 1e-11
        assert su.unwrapped_from_dump
        assert np.abs(su.pos - (st @ a.T + np.asarray(lo))).max() < 1e-11


def test_read_frames_batches_and_checks(tmp_path):
    rng = np.random.default_rng(9)
    a, traj = _trajectory(rng, 500, 7, 15.0, 0.4)
    text = _dump_text(a, traj, "id type x y z")
    p = _write(tmp_path, "b", text)
    frames, _ = _read_all(p)
    with DumpReader(p, 4) as rd:
        buf = np.zeros((3, 500, 3))
        got = []
        while True:
            n, hdr = rd.read_frames(buf, 500, frames[0].cell)
            got.extend(buf[:n].copy())
            if n < 3:
                break
    assert len(got) == 7
    for g, f in zip(got, frames):
        assert np.array_equal(g, f.pos)
    # a box change (app/main.f90:244-247) and an atom-count change with
    # unwrapped columns (app/main.f90:240-243) end the batch
    b2 = text.split("ITEM: TIMESTEP")
    b2[3] = b2[3].replace("%.10f %.10f" % tuple(_bounds(a)[0]),
                          "%.10f %.10f" % (_bounds(a)[0][0], _bounds(a)[0][1] + 0.5))
    p2 = _write(tmp_path, "box", "ITEM: TIMESTEP".join(b2))
    with DumpReader(p2, 2) as rd:
        with pytest.raises(DumpError) as exc:
            rd.read_frames(np.zeros((7, 500, 3)), 500, frames[0].cell)
        assert exc.value.code == 13
    grow = _dump_text(a, [s[:400] for s in traj[:2]], "id type xu yu zu", wrap=False)
    grow += _dump_text(a, traj[2:3], "id type xu yu zu", wrap=False)
    with DumpReader(_write(tmp_path, "grow", grow), 2) as rd:
        with pytest.raises(DumpError) as exc:
            rd.read_frames(np.zeros((4, 400, 3)), 400)
        assert exc.value.code == 12


def test_missing_file_and_closed_reader(tmp_path):
    with pytest.raises(DumpError) as exc:
        DumpReader(tmp_path / "nope.dump")
    assert exc.value.code == 2


def test_read_frames_rejects_unsafe_buffers(tmp_path):
    rng = np.random.default_rng(2)
    a, traj = _trajectory(rng, 50, 2, 10.0, 0.3)
    p = _write(tmp_path, "s", _dump_text(a, traj, "id type x y z"))
    with DumpReader(p, 1) as rd:
        for bad in (np.zeros((2, 50, 3), dtype=np.float32), np.zeros((2, 3, 50)).transpose(0, 2, 1),
                    np.zeros((2, 51, 3)), np.zeros((100, 3))):
            with pytest.raises(ValueError):
                rd.read_frames(bad, 50)
        n, _ = rd.read_frames(np.zeros((2, 50, 3)), 50)
        assert n == 2


def test_background_batch_read_matches_synchronous(tmp_path):
    rng = np.random.default_rng(13)
    a, traj = _trajectory(rng, 3000, 7, 18.0, 0.6)
    p = _write(tmp_path, "bg", _dump_text(a, traj, "id type x y z"))
    frames, _ = _read_all(p)
    with DumpReader(p, 3) as rd:
        got = []
        bufs = [np.zeros((3, 3000, 3)), np.zeros((3, 3000, 3))]
        k = 0
        rd.read_begin(bufs[0], 3000, frames[0].cell)
        while True:
            n = rd.read_end()
            got.extend(bufs[k % 2][:n].copy())
            if n < 3:
                break
            k += 1
            rd.read_begin(bufs[k % 2], 3000, frames[0].cell)
    assert len(got) == 7
    for g, f in zip(got, frames):
        assert np.array_equal(g, f.pos)
    # one batch in flight: a second begin before end is refused
    with DumpReader(p, 2) as rd:
        buf = np.zeros((2, 3000, 3))
        rd.read_begin(buf, 3000)
        with pytest.raises(DumpError):
            rd.read_begin(np.zeros((2, 3000, 3)), 3000)
        assert rd.read_end() == 2
