"""Fortran host (sqcalc_amd/fortran): options, exit codes and scattering
tables on the CPU; the S(q) tables it writes on the GPU.

The exit codes and messages follow `sqcalc static` (app/main.f90: 2 bad
options, 4 unreadable dump, 5 no frame, 6 set-up, 7 malformed frame, 8 atom
count change, 9 box change, 24 id order change).  Dumps are written from the
golden fixtures, the parity reference is the golden table (FP64 within
1e-9) and the Python host on the same plan settings.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import golden_cases, load_golden, read_table, rel_dev, write_dump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sqcalc_amd", "fortran", "sqcalc_hip")
HAVE_FLANG = shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang")


def run(args, **kw):
    return subprocess.run([EXE] + list(args), capture_output=True, text=True, timeout=300, **kw)


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        if HAVE_FLANG:
            pytest.fail("sqcalc_hip is not built (run __graft_entry__.build())")
        pytest.skip("no Fortran compiler and no sqcalc_hip binary")
    return EXE


def test_element_tables_match_reference_data(exe):
    """`sqcalc_hip elements` lists what the Fortran host weights with: every
    tabulated neutron length and IT92 row, against the reference's data
    (tests/golden/elements/elements.json)."""
    import json
    from helpers import GOLDEN
    with open(os.path.join(GOLDEN, "elements", "elements.json")) as fh:
        d = json.load(fh)
    r = run(["elements"])
    assert r.returncode == 0, r.stderr
    neutron = {e["symbol"]: e["neutron_b"] for e in d["elements"] if e["has_neutron"]}
    xray = {sp["label"]: sp["coef"] for sp in d["species"]}
    seen_n, seen_x = set(), set()
    for line in r.stdout.splitlines():
        f = line.split()
        if f[0] == "neutron":
            assert float(f[2]) == pytest.approx(neutron[f[1]], rel=1e-15)
            seen_n.add(f[1])
        elif f[0] == "xray":
            np.testing.assert_allclose([float(x) for x in f[2:]], xray[f[1]], rtol=1e-15)
            seen_x.add(f[1])
    assert seen_n == set(neutron) and len(seen_n) == 86
    assert seen_x == set(xray) and len(seen_x) == 213


@pytest.mark.parametrize("weight,label,text", [
    ("xray", "Es", "no X-ray form factor for Es; use --weight neutron or unit"),
    ("neutron", "Fm", "no neutron scattering length for Fm; use --weight xray or unit"),
])
def test_species_without_data(exe, tmp_path, weight, label, text):
    """The reference's validate_mapping (app/main.f90:515-551): a label that
    resolves but has no data for the weights stops with code 15."""
    d = load_golden("gas_unit_self")
    p = tmp_path / "g.dump"
    write_dump(p, d["frames"], np.ones_like(d["types"]), d["lo"], d["hi"], d["tilt"])
    r = run(["static", "-i", str(p), "-m", "1:" + label, "-w", weight, "--quiet"])
    assert r.returncode == 15, (r.returncode, r.stderr)
    assert text in r.stderr


@pytest.mark.parametrize("args,code,text", [
    ([], 2, "subcommand"),
    (["dynamic"], 2, "subcommand"),
    (["static"], 2, "missing input"),
    (["static", "-i", "x.dump", "--bogus"], 2, "unknown option"),
    (["static", "-i", "x.dump", "--device", "cpu"], 2, "GPU path only"),
    (["static", "-i", "x.dump", "-w", "photon"], 2, "unknown weight"),
    (["static", "-i", "x.dump", "-w", "neutron"], 2, "mapping"),
    # the reference's label failures (src/sqc_elements.f90:241-291)
    (["static", "-i", "x.dump", "-m", "1:Xq"], 2,
     'unknown element symbol "Xq" in the element mapping'),
    (["static", "-i", "x.dump", "-m", "1:Si,2:Fe7+"], 2,
     'element "Fe" has no charge state "Fe7+"; available: Fe, Fe2+, Fe3+'),
    (["static", "-i", "x.dump", "--qmin", "5", "--qmax", "4"], 2, "larger than qmin"),
    (["static", "-i", "x.dump", "--precision", "half"], 2, "unknown precision"),
    (["static", "-i", "x.dump", "--tiles", "8,2"], 2, "--tiles takes four positive integers"),
    (["static", "-i", "x.dump", "--tiles", "8,2,0,1"], 2, "--tiles takes four positive integers"),
    (["static", "-i", "/nonexistent/x.dump"], 4, "cannot open"),
    # --xrd rules of src/sqc_options.f90:620-660
    (["static", "-i", "x.dump", "--xrd", "p.txt"], 2, "--xrd-lambda"),
    (["static", "-i", "x.dump", "--xrd", "p.txt", "--xrd-lambda", "1.54", "--qmax", "5"], 2,
     "do not pass --qmin or --qmax"),
    (["static", "-i", "x.dump", "--xrd-lambda", "1.54"], 2, "belong to --xrd FILE"),
    (["static", "-i", "x.dump", "--xrd", "p.txt", "--xrd-lambda", "1.54", "--xrd-range", "0",
      "90"], 2, "0 < MIN2TH < MAX2TH < 180"),
])
def test_option_errors(exe, args, code, text):
    r = run(args)
    assert r.returncode == code, (r.returncode, r.stderr)
    assert text in r.stderr


def test_empty_dump(exe, tmp_path):
    p = tmp_path / "empty.dump"
    p.write_text("")
    r = run(["static", "-i", str(p)])
    assert r.returncode == 5 and "no complete frame" in r.stderr


def test_setup_without_device(exe, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    d = load_golden("gas_unit_self")
    p = tmp_path / "g.dump"
    write_dump(p, d["frames"], d["types"], d["lo"], d["hi"], d["tilt"])
    r = run(["static", "-i", str(p), "--quiet"])
    assert r.returncode == 6, r.stderr
    assert "device" in r.stderr.lower()


# --------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("tiles", [None, "4,2,4,1"])
def test_golden_tables_through_fortran(exe, hip_device, tmp_path, name, tiles):
    """Default launch shapes, and pinned ones (--tiles: spread tile 4 x 2,
    y-pass chunk 4, one x line per group)."""
    d = load_golden(name)
    scaled = name.startswith("tri")
    p = tmp_path / (name + ".dump")
    write_dump(p, d["frames"], d["types"], d["lo"], d["hi"], d["tilt"], scaled=scaled)
    out = tmp_path / "sq.txt"
    print("case fortran golden", name, tiles, flush=True)
    r = run(["static", "-i", str(p), "-m", d["mapping"], "-w", d["weight"], "--norm", d["norm"],
             "--qmin", repr(d["qmin"]), "--qmax", repr(d["qmax"]), "--nq", str(d["nq"]),
             "--eps", "1e-10", "--no-partials", "--quiet", str(out)]
            + (["--tiles", tiles] if tiles else []))
    assert r.returncode == 0, r.stderr
    header, rows = read_table(out.read_text())
    assert header.strip() == "# q S(q)"
    np.testing.assert_allclose(rows[:, 0], np.round(np.array(d["q"]), 6), atol=1e-12)
    dev = rel_dev(np.array(d["S"]), rows[:, 1])
    print("  max deviation %.2e" % dev)
    assert dev < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("precision,fz", [("double", False), ("single", True)])
def test_partials_match_python_host(exe, hip_device, tmp_path, precision, fz):
    from sqcalc_amd import HipStructureFactor
    from helpers import random_case
    cell, origin, types, frames = random_case(natoms=600, length=18.0, nframes=3, seed=41,
                                              ntypes=3)
    lo, hi = [0.0] * 3, [18.0] * 3
    p = tmp_path / "r.dump"
    write_dump(p, frames, types, lo, hi)
    out = tmp_path / "sq.txt"
    grid = tmp_path / "grid.txt"
    args = ["static", "-i", str(p), "-m", "1:Si,2:O,3:Na", "-w", "xray", "--qmax", "5",
            "--nq", "40", "--precision", precision, "--grid", str(grid), "--quiet", str(out)]
    if fz:
        args.insert(1, "-fz")
    print("case fortran partials", precision, fz, flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    header, rows = read_table(out.read_text())
    # the dump rounds positions to 1e-10
    frames_r = np.round(frames, 10)
    sf = HipStructureFactor(cell, origin, types, {1: "Si", 2: "O", 3: "Na"}, weight="xray",
                            qmax=5.0, nq=40, precision=precision, faber_ziman=fz,
                            want_grid=True)
    sf.accumulate(frames_r)
    ref_header = sf.format_table().splitlines()[0]
    assert header == ref_header
    tol = 1e-12 if precision == "double" else 1e-6
    assert rel_dev(sf.table(), rows) < 1e-11 + tol
    qv, gs = sf.grid_table()
    g = np.loadtxt(grid)
    np.testing.assert_allclose(g[:, :3], np.round(qv, 8), atol=1e-12)
    assert rel_dev(gs, g[:, 3]) < 1e-11 + tol
    sf.close()


@pytest.mark.gpu
@pytest.mark.parametrize("change,code", [("natoms", 25), ("box", 9), ("order", 24),
                                         ("truncated", 7)])
def test_frame_consistency_errors(exe, hip_device, tmp_path, change, code):
    """Exit codes of the reference driver (app/main.f90:225-247).  A changed
    atom count in an x y z dump is caught by the reader's unwrapping state
    first (ierr 11, exit 25), as in the reference."""
    d = load_golden("gas_unit_self")
    frames, types = d["frames"], d["types"]
    p = tmp_path / "bad.dump"
    write_dump(p, frames[:1], types, d["lo"], d["hi"], d["tilt"])
    n = len(types)
    second = tmp_path / "second.dump"
    if change == "natoms":
        write_dump(second, frames[1:2, :-1], types[:-1], d["lo"], d["hi"], d["tilt"])
    elif change == "box":
        hi = [h + 1.0 for h in d["hi"]]
        write_dump(second, frames[1:2], types, d["lo"], hi, d["tilt"])
    elif change == "order":
        ids = [np.concatenate([[2, 1], np.arange(3, n + 1)])]
        write_dump(second, frames[1:2], types, d["lo"], d["hi"], d["tilt"], ids=ids)
    else:
        write_dump(second, frames[1:2], types, d["lo"], d["hi"], d["tilt"])
        text = second.read_text().splitlines()[:-3]
        second.write_text("\n".join(text) + "\n")
    with open(p, "a") as fh:
        fh.write(second.read_text())
    print("case fortran frame error", change, flush=True)
    r = run(["static", "-i", str(p), "--quiet", str(tmp_path / "o.txt")])
    assert r.returncode == code, (r.returncode, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("partials,step,lp", [(True, 0.0, True), (False, 0.0, True),
                                              (True, 0.5, False)])
def test_xrd_table_matches_python_host(exe, hip_device, tmp_path, partials, step, lp):
    """--xrd through the command against HipStructureFactor.format_xrd on the
    frames the reader returned: same header, bins and columns to 1e-9."""
    from sqcalc_amd import HipStructureFactor, xrd
    from sqcalc_amd.dump import DumpReader
    from helpers import random_case
    cell, origin, types, frames = random_case(natoms=500, length=14.0, nframes=3, seed=7)
    p = tmp_path / "x.dump"
    write_dump(p, frames, types, [0.0] * 3, [14.0] * 3)
    out = tmp_path / "xrd.txt"
    args = ["static", "-i", str(p), "-m", "1:Si,2:O", "--xrd", str(out), "--xrd-lambda",
            "1.541838", "--xrd-range", "12", "125", "--quiet", str(tmp_path / "sq.txt")]
    if not partials:
        args.insert(1, "--no-partials")
    if step > 0:
        args[1:1] = ["--xrd-step", str(step)]
    if not lp:
        args.insert(1, "--no-lp")
    print("case fortran xrd", partials, step, lp, flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    with DumpReader(p, 2) as rd:
        read = list(rd)
    qmin, qmax = xrd.q_range(1.541838, 12.0, 125.0)
    sf = HipStructureFactor(read[0].cell, read[0].origin, read[0].type_id, {1: "Si", 2: "O"},
                            weight="xray", qmin=qmin, qmax=qmax, partials=partials,
                            want_grid=True, mode_partials=partials)
    sf.accumulate(np.stack([fr.pos for fr in read]))
    want = sf.format_xrd(1.541838, 12.0, 125.0, step=step, lp=lp).splitlines()
    got = out.read_text().splitlines()
    nhead = 4 if partials else 3
    assert got[:nhead] == want[:nhead]
    assert len(got) == len(want)
    g = np.array([[float(x) for x in l.split()] for l in got[nhead:]])
    w = np.array([[float(x) for x in l.split()] for l in want[nhead:]])
    np.testing.assert_allclose(g[:, 0], w[:, 0], atol=1e-9)
    assert np.abs(g[:, 1:] - w[:, 1:]).max() <= 1e-9 * np.abs(w[:, 1:]).max()
    sf.close()


@pytest.mark.gpu
def test_xrd_fcc_nickel_lines(exe, hip_device, tmp_path):
    """Closed-form powder pattern: a perfect fcc Ni crystal (a = 3.52 A, 20^3
    cells, 32,000 atoms) at Cu K-alpha over 40-80 deg in 0.2 deg bins, the
    case the reference's own suite runs (test/test_sqcalc.py:936-1000).  Only
    the (111), (200) and (220) reflections fall in the range; their bin
    intensities relative to (111) must follow multiplicity x |f_Ni(q)|^2 x LP
    (the structure factor of fcc is 4 f for all-even or all-odd hkl), every
    other bin must stay empty, and with one species the pair column is the
    total."""
    from sqcalc_amd import elements, xrd
    a, side, lam = 3.52, 20, 1.541838
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    n = len(pos)
    p = tmp_path / "ni.dump"
    with open(p, "w") as fh:
        fh.write("ITEM: TIMESTEP\n0\nITEM: NUMBER OF ATOMS\n%d\nITEM: BOX BOUNDS pp pp pp\n" % n)
        fh.write("0 %.10f\n" % (a * side) * 3)
        fh.write("ITEM: ATOMS id type x y z\n")
        fh.write("".join("%d 1 %.10f %.10f %.10f\n" % (i + 1, *pos[i]) for i in range(n)))
    out = tmp_path / "ni.xrd"
    print("case fortran xrd fcc Ni", flush=True)
    r = run(["static", "-i", str(p), "-m", "1:Ni", "-w", "xray", "--xrd", str(out),
             "--xrd-lambda", str(lam), "--xrd-range", "40", "80", "--xrd-step", "0.2",
             "--quiet", str(tmp_path / "sq.txt")])
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert lines[0].startswith("# xrd: lambda 1.541838 A") and "200 bins" in lines[0]
    assert "Lorentz-polarization factor" in lines[1]
    assert lines[3].split() == ["#", "2theta[deg]", "I", "I(Ni-Ni)"]
    tab = np.array([[float(v) for v in l.split()] for l in lines[4:]])
    assert tab.shape == (200, 3)
    assert np.abs(tab[:, 1] - tab[:, 2]).max() <= 1e-9 * np.abs(tab[:, 1]).max()
    c = elements.xray_coefficients("Ni")

    def f_ni(q):
        st = (q / (4.0 * np.pi)) ** 2
        return c[8] + sum(c[k] * np.exp(-c[4 + k] * st) for k in range(4))

    want, got = [], []
    for hkl, mult in (((1, 1, 1), 8), ((2, 0, 0), 6), ((2, 2, 0), 12)):
        q = np.sqrt(sum(h * h for h in hkl)) * 2.0 * np.pi / a
        tth = float(xrd.two_theta_deg(q, lam))
        j = int(np.argmin(np.abs(tab[:, 0] - tth)))
        assert abs(tab[j, 0] - tth) <= 0.1 + 1e-9, (hkl, tth, tab[j, 0])
        want.append(mult * f_ni(q) ** 2 * float(xrd.lorentz_polarization(tth)))
        got.append(tab[j, 1])
    for w, g in zip(want[1:], got[1:]):
        assert abs(g / got[0] - w / want[0]) <= 0.01 * (w / want[0]), (g / got[0], w / want[0])
    assert np.sort(tab[:, 1])[-4] <= 1e-6 * tab[:, 1].max()


@pytest.mark.gpu
@pytest.mark.parametrize("weight", ["neutron", "xray"])
def test_single_species_table_matches_python_host(exe, hip_device, tmp_path, weight):
    """One atom type: the S(q) table and its S(Ni-Ni) column through the
    command against the Python host on the frames the reader returned."""
    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.dump import DumpReader
    from helpers import random_case
    cell, origin, types, frames = random_case(natoms=400, length=15.0, nframes=2, seed=3,
                                              ntypes=1)
    p = tmp_path / "one.dump"
    write_dump(p, frames, types, [0.0] * 3, [15.0] * 3)
    out = tmp_path / "sq.txt"
    print("case fortran single species", weight, flush=True)
    r = run(["static", "-i", str(p), "-m", "1:Ni", "-w", weight, "--qmax", "5", "--nq", "40",
             "--quiet", str(out)])
    assert r.returncode == 0, r.stderr
    header, rows = read_table(out.read_text())
    assert header.split()[:4] == ["#", "q", "S(q)", "S(Ni-Ni)"]
    with DumpReader(p, 2) as rd:
        read = list(rd)
    sf = HipStructureFactor(read[0].cell, read[0].origin, read[0].type_id, {1: "Ni"},
                            weight=weight, qmax=5.0, nq=40)
    sf.accumulate(np.stack([fr.pos for fr in read]))
    ref = sf.table()
    assert np.abs(ref[:, 2]).max() > 0
    assert rel_dev(ref, rows) < 1e-9
    sf.close()


@pytest.mark.gpu
def test_c1_scaled_dump_through_fortran(exe, hip_device, tmp_path):
    """The CPU-runnable configuration as stated: cube L = 20, N = 1000 (333 Si,
    667 O), 10 frames written as a LAMMPS dump with scaled coordinates (xs ys
    zs), unit weights, norm mean, partials, q_max 10, NUFFT tolerance 1e-9,
    through `sqcalc_hip static` in FP64: every column within 1e-9 of the
    oracle's direct sum over the positions the dump holds (the fractional
    coordinates as written, %.10f)."""
    from helpers import oracle_run
    from oracle import oracle as O
    from sqcalc_amd.synthetic import config
    wl = config("c1")
    frames = wl.frames()
    lo, hi = [0.0] * 3, [20.0] * 3
    p = tmp_path / "c1.dump"
    write_dump(p, frames, wl.types, lo, hi, scaled=True)
    cell, origin = O.cell_from_bounds(lo, hi)
    inv = np.linalg.inv(cell)
    held = []
    for f in frames:
        frac = (f - origin) @ inv.T
        frac = np.array([float("%.10f" % v) for v in frac.ravel()]).reshape(frac.shape)
        held.append(frac @ cell.T + origin)
    out = tmp_path / "sq.txt"
    print("case c1 fortran scaled dump", flush=True)
    r = run(["static", "-i", str(p), "-m", "1:Si,2:O", "-w", "unit", "--norm", "mean",
             "--qmax", "10", "--nq", "500", "--eps", "1e-9", "--quiet", str(out)])
    assert r.returncode == 0, r.stderr
    header, rows = read_table(out.read_text())
    assert header.split()[:3] == ["#", "q", "S(q)"] and rows.shape == (500, 5)
    ref = oracle_run(cell, origin, wl.types, wl.mapping, "unit", "mean", 0.0, 10.0, 500,
                     held).table()
    np.testing.assert_allclose(rows[:, 0], np.round(ref[:, 0], 6), atol=1e-12)
    dev = rel_dev(ref[:, 1:], rows[:, 1:])
    print("  max deviation %.2e" % dev)
    assert dev < 1e-9
