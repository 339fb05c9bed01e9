"""The Debye oracle (sqo_debye_positions / sqo_debye_hist of
oracle/sqc_oracle.c, debye_sq / debye_g of oracle/oracle.py) pinned against
the reference's own cell list (oracle/_ref/ref_debye: src/sqc_neighbour_list.f90
compiled where it lies, with the pair loop of debye_accumulate,
src/sqc_debye.f90:208-226), the cases of the reference's
test/test_neighbour_list.f90 (orthogonal, triclinic, slab, cluster, small box
with several images, skin reuse over a rattled trajectory), closed forms
(lattice shells on exact bin edges, the partial sum rule, the ideal gas
S(q) -> 1 of test/test_sqcalc.py:485-491), and the host-side evaluation of
sqcalc_amd.debye against the oracle's.  CPU only."""

import math

import numpy as np
import pytest

from debye_cases import CASES, lattice, make_cell, oracle_hist, run_case, trajectory
from oracle import oracle as O
from oracle import ref_debye as R

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref/ref_debye not built")


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_cell_list(name):
    cell, origin, pbc, types, T, rmax, skin, frames = run_case(name)
    run = oracle_hist(cell, origin, pbc, types, T, rmax, 0.01, skin, frames)
    h, nbuilds, _ = R.ref_hist(cell, origin, types, T, pbc, rmax, 0.01, skin, frames)
    assert run.hist.sum() > 0
    assert np.array_equal(run.hist.astype(np.float64), h)
    if name == "skin_reuse":
        assert nbuilds < len(frames)       # the list was reused: shifts kept across frames


@needs_ref
def test_oracle_matches_reference_unwrapped_crossings():
    """Atoms handed over unwrapped and crossing the box many times: the list's
    smallest-image carry (nl_begin_frame) and the stored shifts."""
    rng = np.random.default_rng(11)
    cell = make_cell((12.0, 13.0, 14.0), (1.0, 0.5, -0.75))
    frames = trajectory(rng, cell, np.zeros(3), 150, 8, 1.2, wrap=False)
    types = rng.integers(1, 3, 150).astype(np.int32)
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, 5.5, 0.02, 0.5, frames)
    h, _, _ = R.ref_hist(cell, np.zeros(3), types, 2, (1, 1, 1), 5.5, 0.02, 0.5, frames)
    assert np.array_equal(run.hist.astype(np.float64), h)


def test_lattice_shells_on_bin_edges():
    """Simple cubic lattice, spacing 2, box 10: every distance sits exactly
    on a bin edge (dr = 0.5).  The fourth shell lies exactly at rmax = 4:
    d2 > rmax^2 is the rejection, so it is kept, but with rmax / dr whole its
    bin int(4 / 0.5) + 1 is past nbins; with dr = 0.3 it lands in the last bin."""
    pos = lattice(5, 2.0) + 0.5
    n = len(pos)
    types = np.ones(n, dtype=np.int32)
    cell = make_cell((10.0, 10.0, 10.0))
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 1, 4.0, 0.5, 1.0, [pos])
    h = run.hist[0, 0]
    assert len(h) == 8
    expect = np.zeros(8, dtype=np.uint64)
    expect[int(2.0 / 0.5)] = 6 * n                          # d = 2
    expect[int(math.sqrt(8.0) / 0.5)] = 12 * n              # d = 2 sqrt 2
    expect[int(math.sqrt(12.0) / 0.5)] = 8 * n              # d = 2 sqrt 3
    assert int(4.0 / 0.5) == 8                              # d = 4 would be bin 9 > nbins
    assert np.array_equal(h, expect)
    # d = 4 exactly at rmax is kept when its bin exists (rmax / dr not whole)
    run2 = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 1, 4.0, 0.3, 1.0, [pos])
    assert run2.hist.shape[2] == 14
    assert int(run2.hist[0, 0][13]) == 6 * n
    assert int(run2.hist.sum()) == (6 + 12 + 8 + 6) * n


@needs_ref
def test_lattice_matches_reference():
    pos = lattice(5, 2.0) + 0.5
    types = (np.arange(len(pos)) % 2 + 1).astype(np.int32)
    cell = make_cell((10.0, 10.0, 10.0))
    for dr, rmax in ((0.5, 4.0), (0.4, 4.0), (0.25, 4.5)):
        run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, rmax, dr, 1.0, [pos])
        h, _, _ = R.ref_hist(cell, np.zeros(3), types, 2, (1, 1, 1), rmax, dr, 1.0, [pos])
        assert np.array_equal(run.hist.astype(np.float64), h)


@pytest.mark.parametrize("name", ["small_box", "triclinic", "slab"])
def test_image_bound_matches_brute_force(name):
    cell, origin, pbc, types, T, rmax, skin, frames = run_case(name, seed=5)
    a = oracle_hist(cell, origin, pbc, types, T, rmax, 0.01, skin, frames, method=1)
    b = oracle_hist(cell, origin, pbc, types, T, rmax, 0.01, skin, frames, method=0)
    assert np.array_equal(a.hist, b.hist)


def test_histogram_symmetry_and_ideal_count():
    cell, origin, pbc, types, T, rmax, skin, frames = run_case("orthogonal", seed=9)
    run = oracle_hist(cell, origin, pbc, types, T, rmax, 0.05, skin, frames[:1])
    h = run.hist.astype(np.int64)
    assert np.array_equal(h, h.transpose(1, 0, 2))           # n_ab = n_ba
    n = len(types)
    expect = n * (n - 1) * 4.0 / 3.0 * math.pi * rmax ** 3 / 8000.0
    assert abs(h.sum() - expect) < 6.0 * math.sqrt(expect)


def minimum_image_distances(pos, cell, rmax):
    s = np.linalg.solve(cell, pos.T).T
    d = s[:, None, :] - s[None, :, :]
    d -= np.rint(d)
    r = np.linalg.norm(d @ cell.T, axis=-1)
    i, j = np.nonzero((r <= rmax) & ~np.eye(len(pos), dtype=bool))
    return i, j, r[i, j]


def test_sq_is_the_debye_sum_of_the_histogram():
    """S(q) from the oracle's histogram and debye_sq against the Debye sum
    over bin centres of an independent minimum-image histogram (rmax below
    half the box: one image per pair), and the sum rule
    S = sum_ab (2 - delta_ab) S_ab under unit weights."""
    rng = np.random.default_rng(21)
    cell = make_cell((20.0, 20.0, 20.0))
    pos = rng.uniform(0.0, 20.0, (400, 3))
    types = rng.integers(1, 3, 400).astype(np.int32)
    rmax, dr = 6.0, 0.01
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, rmax, dr, 1.0, [pos])
    i, j, r = minimum_image_distances(pos, cell, rmax)
    nb = int(math.ceil(rmax / dr))
    href = np.zeros((2, 2, nb))
    np.add.at(href, (types[i] - 1, types[j] - 1, (r / dr).astype(int)), 1.0)
    # a distance within one rounding of a bin edge may land either side
    assert np.abs(run.hist - href).sum() <= 2
    counts = np.bincount(types, minlength=3)[1:]
    for norm in ("n", "self", "mean"):
        num, den, part = O.debye_sq(run.hist, 1, counts, 400, 8000.0, (1, 1, 1), rmax, dr, 0.0,
                                    6.0, 30, "unit", np.ones(2), np.zeros((2, 9)), norm=norm,
                                    correct_cutoff=False)
        s = num / den
        q = (np.arange(30) + 0.5) * 0.2
        rk = (np.arange(nb) + 0.5) * dr
        direct = np.array([np.sum(href.sum(axis=(0, 1)) * np.sin(qq * rk) / (qq * rk))
                           for qq in q])
        assert np.allclose(s, (direct + 400.0) / 400.0, rtol=1e-9, atol=2e-2)
        if norm == "n":
            total = part[0, 0] + 2.0 * part[0, 1] + part[1, 1]
            assert np.allclose(total / 400.0, s, rtol=1e-12, atol=1e-12)


def test_corrected_ideal_gas_tends_to_one():
    """test/test_sqcalc.py:485-491: after debyer's cut-off correction an ideal
    gas has S(q) -> 1; rms(S - 1) for q > 1.5 below 0.1."""
    rng = np.random.default_rng(2)
    cell = make_cell((20.0, 20.0, 20.0))
    frames = [rng.uniform(0.0, 20.0, (1000, 3)) for _ in range(3)]
    types = np.where(np.arange(1000) % 3 == 0, 1, 2).astype(np.int32)
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, 6.0, 0.01, 1.0, frames)
    counts = np.bincount(types, minlength=3)[1:]
    num, den, _ = O.debye_sq(run.hist, 3, counts, 1000, 8000.0, (1, 1, 1), 6.0, 0.01, 0.0, 6.0,
                             30, "unit", np.ones(2), np.zeros((2, 9)), norm="self")
    q = (np.arange(30) + 0.5) * 0.2
    s = num / (3 * den)
    rms = math.sqrt(float(np.mean((s[q > 1.5] - 1.0) ** 2)))
    assert rms < 0.1


def test_host_evaluation_matches_oracle():
    """sqcalc_amd.debye.shell_sums (the product's host-side prepare_output)
    against the oracle's debye_sq for every norm, weights and partials."""
    from sqcalc_amd import debye as D
    rng = np.random.default_rng(4)
    T, nb, nh = 3, 500, 4
    hist = rng.integers(0, 50, (T, T, nb)).astype(np.float64)
    hist = hist + hist.transpose(1, 0, 2)
    counts = np.array([40, 0, 80])
    amp = np.array([4.1491, 0.0, 5.803])
    xcoef = np.array([[6.2915, 3.0353, 1.9891, 1.541, 2.4386, 32.3337, 0.6785, 81.6937, 1.1407],
                      [0.0] * 9,
                      [3.0485, 2.2868, 1.5463, 0.867, 13.2771, 5.7011, 0.3239, 32.9089, 0.2508]])
    for kind in ("unit", "neutron", "xray"):
        def ampl(t, q, kind=kind):
            return float(O.amplitude(kind, amp, xcoef, t, q))
        for norm in ("mean", "self", "n"):
            for cut in (True, False):
                q, num, den, part = D.shell_sums(hist, nh, counts, 9000.0, (1, 1, 1), 5.0, 0.01,
                                                 ampl, norm, True, cut, 0.0, 8.0, 40)
                rnum, rden, rpart = O.debye_sq(hist, nh, counts, 120, 9000.0, (1, 1, 1), 5.0,
                                               0.01, 0.0, 8.0, 40, kind, amp, xcoef, norm=norm,
                                               correct_cutoff=cut)
                assert np.allclose(den, rden, rtol=1e-13)
                assert np.allclose(num, rnum, rtol=1e-11, atol=1e-9 * np.abs(rnum).max())
                assert np.allclose(part, rpart, rtol=1e-11, atol=1e-9 * np.abs(rpart).max())


def test_plan_create_rejects_bad_settings():
    """The set-up checks of debye_setup (src/sqc_debye.f90:95-177) run before
    any device is touched."""
    import ctypes
    from sqcalc_amd import _lib
    lib = _lib.load()
    types = np.array([1, 2, 1, 2], dtype=np.int32)

    def create(**kw):
        cfg = _lib.DebyeConfig()
        cfg.abi_version = kw.get("abi", _lib.ABI_VERSION)
        cfg.natoms = 4
        cfg.ntypes = kw.get("ntypes", 2)
        tt = kw.get("types", types)
        cfg.type_id = tt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        for k in range(3):
            cfg.pbc[k] = kw.get("pbc", (1, 1, 1))[k]
        cell = kw.get("cell", make_cell((10.0, 12.0, 14.0)))
        for i, v in enumerate(cell.ravel(order="F")):
            cfg.cell[i] = v
        cfg.rmax = kw.get("rmax", 4.0)
        cfg.dr = kw.get("dr", 0.01)
        cfg.skin = 1.0
        plan = ctypes.c_void_p()
        rc = lib.sqc_hip_debye_plan_create(ctypes.byref(plan), ctypes.byref(cfg))
        assert not plan.value
        return rc, lib.sqc_hip_last_error().decode()

    assert create(abi=99)[0] == 30
    rc, msg = create(types=np.array([1, 3, 1, 2], dtype=np.int32))
    assert rc == 20 and "outside" in msg
    rc, msg = create(rmax=10.0)
    assert rc == 20 and "smallest periodic box side" in msg
    assert create(rmax=10.0, pbc=(0, 1, 1))[0] != 20 or "smallest" not in create(
        rmax=10.0, pbc=(0, 1, 1))[1]
    rc, msg = create(dr=0.0)
    assert rc == 20 and "bin width" in msg
    rc, msg = create(rmax=-1.0)
    assert rc == 20 and "cutoff" in msg
    rc, msg = create(ntypes=64, types=np.array([1, 64, 1, 2], dtype=np.int32), dr=1e-5, rmax=9.0,
                     pbc=(0, 0, 0))
    assert rc == 2 and "too many bins" in msg
    rc, msg = create(cell=np.zeros((3, 3)))
    assert rc == 20 and "degenerate" in msg
