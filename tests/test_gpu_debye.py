"""Debye pair histograms on the GPU (sqc_hip_debye_*) against the oracle,
which is pinned by the reference's own cell list (tests/test_debye_oracle.py).
Integer counts: bit-exact, every case.  S(q), partials and g(r) from the
histogram against the oracle's debye_sq / debye_g at 1e-12 relative."""

import math

import numpy as np
import pytest

from debye_cases import CASES, lattice, make_cell, oracle_hist, run_case, trajectory
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def hip_hist(cell, origin, pbc, types, T, rmax, dr, skin, frames, device_path=False, **kw):
    from sqcalc_amd.debye import HipDebye
    deb = HipDebye(cell, origin, types, pbc=pbc, rmax=rmax, dr=dr, skin=skin, **kw)
    if device_path:
        import torch
        deb.accumulate_device(torch.tensor(np.asarray(frames), dtype=torch.float64,
                                           device="cuda:0").contiguous())
    else:
        for f in frames:
            deb.accumulate(f)
    h, nf = deb.histogram()
    return deb, h, nf


@pytest.mark.parametrize("name", sorted(CASES))
def test_cases_bit_exact(hip_device, name):
    cell, origin, pbc, types, T, rmax, skin, frames = run_case(name)
    print("case debye", name, len(types), rmax, flush=True)
    run = oracle_hist(cell, origin, pbc, types, T, rmax, 0.01, skin, frames)
    deb, h, nf = hip_hist(cell, origin, pbc, types, T, rmax, 0.01, skin, frames,
                          device_path=(name == "triclinic"))
    assert nf == len(frames)
    assert h.shape == run.hist.shape
    assert np.array_equal(h, run.hist.astype(np.float64))
    deb.close()


def test_lattice_edges_bit_exact(hip_device):
    print("case debye lattice", flush=True)
    pos = lattice(5, 2.0) + 0.5
    types = (np.arange(len(pos)) % 2 + 1).astype(np.int32)
    cell = make_cell((10.0, 10.0, 10.0))
    for dr, rmax in ((0.5, 4.0), (0.3, 4.0), (0.25, 4.5)):
        run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, rmax, dr, 1.0, [pos])
        deb, h, _ = hip_hist(cell, np.zeros(3), (1, 1, 1), types, 2, rmax, dr, 1.0, [pos])
        assert np.array_equal(h, run.hist.astype(np.float64))
        deb.close()


def test_unwrapped_crossings_and_reset(hip_device):
    """Positions handed over unwrapped, crossing the box; reset clears the
    counts while the list positions carry on."""
    print("case debye crossings", flush=True)
    rng = np.random.default_rng(11)
    cell = make_cell((12.0, 13.0, 14.0), (1.0, 0.5, -0.75))
    frames = trajectory(rng, cell, np.zeros(3), 150, 8, 1.2, wrap=False)
    types = rng.integers(1, 3, 150).astype(np.int32)
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, 5.5, 0.02, 0.5, frames)
    deb, h, _ = hip_hist(cell, np.zeros(3), (1, 1, 1), types, 2, 5.5, 0.02, 0.5, frames)
    assert np.array_equal(h, run.hist.astype(np.float64))
    deb.reset()
    assert deb.histogram()[1] == 0 and not deb.histogram()[0].any()
    more = trajectory(rng, cell, np.zeros(3), 150, 2, 1.2, wrap=False)
    for f in more:
        run.add_frame(f)
        deb.accumulate(f)
    # counts of the two frames after the reset: the run's minus its first eight
    h2, nf = deb.histogram()
    assert nf == 2
    assert np.array_equal(h2, (run.hist - h.astype(np.uint64)).astype(np.float64))
    deb.close()


@pytest.mark.parametrize("n", [3000, 50000])
def test_large_bit_exact(hip_device, n):
    """Both tile sizes (256 below about 46k atoms, 1024 above), default rmax
    (half the box) at liquid density, two frames."""
    from sqcalc_amd.debye import HipDebye
    print("case debye large", n, flush=True)
    rng = np.random.default_rng(n)
    L = (n / 0.066) ** (1.0 / 3.0)
    cell = make_cell((L, L, L))
    frames = trajectory(rng, cell, np.zeros(3), n, 2, 0.3)
    types = np.where(np.arange(n) % 3 == 0, 1, 2).astype(np.int32)
    rmax = 0.5 * L if n <= 3000 else 12.0
    deb = HipDebye(cell, np.zeros(3), types, rmax=rmax, dr=0.01)
    assert deb.info.workgroups >= 1
    for f in frames:
        deb.accumulate(f)
    h, _ = deb.histogram()
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, rmax, 0.01, 1.0, frames)
    assert np.array_equal(h, run.hist.astype(np.float64))
    deb.close()


def test_histogram_outside_lds(hip_device):
    """A histogram row too long for the LDS (28,000 bins, 8 types): the
    shared-row path with HBM atomics."""
    from sqcalc_amd.debye import HipDebye
    print("case debye hbm histogram", flush=True)
    rng = np.random.default_rng(8)
    cell = make_cell((30.0, 31.0, 29.0), (1.0, -2.0, 0.5))
    frames = trajectory(rng, cell, np.zeros(3), 2500, 2, 0.5)
    types = rng.integers(1, 9, 2500).astype(np.int32)
    deb = HipDebye(cell, np.zeros(3), types, rmax=14.0, dr=0.0005)
    assert deb.info.hist_in_lds == 0
    for f in frames:
        deb.accumulate(f)
    h, _ = deb.histogram()
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 8, 14.0, 0.0005, 1.0, frames)
    assert np.array_equal(h, run.hist.astype(np.float64))
    deb.close()


def test_positions_outside_a_nonperiodic_box(hip_device):
    from sqcalc_amd import _lib
    from sqcalc_amd.debye import HipDebye
    print("case debye outside slab", flush=True)
    rng = np.random.default_rng(3)
    cell = make_cell((15.0, 15.0, 15.0))
    pos = rng.uniform(1.0, 14.0, (200, 3))
    pos[7, 2] = 13.0          # then 15.5: the smallest-image step carries it out of the box
    types = np.ones(200, dtype=np.int32)
    deb = HipDebye(cell, np.zeros(3), types, pbc=(True, True, False), rmax=5.0)
    deb.accumulate(pos)
    pos2 = pos.copy()
    pos2[7, 2] = 15.5
    with pytest.raises(_lib.SqcHipError, match="non-periodic box"):
        deb.accumulate(pos2)
    with pytest.raises(_lib.SqcHipError, match="non-periodic box"):
        deb.histogram()
    deb.reset()
    deb.accumulate(pos)
    assert deb.histogram()[1] == 1
    deb.close()


def test_structure_factor_and_rdf_against_oracle(hip_device):
    """S(q) with partials (neutron and X-ray weights, every norm), the
    Faber-Ziman columns, g(r) and the pair entropy from the GPU histogram
    against the oracle's evaluation of the oracle's histogram."""
    from sqcalc_amd.debye import HipDebye
    print("case debye S(q)", flush=True)
    rng = np.random.default_rng(5)
    cell = make_cell((20.0, 20.0, 20.0))
    frames = trajectory(rng, cell, np.zeros(3), 1000, 3, 0.2)
    types = np.where(np.arange(1000) % 3 == 0, 1, 2).astype(np.int32)
    run = oracle_hist(cell, np.zeros(3), (1, 1, 1), types, 2, 6.0, 0.01, 1.0, frames)
    counts = np.array([np.count_nonzero(types == 1), np.count_nonzero(types == 2)])
    for weight in ("unit", "neutron", "xray"):
        amp, xcoef = O.weight_params(weight, ["Si", "O"])
        for norm in ("mean", "self", "n"):
            deb = HipDebye(cell, np.zeros(3), types, mapping={1: "Si", 2: "O"}, rmax=6.0,
                           dr=0.01, weight=weight, norm=norm)
            for f in frames:
                deb.accumulate(f)
            rows = deb.structure_factor(0.0, 6.0, 30)
            num, den, part = O.debye_sq(run.hist, 3, counts, 1000, 8000.0, (1, 1, 1), 6.0, 0.01,
                                        0.0, 6.0, 30, weight, amp, xcoef, norm=norm)
            s = num / (3 * den)
            assert np.allclose(rows[:, 1], s, rtol=1e-12, atol=1e-12)
            cols = [part[0, 0], part[0, 1], part[1, 1]]
            for c, ref in zip(rows[:, 2:].T, cols):
                assert np.allclose(c, ref / (3 * 1000), rtol=1e-12, atol=1e-12)
            deb.close()
    deb = HipDebye(cell, np.zeros(3), types, rmax=6.0, dr=0.01, faber_ziman=True)
    for f in frames:
        deb.accumulate(f)
    fz = deb.structure_factor(0.0, 6.0, 30)
    assert np.all(np.abs(fz[-5:, 2:] - 1.0) < 0.5)        # A_ab -> 1 at large q
    r, g_tot, g = deb.rdf()
    r_ref, g_ref = O.debye_g(run.hist, 3, counts, 8000.0, 0.01)
    assert np.allclose(g, g_ref, rtol=1e-12, atol=1e-14)
    assert abs(float(np.mean(g_tot[300:590])) - 1.0) < 0.05  # ideal-gas-like far field
    total, parts, _, _ = deb.pair_entropy()
    assert math.isfinite(total) and set(parts) == {"1-1", "1-2", "2-2"}
    deb.close()
