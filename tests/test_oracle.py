"""CPU checks that pin the FP64 oracle before anything is compared with it.

Each check is chosen so that a plausible slip in the oracle (a wrong sign, a
dropped term, a transposed cell, a shell off by one) fails:

* golden tables written by the reference's own numpy reference
  (tests/golden/, written by tests/golden/make_golden.py from test/ref_sq.py);
* a_i . b_j = 2 pi delta_ij and brute-force enumeration of the mode rule;
* closed forms of the reference suite: single species S = S_aa, the OVITO
  sum rule, a perfect lattice S = N on the Bragg shell, B2 antiphase
  contrast, the high-q limits (test/test_sqcalc.py:755-931, :1164-1260);
* the analytic gradient against central finite differences;
* the CPU restatement of the FINUFFT path against the direct sum.
"""

import math

import numpy as np
import pytest

from helpers import golden_cases, load_golden, oracle_run, random_case, rel_dev
from oracle import oracle as O


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference_numpy(name):
    d = load_golden(name)
    run = oracle_run(d["cell"], d["origin"], d["types"], d["mapping_dict"], d["weight"],
                     d["norm"], d["qmin"], d["qmax"], d["nq"], d["frames"])
    t = run.table()
    np.testing.assert_allclose(t[:, 0], d["q"], rtol=0, atol=5e-7)
    # the fixture carries 13 significant digits
    assert rel_dev(np.array(d["S"]), t[:, 1]) < 2e-12


def test_reciprocal_basis_invariant():
    cell = np.array([[250.0, 10.0, 5.0], [0.0, 250.0, 0.0], [0.0, 0.0, 250.0]])
    b = O.reciprocal(cell)
    np.testing.assert_allclose(cell.T @ b, 2 * np.pi * np.eye(3), atol=1e-12)
    # LAMMPS bounds of that cell give it back
    a, o = O.cell_from_bounds([0.0, 0.0, 0.0], [265.0, 250.0, 250.0], [10.0, 5.0, 0.0])
    np.testing.assert_allclose(a, cell, atol=1e-12)
    np.testing.assert_allclose(o, 0.0, atol=1e-12)


@pytest.mark.parametrize("tilt", [None, (1.5, 0.5, -2.0)])
def test_modes_brute_force(tilt):
    cell, _, _, _ = random_case(natoms=2, length=13.0, tilt=tilt)
    qmin, qmax, nq = 0.4, 4.0, 37
    md = O.modes(cell, qmin, qmax, nq)
    b = 2 * np.pi * np.linalg.inv(cell).T
    m = [max(3, 2 * math.ceil(qmax * np.linalg.norm(cell[:, i]) / (2 * np.pi)) + 1) for i in range(3)]
    assert list(md["m"]) == m
    r = [np.arange(mi) - (mi - 1) // 2 for mi in m]
    h = np.stack(np.meshgrid(r[0], r[1], r[2], indexing="ij"), -1).reshape(-1, 3)
    h = h[np.any(h != 0, axis=1)]
    q = h @ b.T
    ql = np.linalg.norm(q, axis=1)
    keep = (ql >= qmin) & (ql <= qmax)
    assert set(map(tuple, h[keep].tolist())) == set(map(tuple, md["hkl"].tolist()))
    ok = np.abs(((ql[keep] - qmin) / ((qmax - qmin) / nq)) % 1.0 - 0.5) < 0.49
    sh = np.clip(((ql[keep] - qmin) / ((qmax - qmin) / nq)).astype(int), 0, nq - 1) + 1
    lookup = {tuple(k): s for k, s in zip(md["hkl"].tolist(), md["shell"].tolist())}
    mine = np.array([lookup[tuple(k)] for k in h[keep].tolist()])
    assert np.array_equal(mine[ok], sh[ok])
    # the flat index is p1 + m1 (p2 - 1 + m2 (p3 - 1)) with p = h + (m + 1) / 2
    p = md["hkl"] + (np.array(m) + 1) // 2
    g = p[:, 0] + m[0] * (p[:, 1] - 1 + m[1] * (p[:, 2] - 1))
    assert np.array_equal(g, md["gidx"])


def test_separable_phase_equals_cos_sin():
    cell, origin, types, frames = random_case(natoms=120, length=12.0, tilt=(1.0, 0.5, 0.2))
    md = O.modes(cell, 0.0, 5.0, 20)
    sp = np.searchsorted(np.unique(types), types)
    a = O.direct_rho(frames[0], cell, origin, sp, 2, md, method=0)
    b = O.direct_rho(frames[0], cell, origin, sp, 2, md, method=1)
    assert np.abs(a - b).max() < 1e-11


def test_single_species_partial_equals_total():
    cell, origin, _, frames = random_case(natoms=200, length=14.0, nframes=2, seed=4)
    types = np.ones(200, dtype=int)
    t = oracle_run(cell, origin, types, None, "unit", "n", 0.0, 5.0, 40, frames).table()
    assert t.shape[1] == 3
    assert rel_dev(t[:, 1], t[:, 2]) < 1e-12


def test_sum_rule_unit_norm_n():
    cell, origin, types, frames = random_case(natoms=300, length=16.0, nframes=2, seed=6)
    t = oracle_run(cell, origin, types, None, "unit", "n", 0.5, 5.9, 27, frames).table()
    assert np.max(np.abs(t[:, 1] - (t[:, 2] + 2 * t[:, 3] + t[:, 4]))) < 1e-9
    high = t[:, 0] > 3.0
    x1 = np.mean(types == 1)
    assert abs(np.mean(t[high, 2]) - x1) < 0.08
    assert np.mean(np.abs(t[high, 3])) < 0.1


def test_bragg_shell_equals_n():
    # simple cubic lattice, 8^3 atoms in L = 8: the <100> family at |q| = 2 pi,
    # h = (8, 0, 0); 64 has no other representation as a sum of three squares,
    # so the shell holds only the Bragg family
    n = 8
    g = np.arange(n, dtype=float)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * 1.0
    cell = np.diag([8.0] * 3)
    types = np.ones(n ** 3, dtype=int)
    q0 = 2 * np.pi
    t = oracle_run(cell, np.zeros(3), types, None, "unit", "n", q0 - 0.003, q0 + 0.003, 1,
                   [pos]).table()
    assert abs(t[0, 1] - n ** 3) / n ** 3 < 1e-9
    assert abs(t[0, 1] - t[0, 2]) / t[0, 1] < 1e-9


def test_b2_antiphase():
    # B2 (CsCl): corner type 1, centre type 2; (100) is a superlattice peak
    side = 4
    cells = np.array([[i, j, k] for i in range(side) for j in range(side) for k in range(side)],
                     dtype=float)
    pos = np.concatenate([cells, cells + 0.5]) * 3.0
    types = np.array([1] * len(cells) + [2] * len(cells))
    cell = np.diag([3.0 * side] * 3)
    q100 = 2 * np.pi / 3.0
    t = oracle_run(cell, np.zeros(3), types, None, "unit", "n", q100 - 0.01, q100 + 0.01, 1,
                   [pos]).table()
    N = len(types)
    # total vanishes, S_aa = S_bb = N/4, S_ab = -N/4 on the (100) shell
    assert abs(t[0, 1]) < 1e-8 * N
    np.testing.assert_allclose(t[0, 2:], [N / 4, -N / 4, N / 4], rtol=1e-9)


@pytest.mark.parametrize("weight", ["unit", "xray"])
def test_gradient_finite_differences(weight):
    cell, origin, types, frames = random_case(natoms=40, length=9.0, nframes=1, seed=2,
                                              tilt=(0.7, 0.3, -0.4))
    mapping = {1: "Si", 2: "O"}
    run = O.Run(cell, origin, types, mapping, weight=weight, norm="mean", qmax=4.0, nq=12)
    pos = frames[0]
    rng = np.random.default_rng(1)
    target = rng.random(12)
    gP = rng.standard_normal((2, 2, 12)) * np.triu(np.ones((2, 2)))[:, :, None]

    def loss(p):
        rho = O.direct_rho(p, cell, origin, run.species_of, 2, run.md, method=0)
        num, part = O.shell_sums(rho, run.md, weight, run.amp, run.xcoef)
        S = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
        P = part / (run.natoms * np.maximum(run.shell_count, 1))
        return float(np.sum((S - target) ** 2) + np.sum(gP * P)), rho, S

    L0, rho, S = loss(pos)
    gS = 2 * (S - target)
    grad = O.gradient(pos, cell, origin, run.species_of, run.md, rho, weight, run.amp,
                      run.xcoef, gS, run.den, run.shell_count, gP=gP)
    h = 1e-4
    for j, c in [(0, 0), (5, 1), (17, 2), (33, 0)]:
        pp = pos.copy()
        pp[j, c] += h
        pm = pos.copy()
        pm[j, c] -= h
        fd = (loss(pp)[0] - loss(pm)[0]) / (2 * h)
        assert abs(fd - grad[j, c]) <= 1e-6 * max(1.0, abs(fd)), (j, c, fd, grad[j, c])


@pytest.mark.parametrize("weight,partials", [("neutron", True), ("unit", False), ("xray", False)])
def test_cpu_nufft_restatement_matches_direct(weight, partials):
    cell, origin, types, frames = random_case(natoms=250, length=18.0, nframes=1, seed=8,
                                              tilt=(1.0, 0.0, 0.5))
    mapping = {1: "Si", 2: "O"}
    run = O.Run(cell, origin, types, mapping, weight=weight, norm="mean", qmax=5.0, nq=50,
                partials=partials)
    run.add_frame(frames[0])
    num, part, info = O.nufft_frame(run, frames[0], eps=1e-9, upsampfac=1.25, nthreads=4)
    ok = run.den > 0
    assert rel_dev(run.num[ok] / run.den[ok], num[ok] / run.den[ok]) < 1e-8
    if partials:
        scale = np.abs(run.partial).max()
        assert np.abs(part - run.partial).max() / scale < 1e-8
