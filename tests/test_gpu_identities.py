"""Closed forms and identities of the reference suite, on the device path.

* B2 NiAl (test/test_sqcalc.py:1164-1220): on the antiphase (100) family the
  total cancels and the pair terms are +-N/4; on the in-phase (110) family the
  total is N; S = S_aa + 2 S_ab + S_bb; under X-ray weights the superlattice
  line is N ((f_Ni - f_Al) / (f_Ni + f_Al))^2.
* a <-> b symmetry (:785-800): swapping the type labels swaps S_aa and S_bb
  and leaves S and S_ab as they are.
* sparse type ids (:840-876): types 1 and 5 give the same result as 1 and 2;
  the shell sums are exact integer sums, so the same bits.
* ideal gas (:193-198): <S(q)> -> 1 at large q.
* grid table (:270-290): under unit weights and --norm self the per-mode
  table's shell averages are the 1D table (here to FP64 rounding, where the
  reference's NUFFT-vs-table check allows 0.05).
"""

import numpy as np
import pytest

from helpers import random_case, rel_dev

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-9, "single": 1e-5}


def _plan(cell, origin, types, mapping, weight, norm, qmin, qmax, nq, precision, **kw):
    from sqcalc_amd import HipStructureFactor
    return HipStructureFactor(cell, origin, types, mapping, weight=weight, norm=norm, qmin=qmin,
                              qmax=qmax, nq=nq, precision=precision, **kw)


def _b2(side=8, a=2.88):
    c = np.array([[i, j, k] for i in range(side) for j in range(side) for k in range(side)],
                 dtype=float)
    pos = np.concatenate([c, c + 0.5]) * a
    types = np.array([1] * len(c) + [2] * len(c))
    return np.diag([side * a] * 3), pos, types


def _form_factor(label, q):
    from sqcalc_amd import elements
    c = np.asarray(elements.xray_coefficients(label), dtype=float)
    s2 = (q / (4.0 * np.pi)) ** 2
    return c[8] + sum(c[i] * np.exp(-c[4 + i] * s2) for i in range(4))


@pytest.mark.parametrize("precision", ["double", "single"])
def test_b2_superlattice_closed_forms(hip_device, precision):
    cell, pos, types = _b2()
    n = len(types)
    a = 2.88
    q100 = 2.0 * np.pi / a
    q110 = q100 * np.sqrt(2.0)
    print("case b2", precision, flush=True)
    tol = TOL[precision]
    eps = 1e-12 if precision == "double" else 1e-6
    rows = {}
    for label, q, cross in (("odd", q100, -n / 4.0), ("even", q110, n / 4.0)):
        # a window of +-5e-5 holds the six (or twelve) vectors of one family
        sf = _plan(cell, np.zeros(3), types, {1: "Ni", 2: "Al"}, "unit", "mean", q - 5e-5,
                   q + 5e-5, 1, precision, eps=eps)
        sf.accumulate(pos[None])
        row = sf.table()[0]
        sf.close()
        rows[label] = row
        assert abs(row[2] - n / 4.0) <= tol * n and abs(row[4] - n / 4.0) <= tol * n
        assert abs(row[3] - cross) <= tol * n
        assert abs(row[1] - (row[2] + 2.0 * row[3] + row[4])) <= 1e-12 * n
    print("  total (100) %.3e, (110) %.6f of N" % (rows["odd"][1], rows["even"][1] / n))
    assert abs(rows["odd"][1]) <= tol * n
    assert abs(rows["even"][1] - n) <= tol * n
    fni, fal = _form_factor("Ni", q100), _form_factor("Al", q100)
    contrast = ((fni - fal) / (fni + fal)) ** 2
    sf = _plan(cell, np.zeros(3), types, {1: "Ni", 2: "Al"}, "xray", "mean", q100 - 5e-5,
               q100 + 5e-5, 1, precision, eps=eps)
    sf.accumulate(pos[None])
    sx = sf.table()[0, 1]
    sf.close()
    print("  xray S(100) / N %.8f, contrast %.8f" % (sx / n, contrast))
    assert abs(sx - n * contrast) <= tol * n * contrast


def test_type_swap_symmetry(hip_device):
    cell, origin, types, frames = random_case(natoms=1500, length=24.0, nframes=2, seed=81,
                                              tilt=(0.5, -1.0, 0.25))
    args = ("neutron", "mean", 0.0, 6.0, 60, "double")
    sf = _plan(cell, origin, types, {1: "Si", 2: "O"}, *args, eps=1e-11)
    sf.accumulate(frames)
    t = sf.table()
    sf.close()
    swapped = 3 - types
    sw = _plan(cell, origin, swapped, {1: "O", 2: "Si"}, *args, eps=1e-11)
    sw.accumulate(frames)
    u = sw.table()
    sw.close()
    # columns: q, S, S(1-1), S(1-2), S(2-2); type 1 of the swapped run is O
    assert rel_dev(t[:, 1], u[:, 1]) <= 1e-12
    assert rel_dev(t[:, 2], u[:, 4]) <= 1e-12 and rel_dev(t[:, 4], u[:, 2]) <= 1e-12
    assert rel_dev(t[:, 3], u[:, 3]) <= 1e-12


@pytest.mark.parametrize("precision", ["double", "single"])
def test_sparse_type_ids_same_bits(hip_device, precision):
    cell, origin, types, frames = random_case(natoms=2000, length=26.0, nframes=2, seed=83)
    args = ("xray", "mean", 0.0, 7.0, 70, precision)
    sf = _plan(cell, origin, types, {1: "Si", 2: "O"}, *args)
    sf.accumulate(frames)
    n0, p0, _, _ = sf.fetch()
    l0 = sf.column_labels()
    sf.close()
    sparse = np.where(types == 2, 5, 1)
    sp = _plan(cell, origin, sparse, {1: "Si", 5: "O"}, *args)
    sp.accumulate(frames)
    n1, p1, _, _ = sp.fetch()
    l1 = sp.column_labels()
    sp.close()
    assert np.array_equal(n0, n1) and np.array_equal(p0, p1)
    assert len(l0) == len(l1)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_grid_table_shell_averages(hip_device, precision):
    cell, origin, types, frames = random_case(natoms=1200, length=20.0, nframes=2, seed=85)
    qmax, nq = 6.0, 120
    sf = _plan(cell, origin, types, None, "unit", "self", 0.0, qmax, nq, precision,
               want_grid=True)
    sf.accumulate(frames)
    table = sf.table()
    qvec, s = sf.grid_table()
    sf.close()
    qlen = np.linalg.norm(qvec, axis=1)
    shell = np.clip((qlen / (qmax / nq)).astype(int), 0, nq - 1)
    gsum = np.bincount(shell, weights=s, minlength=nq)
    gcnt = np.bincount(shell, minlength=nq)
    avg = np.divide(gsum, gcnt, out=np.zeros(nq), where=gcnt > 0)
    dev = rel_dev(table[:, 1], avg)
    print("  grid shell averages against the table: %.2e" % dev)
    assert dev <= (1e-12 if precision == "double" else 1e-6)


def test_ideal_gas_high_q_limit(hip_device):
    """Uniform random positions: <S(q)> -> 1 at large q (test/test_sqcalc.py:193-198
    allows 0.35; 4000 atoms and 5 frames average about 7000 modes a shell)."""
    rng = np.random.default_rng(87)
    L, n = 30.0, 4000
    cell = np.diag([L] * 3)
    types = np.ones(n, dtype=np.int64)
    frames = rng.random((5, n, 3)) * L
    sf = _plan(cell, np.zeros(3), types, None, "unit", "self", 0.0, 6.0, 120, "single")
    sf.accumulate(frames)
    t = sf.table()
    sf.close()
    mean_s = float(np.mean(t[t[:, 0] > 3.0, 1]))
    print("  ideal gas <S(q)> for q > 3: %.4f" % mean_s)
    assert abs(mean_s - 1.0) < 0.02
