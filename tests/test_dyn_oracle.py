"""The dyn amplitude oracle (sqo_dyn_rho) on the CPU.

Pinned by the reference's own separable phase factors built from
src/sqc_phase.f90 (oracle/_ref/ref_dyn, when built), by its direct cos/sin
branch, and by closed forms: one atom gives exp(i q . r) exactly; a perfect
lattice gives rho = N on its Bragg modes and 0 elsewhere; lattice modes are
blind to a lattice translation of any atom (the reason wrapped and
unwrapped positions agree, test/test_phase.f90); a q line with a base is
exp(i (s0 + i dq) u . r)."""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref_dyn as R


def _case(seed=3, n=300, ns=3):
    rng = np.random.default_rng(seed)
    cell = np.array([[12.0, 0.7, -0.4], [0.0, 11.0, 0.3], [0.0, 0.0, 13.0]])
    sp = rng.integers(0, ns, n).astype(np.int32)
    pos = (rng.random((n, 3)) * 1.4 - 0.2) @ cell.T
    return cell, 2 * np.pi * np.linalg.inv(cell).T, sp, pos


def test_separable_equals_direct():
    from sqcalc_amd.dynamics import grid_modes
    cell, b, sp, pos = _case()
    hkl, _ = grid_modes(cell, 2.5)
    r1 = O.dyn_rho(pos, sp, 3, b, hkl, method=1)
    r0 = O.dyn_rho(pos, sp, 3, b, hkl, method=0)
    assert np.abs(r1 - r0).max() < 1e-12 * np.abs(r0).max()


@pytest.mark.skipif(not R.available(), reason="oracle/_ref/ref_dyn not built (make -C oracle ref)")
@pytest.mark.parametrize("line", [False, True])
def test_matches_reference_phase_module(line):
    from sqcalc_amd.dynamics import grid_modes, line_modes
    cell, b, sp, pos = _case(seed=5)
    base = None
    if line:
        hkl, b, base = line_modes(60, 0.4, 9.0, (1.0, -2.0, 0.5))
    else:
        hkl, _ = grid_modes(cell, 2.5)
    mine = O.dyn_rho(pos, sp, 3, b, hkl, base=base)
    ref = R.ref_rho(pos, sp, 3, b, hkl, base=base)
    assert np.abs(mine - ref).max() < 1e-12 * max(np.abs(ref).max(), 1.0)


def test_single_atom_closed_form():
    from sqcalc_amd.dynamics import grid_modes
    cell, b, _, _ = _case()
    r = np.array([[3.1, -7.2, 40.5]])        # far outside the cell: unwrapped
    hkl, _ = grid_modes(cell, 3.0)
    got = O.dyn_rho(r, np.zeros(1, np.int32), 1, b, hkl)[0]
    want = np.exp(1j * (hkl @ b.T) @ r[0])
    assert np.abs(got - want).max() < 1e-12


def test_perfect_lattice_bragg():
    from sqcalc_amd.synthetic import lattice_sites
    from sqcalc_amd.dynamics import grid_modes
    L, side = 20.0, 5
    cell = np.diag([L] * 3)
    pos = lattice_sites(cell, side)
    b = 2 * np.pi * np.linalg.inv(cell).T
    hkl, _ = grid_modes(cell, 2 * np.pi * side / L * 1.8)
    got = O.dyn_rho(pos, np.zeros(len(pos), np.int32), 1, b, hkl)[0]
    bragg = np.all(hkl % side == 0, axis=1)
    assert bragg.sum() > 0
    n = len(pos)
    # |rho| = N on the Bragg modes (phase exp(i pi h / side) of the half-cell offset)
    assert np.abs(np.abs(got[bragg]) - n).max() < 1e-10 * n
    assert np.abs(got[~bragg]).max() < 1e-10 * n


def test_lattice_translation_invariance():
    from sqcalc_amd.dynamics import grid_modes
    cell, b, sp, pos = _case(seed=9)
    hkl, _ = grid_modes(cell, 2.0)
    shift = np.random.default_rng(2).integers(-3, 4, (len(pos), 3)) @ cell.T
    a = O.dyn_rho(pos, sp, 3, b, hkl)
    c = O.dyn_rho(pos + shift, sp, 3, b, hkl)
    assert np.abs(a - c).max() < 1e-11 * np.abs(a).max()


def test_line_with_base_closed_form():
    from sqcalc_amd.dynamics import line_modes
    _, _, sp, pos = _case(seed=4)
    s0, s1, n = 0.7, 6.0, 40
    hkl, rec, base = line_modes(n, s0, s1, (0.0, 1.0, 1.0))
    got = O.dyn_rho(pos, sp, 3, rec, hkl, base=base)
    u = np.array([0.0, 1.0, 1.0]) / np.sqrt(2.0)
    s = s0 + np.arange(n + 1) * (s1 - s0) / n
    for a in range(3):
        want = np.exp(1j * np.outer(s, pos[sp == a] @ u)).sum(axis=1)
        assert np.abs(got[a] - want).max() < 1e-11 * np.abs(want).max()
