"""Edge positions, coincident atoms and the one-atom closed form.

The atom preparation wraps each fractional coordinate s into [0, 1) as
s - floor(s) and scales it to the fine grid in FP64.  A tiny negative s wraps
to exactly 1.0, and (1 - 2^-53) n rounds to n, so both reach the grid point
n, which is folded back to 0.  A stencil's first point is ceil(u - ns / 2):
on a fine-grid point (even ns) or half-way between two (odd ns), the
stencil's end point sits exactly on the kernel's support edge, and the last
bit of u decides which end is kept.

The edge frame places atoms on those values (cell faces, 1 - 2^-53, +-2^-60,
grid and half-grid points of both parities), many cells away (up to 1000
cells), and on top of each other.  S(q), the pair columns and dL/dr are
compared with the oracle's direct sum, at the tolerances of the other parity
tests (S within 1e-9 / 1e-5 absolute, dL/dr within 1e-8 / 1e-5 of
max |dL/dr| for the FP64 / FP32 paths).  One atom has S(q) = 1 in every
shell with modes under mean normalisation, whatever its position, so its
dL/dr vanishes.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

S_TOL = {"double": 1e-9, "single": 1e-5}
G_TOL = {"double": 1e-8, "single": 1e-5}
EPS = {"double": 1e-12, "single": 1e-7}
MAPPING = {1: "Si", 2: "O"}
# power-of-two cell lengths: L s and pos / L are exact, so the fractional
# coordinates below reach the device unrounded
LENGTHS = (32.0, 32.0, 16.0)
FACES = np.array([0.0, 1.0, 1.0 - 2.0 ** -53, -(2.0 ** -60), 2.0 ** -60, 0.5, 1.0 - 2.0 ** -40])
IMAGES = np.array([0, 0, 0, 0, -1, 1, -7, 3, 1000, -1000])


def _plan(precision, cell, types, partials=True):
    from sqcalc_amd import HipStructureFactor
    return HipStructureFactor(cell, np.zeros(3), types, MAPPING, weight="xray", norm="mean",
                              qmax=5.0, nq=40, precision=precision, partials=partials,
                              eps=EPS[precision])


def _edge_frame(fine, natoms, seed):
    """Fractional coordinates on grid points, half-points, thirds, cell faces
    and their far images; the last 24 atoms sit on top of earlier ones."""
    rng = np.random.default_rng(seed)
    fine = np.asarray(fine, dtype=np.float64)
    k = rng.integers(0, fine.astype(np.int64), size=(natoms, 3)).astype(np.float64)
    off = rng.choice(np.array([0.0, 0.5, -0.5, 1.0 / 3.0, 0.25]), size=(natoms, 3))
    frac = (k + off) / fine
    on_face = rng.random((natoms, 3)) < 0.25
    frac[on_face] = rng.choice(FACES, size=int(on_face.sum()))
    frac += rng.choice(IMAGES, size=(natoms, 3))
    frac[-24:] = frac[rng.integers(0, natoms - 24, 24)]
    return frac * np.array(LENGTHS)


def _oracle(pos, cell, types, run, nsp, gS=None, gP=None):
    rho = O.direct_rho(pos, cell, np.zeros(3), run.species_of, nsp, run.md, method=0)
    num, part = O.shell_sums(rho, run.md, "xray", run.amp, run.xcoef)
    S = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    g = None
    if gS is not None:
        gS = gS(S)
        g = O.gradient(pos, cell, np.zeros(3), run.species_of, run.md, rho, "xray", run.amp,
                       run.xcoef, gS, run.den, run.shell_count, gP=gP)
    return S, g


@pytest.mark.parametrize("precision", ["double", "single"])
def test_edge_positions_forward_and_adjoint(hip_device, precision):
    import torch

    from sqcalc_amd.autograd import structure_factor
    cell = np.diag(LENGTHS)
    natoms, nq, nsp = 600, 40, 2
    types = np.random.default_rng(3).integers(1, 3, natoms)
    sf = _plan(precision, cell, types)
    fine = list(sf.info.fine)
    pos0 = _edge_frame(fine, natoms, seed=11)
    print("case edge positions", precision, "fine", fine, "ns", sf.info.nspread, flush=True)
    rng = np.random.default_rng(12)
    target = rng.random(nq)
    pos = torch.tensor(pos0, device=hip_device, requires_grad=True)
    S, P = structure_factor(pos, sf)
    wP = rng.standard_normal(tuple(P.shape))
    loss = (torch.sum((S - torch.tensor(target, device=hip_device)) ** 2)
            + torch.sum(torch.tensor(wP, device=hip_device) * P))
    loss.backward()
    grad = pos.grad.cpu().numpy()

    run = O.Run(cell, np.zeros(3), types, MAPPING, weight="xray", norm="mean", qmax=5.0, nq=nq,
                partials=True)
    gP = np.zeros((nsp, nsp, nq))
    gP[0, 0], gP[0, 1], gP[1, 1] = wP[0], wP[1], wP[2]
    Sref, gref = _oracle(pos0, cell, types, run, nsp, gS=lambda s: 2 * (s - target), gP=gP)
    dS = float(np.abs(S.detach().cpu().numpy() - Sref).max())
    err = float(np.abs(grad - gref).max() / np.abs(gref).max())
    worst = int(np.abs(grad - gref).max(axis=1).argmax())
    print("  max |dS| = %.2e, max |dg| / max |g| = %.2e (atom %d at fractional %s)"
          % (dS, err, worst, (pos0[worst] / np.array(LENGTHS)).tolist()), flush=True)
    assert dS <= S_TOL[precision]
    assert err <= G_TOL[precision]

    # the same atoms moved into the cell on the host give the same S(q)
    # (the device's wrap against numpy's) within the forward tolerance
    frac = pos0 / np.array(LENGTHS)
    wrapped = (frac - np.floor(frac)) * np.array(LENGTHS)
    S2, _ = structure_factor(torch.tensor(wrapped, device=hip_device), sf)
    dW = float(np.abs(S2.cpu().numpy() - S.detach().cpu().numpy()).max())
    print("  wrapped on the host vs on the device: %.2e" % dW)
    assert dW <= S_TOL[precision]
    sf.close()


@pytest.mark.parametrize("precision", ["double", "single"])
def test_one_atom_closed_form(hip_device, precision):
    """S(q) = 1 in every shell that holds modes, at positions on a face, a
    far image and inside a triclinic cell; S = 0 in shells without modes.
    dL/dr = 0 up to the transform's accuracy: for L = sum_s (S_s - 2)^2 the
    bound is the gradient tolerance times sum_s |dL/dS_s| q_max (each shell's
    dS/dr is a sum of terms of size 2 |q|, which cancel exactly)."""
    import torch

    from sqcalc_amd.autograd import structure_factor
    cell = np.array([[20.0, 1.5, -0.7], [0.0, 21.0, 0.4], [0.0, 0.0, 19.0]])
    from sqcalc_amd import HipStructureFactor
    sf = HipStructureFactor(cell, np.zeros(3), [1], {1: "O"}, weight="xray", norm="mean",
                            qmax=5.0, nq=40, precision=precision, partials=False,
                            eps=EPS[precision])
    run = O.Run(cell, np.zeros(3), [1], {1: "O"}, weight="xray", norm="mean", qmax=5.0, nq=40,
                partials=False)
    has = run.den > 0
    assert has.sum() > 30
    for p in ([0.0, 0.0, 0.0], [20.0, 21.0, 19.0], [-1e-17, 3.3, 7.25],
              [20000.5, -19999.0, 40.1], [10.0, 10.5, 9.5]):
        pos = torch.tensor([p], dtype=torch.float64, device=hip_device, requires_grad=True)
        S, _ = structure_factor(pos, sf)
        dev = float(np.abs(S.detach().cpu().numpy()[has] - 1.0).max())
        zero = float(np.abs(S.detach().cpu().numpy()[~has]).max()) if (~has).any() else 0.0
        loss = torch.sum((S[torch.tensor(has, device=hip_device)] - 2.0) ** 2)
        loss.backward()
        g = float(np.abs(pos.grad.cpu().numpy()).max())
        print("  one atom at %s: max |S - 1| = %.2e, |S| off-shell %.1e, max |dL/dr| = %.2e"
              % (p, dev, zero, g), flush=True)
        assert dev <= S_TOL[precision]
        assert zero == 0.0
        assert g <= G_TOL[precision] * 2 * int(has.sum()) * 5.0
    sf.close()


CELLS = {
    # tilts at LAMMPS's limit of half a box length
    "skewed": (np.array([[24.0, 12.0, -12.0], [0.0, 22.0, 11.0], [0.0, 0.0, 20.0]]), 1500, 5.0),
    # a 2 ns fine grid: every stencil covers half an axis, spans several z
    # slabs and wraps through 0
    "tiny": (np.array([[4.1, 0.3, 0.0], [0.0, 3.9, -0.2], [0.0, 0.0, 4.3]]), 9, 9.0),
}


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("name", list(CELLS))
def test_cell_shapes_forward_and_adjoint(hip_device, precision, name):
    """A cell skewed to LAMMPS's tilt limit, and a cell so small that the fine
    grid sits at its 2 ns floor: S(q), the pair columns and dL/dr against the
    oracle, at the tolerances above."""
    import torch

    from sqcalc_amd.autograd import structure_factor
    cell, natoms, qmax = CELLS[name]
    nq, nsp = 30, 2
    rng = np.random.default_rng(21)
    types = rng.integers(1, 3, natoms)
    types[:2] = [1, 2]
    pos0 = (rng.random((natoms, 3)) * 1.4 - 0.2) @ cell.T
    from sqcalc_amd import HipStructureFactor
    sf = HipStructureFactor(cell, np.zeros(3), types, MAPPING, weight="xray", norm="mean",
                            qmax=qmax, nq=nq, precision=precision, partials=True,
                            eps=EPS[precision])
    print("case cell", name, precision, "fine", list(sf.info.fine), "modes", sf.info.nmodes,
          "ns", sf.info.nspread, flush=True)
    target = rng.random(nq)
    pos = torch.tensor(pos0, device=hip_device, requires_grad=True)
    S, P = structure_factor(pos, sf)
    wP = rng.standard_normal(tuple(P.shape))
    loss = (torch.sum((S - torch.tensor(target, device=hip_device)) ** 2)
            + torch.sum(torch.tensor(wP, device=hip_device) * P))
    loss.backward()
    grad = pos.grad.cpu().numpy()

    run = O.Run(cell, np.zeros(3), types, MAPPING, weight="xray", norm="mean", qmax=qmax, nq=nq,
                partials=True)
    assert len(run.md["qlen"]) == sf.info.nmodes
    gP = np.zeros((nsp, nsp, nq))
    gP[0, 0], gP[0, 1], gP[1, 1] = wP[0], wP[1], wP[2]
    Sref, gref = _oracle(pos0, cell, types, run, nsp, gS=lambda s: 2 * (s - target), gP=gP)
    dS = float(np.abs(S.detach().cpu().numpy() - Sref).max())
    err = float(np.abs(grad - gref).max() / np.abs(gref).max())
    print("  max |dS| = %.2e, max |dg| / max |g| = %.2e" % (dS, err), flush=True)
    assert dS <= S_TOL[precision]
    assert err <= G_TOL[precision]
    sf.close()
