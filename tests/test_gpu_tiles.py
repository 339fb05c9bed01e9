"""Pinned launch shapes (sqc_hip_set_tiles) against the FP64 oracle.

The autotuner picks one spread tile per plan, so most tile shapes otherwise
run only as its timing trials.  Here every tile shape the plan accepts is
pinned (2 x 1 only as a refused shape), and S(q), the pair columns and dL/dr of the same frame are compared
with the oracle.  The grid sizes are not among the compiled line lengths,
so the 8 x 2 tile runs the spreader's and the gather's compiled-tile
instances with the line length read at run time (the gather's
register-resident rounds), in both precisions.  Tolerances as in the
forward and adjoint parity tests: S within 1e-9 (FP64) / 1e-5 (FP32)
absolute, dL/dr within 1e-8 / 1e-5 of max |dL/dr|.
"""

import numpy as np
import pytest

from helpers import random_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# every tile of the plan-side list but 2 x 1, whose bin window (more than
# 128 bins at these plans' ns of 12 and 13) the plans refuse
TILES = [(16, 2), (8, 2), (16, 1), (4, 2), (8, 1), (4, 1)]
S_TOL = {"double": 1e-9, "single": 1e-5}
G_TOL = {"double": 1e-8, "single": 1e-5}
MAPPING = {1: "Si", 2: "O", 3: "Na"}


def _case():
    cell, origin, types, frames = random_case(natoms=2000, length=30.0, nframes=1, seed=77,
                                              tilt=(1.5, -0.7, 0.4), ntypes=3)
    return cell, origin, types, frames[0]


def _plan(precision, cell, origin, types):
    from sqcalc_amd import HipStructureFactor
    # the gradient of 2000 atoms reaches 1e-8 (FP64) / 1e-5 (FP32) of its
    # maximum at eps 1e-12 / 1e-7 (the FP32 gradient plans' eps)
    eps = 1e-12 if precision == "double" else 1e-7
    return HipStructureFactor(cell, origin, types, MAPPING, weight="xray", norm="mean",
                              qmax=5.0, nq=40, precision=precision, partials=True, eps=eps)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_pinned_shapes_are_kept_and_bad_ones_refused(hip_device, precision):
    from sqcalc_amd._lib import SqcHipError
    cell, origin, types, pos = _case()
    sf = _plan(precision, cell, origin, types)
    for tile in [(8, 2), (4, 2), (16, 1), (8, 1)]:
        sf.set_tiles(tile[0], tile[1], 4, 2)
        assert list(sf.info.tile) == [tile[0], tile[1], 4] and sf.info.x_lines == 2
    for bad in [(3, 2, 4, 2), (8, 2, 5, 2), (8, 2, 4, 0), (8, 2, 4, 9), (32, 2, 4, 2),
                (2, 1, 4, 2)]:
        with pytest.raises(SqcHipError):
            sf.set_tiles(*bad)
    # a refused shape leaves the plan as it was
    assert list(sf.info.tile) == [8, 1, 4] and sf.info.x_lines == 2
    sf.close()


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("tile", TILES, ids=["%dx%d" % t for t in TILES])
def test_pinned_tile_forward_and_adjoint(hip_device, precision, tile):
    import torch

    from sqcalc_amd.autograd import structure_factor

    cell, origin, types, pos0 = _case()
    sf = _plan(precision, cell, origin, types)
    sf.set_tiles(tile[0], tile[1], 8, 2)
    print("case pinned tile", precision, tile, list(sf.info.fine), sf.info.nspread, flush=True)
    nq, nsp = 40, 3
    rng = np.random.default_rng(9)
    target = torch.tensor(rng.random(nq), device=hip_device)
    pos = torch.tensor(pos0, device=hip_device, requires_grad=True)
    S, P = structure_factor(pos, sf)
    wP = torch.tensor(rng.standard_normal(tuple(P.shape)), device=hip_device)
    loss = torch.sum((S - target) ** 2) + torch.sum(wP * P)
    loss.backward()
    grad = pos.grad.cpu().numpy()

    run = O.Run(cell, origin, types, MAPPING, weight="xray", norm="mean", qmax=5.0, nq=nq,
                partials=True)
    rho = O.direct_rho(pos0, cell, origin, run.species_of, nsp, run.md, method=0)
    num, _ = O.shell_sums(rho, run.md, "xray", run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    dS = float(np.abs(S.detach().cpu().numpy() - Sref).max())
    print("  max |dS| = %.2e" % dS)
    assert dS <= S_TOL[precision]
    gS = 2 * (Sref - target.cpu().numpy())
    gP = np.zeros((nsp, nsp, nq))
    w = wP.cpu().numpy()
    k = 0
    for a in range(nsp):
        for b in range(a, nsp):
            gP[a, b] = w[k]
            k += 1
    gref = O.gradient(pos0, cell, origin, run.species_of, run.md, rho, "xray", run.amp,
                      run.xcoef, gS, run.den, run.shell_count, gP=gP)
    err = np.abs(grad - gref).max() / np.abs(gref).max()
    print("  max |dg| / max |g| = %.2e" % err)
    assert err <= G_TOL[precision]
    sf.close()
