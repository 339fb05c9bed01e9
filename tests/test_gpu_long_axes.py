"""Fine-grid axes longer than 1024 points.

The plain FP64 path (sigma 2) needs them for boxes beyond about 160 A at
q_max 10: the reference's own limit is 4e8 reciprocal grid points
(src/sqc_structure_factor.f90:63-64, 339-345), m <= 736, which at sigma 2 is a
fine axis of up to 1500 points.  Lines of more than 1024 points are
transformed out of place through a scratch line per wave, and the transform
buffers are indexed with 64-bit offsets.  Checked against the oracle's
direct sum on the lowest shells of the full-range plan (both plans bin from
q = 0 with the same dq, so their shells coincide).
"""

import numpy as np
import pytest

from helpers import oracle_run, rel_dev

pytestmark = pytest.mark.gpu

SIO = {1: "Si", 2: "O"}


def _case(lengths, natoms, seed, nframes=1):
    rng = np.random.default_rng(seed)
    cell = np.diag(lengths).astype(np.float64)
    cell[0, 1], cell[0, 2] = 1.5, -2.0  # xy, xz tilts
    types = rng.integers(1, 3, natoms)
    frames = np.array([rng.random((natoms, 3)) @ cell.T for _ in range(nframes)])
    return cell, np.zeros(3), types, frames


@pytest.mark.parametrize("precision,upsampfac,lengths,long_axes", [
    ("double", 0.0, (170.0, 170.0, 170.0), 3),
    ("single", 2.0, (170.0, 70.0, 80.0), 1),
    ("double", 0.0, (60.0, 80.0, 175.0), 1),
])
def test_long_axes_low_shells_against_oracle(hip_device, precision, upsampfac, lengths, long_axes):
    from sqcalc_amd import HipStructureFactor
    cell, origin, types, frames = _case(lengths, 1500, seed=61)
    qmax, nq, nlow = 10.0, 500, 30  # dq = 0.02: the 30 lowest shells are q < 0.6
    print("case long axes", precision, upsampfac, lengths, flush=True)
    sf = HipStructureFactor(cell, origin, types, SIO, weight="neutron", norm="mean", qmax=qmax,
                            nq=nq, precision=precision, partials=True, max_batch=1,
                            upsampfac=upsampfac, eps=1e-10 if precision == "double" else 1e-6)
    fine = list(sf.info.fine)
    print("  modes %s fine %s" % (list(sf.info.modes), fine), flush=True)
    assert sum(n > 1024 for n in fine) == long_axes
    sf.accumulate(frames)
    got = sf.table()
    sf.close()
    ref = oracle_run(cell, origin, types, SIO, "neutron", "mean", 0.0, qmax * nlow / nq, nlow,
                     frames).table()
    np.testing.assert_allclose(got[:nlow, 0], ref[:, 0], atol=1e-12)
    dev = rel_dev(ref[:, 1:], got[:nlow, 1:])
    print("  max deviation on %d shells %.2e" % (nlow, dev))
    assert dev <= (1e-9 if precision == "double" else 1e-6)


def test_long_axis_adjoint_against_oracle(hip_device):
    """dL/dr through the long-line inverse transforms: an FP64 plan with one
    axis of more than 1024 points, a loss on its lowest shells, against the
    oracle's analytic gradient."""
    import torch

    from oracle import oracle as O
    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import structure_factor

    cell, origin, types, frames = _case((172.0, 30.0, 34.0), 400, seed=67)
    qmax, nq = 10.0, 500
    sf = HipStructureFactor(cell, origin, types, SIO, weight="neutron", norm="mean", qmax=qmax,
                            nq=nq, precision="double", partials=False, max_batch=1, eps=1e-11)
    assert sf.info.fine[0] > 1024
    nlow = 20
    w = np.zeros(nq)
    w[:nlow] = 1.0
    rng = np.random.default_rng(3)
    target = rng.random(nq)
    pos = torch.tensor(frames[0], device=hip_device, requires_grad=True)
    S, _ = structure_factor(pos, sf)
    wt = torch.tensor(w, device=hip_device)
    loss = torch.sum(wt * (S - torch.tensor(target, device=hip_device)) ** 2)
    loss.backward()
    grad = pos.grad.cpu().numpy()
    run = O.Run(cell, origin, types, SIO, weight="neutron", norm="mean", qmax=qmax * nlow / nq,
                nq=nlow, partials=False)
    rho = O.direct_rho(frames[0], cell, origin, run.species_of, 2, run.md, method=0)
    num, _ = O.shell_sums(rho, run.md, "neutron", run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    gS = 2 * (Sref - target[:nlow])
    gref = O.gradient(frames[0], cell, origin, run.species_of, run.md, rho, "neutron", run.amp,
                      run.xcoef, gS, run.den, run.shell_count)
    err = np.abs(grad - gref).max() / np.abs(gref).max()
    print("  max |dg| / max |g| = %.2e" % err)
    assert err <= 1e-8
    sf.close()
