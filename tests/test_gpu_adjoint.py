"""GPU adjoint: dL/dr from the autograd entry point against the FP64 oracle.

The oracle's analytic gradient is itself pinned by central finite
differences (tests/test_oracle.py).  Tolerance: max |g - g_ref| / max |g_ref|
below 1e-8 (FP64 path) and 1e-5 (FP32 path); the FP32 bound is the
gradient's relative accuracy target.
"""

import numpy as np
import pytest

from helpers import random_case
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-8, "single": 1e-5}


MAPPINGS = {2: {1: "Si", 2: "O"}, 4: {1: "Si", 2: "O", 3: "Na", 4: "Ca"},
            10: {1: "Si", 2: "O", 3: "Na", 4: "Ca", 5: "Al", 6: "Mg", 7: "Fe", 8: "K", 9: "C",
                 10: "N"}}


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("weight,partials,nspecies", [("neutron", False, 2), ("unit", True, 2),
                                                      ("xray", True, 2), ("xray", False, 2),
                                                      ("xray", True, 4), ("neutron", True, 4),
                                                      ("neutron", False, 10), ("xray", True, 10),
                                                      ("neutron", True, 10)])
def test_adjoint_matches_oracle(hip_device, precision, weight, partials, nspecies):
    """Four species run the x pass and its adjoint with the field count read
    at run time (the compiled instances hold one and two fields); ten run the
    any-field-count instances, with the weights read from the plan's table
    (and, for one weighted field, the spreader's and the slot sum's
    per-species strengths as well)."""
    _adjoint_case(hip_device, precision, weight, partials, nspecies, "mean")


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("norm", ["self", "n"])
def test_adjoint_other_norms(hip_device, precision, norm):
    """The self and n normalisations divide S by another denominator (the
    pair columns keep N times the shell's mode count): X-ray weights, four
    species, pair columns."""
    _adjoint_case(hip_device, precision, "xray", True, 4, norm)


def _adjoint_case(hip_device, precision, weight, partials, nspecies, norm):
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import structure_factor

    cell, origin, types, frames = random_case(natoms=300, length=15.0, nframes=1, seed=21,
                                              tilt=(1.0, -0.5, 0.3), ntypes=nspecies)
    mapping = MAPPINGS[nspecies]
    qmax, nq = 4.5, 30
    print("case adjoint", precision, weight, partials, nspecies, norm, flush=True)
    eps = 1e-11 if precision == "double" else 1e-6
    sf = HipStructureFactor(cell, origin, types, mapping, weight=weight, norm=norm, qmax=qmax,
                            nq=nq, precision=precision, partials=partials, eps=eps)
    rng = np.random.default_rng(4)
    target = torch.tensor(rng.random(nq), device=hip_device)
    pos = torch.tensor(frames[0], device=hip_device, requires_grad=True)
    S, P = structure_factor(pos, sf)
    loss = torch.sum((S - target) ** 2)
    wP = None
    if partials:
        wP = torch.tensor(rng.standard_normal(tuple(P.shape)), device=hip_device)
        loss = loss + torch.sum(wP * P)
    loss.backward()
    grad = pos.grad.cpu().numpy()

    run = O.Run(cell, origin, types, mapping, weight=weight, norm=norm, qmax=qmax, nq=nq,
                partials=partials)
    rho = O.direct_rho(frames[0], cell, origin, run.species_of, nspecies, run.md, method=0)
    num, _ = O.shell_sums(rho, run.md, weight, run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    # 1e-9 / 1e-5 of S's scale (about 1 under mean and self normalisation,
    # N times larger under n)
    scale = max(1.0, float(np.abs(Sref).max()))
    np.testing.assert_allclose(S.detach().cpu().numpy(), Sref, rtol=0,
                               atol=(1e-9 if precision == "double" else 1e-5) * scale)
    gS = 2 * (Sref - target.cpu().numpy())
    gP = None
    if partials:
        gP = np.zeros((nspecies, nspecies, nq))
        w = wP.cpu().numpy()
        k = 0
        for a in range(nspecies):
            for b in range(a, nspecies):
                gP[a, b] = w[k]
                k += 1
    gref = O.gradient(frames[0], cell, origin, run.species_of, run.md, rho, weight, run.amp,
                      run.xcoef, gS, run.den, run.shell_count, gP=gP)
    err = np.abs(grad - gref).max() / np.abs(gref).max()
    print("  max |dg| / max |g| = %.2e" % err)
    assert err <= TOL[precision]
    sf.close()


@pytest.mark.parametrize("precision", ["double", "single"])
def test_adjoint_dense_windows(hip_device, precision):
    """Gather windows that span many staging rounds (6,000 atoms in 12^3 A^3).
    FP64 pins the gather to 1e-8; in FP32 the tiny dense grid amplifies
    rounding ~N / sqrt(N) (see the dense forward test), bound 1e-4."""
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import structure_factor

    cell, origin, types, frames = random_case(natoms=6000, length=12.0, nframes=1, seed=52)
    print("case adjoint dense", precision, flush=True)
    qmax, nq = 2.5, 25
    eps = 1e-11 if precision == "double" else 1e-6
    sf = HipStructureFactor(cell, origin, types, {1: "Si", 2: "O"}, weight="neutron",
                            norm="mean", qmax=qmax, nq=nq, precision=precision, partials=False,
                            eps=eps)
    rng = np.random.default_rng(9)
    target = torch.tensor(rng.random(nq), device=hip_device)
    pos = torch.tensor(frames[0], device=hip_device, requires_grad=True)
    S, _ = structure_factor(pos, sf)
    torch.sum((S - target) ** 2).backward()
    grad = pos.grad.cpu().numpy()
    run = O.Run(cell, origin, types, {1: "Si", 2: "O"}, weight="neutron", norm="mean",
                qmax=qmax, nq=nq, partials=False)
    rho = O.direct_rho(frames[0], cell, origin, run.species_of, 2, run.md, method=1)
    num, _ = O.shell_sums(rho, run.md, "neutron", run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    gref = O.gradient(frames[0], cell, origin, run.species_of, run.md, rho, "neutron", run.amp,
                      run.xcoef, 2 * (Sref - target.cpu().numpy()), run.den, run.shell_count)
    err = np.abs(grad - gref).max() / np.abs(gref).max()
    print("  max |dg| / max |g| = %.2e" % err)
    assert err <= (TOL[precision] if precision == "double" else 1e-4)
    sf.close()


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("partials", [False, True])
def test_loss_step_under_graph_capture(hip_device, precision, partials):
    """The forward and adjoint entry points order their work on the caller's
    stream with events and make no host synchronisation, so a whole loss step
    (forward, loss, backward) can be captured with torch.cuda.graph and
    replayed per frame.  Each replay, with eager work and a synchronisation
    in between, must give the eager gradient of its frame: 1e-10 relative in
    FP64; in FP32 the FP64 shell sums' atomic order can flip single-precision
    roundings downstream, so the gradient tolerance applies.  With partials
    the loss adds a weighted sum of the pair columns, so the pair gradients
    go through the captured backward as well."""
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import sq_matching_loss, structure_factor

    cell, origin, types, frames = random_case(natoms=800, length=16.0, nframes=4, seed=5,
                                              tilt=(0.8, 0.0, -0.4))
    print("case adjoint graph capture", precision, partials, flush=True)
    eps = 1e-11 if precision == "double" else 1e-6
    sf = HipStructureFactor(cell, origin, types, {1: "Si", 2: "O"}, weight="neutron",
                            norm="mean", qmax=4.0, nq=40, precision=precision, partials=partials,
                            eps=eps, tune=True)
    rng = np.random.default_rng(1)
    target = torch.tensor(rng.random(40), device=hip_device)
    wP = torch.tensor(rng.standard_normal((3, 40)), device=hip_device)

    def loss_of(x):
        if not partials:
            return sq_matching_loss(x, sf, target)
        S, P = structure_factor(x, sf)
        return torch.sum((S - target) ** 2) + torch.sum(wP * P)

    pos = torch.tensor(frames, device=hip_device)
    ref = []
    for k in range(len(frames)):
        x = pos[k].clone().requires_grad_(True)
        loss_of(x).backward()
        ref.append(x.grad.clone())
    static = pos[0].clone().requires_grad_(True)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            static.grad = None
            loss_of(static).backward()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    static.grad = None
    with torch.cuda.graph(graph):
        loss_of(static).backward()
    tol = 1e-10 if precision == "double" else TOL["single"]
    for rep in range(3):
        for k in range(len(frames)):
            with torch.no_grad():
                static.copy_(pos[k])
            graph.replay()
            junk = (pos[k] * 2.0).sum()
            torch.cuda.synchronize()
            dev = float((static.grad - ref[k]).abs().max() / ref[k].abs().max())
            assert dev < tol, (rep, k, dev)
            assert float(junk) != 0.0
    sf.close()


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("weight", ["xray", "neutron"])
def test_single_species(hip_device, precision, weight):
    """One species has one field.  Under X-ray weights its strength is 1 and
    f(|q|) enters in the x-pass epilogue and the adjoint's amplification;
    under neutron weights b rides in the strength.  The S(q) table with its
    single pair column and dL/dr must match the oracle."""
    import torch

    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.autograd import sq_matching_loss

    cell, origin, types, frames = random_case(natoms=250, length=13.0, nframes=2, seed=9,
                                              ntypes=1, tilt=(0.6, 0.2, -0.3))
    mapping = {1: "Ni"}
    qmax, nq = 5.0, 40
    print("case single species", weight, precision, flush=True)
    eps = 1e-11 if precision == "double" else 1e-6
    sf = HipStructureFactor(cell, origin, types, mapping, weight=weight, norm="mean", qmax=qmax,
                            nq=nq, precision=precision, eps=eps, want_grid=True)
    sf.accumulate(frames)
    run = O.Run(cell, origin, types, mapping, weight=weight, norm="mean", qmax=qmax, nq=nq)
    for f in frames:
        run.add_frame(f)
    ref = run.table()
    got = sf.table()
    scale = np.maximum(np.abs(ref[:, 1:]), 1.0)
    assert got.shape[1] == 3          # q, S, S(Ni-Ni)
    assert np.abs(got[:, 2]).max() > 0
    assert np.abs(got[:, 1:] - ref[:, 1:]).max() / scale.max() < (1e-9 if precision == "double"
                                                                   else 1e-6)
    rng = np.random.default_rng(2)
    target = torch.tensor(rng.random(nq), device=hip_device)
    pos = torch.tensor(frames[0], device=hip_device, requires_grad=True)
    sq_matching_loss(pos, sf, target).backward()
    rho = O.direct_rho(frames[0], cell, origin, run.species_of, 1, run.md)
    num, _ = O.shell_sums(rho, run.md, weight, run.amp, run.xcoef, partials=False)
    S = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1.0), 0.0)
    gS = 2.0 * (S - target.cpu().numpy())
    g_ref = O.gradient(frames[0], cell, origin, run.species_of, run.md, rho, weight, run.amp,
                       run.xcoef, gS, run.den, run.shell_count)
    g = pos.grad.cpu().numpy()
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print("  gradient deviation %.2e" % err)
    assert err < TOL[precision]
    sf.close()
