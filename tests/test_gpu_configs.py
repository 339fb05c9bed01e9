"""Parity at the sizes of the benchmark configurations.

The FP64 oracle cannot sum every mode at the larger sizes, so these tests
tie the full-size runs to it through properties that hold at any size:

c1  L=20, N=1000, unit, partials, q_max 10, 10 frames: GPU FP64 (1e-9) and
    FP32 (1e-6) against the oracle's direct sum over all 135k modes.
c2  L=113, N=96,000, neutron, partials, q_max 10:
    - the full-range FP64 plan against the oracle on every shell below
      q = 0.8 (both plans bin from q = 0 with dq = 0.02, so the shells are the
      same), 1e-9;
    - FP32 against FP64 on every column, 1e-6;
    - the shell sums against the pair sums: num = sum_{a<=b} (2 - d_ab) b_a b_b
      P_ab holds mode by mode, so to the rounding of the shell sums (FP64
      plans 1e-12, FP32 plans' one-word sums 1e-9);
    - the perfect 46^3 lattice: S of a shell is N * (Bragg modes in the
      shell) / (modes in the shell), every other mode giving exactly zero.
c3  L=80, N=32,000, 60 frames: graph replays against kernel-by-kernel
    launches (1e-12); the autotuned graph plan over 100 frames against the
    oracle on its 25 lowest shells (1e-6).
c4  L=44, N=8,000, 3 species: dL/dr of sum_s (S_s - T_s)^2 against the
    oracle's analytic gradient (1e-8 FP64, 1e-5 FP32 at eps 1e-7) and against
    central differences of the GPU FP64 forward on 64 coordinates (h = 1e-4 A).
c5  triclinic 250 A, N = 10^6, 4 species, xray, 10 partials, q_max 4: FP32
    against FP64 on every column, 1e-6; the FP64 plan against the oracle's
    direct sum on the 25 shells below q = 0.2 (dq = 0.008 in both), 1e-9;
    dL/dr of a loss on those shells (S and all ten pair columns) against the
    oracle's analytic gradient for every atom (1e-8 FP64, 1e-5 FP32), and
    the FP32 gradient of the full-range loss against the FP64 one (1e-5).
Tolerances use the reference suite's metric max |d| / max(|ref|, 1).
"""

import numpy as np
import pytest

from helpers import rel_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _plan(wl, precision, **kw):
    from sqcalc_amd import HipStructureFactor
    args = dict(weight=wl.weight, norm=wl.norm, qmin=wl.qmin, qmax=wl.qmax, nq=wl.nq,
                precision=precision)
    args.update(kw)
    return HipStructureFactor(wl.cell, wl.origin, wl.types, wl.mapping, **args)


def _device(frames):
    import torch
    return torch.tensor(np.asarray(frames), dtype=torch.float64, device="cuda:0").contiguous()


def test_c1_against_oracle(hip_device):
    from sqcalc_amd.synthetic import config
    wl = config("c1")
    frames = wl.frames()
    print("case c1 oracle", flush=True)
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmax=wl.qmax, nq=wl.nq)
    for f in frames:
        run.add_frame(f)
    ref = run.table()
    for precision, eps, tol in (("double", 1e-10, 1e-9), ("single", None, 1e-6)):
        sf = _plan(wl, precision, eps=eps)
        sf.accumulate(frames)
        dev = rel_dev(ref[:, 1:], sf.table()[:, 1:])
        print("  %s: max deviation %.2e" % (precision, dev))
        assert dev < tol
        sf.close()


@pytest.fixture(scope="module")
def c2_frame():
    from sqcalc_amd.synthetic import config
    wl = config("c2")
    return wl, wl.frames(0, 1)


def test_c2_low_shells_against_oracle(hip_device, c2_frame):
    wl, frames = c2_frame
    print("case c2 low shells", flush=True)
    sf = _plan(wl, "double", eps=1e-10)
    sf.accumulate_device(_device(frames))
    full = sf.table()
    sf.close()
    nlow = 40  # q < 0.8 at dq = 0.02
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=0.0, qmax=0.8, nq=nlow)
    run.add_frame(frames[0])
    ref = run.table()
    np.testing.assert_allclose(full[:nlow, 0], ref[:, 0], atol=1e-12)
    dev = rel_dev(ref[:, 1:], full[:nlow, 1:])
    print("  max deviation on %d shells: %.2e" % (nlow, dev))
    assert dev < 1e-9


def test_c2_single_against_double_and_sum_rule(hip_device, c2_frame):
    wl, frames = c2_frame
    print("case c2 fp32 vs fp64", flush=True)
    pos = _device(frames)
    out = {}
    for precision in ("double", "single"):
        sf = _plan(wl, precision, tune=True)
        sf.accumulate_device(pos)
        num, part, _, _ = sf.fetch()
        out[precision] = (sf.table(), num, part, sf.amp.copy())
        sf.close()
    dev = rel_dev(out["double"][0][:, 1:], out["single"][0][:, 1:])
    print("  fp32 vs fp64 max deviation %.2e" % dev)
    assert dev < 1e-6
    for precision in ("double", "single"):
        _, num, part, b = out[precision]
        S = len(b)
        recon = np.zeros_like(num)
        for a in range(S):
            for c in range(a, S):
                recon += (1.0 if a == c else 2.0) * b[a] * b[c] * part[a, c]
        err = np.abs(recon - num).max() / np.abs(num).max()
        print("  %s: sum rule %.2e" % (precision, err))
        # num and the pair sums are exact integer sums of each mode's value
        # rounded to the plan's quantum: two words (FP64 plans) carry a value
        # far below FP64 rounding, one word (FP32 plans) to about 1e-8 of a
        # typical mode here
        assert err < (1e-12 if precision == "double" else 1e-9)


@pytest.mark.parametrize("precision,tol", [("double", 1e-9), ("single", 1e-6)])
def test_c2_perfect_lattice_bragg(hip_device, precision, tol):
    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd.synthetic import lattice_sites
    L, side = 113.0, 46
    cell = np.diag([L] * 3)
    pos = lattice_sites(cell, side)
    n = len(pos)
    types = np.ones(n, dtype=np.int64)
    print("case c2 bragg", precision, flush=True)
    sf = HipStructureFactor(cell, np.zeros(3), types, None, weight="unit", norm="mean",
                            qmax=10.0, nq=500, precision=precision, partials=False)
    sf.accumulate_device(_device(pos[None]))
    got = sf.table()[:, 1]
    # closed form: |rho| = N on the lattice's reciprocal points (h = 0 mod 46),
    # exactly 0 elsewhere; S of a shell = N * (Bragg modes) / (modes)
    K = int(sf.info.nmodes)
    hkl = np.zeros((K, 3), dtype=np.int32)
    qvec = np.zeros((K, 3))
    assert sf.lib.sqc_hip_modes(sf.plan, hkl.ctypes.data, qvec.ctypes.data, None) == 0
    bragg = np.all(hkl % side == 0, axis=1)
    dq = 10.0 / 500
    shell = np.minimum((np.linalg.norm(qvec[bragg], axis=1) / dq).astype(int), 499)
    nb = np.bincount(shell, minlength=500)
    ref = np.where(sf.shell_count > 0, n * nb / np.maximum(sf.shell_count, 1), 0.0)
    assert nb.sum() > 0
    dev = rel_dev(ref, got)
    print("  %d Bragg modes in %d shells, max deviation %.2e" % (nb.sum(), (nb > 0).sum(), dev))
    assert dev < tol

    # dL/dr of L = sum_s (S_s - S_s / 2)^2 vanishes on a perfect Bravais
    # lattice: e^{i q r_j} = 1 at every Bragg mode, so each mode's term
    # 2 Re(conj(rho) i q e^{i q r_j}) / den = Re(2 i N q) / den is 0, and
    # rho = 0 at every other mode.  The residue is held to the gradient
    # tolerance (1e-8 / 1e-5) times the sum of the magnitudes that cancel,
    # sum over Bragg modes of |dL/dS_s| 2 |q| / count_s (den = N count_s).
    import torch

    from sqcalc_amd.autograd import structure_factor
    pt = torch.tensor(pos, device=hip_device, requires_grad=True)
    S, _ = structure_factor(pt, sf)
    loss = torch.sum((S - 0.5 * S.detach()) ** 2)
    loss.backward()
    g = np.abs(pt.grad.cpu().numpy()).max()
    gS = np.abs(S.detach().cpu().numpy())
    qb = np.linalg.norm(qvec[bragg], axis=1)
    scale = float(np.sum(gS[shell] * 2 * qb / sf.shell_count[shell]))
    print("  lattice gradient: max |dL/dr| = %.2e of %.2e" % (g, scale))
    assert g <= (1e-8 if precision == "double" else 1e-5) * scale
    sf.close()


def test_c3_graph_replay(hip_device):
    from sqcalc_amd.synthetic import config
    wl = config("c3")
    frames = wl.frames(0, 60)
    pos = _device(frames)
    print("case c3 graph", flush=True)
    res = []
    # the same (default) tiles for both, so that only the launch mode differs
    for kw in (dict(max_batch=1), dict(max_batch=25)):
        sf = _plan(wl, "single", **kw)
        sf.accumulate_device(pos)
        num, part, _, nfr = sf.fetch()
        assert nfr == 60
        res.append((num, part))
        sf.close()
    np.testing.assert_allclose(res[1][0], res[0][0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(res[1][1], res[0][1], rtol=1e-12,
                               atol=1e-12 * np.abs(res[0][1]).max())


def test_c3_autotuned_graph_low_shells_against_oracle(hip_device):
    """The batched-trajectory plan as the benchmark runs it (autotuned tiles,
    captured graphs of 50 frames, FP32): 100 frames of the random walk, the
    25 lowest shells (q < 0.4) of every column against the oracle's direct
    sum over the same frames, 1e-6."""
    from sqcalc_amd.synthetic import config
    wl = config("c3")
    frames = wl.frames(0, 100)
    print("case c3 tuned graph vs oracle", flush=True)
    sf = _plan(wl, "single", tune=True, max_batch=50)
    sf.accumulate_device(_device(frames))
    got = sf.table()
    _, _, _, nfr = sf.fetch()
    assert nfr == 100
    sf.close()
    nlow = 25
    dq = (wl.qmax - wl.qmin) / wl.nq
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=wl.qmin, qmax=wl.qmin + nlow * dq, nq=nlow)
    for f in frames:
        run.add_frame(f)
    ref = run.table()
    np.testing.assert_allclose(got[:nlow, 0], ref[:, 0], atol=1e-12)
    dev = rel_dev(ref[:, 1:], got[:nlow, 1:])
    print("  fp32 graph vs oracle on %d shells: %.2e" % (nlow, dev))
    assert dev < 1e-6


@pytest.fixture(scope="module")
def c4_case():
    from sqcalc_amd.synthetic import config
    wl = config("c4")
    tgt = config("c4_target")
    return wl, wl.frames()[0], tgt.frames()[0]


@pytest.mark.parametrize("precision,tol", [("double", 1e-8), ("single", 1e-5)])
def test_c4_gradient(hip_device, c4_case, precision, tol):
    import torch

    from sqcalc_amd.autograd import structure_factor
    wl, frame, target_frame = c4_case
    print("case c4 gradient", precision, flush=True)
    # FP32 gradient plans run at eps 1e-7 (ns = 12): the kernel derivative's
    # aliasing error sets the gradient's accuracy (measured here: 1.4e-5 at
    # eps 1e-5, 7.7e-6 at 1e-6, 2.4e-7 at 1e-7)
    eps = 1e-11 if precision == "double" else 1e-7
    sf = _plan(wl, precision, eps=eps, partials=False)
    st = _plan(wl, "double", eps=1e-11, partials=False)
    st.accumulate(target_frame)
    T = st.table()[:, 1]
    st.close()
    target = torch.tensor(T, device=hip_device)
    pos = torch.tensor(frame, device=hip_device, requires_grad=True)
    S, _ = structure_factor(pos, sf)
    loss = torch.sum((S - target) ** 2)
    loss.backward()
    g = pos.grad.cpu().numpy()
    # the oracle's analytic FP64 gradient
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmax=wl.qmax, nq=wl.nq, partials=False)
    rho = O.direct_rho(frame, wl.cell, wl.origin, run.species_of, len(run.amp), run.md, method=1)
    num, _ = O.shell_sums(rho, run.md, wl.weight, run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    gref = O.gradient(frame, wl.cell, wl.origin, run.species_of, run.md, rho, wl.weight, run.amp,
                      run.xcoef, 2 * (Sref - T), run.den, run.shell_count)
    err = np.abs(g - gref).max() / np.abs(gref).max()
    print("  adjoint vs analytic: %.2e" % err)
    assert err < tol
    if precision == "double":
        # central differences of the FP64 forward on 64 random coordinates
        rng = np.random.default_rng(64)
        picks = rng.choice(frame.size, 64, replace=False)
        h = 1e-4
        base = frame.reshape(-1).copy()
        fd = np.zeros(64)
        for k, c in enumerate(picks):
            vals = []
            for sgn in (1.0, -1.0):
                x = base.copy()
                x[c] += sgn * h
                p = torch.tensor(x.reshape(frame.shape), device=hip_device)
                Sx, _ = structure_factor(p, sf)
                vals.append(float(torch.sum((Sx - target) ** 2)))
            fd[k] = (vals[0] - vals[1]) / (2 * h)
        gflat = g.reshape(-1)[picks]
        errfd = np.abs(fd - gflat).max() / np.abs(g).max()
        print("  adjoint vs central differences: %.2e" % errfd)
        assert errfd < 1e-6
    sf.close()


def test_c5_single_against_double(hip_device):
    from sqcalc_amd.synthetic import config
    wl = config("c5")
    frames = wl.frames(0, 1)
    print("case c5 fp32 vs fp64", flush=True)
    pos = _device(frames)
    tables = {}
    for precision in ("double", "single"):
        sf = _plan(wl, precision)
        sf.accumulate_device(pos)
        tables[precision] = sf.table()
        assert sf.table().shape[1] == 2 + 10
        sf.close()
    dev = rel_dev(tables["double"][:, 1:], tables["single"][:, 1:])
    print("  fp32 vs fp64 max deviation %.2e" % dev)
    assert dev < 1e-6
    # the lowest shells at full size against the direct sum (about 2,100 modes)
    nlow = 25
    dq = (wl.qmax - wl.qmin) / wl.nq
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=wl.qmin, qmax=wl.qmin + nlow * dq, nq=nlow)
    run.add_frame(frames[0])
    ref = run.table()
    full = tables["double"]
    np.testing.assert_allclose(full[:nlow, 0], ref[:, 0], atol=1e-12)
    dev = rel_dev(ref[:, 1:], full[:nlow, 1:])
    print("  fp64 vs oracle on %d shells: %.2e" % (nlow, dev))
    assert dev < 1e-9


@pytest.fixture(scope="module")
def c5_adjoint_case():
    """c5's first frame, the direct-sum oracle on the 25 lowest shells
    (q < 0.2, about 2,100 modes) and the oracle's analytic gradient of
        L = sum_{s < 25} (S_s - T_s)^2 + sum_{pairs, s < 25} wP_ab,s P_ab,s
    with T = S_ref / 2 and a seeded random pair weight wP (0 on every other
    shell), so every one of the 10^6 atoms gets a gradient from all four
    species and all ten pairs."""
    from sqcalc_amd.synthetic import config
    wl = config("c5")
    frame = wl.frames(0, 1)[0]
    nlow = 25
    dq = (wl.qmax - wl.qmin) / wl.nq
    run = O.Run(wl.cell, wl.origin, wl.types, wl.mapping, weight=wl.weight, norm=wl.norm,
                qmin=wl.qmin, qmax=wl.qmin + nlow * dq, nq=nlow)
    S = len(run.species_type)
    rho = O.direct_rho(frame, wl.cell, wl.origin, run.species_of, S, run.md, method=1)
    num, _ = O.shell_sums(rho, run.md, wl.weight, run.amp, run.xcoef)
    Sref = np.where(run.den > 0, num / np.where(run.den > 0, run.den, 1), 0)
    T = 0.5 * Sref
    wP = np.random.default_rng(55).standard_normal((S * (S + 1) // 2, nlow))
    gP = np.zeros((S, S, nlow))
    k = 0
    for a in range(S):
        for b in range(a, S):
            gP[a, b] = wP[k]
            k += 1
    gref = O.gradient(frame, wl.cell, wl.origin, run.species_of, run.md, rho, wl.weight,
                      run.amp, run.xcoef, 2 * (Sref - T), run.den, run.shell_count, gP=gP)
    return wl, frame, nlow, T, wP, gref


@pytest.mark.parametrize("precision,eps", [("double", 1e-13), ("single", 1e-7)])
def test_c5_adjoint_low_shells_against_oracle(hip_device, c5_adjoint_case, precision, eps):
    """The forward + adjoint of the 10^6-atom triclinic X-ray workload with
    its four fields and ten pair columns (the x pass and its adjoint with
    the field count read at run time) against the oracle's gradient, every
    atom: max |g - g_ref| / max |g_ref| below 1e-8 (FP64) and 1e-5 (FP32).
    The plan covers the whole q range; the loss reads its 25 lowest shells.

    A loss confined to q < 0.2 is the hard case for the kernel-derivative
    gather: the gradient is of order q |V| while the derivative's aliasing
    error is of order eps |V| / h (h the grid spacing), about 10^4 eps
    relative here (measured in FP64: 1.6e-7 at eps 1e-11, 1.6e-9 at 1e-13;
    in FP32: 3.0e-4 at 1e-6, 8.3e-6 at 1e-7, where ns = 12).  Hence the
    tolerances of the gradient plans: eps 1e-13 in FP64, 1e-7 in FP32."""
    import torch

    from sqcalc_amd.autograd import structure_factor
    wl, frame, nlow, T, wP, gref = c5_adjoint_case
    print("case c5 adjoint", precision, flush=True)
    sf = _plan(wl, precision, eps=eps)
    assert sf.info.nfields == 4 and sf.info.ncomp == 11
    pos = torch.tensor(frame, device=hip_device, requires_grad=True)
    S, P = structure_factor(pos, sf)
    Tt = torch.tensor(T, device=hip_device)
    wPt = torch.tensor(wP, device=hip_device)
    loss = torch.sum((S[:nlow] - Tt) ** 2) + torch.sum(wPt * P[:, :nlow])
    loss.backward()
    g = pos.grad.cpu().numpy()
    err = np.abs(g - gref).max() / np.abs(gref).max()
    print("  adjoint vs analytic on %d atoms: %.2e" % (len(frame), err))
    assert err < (1e-8 if precision == "double" else 1e-5)
    sf.close()


def test_c5_adjoint_single_against_double_full_range(hip_device, c5_adjoint_case):
    """The S(q)-matching loss over all 500 shells with the pair columns:
    the FP32 adjoint (eps 1e-7, ns = 12) against the FP64 one (eps 1e-13),
    1e-5 relative.  The Bragg shells of the jittered lattice make this the
    harder case for FP32: measured 1.5e-3 at eps 1e-5, 4.6e-5 at 1e-6,
    8.0e-7 at 1e-7.

    Both gradients also keep the exact translation invariance of S and the
    pair columns (a rigid shift changes only the phase of every rho_a(q)):
    sum_j dL/dr_j = 0.  Per-atom errors of relative size d add up to about
    d sqrt(N) max |g|, so |sum_j g_j| / (sqrt(N) max |g|) is held to the
    gradient tolerances, 1e-8 (FP64) and 1e-5 (FP32)."""
    import torch

    from sqcalc_amd.autograd import structure_factor
    wl, frame, _, _, _, _ = c5_adjoint_case
    print("case c5 adjoint fp32 vs fp64", flush=True)
    grads = {}
    target = None
    wP = None
    for precision, eps in (("double", 1e-13), ("single", 1e-7)):
        sf = _plan(wl, precision, eps=eps)
        pos = torch.tensor(frame, device=hip_device, requires_grad=True)
        S, P = structure_factor(pos, sf)
        if target is None:
            target = 0.5 * S.detach().clone()
            wP = torch.tensor(np.random.default_rng(56).standard_normal(tuple(P.shape)),
                              device=hip_device)
        loss = torch.sum((S - target) ** 2) + torch.sum(wP * P)
        loss.backward()
        grads[precision] = pos.grad.cpu().numpy()
        sf.close()
    err = np.abs(grads["single"] - grads["double"]).max() / np.abs(grads["double"]).max()
    print("  fp32 vs fp64 gradient %.2e" % err)
    for precision, tol in (("double", 1e-8), ("single", 1e-5)):
        g = grads[precision]
        drift = np.abs(g.sum(axis=0)).max() / (np.sqrt(len(g)) * np.abs(g).max())
        print("  %s: |sum_j g_j| / (sqrt(N) max |g|) = %.2e" % (precision, drift))
        assert drift < tol
    assert err < 1e-5
