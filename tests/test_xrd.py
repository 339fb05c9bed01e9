"""Powder XRD: the oracle restatement and the product's binning, on the CPU.

Pins:
  - the reference's own known answers (test/test_xrd.f90:28-60): the first
    fcc Ni reflections at Cu K-alpha land on the textbook angles, the 45 deg
    line inverts, two-theta clips at 180 deg, LP(90 deg) = 2 sqrt(2) and LP
    grows towards both ends;
  - the pair sum rule I = sum_a I_aa + 2 sum_{a<b} I_ab, which holds only if
    the pair columns carry w_a w_b and the same Lorentz-polarization weight
    as the total;
  - every kept mode lands in exactly one bin, at floor((2theta - min)/step);
  - the product's pattern (sqcalc_amd.xrd, fed the per-mode frame sums the
    device keeps) equals the oracle's per-frame accumulation.
"""

import math

import numpy as np
import pytest

from helpers import random_case
from oracle import oracle as O
from oracle.xrd_oracle import Xrd, lorentz_polarization, two_theta_deg
from sqcalc_amd import xrd as X

LAM = 1.541838


@pytest.mark.parametrize("tth_fn,lp_fn", [(two_theta_deg, lorentz_polarization),
                                          (lambda q, l: float(X.two_theta_deg(q, l)),
                                           lambda t: float(X.lorentz_polarization(t)))])
def test_reference_known_answers(tth_fn, lp_fn):
    a = 3.52
    for hkl, want in (((1, 1, 1), 44.585405), ((2, 0, 0), 51.955510), ((2, 2, 0), 76.552967)):
        q = math.sqrt(sum(h * h for h in hkl)) * 2.0 * math.pi / a
        assert abs(tth_fn(q, LAM) - want) / want < 1e-7
    q = 4.0 * math.pi * math.sin(math.pi / 8.0) / LAM
    assert abs(tth_fn(q, LAM) - 45.0) < 1e-9
    assert tth_fn(1.0e9, LAM) == 180.0
    assert abs(lp_fn(90.0) - 2.0 * math.sqrt(2.0)) < 1e-12
    assert lp_fn(1.0) > lp_fn(90.0) and lp_fn(179.0) > lp_fn(90.0)


def _case(weight, seed, lp=True, step=0.0):
    cell, origin, types, frames = random_case(natoms=120, length=9.0, nframes=2, seed=seed,
                                              ntypes=3, tilt=(0.7, -0.4, 0.3))
    mapping = {1: "Si", 2: "O", 3: "Na"}
    qmin, qmax = X.q_range(LAM, 15.0, 120.0)
    run = O.Run(cell, origin, types, mapping, weight=weight, norm="mean", qmin=qmin, qmax=qmax,
                nq=50)
    xr = Xrd(run, LAM, 15.0, 120.0, step=step, lp=lp)
    for f in frames:
        xr.add_frame(f)
    return run, xr, frames


@pytest.mark.parametrize("weight,lp,step", [("xray", True, 0.0), ("neutron", False, 0.0),
                                            ("unit", True, 0.35)])
def test_oracle_sum_rule_and_bins(weight, lp, step):
    run, xr, _ = _case(weight, 3, lp, step)
    K = len(run.md["qlen"])
    assert sum(xr.count) == K
    tth, total, pairs = xr.table()
    S = len(run.species_type)
    k, acc = 0, np.zeros_like(total)
    for a in range(S):
        for b in range(a, S):
            acc += pairs[k] if a == b else 2.0 * pairs[k]
            k += 1
    assert np.abs(acc - total).max() <= 1e-10 * np.abs(total).max()
    # interior modes sit at floor((2theta - min) / step)
    t = np.array([two_theta_deg(float(q), LAM) for q in run.md["qlen"]])
    inside = (t > 15.0 + 1e-9) & (t < 120.0 - 1e-9)
    want = np.floor((t[inside] - 15.0) / xr.step).astype(int)
    assert np.array_equal(np.array(xr.bin)[inside], want)
    if step > 0:
        assert xr.bins == math.ceil((120.0 - 15.0) / step)


@pytest.mark.parametrize("weight", ["xray", "neutron"])
def test_product_pattern_matches_oracle(weight):
    run, xr, frames = _case(weight, 8)
    S = len(run.species_type)
    K = len(run.md["qlen"])
    gnum = np.zeros(K)
    gpart = np.zeros((S * (S + 1) // 2, K))
    amps = np.array([O.amplitude(run.weight, run.amp, run.xcoef, a, run.md["qlen"])
                     for a in range(S)])
    for f in frames:
        rho = O.direct_rho(f, run.cell, run.origin, run.species_of, S, run.md)
        t = (amps * rho).sum(axis=0)
        gnum += (t * np.conj(t)).real
        k = 0
        for a in range(S):
            for b in range(a, S):
                gpart[k] += (rho[a] * np.conj(rho[b])).real
                k += 1
    pat = X.pattern(gnum, gpart, amps, run.md["qlen"], len(frames), run.natoms, LAM, 15.0, 120.0,
                    cell=run.cell)
    tth, total, pairs = xr.table()
    assert pat["bins"] == xr.bins and abs(pat["step"] - xr.step) < 1e-15
    np.testing.assert_allclose(pat["two_theta"], tth, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(pat["count"], xr.count)
    scale = np.abs(total).max()
    assert np.abs(pat["intensity"] - total).max() <= 1e-12 * scale
    for p, q in zip(pat["pairs"], pairs):
        assert np.abs(p - q).max() <= 1e-12 * scale


def test_header_format_matches_fortran():
    pat = {"step": 0.12345678, "bins": 7, "two_theta": np.array([0.25]),
           "intensity": np.array([1.234567890123e-5]), "pairs": []}
    pat["bins"] = 1
    text = X.format_pattern(pat, 0.5, 1.0, 179.0, 1, True, [])
    first = text.splitlines()[0]
    # amdflang's f0.d drops the leading zero, as the reference build prints it
    assert first == "# xrd: lambda .500000 A, 2theta 1.0000 to 179.0000 deg, step .12346 deg, " \
                    "1 bins, 1 frame"
    assert text.splitlines()[-1] == "      0.2500    1.234567890123E-05"
