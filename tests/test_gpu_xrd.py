"""Powder XRD from the device's per-mode sums against the oracle.

The plan keeps, per kept mode, the frame sums of the weighted |rho|^2
(want_grid) and of Re rho_a rho_b* (mode_partials); HipStructureFactor.xrd
bins them by two-theta.  The pattern and its pair columns must match the
oracle's per-frame accumulation (src/sqc_structure_factor.f90:538-643) to
1e-9 of the largest bin in FP64 and 1e-5 in FP32, and the table must keep the
pair sum rule.
"""

import numpy as np
import pytest

from helpers import random_case
from oracle import oracle as O
from oracle.xrd_oracle import Xrd

pytestmark = pytest.mark.gpu

LAM = 1.541838


@pytest.mark.parametrize("precision,weight", [("double", "xray"), ("single", "xray"),
                                              ("double", "neutron")])
def test_device_xrd_matches_oracle(hip_device, precision, weight):
    from sqcalc_amd import HipStructureFactor
    from sqcalc_amd import xrd as X
    cell, origin, types, frames = random_case(natoms=400, length=12.0, nframes=3, seed=23,
                                              ntypes=2, tilt=(0.9, -0.5, 0.4))
    mapping = {1: "Si", 2: "O"}
    qmin, qmax = X.q_range(LAM, 10.0, 130.0)
    print("case xrd", precision, weight, flush=True)
    sf = HipStructureFactor(cell, origin, types, mapping, weight=weight, norm="mean", qmin=qmin,
                            qmax=qmax, nq=80, precision=precision,
                            eps=1e-10 if precision == "double" else None, want_grid=True,
                            mode_partials=True, max_batch=2)
    sf.accumulate(frames)
    pat = sf.xrd(LAM, 10.0, 130.0)
    run = O.Run(cell, origin, types, mapping, weight=weight, norm="mean", qmin=qmin, qmax=qmax,
                nq=80)
    xr = Xrd(run, LAM, 10.0, 130.0)
    for f in frames:
        xr.add_frame(f)
    tth, total, pairs = xr.table()
    assert pat["bins"] == xr.bins and pat["nframes"] == 3
    np.testing.assert_array_equal(pat["count"], xr.count)
    tol = 1e-9 if precision == "double" else 1e-5
    scale = np.abs(total).max()
    assert np.abs(pat["intensity"] - total).max() <= tol * scale
    assert len(pat["pairs"]) == len(pairs) == 3
    for p, q in zip(pat["pairs"], pairs):
        assert np.abs(p - q).max() <= tol * scale
    rule = pat["pairs"][0] + 2.0 * pat["pairs"][1] + pat["pairs"][2]
    assert np.abs(rule - pat["intensity"]).max() <= tol * scale
    text = sf.format_xrd(LAM, 10.0, 130.0)
    lines = text.splitlines()
    assert lines[0].startswith("# xrd: lambda 1.541838 A, 2theta 10.0000 to 130.0000 deg")
    assert lines[3] == "# 2theta[deg] I I(Si-Si) I(Si-O) I(O-O)"
    assert len(lines) == 4 + pat["bins"]
    # reset clears the per-mode pair sums too
    sf.reset()
    sf.accumulate(frames[:1])
    one = sf.xrd(LAM, 10.0, 130.0)
    assert one["nframes"] == 1
    sf.close()
