"""Bit-reproducible shell sums.

Every value the x pass adds into a shell is split into two parts that are
exact multiples of the plan's quanta, so all additions of them are exact
and their order (lanes, waves, workgroups, LDS or HBM atomics) cannot change
a bit.  The reference sums per OpenMP thread and then in thread order
(src/sqc_structure_factor.f90:604-633), which is reproducible for a fixed
thread count; here the sums are reproducible for any schedule.
"""

import numpy as np
import pytest

from helpers import oracle_run, random_case, rel_dev

pytestmark = pytest.mark.gpu

SIO = {1: "Si", 2: "O"}


def _plan(cell, origin, types, mapping, weight, precision, nq, qmax, **kw):
    from sqcalc_amd import HipStructureFactor
    return HipStructureFactor(cell, origin, types, mapping, weight=weight, norm="mean", qmin=0.0,
                              qmax=qmax, nq=nq, precision=precision, **kw)


def _sums(cell, origin, types, frames, precision, one_by_one=False, **kw):
    sf = _plan(cell, origin, types, SIO, "neutron", precision, 400, 8.0, partials=True, **kw)
    if one_by_one:
        for f in frames:
            sf.accumulate(f)
    else:
        sf.accumulate(frames)
    num, part, _, nfr = sf.fetch()
    sf.close()
    assert nfr == len(frames)
    return num, part


@pytest.mark.parametrize("precision", ["double", "single"])
def test_identical_plans_give_identical_bits(hip_device, precision):
    """Two plans of one workload in one process, and the same frames fed as
    one captured batch or frame by frame: num and every pair sum agree to the
    last bit.  Thousands of workgroups add into the same shells here."""
    cell, origin, types, frames = random_case(natoms=24000, length=62.0, nframes=3, seed=41)
    print("case identical plans", precision, flush=True)
    a = _sums(cell, origin, types, frames, precision)
    b = _sums(cell, origin, types, frames, precision)
    c = _sums(cell, origin, types, frames, precision, one_by_one=True, max_batch=1)
    for x, y in ((a, b), (a, c)):
        assert np.array_equal(x[0], y[0])
        assert np.array_equal(x[1], y[1])
    assert np.count_nonzero(a[0]) > 300  # of 400 shells; the lowest hold no mode


@pytest.mark.parametrize("precision", ["double", "single"])
def test_hbm_shell_sums_identical_bits(hip_device, precision):
    """8 species, 600 shells: the sums do not fit the LDS and every workgroup
    adds into its own HBM row with global atomics; two runs agree bit for bit
    and both match the oracle."""
    cell, origin, types, frames = random_case(natoms=480, length=12.0, nframes=2, seed=17,
                                              ntypes=8, tilt=(0.5, 0.0, -0.3))
    print("case hbm sums", precision, flush=True)
    eps = 1e-10 if precision == "double" else 1e-6
    runs = []
    for _ in range(2):
        sf = _plan(cell, origin, types, None, "unit", precision, 600, 4.8, eps=eps)
        assert sf.info.shell_sums_in_hbm == 1
        sf.accumulate(frames)
        runs.append((sf.fetch(), sf.table()))
        sf.close()
    (n0, p0, _, _), t0 = runs[0]
    (n1, p1, _, _), _ = runs[1]
    assert np.array_equal(n0, n1) and np.array_equal(p0, p1)
    ref = oracle_run(cell, origin, types, None, "unit", "mean", 0.0, 4.8, 600, frames).table()
    dev = max(rel_dev(ref[:, c], t0[:, c]) for c in range(1, ref.shape[1]))
    print("  worst column vs oracle %.2e" % dev)
    assert dev <= (1e-9 if precision == "double" else 1e-6)


def test_exact_sums_keep_fp64_accuracy(hip_device):
    """The split loses nothing the FP64 path needs: a 2-species neutron
    workload with partials against the direct-sum oracle at 1e-9 (the FP64
    bound), with shells of hundreds of modes each."""
    cell, origin, types, frames = random_case(natoms=3000, length=30.0, nframes=2, seed=5)
    print("case exact sums accuracy", flush=True)
    qmax, nq = 4.0, 40
    sf = _plan(cell, origin, types, SIO, "neutron", "double", nq, qmax, partials=True, eps=1e-11)
    sf.accumulate(frames)
    got = sf.table()
    sf.close()
    ref = oracle_run(cell, origin, types, SIO, "neutron", "mean", 0.0, qmax, nq, frames).table()
    dev = max(rel_dev(ref[:, c], got[:, c]) for c in range(1, ref.shape[1]))
    print("  worst column vs oracle %.2e" % dev)
    assert dev <= 1e-9
