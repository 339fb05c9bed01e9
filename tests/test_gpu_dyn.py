"""dyn amplitudes on the matrix cores (sqc_hip_dyn_*) against the FP64
oracle, which is itself pinned by the reference's phase module and closed
forms (tests/test_dyn_oracle.py).  Bound: max |rho - rho_ref| below 1e-10
of max |rho_ref| (FP64; the north-star bar is 1e-9)."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _random(seed, n, ns, cell):
    rng = np.random.default_rng(seed)
    types = rng.integers(1, ns + 1, n)
    pos = (rng.random((2, n, 3)) * 1.6 - 0.3) @ cell.T      # unwrapped, outside the cell too
    return types, pos


def _check(got, ref, tol=1e-10):
    dev = np.abs(got - ref).max() / max(np.abs(ref).max(), 1.0)
    print("  max deviation %.2e of max |rho| %.3e" % (dev, np.abs(ref).max()))
    assert dev < tol


@pytest.mark.parametrize("ns", [1, 3])
def test_lattice_modes_against_oracle(hip_device, ns):
    from sqcalc_amd.dynamics import HipDynAmplitudes, grid_modes
    cell = np.array([[17.0, 1.2, -0.8], [0.0, 15.5, 0.6], [0.0, 0.0, 19.0]])
    types, frames = _random(11 + ns, 1500, ns, cell)
    b = 2 * np.pi * np.linalg.inv(cell).T
    hkl, _ = grid_modes(cell, 3.0)
    print("case dyn lattice", ns, len(hkl), flush=True)
    dy = HipDynAmplitudes(b, types, hkl)
    got = dy.rho(frames)
    sp = np.searchsorted(np.unique(types), types).astype(np.int32)
    for f in range(len(frames)):
        _check(got[f], O.dyn_rho(frames[f], sp, ns, b, hkl))
    dy.close()


def test_q_line_with_base(hip_device):
    from sqcalc_amd.dynamics import HipDynAmplitudes, line_modes
    cell = np.diag([14.0, 16.0, 15.0])
    types, frames = _random(21, 900, 2, cell)
    hkl, rec, base = line_modes(150, 0.5, 12.0, (1.0, 1.0, 0.0))
    print("case dyn q line", flush=True)
    dy = HipDynAmplitudes(rec, types, hkl, base=base)
    got = dy.rho(frames[0])[0]
    sp = np.searchsorted(np.unique(types), types).astype(np.int32)
    _check(got, O.dyn_rho(frames[0], sp, 2, rec, hkl, base=base))
    dy.close()


def test_arbitrary_mode_list_and_device_path(hip_device):
    """A scattered list with repeats and both signs, the device entry point,
    and bit-identical repeats."""
    import torch

    from sqcalc_amd.dynamics import HipDynAmplitudes
    cell = np.diag([11.0, 13.0, 12.0])
    types, frames = _random(31, 700, 2, cell)
    rng = np.random.default_rng(5)
    hkl = rng.integers(-9, 10, (777, 3)).astype(np.int32)
    hkl[100:110] = hkl[0]
    b = 2 * np.pi * np.linalg.inv(cell).T
    print("case dyn scattered modes", flush=True)
    dy = HipDynAmplitudes(b, types, hkl)
    pos = torch.tensor(frames, device=hip_device)
    a = dy.rho_device(pos)
    c = dy.rho_device(pos)
    torch.cuda.synchronize()
    assert torch.equal(a, c)
    sp = np.searchsorted(np.unique(types), types).astype(np.int32)
    for f in range(2):
        _check(a[f].cpu().numpy(), O.dyn_rho(frames[f], sp, 2, b, hkl))
    dy.close()


def test_benchmark_geometry_subset(hip_device):
    """The 96,000-atom workload (L = 113, Si/O) with every lattice mode up to
    |q| = 2 (about 97,600 modes, one of each +- pair): 2,000 seeded modes
    against the oracle."""
    from sqcalc_amd.dynamics import HipDynAmplitudes, grid_modes
    from sqcalc_amd.synthetic import config
    wl = config("c2")
    frame = wl.frames(0, 1)[0]
    b = 2 * np.pi * np.linalg.inv(wl.cell).T
    hkl, _ = grid_modes(wl.cell, 2.0)
    print("case dyn c2 geometry", len(hkl), flush=True)
    dy = HipDynAmplitudes(b, wl.types, hkl)
    info = dy.info
    print("  items %d chunks %d chunk atoms %d column axis %d" % (info.nitems, info.nchunks,
                                                               info.chunk_atoms, info.column_axis))
    got = dy.rho(frame)[0]
    pick = np.sort(np.random.default_rng(7).choice(len(hkl), 2000, replace=False))
    sp = np.searchsorted(np.unique(wl.types), wl.types).astype(np.int32)
    _check(got[:, pick], O.dyn_rho(frame, sp, 2, b, hkl[pick]))
    dy.close()
