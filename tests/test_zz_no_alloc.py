"""The no-allocation adjoint (sqc_hip_adjoint_device_noalloc) through the C
ABI against sqc_hip_adjoint_device on the same inputs.

Both run the same kernels in the same order, so the gradients must agree
bit for bit.  Two plans are called in a row, the second with fewer atoms, so
the shared device scratch still holds the first call's slots when the second
call starts; the no-allocation call runs first on each plan, and the plan's
device footprint must not change over it.  A plan whose slots exceed the
scratch is refused before any launch."""

import ctypes

import numpy as np
import pytest

from helpers import random_case

pytestmark = pytest.mark.gpu


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _plan(natoms, seed, precision, partials, length=14.0, qmax=4.0, nq=30):
    from sqcalc_amd import HipStructureFactor
    cell, origin, types, frames = random_case(natoms=natoms, length=length, nframes=1,
                                              seed=seed, tilt=(0.7, -0.2, 0.4))
    sf = HipStructureFactor(cell, origin, types, {1: "Si", 2: "O"}, weight="neutron",
                            norm="mean", qmax=qmax, nq=nq, precision=precision,
                            partials=partials, eps=1e-11 if precision == "double" else 1e-6)
    return sf, frames[0]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("partials", [False, True])
def test_noalloc_adjoint_matches_adjoint(hip_device, precision, partials):
    import torch

    from sqcalc_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(8)
    for natoms, seed in ((3000, 31), (700, 32)):
        print("case noalloc adjoint", precision, partials, natoms, flush=True)
        sf, frame = _plan(natoms, seed, precision, partials)
        info = sf.info
        nq = sf.nq
        pos = torch.tensor(frame, device=hip_device).contiguous()
        out = torch.empty((info.ncomp, nq), dtype=torch.float64, device=hip_device)
        cdt = torch.complex64 if precision == "single" else torch.complex128
        rho = torch.empty((info.nfields, info.rho_len), dtype=cdt, device=hip_device)
        _lib.check(lib.sqc_hip_forward_device(sf.plan, pos.data_ptr(), 1, out.data_ptr(),
                                              rho.data_ptr(), _stream()), "forward")
        gS = torch.tensor(rng.standard_normal(nq), device=hip_device)
        S = len(sf.species_type)
        gP = torch.tensor(rng.standard_normal((S, S, nq)), device=hip_device) if partials else None
        gp_ptr = gP.data_ptr() if gP is not None else None
        before = sf.info.device_bytes
        g_new = torch.full_like(pos, float("nan"))
        _lib.check(lib.sqc_hip_adjoint_device_noalloc(sf.plan, pos.data_ptr(), rho.data_ptr(),
                                                      gS.data_ptr(), gp_ptr, g_new.data_ptr(),
                                                      _stream()), "adjoint_noalloc")
        torch.cuda.synchronize()
        assert sf.info.device_bytes == before
        g_ref = torch.full_like(pos, float("nan"))
        _lib.check(lib.sqc_hip_adjoint_device(sf.plan, pos.data_ptr(), rho.data_ptr(),
                                              gS.data_ptr(), gp_ptr, g_ref.data_ptr(),
                                              _stream()), "adjoint")
        torch.cuda.synchronize()
        assert sf.info.device_bytes > before        # this one allocated its slots
        assert bool(torch.isfinite(g_ref).all())
        assert torch.equal(g_new, g_ref), float((g_new - g_ref).abs().max())
        sf.close()


def test_noalloc_refuses_plans_beyond_the_scratch(hip_device):
    import torch

    from sqcalc_amd import _lib
    lib = _lib.load()
    # 200,000 atoms at q_max 1 in a 30 A box: 3 x 200,000 x sx x sy doubles
    # exceed the scratch; the call is refused before any launch
    print("case noalloc refusal", flush=True)
    sf, frame = _plan(200000, 33, "single", False, length=30.0, qmax=1.0, nq=10)
    pos = torch.tensor(frame, device=hip_device).contiguous()
    g = torch.empty_like(pos)
    gS = torch.zeros(10, dtype=torch.float64, device=hip_device)
    rho = torch.empty((1, sf.info.rho_len), dtype=torch.complex64, device=hip_device)
    rc = lib.sqc_hip_adjoint_device_noalloc(sf.plan, pos.data_ptr(), rho.data_ptr(),
                                            gS.data_ptr(), None, g.data_ptr(), _stream())
    assert rc == 2
    assert b"no-allocation scratch" in lib.sqc_hip_last_error()
    sf.close()
